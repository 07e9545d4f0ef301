// HIP kernels for gfx950: finpar/OptionPricing.fut's `main`.
//
// One lane per Monte-Carlo path.  A workgroup takes 256 consecutive Sobol
// indices at a time (a grid-stride loop over such blocks, for a grid fixed
// by the path count alone), so each wave owns an aligned block of 64 and
// the Gray code of index n = c + lane (c a multiple of 64) splits into
// G(c) ^ G(lane).  The Sobol number is linear in the Gray code over GF(2)
// (:17-27), so its high part, one value per wave and dimension, is
// computed once per block into LDS, and a lane adds the low part from its
// 6 low Gray-code bits.  A path runs the reference's pipeline in f32 with
// its order of operations: the inverse Gaussian (:40-140) of each Sobol
// number as the Brownian bridge (:145-183) consumes it, one bridge row per
// underlying, then per market model the Black-Scholes path (:190-225),
// with the bridge's differences taken on the fly, streamed date by date
// into the contract's payoff (:230-293).  The shapes are fixed by the
// contract (:289-293), so each contract is its own instantiation with
// compile-time sizes: contracts 1 and 2 keep their rows in registers,
// contract 3 (367 dates x 3 underlyings) in private memory.  Payoffs
// accumulate in f64 per lane, then per workgroup (a fixed tree in LDS),
// then over the workgroups in a second launch, so a result depends only on
// the inputs (the reference leaves the order of this sum, :318, to its
// backend).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "op_internal.h"

namespace {

constexpr int kThreads = 256, kWaves = kThreads / 64;

template <int C>
struct Shape;
template <>
struct Shape<1> {
  static constexpr int D = 1, U = 1;
};
template <>
struct Shape<2> {
  static constexpr int D = 5, U = 3;
};
template <>
struct Shape<3> {
  static constexpr int D = 367, U = 3;
};

// OptionPricing.fut:40-56 (polyAppr), Horner as written, IEEE division.
__device__ __forceinline__ float poly_appr(float x, const float (&a)[8], const float (&b)[8]) {
  const float num =
      x * (x * (x * (x * (x * (x * (x * a[7] + a[6]) + a[5]) + a[4]) + a[3]) + a[2]) + a[1]) +
      a[0];
  const float den =
      x * (x * (x * (x * (x * (x * (x * b[7] + b[6]) + b[5]) + b[4]) + b[3]) + b[2]) + b[1]) +
      b[0];
  return num / den;
}

// OptionPricing.fut:121-136 (ugaussianEl) with the coefficients of
// smallcase (:58-79), intermediate (:81-99) and tail (:101-119), in three
// parts: the branch test on dp = p - 0.5, the central value, and the tail
// value before its sign (a function of pp alone).
constexpr float kSa[8] = {3.387132872796366608f, 133.14166789178437745f,
                          1971.5909503065514427f, 13731.693765509461125f,
                          45921.953931549871457f, 67265.770927008700853f,
                          33430.575583588128105f, 2509.0809287301226727f};
constexpr float kSb[8] = {1.0f, 42.313330701600911252f, 687.1870074920579083f,
                          5394.1960214247511077f, 21213.794301586595867f,
                          39307.89580009271061f, 28729.085735721942674f,
                          5226.495278852854561f};
constexpr float kIa[8] = {1.42343711074968357734f, 4.6303378461565452959f,
                          5.7694972214606914055f, 3.64784832476320460504f,
                          1.27045825245236838258f, 0.24178072517745061177f,
                          0.0227238449892691845833f, 7.7454501427834140764e-4f};
constexpr float kIb[8] = {1.0f, 2.05319162663775882187f, 1.6763848301838038494f,
                          0.68976733498510000455f, 0.14810397642748007459f,
                          0.0151986665636164571966f, 5.475938084995344946e-4f,
                          1.05075007164441684324e-9f};
constexpr float kTa[8] = {6.6579046435011037772f, 5.4637849111641143699f,
                          1.7848265399172913358f, 0.29656057182850489123f,
                          0.026532189526576123093f, 0.0012426609473880784386f,
                          2.71155556874348757815e-5f, 2.01033439929228813265e-7f};
constexpr float kTb[8] = {1.0f, 0.59983220655588793769f, 0.13692988092273580531f,
                          0.0148753612908506148525f, 7.868691311456132591e-4f,
                          1.8463183175100546818e-5f, 1.4215117583164458887e-7f,
                          2.04426310338993978564e-5f};

__device__ __forceinline__ bool ugaussian_central(float dp) {
  return (dp < 0.0f && 0.0f - dp <= 0.425f) || (0.0f <= dp && dp <= 0.425f);
}
__device__ __forceinline__ float ugaussian_small(float dp) {
  return dp * poly_appr(0.180625f - dp * dp, kSa, kSb);
}
__device__ __forceinline__ float ugaussian_pp(float dp) {
  return dp < 0.0f ? dp + 0.5f : 0.5f - dp;
}
__device__ __forceinline__ float ugaussian_tail(float pp) {
  const float r = sqrtf(-logf(pp));
  return r <= 5.0f ? poly_appr(r - 1.6f, kIa, kIb) : poly_appr(r - 5.0f, kTa, kTb);
}

__device__ __forceinline__ float ugaussian(float p) {
  const float dp = p - 0.5f;
  if (ugaussian_central(dp)) return ugaussian_small(dp);
  const float x = ugaussian_tail(ugaussian_pp(dp));
  return dp < 0.0f ? 0.0f - x : x;
}

// ugaussian of U numbers per lane, with the tail cases of the whole wave
// packed together: about 15 % of the numbers take the tail branch, so a
// wave would run the tail's log, sqrt, polynomial and division once per
// number with most lanes idle.  Here the tail inputs of all U numbers are
// written to LDS (`buf`, U * 64 floats of this wave) at slots given by a
// ballot, evaluated 64 at a time, and read back; every value goes through
// the same operations as in ugaussian().  All 64 lanes must be active.
template <int U>
__device__ __forceinline__ void ugaussian_wave(const float (&p)[U], float (&out)[U],
                                               float *buf, int lane) {
  float dp[U];
  bool tail[U];
  int slot[U];
  int total = 0;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    dp[u] = p[u] - 0.5f;
    tail[u] = !ugaussian_central(dp[u]);
    out[u] = ugaussian_small(dp[u]);  // replaced below for the tail lanes
    const uint64_t mask = __ballot(tail[u]);
    slot[u] = total + (int)__builtin_amdgcn_mbcnt_hi(
                          (uint32_t)(mask >> 32),
                          __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    total += __popcll(mask);
  }
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (tail[u]) buf[slot[u]] = ugaussian_pp(dp[u]);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int t = lane; t < total; t += 64) buf[t] = ugaussian_tail(buf[t]);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (tail[u]) {
      const float x = buf[slot[u]];
      out[u] = dp[u] < 0.0f ? 0.0f - x : x;
    }
  }
  // the next call's writes wait for these reads
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// OptionPricing.fut:230-236 (fminPayoff)
__device__ __forceinline__ float fmin_payoff(float x0, float x1, float x2) {
  const float a = x0 / 3758.05f, b = x1 / 11840.0f, c = x2 / 1200.0f;
  return a < b ? (a < c ? a : c) : (b < c ? b : c);
}

// One market model set (OptionPricing.fut:299-304), row-major; the kernel's
// __restrict__ parameters, so that wave-uniform reads of them become scalar
// loads.
struct Market {
  const float *cs, *vols, *drifts, *sts, *detvals, *discts;
  int K;
};

// The payoff of one path under model m (Black-Scholes :190-225 streamed
// date by date into payoff1-3, :238-293).  rows: the bridge rows [U][D]
// before the differences of :178-182, which are taken here.
template <int C>
__device__ __forceinline__ float model_payoff(const Market &mk, int m,
                                              const float (*rows)[Shape<C>::D]) {
  constexpr int D = Shape<C>::D, U = Shape<C>::U;
  const float *cs = mk.cs + m * U * U, *vols = mk.vols + m * D * U;
  const float *drifts = mk.drifts + m * D * U, *sts = mk.sts + m * U;
  const float *disc = mk.discts + m * mk.K;
  float acc[U], price[U], prev[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    acc[u] = 1.0f;
    price[u] = 0.0f;
    prev[u] = 0.0f;
  }
  float xss[C == 2 ? D * U : 1];
  bool cond = false;
  for (int d = 0; d < D; ++d) {
    float z[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const float r = rows[u][d];
      z[u] = d == 0 ? r : r - prev[u];  // bbrow[i] - bbrow[i - 1] (:181)
      prev[u] = r;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      float noise = 0.0f;  // f32.sum (:194-197), from the left
#pragma unroll
      for (int k = 0; k <= u; ++k) noise = noise + z[k] * cs[u * U + k];
      const float c = drifts[d * U + u] + noise * vols[d * U + u];
      acc[u] = acc[u] * expf(c);       // the scan of (*) from 1 (:218)
      price[u] = sts[u] * acc[u];      // (:219)
    }
    if constexpr (C == 2) {
#pragma unroll
      for (int u = 0; u < U; ++u) xss[d * U + u] = price[u];
    }
    if constexpr (C == 3)
      cond = cond || price[0] <= 2630.6349999999998f || price[1] <= 8288.0f ||
             price[2] <= 840.0f;
  }
  if constexpr (C == 1) {  // payoff1 (:240-244)
    const float amount = (price[0] - 4000.0f) * mk.detvals[m];
    const float amount0 = 0.0f < amount ? amount : 0.0f;
    return amount0 * disc[0];
  } else if constexpr (C == 2) {  // payoff2 (:246-265)
    int date;
    float amount;
    if (1.0f <= fmin_payoff(xss[0], xss[1], xss[2])) {
      date = 0;
      amount = 1150.0f;
    } else if (1.0f <= fmin_payoff(xss[3], xss[4], xss[5])) {
      date = 1;
      amount = 1300.0f;
    } else if (1.0f <= fmin_payoff(xss[6], xss[7], xss[8])) {
      date = 2;
      amount = 1450.0f;
    } else if (1.0f <= fmin_payoff(xss[9], xss[10], xss[11])) {
      date = 3;
      amount = 1600.0f;
    } else {
      const float x50 = fmin_payoff(xss[12], xss[13], xss[14]);
      date = 4;
      amount = 1.0f <= x50 ? 1750.0f : (0.75f < x50 ? 1000.0f : x50 * 1000.0f);
    }
    return amount * disc[date];
  } else {  // payoff3 (:267-283); price[] holds date 366
    const float price1 = 100.0f * disc[0];
    const bool goto40 =
        cond && (price[0] < 3758.05f || price[1] < 11840.0f || price[2] < 1200.0f);
    const float amount = goto40 ? 1000.0f * fmin_payoff(price[0], price[1], price[2]) : 1000.0f;
    const float price2 = amount * disc[1];
    return price1 + price2;
  }
}

template <int C>
__global__ __launch_bounds__(kThreads) void op_paths_kernel(
    const int32_t *__restrict__ dv, const int32_t *__restrict__ bbi,
    const float *__restrict__ bbd, const float *__restrict__ cs,
    const float *__restrict__ vols, const float *__restrict__ drifts,
    const float *__restrict__ sts, const float *__restrict__ detvals,
    const float *__restrict__ discts, int bits, int K, int M, int64_t N,
    const int32_t *__restrict__ bb_ok, double *__restrict__ partial) {
  const Market mk = {cs, vols, drifts, sts, detvals, discts, K};
  constexpr int D = Shape<C>::D, U = Shape<C>::U, DIMS = D * U;
  // The high part of the Sobol numbers: for contract 1's one dimension
  // each wave makes its own from its wave-uniform Gray code (scalar loads
  // and scalar ALU, no barrier); for more dimensions the workgroup
  // tabulates it in LDS (per-wave scalar evaluation of contract 2's 15
  // measured 1.9x slower).
  constexpr bool kLdsHigh = DIMS > 1;
  __shared__ int32_t xhi[kWaves][kLdsHigh ? DIMS : 1];
  __shared__ float tailbuf[kWaves][U * 64];
  __shared__ double red[kThreads];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t glo = (uint32_t)(lane ^ (lane >> 1));
  double sum[OP_MAX_MODELS];
#pragma unroll
  for (int m = 0; m < OP_MAX_MODELS; ++m) sum[m] = 0.0;
  // x / 2 ** num_bits (:31-35): for any i32 x and bits <= 32 the quotient
  // is a power-of-two scaling of (float)x with no underflow, hence exact,
  // and so is the product with the exact reciprocal 2 ** -num_bits: same
  // bits, one multiply instead of an IEEE division
  const float inv_divisor = ldexpf(1.0f, -bits);
  float rows[U][D];
  const bool zero_free = __builtin_amdgcn_readfirstlane(*bb_ok) != 0;
  // blocks of kThreads Sobol indices covering 0 .. N (index 0 is no path)
  const int64_t nblk = N / kThreads + 1;
  for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int64_t n0 = blk * kThreads;
    // grayCode (:12) of this wave's 64-aligned base (bits 0-4 are zero)
    const uint32_t cw = (uint32_t)(n0 + 64 * wave);
    const uint32_t ghi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(cw ^ (cw >> 1)));
    if constexpr (kLdsHigh) {
      __syncthreads();  // the previous block's readers of xhi are done
      for (int e = tid; e < kWaves * DIMS; e += kThreads) {
        const int w = e / DIMS, d = e - w * DIMS;
        const uint32_t c = (uint32_t)(n0 + 64 * w);
        const uint32_t g = c ^ (c >> 1);
        const int32_t *v = dv + d * OP_DV_STRIDE;
        int32_t x = 0;
#pragma unroll
        for (int b = 0; b < OP_DV_STRIDE; ++b) x ^= ((g >> b) & 1u) ? v[b] : 0;
        xhi[w][d] = x;
      }
      __syncthreads();
    }
    const int64_t n = n0 + tid;  // Sobol index of path n - 1 (:314)
    const bool live = n >= 1 && n <= N;
    // the Sobol number of dimension d of this path (:29-35)
    auto sobol = [&](int d) {
      const int32_t *v = dv + d * OP_DV_STRIDE;
      int32_t x = 0;
      if constexpr (kLdsHigh) {
        x = xhi[wave][d];
      } else {
#pragma unroll
        for (int b = 5; b < OP_DV_STRIDE; ++b) x ^= ((ghi >> b) & 1u) ? v[b] : 0;
      }
#pragma unroll
      for (int b = 0; b < 6; ++b) x ^= ((glo >> b) & 1u) ? v[b] : 0;
      return (float)x * inv_divisor;
    };
    // the Gaussians (:121-140) of dimensions i * U .. i * U + U - 1
    auto gauss_date = [&](int i, float(&z)[U]) {
      float p[U];
#pragma unroll
      for (int u = 0; u < U; ++u) p[u] = sobol(i * U + u);
      if constexpr (U > 1) {
        ugaussian_wave<U>(p, z, tailbuf[wave], lane);
      } else {
#pragma unroll
        for (int u = 0; u < U; ++u) z[u] = ugaussian(p[u]);
      }
    };
    // the Brownian bridge of each underlying (:145-177) over the Gaussians
    // of dimensions i * U + u (the [num_dates][num_und] reshape, :311)
    // (the reference reads bb_inds unchecked, :156; indexes are clamped to
    // the row so that no input makes an access out of bounds)
    const int32_t *bi = bbi, *li = bbi + D, *ri = bbi + 2 * D;
    const float *sd = bbd, *lw = bbd + D, *rw = bbd + 2 * D;
    auto clamp = [](int i) { return i < 0 ? 0 : (i >= D ? D - 1 : i); };
    // The U rows are independent chains: each step reads all of them,
    // then computes, then writes, so a lane overlaps U chains' memory
    // latency (a step's own reads precede its write, as in :157-167).
    // The rows start at zero (:153) only where that can be read: not when
    // the bridge writes every date once and reads only dates written at an
    // earlier step (op_bridge_check_kernel), the datasets' tables included.
    if (!zero_free) {
      for (int i = 0; i < D; ++i) {
#pragma unroll
        for (int u = 0; u < U; ++u) rows[u][i] = 0.0f;
      }
    }
    {
      const int l0 = clamp(bi[0] - 1);
      float z0[U];
      gauss_date(0, z0);
#pragma unroll
      for (int u = 0; u < U; ++u) rows[u][l0] = sd[0] * z0[u];
    }
    for (int i = 1; i < D; ++i) {
      const int j = li[i] - 1, k = clamp(ri[i] - 1), l = clamp(bi[i] - 1), jc = clamp(j);
      float wk[U], rj[U], zi[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        wk[u] = rows[u][k];
        rj[u] = rows[u][jc];
      }
      gauss_date(i, zi);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float tmp = rw[i] * wk[u] + sd[i] * zi[u];
        rows[u][l] = j == -1 ? tmp : tmp + lw[i] * rj[u];
      }
    }
    for (int m = 0; m < M; ++m) {
      const float pay = model_payoff<C>(mk, m, rows);
      if (live) sum[m] += (double)pay;
    }
  }
  // the workgroup's sums, model by model, by a fixed tree
  for (int m = 0; m < M; ++m) {
    red[tid] = sum[m];
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    if (tid == 0) partial[(int64_t)blockIdx.x * M + m] = red[0];
    __syncthreads();
  }
}

// Whether the bridge tables make the rows' zero start unobservable: with
// the path kernel's clamped indexes, step i (of D) writes date l_i and, for
// i >= 1, reads dates k_i and (when li != 0) j_i.  When every date is written
// by exactly one step and every read date by an earlier step, each read sees
// that step's value and every date ends written, with or without the zeros.
// One workgroup; *ok = 1 when that holds, 0 otherwise or past kCheckDates.
constexpr int kCheckDates = 4096;
__global__ __launch_bounds__(kThreads) void op_bridge_check_kernel(const int32_t *__restrict__ bbi,
                                                                   int D, int32_t *__restrict__ ok) {
  __shared__ int32_t first[kCheckDates], count[kCheckDates];
  __shared__ int32_t bad;
  const int tid = threadIdx.x;
  if (D > kCheckDates) {
    if (tid == 0) *ok = 0;
    return;  // uniform: no barrier passed yet
  }
  for (int p = tid; p < D; p += kThreads) {
    first[p] = 0x7fffffff;
    count[p] = 0;
  }
  if (tid == 0) bad = 0;
  __syncthreads();
  const int32_t *bi = bbi, *li = bbi + D, *ri = bbi + 2 * D;
  auto clamp = [D](int i) { return i < 0 ? 0 : (i >= D ? D - 1 : i); };
  for (int i = tid; i < D; i += kThreads) {
    const int l = clamp(bi[i] - 1);
    atomicMin(&first[l], i);
    atomicAdd(&count[l], 1);
  }
  __syncthreads();
  int b = 0;
  for (int i = tid; i < D; i += kThreads) {
    if (count[i] != 1) b = 1;
    if (i >= 1) {
      if (first[clamp(ri[i] - 1)] >= i) b = 1;
      const int j = li[i] - 1;
      if (j != -1 && first[clamp(j)] >= i) b = 1;
    }
  }
  if (b) atomicOr(&bad, 1);
  __syncthreads();
  if (tid == 0) *ok = bad ? 0 : 1;
}

// The mean per model: the workgroups' partial sums by a fixed tree, / N.
__global__ __launch_bounds__(kThreads) void op_mean_kernel(const double *partial, int nblocks,
                                                           int M, int64_t N, float *out) {
  __shared__ double red[kThreads];
  for (int m = 0; m < M; ++m) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kThreads) s += partial[(int64_t)b * M + m];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = kThreads / 2; k > 0; k >>= 1) {
      if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[m] = (float)(red[0] / (double)N);
    __syncthreads();
  }
}

// dir_vs [dims][bits] -> [dims][OP_DV_STRIDE], zero past bits
__global__ void op_pad_kernel(const int32_t *in, int32_t *out, int dims, int bits) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < dims * OP_DV_STRIDE) {
    const int d = i / OP_DV_STRIDE, b = i % OP_DV_STRIDE;
    out[i] = b < bits ? in[d * bits + b] : 0;
  }
}

}  // namespace

extern "C" {

int op_blocks(int64_t paths) {
  const int64_t b = paths / kThreads + 1;
  return (int)(b < OP_MAX_BLOCKS ? b : OP_MAX_BLOCKS);
}

int op_launch(void *stream, int contract, const op_args *A, const int32_t *dir_vs,
              int32_t *dv, double *partial, float *out) {
  hipStream_t st = (hipStream_t)stream;
  const int dims = contract == 1 ? 1 : contract == 2 ? 15 : contract == 3 ? 1101 : 0;
  if (dims == 0 || A->M < 1 || A->bits < 0 || A->bits > OP_DV_STRIDE || A->N < 1)
    return (int)hipErrorInvalidValue;
  const int np = dims * OP_DV_STRIDE;
  op_pad_kernel<<<(np + 255) / 256, 256, 0, st>>>(dir_vs, dv, dims, A->bits);
  const int nblocks = op_blocks(A->N);
  const int D = contract == 1 ? 1 : contract == 2 ? 5 : 367, U = contract == 1 ? 1 : 3;
  int32_t *bb_ok = dv + np;  // the bridge check's flag, after the padded vectors
  op_bridge_check_kernel<<<1, kThreads, 0, st>>>(A->bbi, D, bb_ok);
  // the models in chunks of OP_MAX_MODELS (a lane's sums are registers): each
  // chunk reruns the paths with its models' market data and writes its means
  for (int m0 = 0; m0 < A->M; m0 += OP_MAX_MODELS) {
    const int mc = A->M - m0 < OP_MAX_MODELS ? A->M - m0 : OP_MAX_MODELS;
#define OP_PATHS_ARGS                                                                   \
  dv, A->bbi, A->bbd, A->cs + (int64_t)m0 * U * U, A->vols + (int64_t)m0 * D * U,       \
      A->drifts + (int64_t)m0 * D * U, A->sts + (int64_t)m0 * U, A->detvals + m0,       \
      A->discts + (int64_t)m0 * A->K, A->bits, A->K, mc, A->N, bb_ok, partial
    switch (contract) {
      case 1: op_paths_kernel<1><<<nblocks, kThreads, 0, st>>>(OP_PATHS_ARGS); break;
      case 2: op_paths_kernel<2><<<nblocks, kThreads, 0, st>>>(OP_PATHS_ARGS); break;
      default: op_paths_kernel<3><<<nblocks, kThreads, 0, st>>>(OP_PATHS_ARGS); break;
    }
#undef OP_PATHS_ARGS
    op_mean_kernel<<<1, kThreads, 0, st>>>(partial, nblocks, mc, A->N, out + m0);
  }
  return (int)hipGetLastError();
}

}  // extern "C"
