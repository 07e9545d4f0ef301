// HIP kernels for gfx950 (MI355X): the LocVolCalib ADI time stepper.
//
// Reference: finpar/LocVolCalib.fut.  One time step of `rollback`
// (:180-209) is two kernels here:
//
//   lvc_xsweep  explicit x and y stencils (:195-200), the implicit x solves
//               (:202, tridagPar :68-124) and the combination
//               y = dtInv*u - 0.5*v (:204-207); reads R, writes Y.
//   lvc_ysweep  the implicit y solves (:208-209); reads Y, writes R.
//
// The reference's transposes (:198, :200, :206, :209) never materialise: the
// x sweep reads rows of R in place and the y sweep stages column panels of Y
// through LDS.
//
// Arithmetic form.  MuX = MuY = 0 (:58, :60), so only the second-derivative
// taps remain, and every interior row of the derivative operator sums to zero
// (:35).  The stencils are therefore evaluated in difference form,
//     a*(r[i-1] - r[i]) + c*(r[i+1] - r[i]),   a = -0.25*var*dd0, c = -0.25*var*dd2,
// which is the reference's operator with its centre tap replaced by the exact
// -(dd0 + dd2), and the implicit systems are
//     s*x[i] + a*(x[i-1] - x[i]) + c*(x[i+1] - x[i]) = d[i],   s = dtInv.
// Both avoid the cancellation of large taps against nearly equal values,
// which is what makes an FP32 state accurate.  The systems are M-matrices
// (a, c <= 0), solved without pivoting.
//
// Tridiagonal solve (one system = one row or column of length n): the system
// is cut into segments of SEG consecutive unknowns, one segment per lane, and
// the `lanes` (a power of two >= ceil(n/SEG)) lanes of a system are
// consecutive lanes of one wavefront.  Each lane eliminates its segment
// locally (Thomas on the first SEG-1 rows, written so that every pivot is a
// sum of positive terms), which leaves one unknown per lane (its last) in a
// tridiagonal reduced system; that is solved by parallel cyclic reduction
// with wavefront shuffles, and each lane then back-substitutes.  Lanes past
// the end of the system hold identity rows.
//
// Precision: T = float or double for the state and the sweeps.  REF adds one
// iterative-refinement pass to each implicit solve: the residual of the FP32
// solution is formed in FP64 (with the FP64 taps and the same rounded time
// step dtInv that the explicit part and the combination use, so that the
// scheme stays consistent), solved again in FP32 with the same
// factorisation, and added in FP64.

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "lvc_internal.h"

namespace {

// ---------------------------------------------------------------------------
// Setup: initGrid (:12-23), initOperator (:26-44) taps, the VarX factor
// x^(2 beta) of updateParams (:61-65) and setPayoff (:46-47).  Computed in
// FP64 with multiply-add contraction disabled, so the grid is bit-identical
// to a C restatement of the same FP64 lines.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void lvc_setup_kernel(lvc_problem *probs,
                                                        double *geo, T *R) {
#pragma clang fp contract(off)
  lvc_problem &pr = probs[blockIdx.x];
  const int nx = pr.nx, ny = pr.ny, nt = pr.nt;
  const double s0 = pr.s0, t = pr.t, alpha = pr.alpha, nu = pr.nu, beta = pr.beta;
  double *x = geo + pr.geo_off;
  double *ax = x + nx, *cx = ax + nx, *y = cx + nx, *ay = y + ny, *cy = ay + ny;

  const double logAlpha = log(alpha);
  const double stdX = 20.0 * alpha * s0 * sqrt(t);
  const double stdY = 10.0 * nu * sqrt(t);
  const double dx = stdX / (double)nx, dy = stdY / (double)ny;
  const int xi = (int)(s0 / dx);  // i32.f64: truncation toward zero
  const int yi = ny / 2;
  for (int i = threadIdx.x; i < nx; i += blockDim.x)
    x[i] = (double)i * dx - (double)xi * dx + s0;
  for (int j = threadIdx.x; j < ny; j += blockDim.x)
    y[j] = (double)j * dy - (double)yi * dy + logAlpha;
  if (threadIdx.x == 0) {
    pr.xindex = xi;
    pr.yindex = yi;
    pr.k = nt - 2;
    pr.status = (xi >= 0 && xi < nx) ? 0 : 3;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nx; i += blockDim.x) {
    double d0 = 0.0, d2 = 0.0;
    if (i >= 1 && i <= nx - 2) {
      const double dl = x[i] - x[i - 1], du = x[i + 1] - x[i];
      d0 = 2.0 / dl / (dl + du);
      d2 = 2.0 / du / (dl + du);
    }
    const double X = exp(2.0 * (beta * log(x[i])));
    ax[i] = -0.25 * X * d0;
    cx[i] = -0.25 * X * d2;
  }
  for (int j = threadIdx.x; j < ny; j += blockDim.x) {
    double d0 = 0.0, d2 = 0.0;
    if (j >= 1 && j <= ny - 2) {
      const double dl = y[j] - y[j - 1], du = y[j + 1] - y[j];
      d0 = 2.0 / dl / (dl + du);
      d2 = 2.0 / du / (dl + du);
    }
    ay[j] = -0.25 * nu * nu * d0;
    cy[j] = -0.25 * nu * nu * d2;
  }
  T *res = R + pr.state_off;
  const int64_t N = (int64_t)nx * ny;
  for (int64_t e = threadIdx.x; e < N; e += blockDim.x) {
    const double v = x[e % nx] - pr.strike;
    res[e] = (T)(v > 0.0 ? v : 0.0);
  }
}

// Timeline point k of initGrid (:14), t * k / (numT - 1): a multiply and a
// divide, bit-identical to a C evaluation of the same expression.
__device__ __forceinline__ double timeline(double t, int k, int nt) {
  return t * (double)k / ((double)nt - 1.0);
}

// ---------------------------------------------------------------------------
// Segment solver shared by both sweeps.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T shfl_up_g(T v, int d, int width) {
  return __shfl_up(v, d, width);
}
template <typename T>
__device__ __forceinline__ T shfl_down_g(T v, int d, int width) {
  return __shfl_down(v, d, width);
}

constexpr int kMaxLevels = 6;  // PCR levels: lanes per system <= 64

template <typename T, int M>
struct SegSys {
  T c[M];    // super-diagonal taps of the segment's rows
  T m[M];    // elimination multipliers (rows 1..M-2)
  T rb[M];   // reciprocal pivots (rows 0..M-2)
  T v[M];    // response of rows 0..M-2 to the previous lane's last unknown
  T w[M];    // response of rows 0..M-2 to this lane's last unknown
  T aL, cL;  // taps of row M-1 (the lane's reduced row)
  T k1[kMaxLevels], k2[kMaxLevels];
  T rB;      // reciprocal of the reduced pivot after PCR
  int seg, lanes, levels;

  // a, c: taps of rows 0..M-1 of this segment; s: diagonal shift per row
  // (dtInv for rows of the system, 1 for padding rows past its end).
  __device__ __forceinline__ void prepare(const T (&a)[M], const T (&cc)[M],
                                          const T (&s)[M], int seg_, int lanes_) {
    seg = seg_;
    lanes = lanes_;
#pragma unroll
    for (int k = 0; k < M; ++k) c[k] = cc[k];
    aL = a[M - 1];
    cL = c[M - 1];
    // Local factorisation of rows 0..M-2: g = pivot + |c|, every update a sum
    // of positive terms for an M-matrix.
    T g = s[0] - a[0];
    rb[0] = T(1) / (g - c[0]);
    m[0] = T(0);
#pragma unroll
    for (int k = 1; k < M - 1; ++k) {
      m[k] = a[k] * rb[k - 1];
      g = s[k] - m[k] * g;
      rb[k] = T(1) / (g - c[k]);
    }
    // Homogeneous responses v (to x[first-1]) and w (to x[last]).
    v[0] = -a[0];
#pragma unroll
    for (int k = 1; k < M - 1; ++k) v[k] = -m[k] * v[k - 1];
    v[M - 2] = v[M - 2] * rb[M - 2];
    w[M - 2] = -c[M - 2] * rb[M - 2];
#pragma unroll
    for (int k = M - 3; k >= 0; --k) {
      v[k] = (v[k] - c[k] * v[k + 1]) * rb[k];
      w[k] = -c[k] * w[k + 1] * rb[k];
    }
    // Reduced row of this lane: A X[seg-1] + B X[seg] + C X[seg+1] = D.
    const bool last = (seg == lanes - 1);
    T vn = shfl_down_g(v[0], 1, lanes), wn = shfl_down_g(w[0], 1, lanes);
    if (last) {
      vn = T(0);
      wn = T(0);
    }
    T A = aL * v[M - 2];
    T B = s[M - 1] - aL * (T(1) - w[M - 2]) - cL * (T(1) - vn);
    T C = cL * wn;
    // PCR on the matrix; the multipliers are kept for the right-hand sides.
    levels = 0;
#pragma unroll
    for (int l = 0; l < kMaxLevels; ++l) {
      const int st = 1 << l;
      if (st < lanes) {
        T Am = shfl_up_g(A, st, lanes), Bm = shfl_up_g(B, st, lanes);
        T Cm = shfl_up_g(C, st, lanes);
        T Ap = shfl_down_g(A, st, lanes), Bp = shfl_down_g(B, st, lanes);
        T Cp = shfl_down_g(C, st, lanes);
        if (seg < st) {
          Am = T(0);
          Bm = T(1);
          Cm = T(0);
        }
        if (seg + st >= lanes) {
          Ap = T(0);
          Bp = T(1);
          Cp = T(0);
        }
        const T q1 = A / Bm, q2 = C / Bp;
        k1[l] = q1;
        k2[l] = q2;
        A = -Am * q1;
        C = -Cp * q2;
        B = B - Cm * q1 - Ap * q2;
        levels = l + 1;
      }
    }
    rB = T(1) / B;
  }

  // Solves for one right-hand side d (rows 0..M-1 of this segment).
  __device__ __forceinline__ void solve(const T (&d)[M], T (&x)[M]) const {
    x[0] = d[0];
#pragma unroll
    for (int k = 1; k < M - 1; ++k) x[k] = d[k] - m[k] * x[k - 1];
    x[M - 2] = x[M - 2] * rb[M - 2];
#pragma unroll
    for (int k = M - 3; k >= 0; --k) x[k] = (x[k] - c[k] * x[k + 1]) * rb[k];
    T un = shfl_down_g(x[0], 1, lanes);
    if (seg == lanes - 1) un = T(0);
    T D = d[M - 1] - aL * x[M - 2] - cL * un;
#pragma unroll
    for (int l = 0; l < kMaxLevels; ++l) {
      if (l < levels) {
        const int st = 1 << l;
        T Dm = shfl_up_g(D, st, lanes), Dp = shfl_down_g(D, st, lanes);
        if (seg < st) Dm = T(0);
        if (seg + st >= lanes) Dp = T(0);
        D = D - Dm * k1[l] - Dp * k2[l];
      }
    }
    const T X = D * rB;
    T XL = shfl_up_g(X, 1, lanes);
    if (seg == 0) XL = T(0);
#pragma unroll
    for (int k = 0; k < M - 1; ++k) x[k] = x[k] + v[k] * XL + w[k] * X;
    x[M - 1] = X;
  }

  // Values of the neighbouring unknowns x[first-1] and x[last+1] (0 outside
  // the system).
  __device__ __forceinline__ void halo(const T (&x)[M], T &xl, T &xr) const {
    xl = shfl_up_g(x[M - 1], 1, lanes);
    xr = shfl_down_g(x[0], 1, lanes);
    if (seg == 0) xl = T(0);
    if (seg == lanes - 1) xr = T(0);
  }
};

// One refinement pass: r = d - A x in FP64 with the FP64 taps, e = A^-1 r
// with the FP32 factorisation, x += e in FP64.
template <typename T, int M>
__device__ __forceinline__ void refine(const SegSys<T, M> &sys, const T (&d)[M],
                                       const double (&a64)[M],
                                       const double (&c64)[M],
                                       const double (&s64)[M], T (&x)[M]) {
  T xl, xr;
  sys.halo(x, xl, xr);
  T r[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const double xk = (double)x[k];
    const double xm = (double)(k > 0 ? x[k - 1] : xl);
    const double xp = (double)(k < M - 1 ? x[k + 1] : xr);
    const double ax = s64[k] * xk + a64[k] * (xm - xk) + c64[k] * (xp - xk);
    r[k] = (T)((double)d[k] - ax);
  }
  T e[M];
  sys.solve(r, e);
#pragma unroll
  for (int k = 0; k < M; ++k) x[k] = (T)((double)x[k] + (double)e[k]);
}

// ---------------------------------------------------------------------------
// x sweep: one workgroup = `count` rows of one problem, `lanes` lanes per row.
// ---------------------------------------------------------------------------
template <typename T, int M, bool REF>
__global__ __launch_bounds__(LVC_BLOCK) void lvc_xsweep_kernel(
    const lvc_problem *__restrict__ probs, const lvc_tile *__restrict__ tiles,
    const double *__restrict__ geo, const T *__restrict__ R, T *__restrict__ Y) {
  const lvc_tile tile = tiles[blockIdx.x];
  const lvc_problem &pr = probs[tile.problem];
  const int kt = pr.k;
  if (kt < 0) return;  // problem finished: whole workgroup exits
  const int G = tile.lanes;
  const int seg = threadIdx.x & (G - 1);
  const int rloc = threadIdx.x / G;
  if (rloc >= tile.count) return;  // whole lane group exits together
  const int nx = pr.nx, ny = pr.ny;
  const int j = tile.first + rloc;
  const double *gx = geo + pr.geo_off;
  const double *ax = gx + nx, *cx = ax + nx, *gy = cx + nx;
  const double *ay = gy + ny, *cy = ay + ny;

  const double tnow = timeline(pr.t, kt, pr.nt), tnext = timeline(pr.t, kt + 1, pr.nt);
  const double dt64 = 1.0 / (tnext - tnow);
  const T dt = (T)dt64;
  const double nu = pr.nu;
  const double E = exp(2.0 * gy[j] - nu * nu * tnow);  // exp(2y - nu^2 t)
  const T ayj = (T)ay[j], cyj = (T)cy[j];

  const T *row = R + pr.state_off + (int64_t)j * nx;
  const int i0 = seg * M;
  T rc[M], rm[M], rp[M];
  T rl = (i0 > 0 && i0 - 1 < nx) ? row[i0 - 1] : T(0);
  T rr = (i0 + M < nx) ? row[i0 + M] : T(0);
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const int i = i0 + k;
    const bool in = i < nx;
    rc[k] = in ? row[i] : T(0);
    rm[k] = (in && j > 0) ? row[i - nx] : T(0);
    rp[k] = (in && j < ny - 1) ? row[i + nx] : T(0);
  }
  T a[M], c[M], s[M], U[M], V[M];
  double a64[M], c64[M], s64[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const int i = i0 + k;
    const bool in = i < nx;
    a64[k] = in ? ax[i] * E : 0.0;
    c64[k] = in ? cx[i] * E : 0.0;
    s64[k] = in ? (double)dt : 1.0;  // the time step the sweep uses
    a[k] = (T)a64[k];
    c[k] = (T)c64[k];
    s[k] = in ? dt : T(1);
    const T r0 = rc[k];
    const T rw = k > 0 ? rc[k - 1] : rl;
    const T re = k < M - 1 ? rc[k + 1] : rr;
    const T ux = a[k] * (rw - r0) + c[k] * (re - r0);            // 0.5 * explicitX
    const T vy = T(-2) * (ayj * (rm[k] - r0) + cyj * (rp[k] - r0));  // explicitY
    V[k] = vy;
    U[k] = dt * r0 - ux + vy;
  }
  SegSys<T, M> sys;
  sys.prepare(a, c, s, seg, G);
  T x[M];
  sys.solve(U, x);
  if (REF) refine(sys, U, a64, c64, s64, x);
  T *out = Y + pr.state_off + (int64_t)j * nx;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const int i = i0 + k;
    if (i < nx) out[i] = dt * x[k] - T(0.5) * V[k];
  }
}

// ---------------------------------------------------------------------------
// y sweep: one workgroup = `count` columns of one problem, `lanes` lanes per
// column; the column panel is staged through LDS (coalesced row loads and
// stores), stored segment-interleaved so that lanes of one column read
// different banks.
// ---------------------------------------------------------------------------
template <typename T, int M, bool REF>
__global__ __launch_bounds__(LVC_BLOCK) void lvc_ysweep_kernel(
    lvc_problem *probs, const lvc_tile *__restrict__ tiles,
    const double *__restrict__ geo, const T *__restrict__ Y, T *__restrict__ R) {
  __shared__ T lds[M * (LVC_BLOCK + 64)];
  const lvc_tile tile = tiles[blockIdx.x];
  const lvc_problem &pr = probs[tile.problem];
  const int kt = pr.k;
  if (kt < 0) return;
  const int G = tile.lanes;
  const int cols = LVC_BLOCK / G;  // columns per workgroup
  const int CP = cols + 1;         // LDS pitch
  const int nx = pr.nx, ny = pr.ny;
  const int ncol = tile.count, col0 = tile.first;
  const int rows_pad = G * M;
  auto lidx = [&](int r, int cc) { return ((r % M) * G + r / M) * CP + cc; };

  const T *src = Y + pr.state_off;
  for (int e = threadIdx.x; e < rows_pad * cols; e += LVC_BLOCK) {
    const int r = e / cols, cc = e % cols;
    T val = T(0);
    if (r < ny && cc < ncol) val = src[(int64_t)r * nx + col0 + cc];
    lds[lidx(r, cc)] = val;
  }
  __syncthreads();

  const double *gx = geo + pr.geo_off;
  const double *gy = gx + 3 * nx, *ay = gy + ny, *cy = ay + ny;
  const double tnow = timeline(pr.t, kt, pr.nt), tnext = timeline(pr.t, kt + 1, pr.nt);
  const double dt64 = 1.0 / (tnext - tnow);
  const T dt = (T)dt64;
  const int q = threadIdx.x & (G - 1);
  const int cc = threadIdx.x / G;
  const int j0 = q * M;
  T a[M], c[M], s[M], d[M];
  double a64[M], c64[M], s64[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const int jj = j0 + k;
    const bool in = jj < ny;
    a64[k] = in ? ay[jj] : 0.0;
    c64[k] = in ? cy[jj] : 0.0;
    s64[k] = in ? (double)dt : 1.0;  // the time step the sweep uses
    a[k] = (T)a64[k];
    c[k] = (T)c64[k];
    s[k] = in ? dt : T(1);
    d[k] = lds[lidx(jj, cc)];
  }
  SegSys<T, M> sys;
  sys.prepare(a, c, s, q, G);
  T x[M];
  sys.solve(d, x);
  if (REF) refine(sys, d, a64, c64, s64, x);
  __syncthreads();  // every lane has read its panel values
#pragma unroll
  for (int k = 0; k < M; ++k) lds[lidx(j0 + k, cc)] = x[k];
  __syncthreads();
  T *dst = R + pr.state_off;
  for (int e = threadIdx.x; e < ny * cols; e += LVC_BLOCK) {
    const int r = e / cols, c2 = e % cols;
    if (c2 < ncol) dst[(int64_t)r * nx + col0 + c2] = lds[lidx(r, c2)];
  }
  // The last of the problem's y workgroups to finish advances its time index
  // (every workgroup of the problem has read `k` before it counts itself).
  if (threadIdx.x == 0) {
    lvc_problem *pw = const_cast<lvc_problem *>(&pr);
    const int done = atomicAdd(&pw->ydone, 1) + 1;
    if (done == pw->ytiles) {
      pw->ydone = 0;
      pw->k = kt - 1;
    }
  }
}

// value (:236): the price at (yindex, xindex).
template <typename T, typename O>
__global__ void lvc_gather_kernel(const lvc_problem *probs, int n, const T *R,
                                  O *out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const lvc_problem &pr = probs[p];
  if (pr.status != 0) {
    out[p] = (O)NAN;
    return;
  }
  out[p] = (O)R[pr.state_off + (int64_t)pr.yindex * pr.nx + pr.xindex];
}

template <typename T, int M, bool REF>
int launch_sweeps(hipStream_t st, bool xs, const lvc_problem *probs,
                  const lvc_tile *tiles, int ntiles, const double *geo,
                  const void *in, void *out) {
  if (ntiles <= 0) return 0;
  if (xs)
    lvc_xsweep_kernel<T, M, REF><<<ntiles, LVC_BLOCK, 0, st>>>(
        probs, tiles, geo, (const T *)in, (T *)out);
  else
    lvc_ysweep_kernel<T, M, REF><<<ntiles, LVC_BLOCK, 0, st>>>(
        const_cast<lvc_problem *>(probs), tiles, geo, (const T *)in, (T *)out);
  return (int)hipGetLastError();
}

template <typename T, bool REF>
int dispatch_seg(hipStream_t st, bool xs, int seg, const lvc_problem *probs,
                 const lvc_tile *tiles, int ntiles, const double *geo,
                 const void *in, void *out) {
  switch (seg) {
    case 4: return launch_sweeps<T, 4, REF>(st, xs, probs, tiles, ntiles, geo, in, out);
    case 8: return launch_sweeps<T, 8, REF>(st, xs, probs, tiles, ntiles, geo, in, out);
    case 16: return launch_sweeps<T, 16, REF>(st, xs, probs, tiles, ntiles, geo, in, out);
    default: return (int)hipErrorInvalidValue;
  }
}

int dispatch(hipStream_t st, bool xs, int precision, int seg,
             const lvc_problem *probs, const lvc_tile *tiles, int ntiles,
             const double *geo, const void *in, void *out) {
  switch (precision) {
    case LVC_FP32: return dispatch_seg<float, false>(st, xs, seg, probs, tiles, ntiles, geo, in, out);
    case LVC_FP32_REFINE: return dispatch_seg<float, true>(st, xs, seg, probs, tiles, ntiles, geo, in, out);
    case LVC_FP64: return dispatch_seg<double, false>(st, xs, seg, probs, tiles, ntiles, geo, in, out);
    default: return (int)hipErrorInvalidValue;
  }
}

}  // namespace

extern "C" {

int lvc_seg_supported(int seg) { return seg == 4 || seg == 8 || seg == 16; }

int lvc_launch_setup(void *stream, int precision, lvc_problem *probs, int nprob,
                     double *geo, void *R) {
  if (nprob <= 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  if (precision == LVC_FP64)
    lvc_setup_kernel<double><<<nprob, 256, 0, st>>>(probs, geo, (double *)R);
  else
    lvc_setup_kernel<float><<<nprob, 256, 0, st>>>(probs, geo, (float *)R);
  return (int)hipGetLastError();
}

int lvc_launch_xsweep(void *stream, int precision, int seg,
                      const lvc_problem *probs, const lvc_tile *tiles,
                      int ntiles, const double *geo, const void *R, void *Y) {
  return dispatch((hipStream_t)stream, true, precision, seg, probs, tiles,
                  ntiles, geo, R, Y);
}

int lvc_launch_ysweep(void *stream, int precision, int seg,
                      const lvc_problem *probs, const lvc_tile *tiles,
                      int ntiles, const double *geo, const void *Y, void *R) {
  return dispatch((hipStream_t)stream, false, precision, seg, probs, tiles,
                  ntiles, geo, Y, R);
}

int lvc_launch_gather(void *stream, int precision, lvc_problem *probs,
                      int nprob, const void *R, void *out, int out_f64) {
  if (nprob <= 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  const int nb = (nprob + 255) / 256;
  if (precision == LVC_FP64) {
    if (out_f64)
      lvc_gather_kernel<double, double><<<nb, 256, 0, st>>>(probs, nprob, (const double *)R, (double *)out);
    else
      lvc_gather_kernel<double, float><<<nb, 256, 0, st>>>(probs, nprob, (const double *)R, (float *)out);
  } else {
    if (out_f64)
      lvc_gather_kernel<float, double><<<nb, 256, 0, st>>>(probs, nprob, (const float *)R, (double *)out);
    else
      lvc_gather_kernel<float, float><<<nb, 256, 0, st>>>(probs, nprob, (const float *)R, (float *)out);
  }
  return (int)hipGetLastError();
}

}  // extern "C"
