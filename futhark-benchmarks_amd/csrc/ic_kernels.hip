// InterestCalib (finpar/InterestCalib.fut) kernels for MI355X (gfx950).
//
// The calibration evaluates, per MCMC iteration, every genome of the
// population on every swaption quote (evalGenomeOnSwap, :484-602): a G2++
// model price from three Brent root finds over the swaption's schedule and
// a Gauss-Hermite sum over it.  That work is independent per (genome,
// swaption) and bounded by the schedule length n (2..60 periods in the
// datasets), so one evaluation runs on a group of lanes holding K schedule
// periods each (K = 1, 2, 4 or 8, a template parameter; the group is a power
// of two of 1..64 lanes), a wavefront holds 64/lanes evaluations of the same
// swaption for consecutive genomes, and every sum, maximum and Brent
// iteration over the schedule is a sum over the lane's K periods followed by
// a butterfly within the group.  The butterfly hands every lane of the group
// the same bits, so a group's control flow (the root finder's branches and
// iteration counts) stays uniform within the group.  An iteration's
// evaluations (25-50K on the datasets) fill only a few wavefronts per SIMD,
// so a launch lasts about as long as its slowest wavefront: the host keeps
// K at 2 (1 for one-period swaptions) up to 128 periods, which measured
// faster than packing more periods into fewer, longer-running lanes, and
// every class of such swaptions runs in one launch.
//
// What does not depend on the genome (dates, discount factors, the Black
// quote, each period's c_i * zc(end) / zc(maturity)) is computed once per
// call by ic_setup_kernel into per-swaption and per-period tables.  The
// log-likelihood terms land in [swaption][genome] order; ic_accept_kernel
// sums them per genome in a fixed order and accepts or rejects, and
// ic_propose_kernel draws the next proposals, so an iteration is three
// launches, each spread over the device.
//
// Arithmetic is f32 in the reference's order of operations, compiled
// without contraction; expf/logf/sqrtf are the device library's, so results
// differ from a host libm in the last bits.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "ic_internal.h"

namespace {

constexpr float kEps0 = 1.0e-3f;  // :20
constexpr float kEps = 1.0e-5f;   // :21
constexpr float kCauchyOffs = 5.0f;
constexpr float kR = 0.03f;       // :27
constexpr float kPi = 3.14159265358979323846f;

// ---- the calendar (lib/github.com/diku-dk/date/date.fut:48-150) ----------
// Integer / and % of Futhark round towards negative infinity.
__device__ inline int32_t fdiv(int32_t a, int32_t b) {
  int32_t q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
  return q;
}
__device__ inline int32_t fmod_(int32_t a, int32_t b) {
  int32_t r = a % b;
  if (r != 0 && ((r < 0) != (b < 0))) r += b;
  return r;
}

constexpr int32_t kMinutesInDay = 24 * 60;
constexpr int32_t kMinutesToNoon = 12 * 60;

// minutes since the epoch of (y, m, d, hour, minute), date.fut:65-75
__device__ int32_t date_of_gregorian(int32_t y, int32_t m, int32_t d, int32_t hr, int32_t mn) {
  int32_t j;
  if (m == 1 || m == 2)
    j = fdiv(1461 * (y + 4800 - 1), 4) + fdiv(367 * (m + 10), 12) -
        fdiv(3 * fdiv(y + 4900 - 1, 100), 4);
  else
    j = fdiv(1461 * (y + 4800), 4) + fdiv(367 * (m - 2), 12) - fdiv(3 * fdiv(y + 4900, 100), 4);
  return (j + d - 32075 - 2444238) * kMinutesInDay + hr * 60 + mn;
}

__device__ int32_t end_of_month(int32_t y, int32_t m) {  // date.fut:93-100
  const bool leap = fmod_(y, 4) == 0 && (fmod_(y, 400) == 0 || fmod_(y, 100) != 0);
  if (m == 2) return leap ? 29 : 28;
  if (m == 4 || m == 6 || m == 9 || m == 11) return 30;
  return 31;
}

// add_months (date.fut:128-137) over gregorian_of_date (:77-91)
__device__ int32_t add_months(int32_t t, int32_t nbmonths) {
  const int32_t jul = fdiv(t, kMinutesInDay);
  int32_t l = jul + 68569 + 2444238;
  const int32_t n = fdiv(4 * l, 146097);
  l = l - fdiv(146097 * n + 3, 4);
  const int32_t i = fdiv(4000 * (l + 1), 1461001);
  l = l - fdiv(1461 * i, 4) + 31;
  const int32_t j = fdiv(80 * l, 2447);
  const int32_t day = l - fdiv(2447 * j, 80);
  l = fdiv(j, 11);
  const int32_t month = j + 2 - 12 * l;
  const int32_t year = 100 * (n - 49) + i + l;
  const int32_t daytime = fmod_(t, kMinutesInDay);
  int32_t hr = 12, mn = 0;
  if (daytime != kMinutesToNoon) {
    hr = fdiv(daytime, 60);
    mn = fmod_(daytime, 60);
  }
  int32_t m = month + nbmonths;
  int32_t y = year + fdiv(m - 1, 12);
  m = fmod_(m - 1, 12) + 1;
  if (m <= 0) {
    y -= 1;
    m += 12;
  }
  const int32_t eom = end_of_month(y, m);
  return date_of_gregorian(y, m, day < eom ? day : eom, hr, mn);
}

// date_act_365(t1, t2) = f32.f64 (diff_dates t2 t1) (:31, date.fut:139-143)
__device__ inline float act365(int32_t t1, int32_t t2) {
  const double inv_minutes_in_day = 1.0 / (double)kMinutesInDay;
  return (float)((double)(t1 - t2) * inv_minutes_in_day / 365.0);
}

__device__ inline int32_t today() { return date_of_gregorian(2012, 1, 1, 12, 0); }

__device__ inline float zc(int32_t t) { return expf(-kR * act365(t, today())); }  // :48

// ---- the math module (:56-101) --------------------------------------------
// Abramowitz & Stegun 7.1.26 without its exp factor
__device__ inline float erf_poly(float x) {
  const float p = 0.3275911f;
  const float a1 = 0.254829592f, a2 = -0.284496736f, a3 = 1.421413741f, a4 = -1.453152027f,
              a5 = 1.061405429f;
  const float t = 1.0f / (1.0f + p * x);
  const float t2 = t * t;
  const float t3 = t * t2;
  const float t4 = t2 * t2;
  const float t5 = t2 * t3;
  return a1 * t + a2 * t2 + a3 * t3 + a4 * t4 + a5 * t5;
}

__device__ inline float ugaussian_P(float x) {  // :60-93
  const float u = x / sqrtf(2.0f);
  const float e = u < 0.0f ? -(1.0f - erf_poly(-u) * expf(-(u * u)))
                           : 1.0f - erf_poly(u) * expf(-(u * u));
  return 0.5f * (1.0f + e);
}

__device__ inline float ugaussian_P_exp(float x, float exp_factor) {  // :95-101
  const float u = fabsf(x / sqrtf(2.0f));
  const float e = erf_poly(u);
  const float res = 0.5f * e * expf(exp_factor - u * u);
  return 0.0f <= x ? expf(exp_factor) - res : res;
}

__device__ inline float log_likelihood(float y_ref, float y) {  // Cauchy, :606-625
  const float gamma = (fabsf(y_ref) / 50.0f) * kCauchyOffs + 0.01f;
  const float x = (y - y_ref) / gamma;
  const float pdf = 1.0f / (kPi * gamma * (1.0f + x * x));
  return logf(pdf + 1.0e-20f);
}

// ---- Sobol numbers (:190-202) ------------------------------------------------
// Unrolled over the largest vector so the direction numbers load once, up
// front, rather than one scalar load and wait per bit.
__device__ inline float sobol(const ic_dirvec &dv, int32_t n) {
  const float norm_fact = 1.0f / ((float)((int64_t)1 << dv.bits) + 1.0f);
  const int32_t n_gray = (n >> 1) ^ n;
  int32_t res = 0;
#pragma unroll
  for (int i = 0; i < IC_MAX_BITS; i++) {
    const int32_t t = (int32_t)(1u << i);
    if (i < dv.bits && (n_gray & t) == t) res ^= dv.dir[i];
  }
  return (float)res * norm_fact;
}

// ---- genomes (:220-320) -------------------------------------------------------
__device__ inline float gene_min(int j) { return j == 2 ? kEps0 - 1.0f : kEps0; }
__device__ inline float gene_max(int j) { return j < 3 ? 1.0f - kEps0 : 0.2f; }
constexpr float kAmplRatio = 0.005f;  // moves_unif_ampl_ratio, :248

// ---- G2++ (:322-413) -------------------------------------------------------------
__device__ inline float b_fun(float z, float tau) { return (1.0f - expf(-z * tau)) / z; }

__device__ inline float t_fun(float sigma, float x, float tau) {
  const float expxtau = expf(-x * tau);
  const float exp2xtau = expxtau * expxtau;
  return sigma * sigma / (x * x) *
         (tau + 2.0f / x * expxtau - 1.0f / (2.0f * x) * exp2xtau - 3.0f / (2.0f * x));
}

struct Genome {
  float a, b, rho, nu, sigma;
};

// bigv (:336-352): V(tau) and B_a(tau), B_b(tau)
__device__ inline float bigv(const Genome &g, float tau, float *ba, float *bb) {
  const float sigma = g.sigma == 0.0f ? 1.0e-10f : g.sigma;
  *ba = b_fun(g.a, tau);
  *bb = b_fun(g.b, tau);
  const float t1 = t_fun(sigma, g.a, tau);
  const float t2 = t_fun(g.nu, g.b, tau);
  const float t3 = 2.0f * g.rho * g.nu * sigma / (g.a * g.b) *
                   (tau - *ba - *bb + b_fun(g.a + g.b, tau));
  return t1 + t2 + t3;
}

// bigmx / bigmy (:364-413) over (today, maturity, today, maturity): ts =
// tmat0 = tmats = t0, tmatt = s0 = 0
__device__ inline float bigmx(const Genome &g, float tmat0) {
  const float a = g.a, b = g.b, rho = g.rho, nu = g.nu, sigma = g.sigma;
  const float ts = tmat0, tmatt = 0.0f, tmats = tmat0, t0 = tmat0, s0 = 0.0f;
  const float tmp1 = (sigma * sigma) / (a * a) + (sigma * rho * nu) / (a * b);
  const float tmp2 = 1.0f - expf(-a * ts);
  const float tmp3 = sigma * sigma / (2.0f * a * a);
  const float tmp4 = rho * sigma * nu / (b * (a + b));
  const float tmp5 = expf(-a * tmatt) - expf(-a * (tmats + ts));
  const float tmp6 = expf(-b * tmatt) - expf(-b * tmat0 - a * t0 + (a + b) * s0);
  return tmp1 * tmp2 - (tmp3 * tmp5) - (tmp4 * tmp6);
}

__device__ inline float bigmy(const Genome &g, float tmat0) {
  const float a = g.a, b = g.b, rho = g.rho, nu = g.nu, sigma = g.sigma;
  const float ts = tmat0, tmatt = 0.0f, tmats = tmat0, t0 = tmat0, s0 = 0.0f;
  const float tmp1 = nu * nu / (b * b) + sigma * rho * nu / (a * b);
  const float tmp2 = 1.0f - expf(-b * ts);
  const float tmp3 = nu * nu / (2.0f * b * b);
  const float tmp4 = sigma * rho * nu / (a * (a + b));
  const float tmp5 = expf(-b * tmatt) - expf(-b * (tmats + ts));
  const float tmp6 = expf(-a * tmatt) - expf(-a * tmat0 - b * t0 + (a + b) * s0);
  return tmp1 * tmp2 - (tmp3 * tmp5) - (tmp4 * tmp6);
}

// ---- group reductions ---------------------------------------------------------------
// Butterflies over the `lanes` lanes of an evaluation group (a power of two,
// wave-uniform); every lane of the group ends with the same bits.
__device__ inline float group_sum(float v, int lanes) {
  for (int m = 1; m < lanes; m <<= 1) v += __shfl_xor(v, m, 64);
  return v;
}
__device__ inline float group_max(float v, int lanes) {
  for (int m = 1; m < lanes; m <<= 1) v = fmaxf(v, __shfl_xor(v, m, 64));
  return v;
}

// One lane's schedule periods (K per lane) and their genome-dependent values.
template <int K>
struct Periods {
  float bai[K], bbi[K], aici[K], log_aici[K], scale[K], t1cs[K], cs[K];
  bool valid[K];
};

// sum over the group of scales * exp(-bbi * y), minus 1 (:108-117, fid 1)
template <int K>
__device__ inline float to_solve(const Periods<K> &P, const float (&ups)[K], float y, int lanes) {
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < K; k++)
    if (P.valid[k]) s += ups[k] * expf(-P.bbi[k] * y);
  return group_sum(s, lanes) - 1.0f;
}

// rootFinding_Brent (:124-185); the loop ends where the reference's
// iterations stop changing its state
template <int K>
__device__ float brent(const Periods<K> &P, const float (&ups)[K], int lanes, float lb, float ub,
                       float tol, int32_t iter_max, float *err) {
  tol = tol <= 0.0f ? 1.0e-9f : tol;
  iter_max = iter_max <= 0 ? 10000 : iter_max;
  float a = lb, b = ub;
  float fa = to_solve(P, ups, a, lanes), fb = to_solve(P, ups, b, lanes);
  if (0.0f <= fa * fb) {
    *err = 0.0f <= a ? INFINITY : -INFINITY;
    return 0.0f;
  }
  if (fabsf(fa) < fabsf(fb)) {
    float t = fa; fa = fb; fb = t;
    t = a; a = b; b = t;
  }
  float c = a, fc = fa, d = 0.0f;
  bool mflag = true;
  for (int32_t i = 0; i < iter_max; i++) {
    if (fb == 0.0f || fabsf(b - a) < tol) break;
    float s;
    if (fa == fc || fb == fc) {
      s = b - fb * (b - a) / (fb - fa);
    } else {
      const float s1 = (a * fb * fc) / ((fa - fb) * (fa - fc));
      const float s2 = (b * fa * fc) / ((fb - fa) * (fb - fc));
      const float s3 = (c * fa * fb) / ((fc - fa) * (fc - fb));
      s = s1 + s2 + s3;
    }
    if (!((3.0f * a + b) / 4.0f <= s && s <= b) ||
        (mflag && fabsf(b - c) / 2.0f <= fabsf(s - b)) ||
        (!mflag && fabsf(c - d) / 2.0f <= fabsf(s - b)) ||
        (mflag && fabsf(b - c) <= fabsf(tol)) || (!mflag && fabsf(c - d) <= fabsf(tol))) {
      mflag = true;
      s = (a + b) / 2.0f;
    } else {
      mflag = false;
    }
    const float fs = to_solve(P, ups, s, lanes);
    d = c;
    c = b;
    fc = fb;
    if (fa * fs < 0.0f) {
      b = s;
      fb = fs;
    } else {
      a = s;
      fa = fs;
    }
    if (fabsf(fa) < fabsf(fb)) {
      float t = a; a = b; b = t;
      t = fa; fa = fb; fb = t;
    }
  }
  *err = fb;
  return b;
}

struct Scals {
  float b, sigmax, sigmay, rhoxy, rhoxyc, rhoxycs, mux, muy, bb_first, bb_last;
};

// exactYhat (:418-481); bb_first and bb_last are bbi[0] and bbi[n-1]
template <int K>
__device__ float exact_yhat(const Periods<K> &P, int n, int lanes, const Scals &sc, float x) {
  const float k = -3.71901648545568f;
  float ups[K];
  float up = 0.0f, lo = -INFINITY;
#pragma unroll
  for (int q = 0; q < K; q++) {
    ups[q] = 0.0f;
    if (P.valid[q]) {
      const float baix = P.bai[q] * x;
      ups[q] = P.aici[q] * expf(-baix);
      up += ups[q];
      lo = fmaxf(lo, (P.log_aici[q] - baix) / P.bbi[q]);
    }
  }
  up = group_sum(up, lanes);
  lo = group_max(lo, lanes);
  if (n < 2) return lo;
  const float log_s = logf(up);
  float tmp = log_s / sc.bb_last;
  if (!(tmp <= 0.0f)) {
    tmp = log_s / sc.bb_first;
    tmp = 0.0f <= tmp ? tmp : -INFINITY;
  }
  const float yl = lo - kEps, yu = tmp + kEps;
  const float y0 = sc.sigmay * (sc.rhoxy * (x - sc.mux) / sc.sigmax + k * sc.rhoxycs) -
                   sc.rhoxyc / sc.b + sc.muy;
  const float y1 = sc.sigmay * (sc.rhoxy * (x - sc.mux) / sc.sigmax - k * sc.rhoxycs) + sc.muy;
  if (y1 <= yl) return y1 + 1.0f;
  if (yu <= y0) return y0 - 1.0f;
  float err;
  const float root = brent(P, ups, lanes, fmaxf(yl, y0), fminf(yu, y1), 1.0e-4f, 1000, &err);
  if (err == -INFINITY) return y0 - 1.0f;
  if (err == INFINITY) return y1 + 1.0f;
  return root;
}

}  // namespace

// ---- per-call setup: what does not depend on the genome (:486-525, :533-547)
// One workgroup per swaption, its periods spread over the threads; the
// quote's level is summed in period order by one thread, as the reference's
// sequential reduction.
constexpr int kSetupThreads = 64;

__global__ __launch_bounds__(kSetupThreads) void ic_setup_kernel(
    const float *__restrict__ swp, int64_t stride, int32_t nswap, ic_swaption *__restrict__ sw,
    ic_period *__restrict__ per) {
  __shared__ float a12s[IC_MAX_PERIODS];
  __shared__ int32_t a1s[IC_MAX_PERIODS], a2s[IC_MAX_PERIODS];
  __shared__ float strike_sh;
  const int32_t s = blockIdx.x;
  if (s >= nswap) return;
  const float *row = swp + (int64_t)s * stride;
  const ic_swaption S0 = sw[s];
  const int32_t n = S0.n;
  const int32_t t_today = today();
  const float swap_freq = row[1];
  const int32_t maturity = add_months(t_today, (int32_t)(row[0] * 12.0f));  // add_years, :33
  const float tmat0 = act365(maturity, t_today);
  for (int32_t i = threadIdx.x; i < n; i += kSetupThreads) {
    const int32_t a1 = add_months(maturity, (int32_t)(swap_freq * (float)(int64_t)i));
    const int32_t a2 = add_months(a1, (int32_t)swap_freq);
    a1s[i] = a1;
    a2s[i] = a2;
    a12s[i] = zc(a2) * act365(a2, a1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float lvl = 0.0f;
    int32_t t0 = date_of_gregorian(2229, 12, 31, 12, 0), tn = date_of_gregorian(1980, 1, 1, 12, 0);
    for (int32_t i = 0; i < n; i++) {
      lvl = lvl + a12s[i];
      t0 = t0 < a1s[i] ? t0 : a1s[i];
      tn = tn > a2s[i] ? tn : a2s[i];
    }
    const float strike = (zc(t0) - zc(tn)) / lvl;
    const float d1 = 0.5f * row[3] * tmat0;
    ic_swaption S = S0;
    S.quote = lvl * strike * (ugaussian_P(d1) - ugaussian_P(-d1));
    S.tmat0 = tmat0;
    S.zc_mat = zc(maturity);
    sw[s] = S;
    strike_sh = strike;
  }
  __syncthreads();
  const float strike = strike_sh;
  const float zc_mat = zc(maturity);
  for (int32_t i = threadIdx.x; i < n; i += kSetupThreads) {
    const int32_t beg = a1s[i], end = a2s[i];  // the same dates as the quote's (:533-536)
    const float res = act365(end, beg) * strike;
    const float cii = i == n - 1 ? 1.0f + res : res;
    ic_period p;
    p.tod1 = act365(end, t_today);
    p.tod2 = act365(end, maturity);
    p.fact = cii * zc(end) / zc_mat;
    p.log_fact = logf(p.fact);
    per[S0.off + i] = p;
  }
}

__global__ void ic_init_kernel(int32_t pop, ic_dirvec dv, float *__restrict__ genomes,
                               float *__restrict__ props) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pop) return;
  for (int j = 0; j < 5; j++) {  // initGenome (:233-238) of sobolInd at 5i+1 .. 5i+5
    const float r01 = sobol(dv, 5 * i + 1 + j);
    const float g = r01 * (gene_max(j) - gene_min(j)) + gene_min(j);
    genomes[5 * i + j] = g;
    props[5 * i + j] = g;
  }
}

// ---- the evaluation ------------------------------------------------------------------
struct EvalArgs {
  const ic_swaption *sw;
  const ic_period *per;
  const ic_work *work;
  int32_t nwork;
  const float *genomes;
  int32_t npop;
  const float *hx, *hw;
  int32_t nherm, nswap, mode;
  float *out0, *out1;
};

// One wavefront's evaluations, K schedule periods a lane.
template <int K>
__device__ __forceinline__ void eval_wave(const EvalArgs &A, const ic_work wk, const ic_swaption S,
                                          const int lane) {
  const int lanes = S.lanes;
  const int n = S.n;
  const int j = lane & (lanes - 1);
  int32_t gi = wk.genome0 + lane / lanes;
  const bool store = gi < A.npop && j == 0;
  gi = gi < A.npop ? gi : A.npop - 1;
  const Genome g = {A.genomes[5 * gi], A.genomes[5 * gi + 1], A.genomes[5 * gi + 2],
                    A.genomes[5 * gi + 3], A.genomes[5 * gi + 4]};
  const float tmat0 = S.tmat0;

  // :528-546
  float dum1, dum2;
  const float v0_mat = bigv(g, tmat0, &dum1, &dum2);
  const float mux = 0.0f - bigmx(g, tmat0);
  const float muy = 0.0f - bigmy(g, tmat0);
  const float sqrt_bfun_a = sqrtf(b_fun(2.0f * g.a, tmat0));
  const float sqrt_bfun_b = sqrtf(b_fun(2.0f * g.b, tmat0));
  const float rhoxy = g.rho * b_fun(g.a + g.b, tmat0) / (sqrt_bfun_a * sqrt_bfun_b);
  const float sigmax = g.sigma * sqrt_bfun_a;
  const float sigmay = g.nu * sqrt_bfun_b;
  const float rhoxyc = 1.0f - rhoxy * rhoxy;
  const float rhoxycs = sqrtf(rhoxyc);
  const float sigmay_rhoxycs = sigmay * rhoxycs;
  const float t4 = (rhoxy * sigmay) / sigmax;

  // the schedule's periods (:549-583), period j + lanes * k on this lane
  Periods<K> P;
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int i = j + lanes * k;
    P.valid[k] = i < n;
    P.bai[k] = P.bbi[k] = P.aici[k] = P.log_aici[k] = P.scale[k] = P.t1cs[k] = P.cs[k] = 0.0f;
    if (P.valid[k]) {
      const ic_period pe = A.per[S.off + i];
      float bai, bbi;
      const float v0_end = bigv(g, pe.tod1, &dum1, &dum2);
      const float vt_end = bigv(g, pe.tod2, &bai, &bbi);
      const float expo_aici = 0.5f * (vt_end - v0_end + v0_mat);
      P.bai[k] = bai;
      P.bbi[k] = bbi;
      P.aici[k] = pe.fact * expf(expo_aici);
      P.log_aici[k] = pe.log_fact + expo_aici;
      P.scale[k] = 0.0f - (bai + bbi * t4);
      P.cs[k] = pe.fact;
      P.t1cs[k] = bbi * (mux * t4 - (muy - 0.5f * rhoxyc * sigmay * sigmay * bbi)) + expo_aici;
    }
  }
  Scals sc;
  sc.b = g.b;
  sc.sigmax = sigmax;
  sc.sigmay = sigmay;
  sc.rhoxy = rhoxy;
  sc.rhoxyc = rhoxyc;
  sc.rhoxycs = rhoxycs;
  sc.mux = mux;
  sc.muy = muy;
  sc.bb_first = n > 0 ? b_fun(g.b, A.per[S.off].tod2) : 0.0f;
  sc.bb_last = n > 0 ? b_fun(g.b, A.per[S.off + n - 1].tod2) : 0.0f;

  // :586-591
  const float eps = 0.5f * sigmax;
  const float f = exact_yhat(P, n, lanes, sc, mux);
  const float gg = exact_yhat(P, n, lanes, sc, mux + eps);
  const float h = exact_yhat(P, n, lanes, sc, mux - eps);
  const float df = 0.5f * (gg - h) / eps;
  const float sqrt2sigmax = sqrtf(2.0f) * sigmax;
  const float t2 = rhoxy / (sigmax * rhoxycs);

  // Gauss-Hermite (:594-620)
  float accum = 0.0f;
  for (int q = 0; q < A.nherm; q++) {
    const float x_quad = A.hx[q], w_quad = A.hw[q];
    const float x = sqrt2sigmax * x_quad + mux;
    const float yhat_x = f + df * (x - mux);
    const float h1 = ((yhat_x - muy) / sigmay_rhoxycs) - t2 * (x - mux);
    float part = 0.0f;
#pragma unroll
    for (int k = 0; k < K; k++)
      if (P.valid[k]) {
        const float h2 = h1 + P.bbi[k] * sigmay_rhoxycs;
        const float expo_aici = P.t1cs[k] + P.scale[k] * x;
        part += P.cs[k] * ugaussian_P_exp(-h2, expo_aici);
      }
    const float accum1 = group_sum(part, lanes);
    const float tmp = sqrtf(2.0f) * x_quad;
    const float t1 = expf(-0.5f * tmp * tmp);
    accum = accum + w_quad * t1 * (ugaussian_P(-h1) - accum1);
  }
  const float price = S.zc_mat * (accum / sqrtf(kPi));
  if (!store) return;
  const int32_t s = wk.swaption;
  if (A.mode == IC_OUT_TERMS) {
    A.out0[(int64_t)s * A.npop + gi] = log_likelihood(S.quote, price);
  } else if (A.mode == IC_OUT_PRICES) {
    A.out0[(int64_t)gi * A.nswap + s] = S.quote;
    A.out1[(int64_t)gi * A.nswap + s] = price;
  } else {  // makeSummary (:205-213)
    const float err_ratio = (price - S.quote) / S.quote;
    A.out0[3 * s] = 10000.0f * price;
    A.out0[3 * s + 1] = 10000.0f * S.quote;
    A.out0[3 * s + 2] = 100.0f * fabsf(err_ratio);
  }
}

// Swaptions of 1 or 2 periods a lane (up to 128 periods) in one launch, so
// the classes share the device; a class is uniform per wavefront.
__global__ __launch_bounds__(256) void ic_eval_kernel(EvalArgs A) {
  const int32_t wv = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (wv >= A.nwork) return;  // wave-uniform
  const ic_work wk = A.work[wv];
  const ic_swaption S = A.sw[wk.swaption];
  if (S.per_lane == 1)
    eval_wave<1>(A, wk, S, threadIdx.x & 63);
  else
    eval_wave<2>(A, wk, S, threadIdx.x & 63);
}

// Longer schedules (4 or 8 periods a lane, up to 512 periods): more registers.
__global__ __launch_bounds__(256) void ic_eval_long_kernel(EvalArgs A) {
  const int32_t wv = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (wv >= A.nwork) return;
  const ic_work wk = A.work[wv];
  const ic_swaption S = A.sw[wk.swaption];
  if (S.per_lane == 4)
    eval_wave<4>(A, wk, S, threadIdx.x & 63);
  else
    eval_wave<8>(A, wk, S, threadIdx.x & 63);
}

// ---- accept / propose ------------------------------------------------------------------
struct StepArgs {
  int32_t pop, nswap;
  const float *terms;
  float *genomes, *props, *logliks, *fb;
  int32_t init, acc, it, conv;
  ic_move next;
  ic_dirvec dv;
  ic_trace trace;
};

constexpr int kSumLanes = 32;  // threads per genome of the accept kernel

// The log-likelihood of each genome (:639-645) and the accept step
// (:733-745).  Accepting needs only the genome's own terms, so the genomes
// spread over the device: 32 lanes per genome, lane t summing the swaptions
// t, t + 32, ... in order (one memory round trip on the datasets), the 32
// partial sums added pairwise with the same bits on every lane.
__global__ __launch_bounds__(256) void ic_accept_kernel(StepArgs A) {
  const int32_t pop = A.pop;
  const int t = threadIdx.x & (kSumLanes - 1);
  const int32_t i = (int32_t)(blockIdx.x * (256 / kSumLanes) + threadIdx.x / kSumLanes);
  float s = 0.0f;
  if (i < pop) {
    // this lane's terms w = t, t + kSumLanes, ... in that order, eight loads
    // issued before their sums (a lane has ceil(nswap / kSumLanes) terms: 7
    // for 196 swaptions, so one batch)
    const float *col = A.terms + i;
    for (int32_t w = t; w < A.nswap; w += 8 * kSumLanes) {
      float v[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const int32_t wk = w + k * kSumLanes;
        v[k] = wk < A.nswap ? col[(int64_t)wk * pop] : 0.0f;
      }
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (w + k * kSumLanes < A.nswap) s += v[k];
    }
  }
  for (int m = 1; m < kSumLanes; m <<= 1) s += __shfl_xor(s, m, 64);
  if (i >= pop || t != 0) return;
  if (A.init) {
    A.logliks[i] = s;
  } else {
    const float acceptance = fminf(1.0f, expf(s - A.logliks[i]) * A.fb[i]);
    const float r = sobol(A.dv, A.acc + i);
    if (A.trace.newl) A.trace.newl[(int64_t)A.it * pop + i] = s;
    if (r < acceptance) {
      // genomes and proposals are separate arrays: the five loads first
      float g[5];
      for (int j = 0; j < 5; j++) g[j] = A.props[5 * i + j];
      for (int j = 0; j < 5; j++) A.genomes[5 * i + j] = g[j];
      A.logliks[i] = s;
    }
  }
  if (A.trace.genomes) {
    const int64_t row = (int64_t)(A.it + 1) * pop + i;
    float g[5];
    for (int j = 0; j < 5; j++) g[j] = A.genomes[5 * i + j];
    for (int j = 0; j < 5; j++) A.trace.genomes[5 * row + j] = g[j];
    A.trace.logliks[row] = A.logliks[i];
  }
}

// The next iteration's proposals (:655-724), one thread per genome, after
// every genome's accept step (differential evolution reads other genomes).
__global__ __launch_bounds__(256) void ic_propose_kernel(StepArgs A) {
  const int32_t pop = A.pop;
  const int32_t i = (int32_t)(blockIdx.x * 256 + threadIdx.x);
  if (i >= pop) return;
  const ic_move mv = A.next;
  float out[5], fb = 1.0f;
  if (mv.type == 1 || mv.type == 2) {  // mutate_dims_all / mutate_dims_one (:278-306)
    for (int j = 0; j < 5; j++) {
      const float r01 = sobol(A.dv, 5 * i + mv.base + j);
      const float ampl = mv.type == 1 || j == mv.dim ? kAmplRatio : 0.0f;
      const float g_min = gene_min(j), g_max = gene_max(j);
      const float gene = A.genomes[5 * i + j], prop = A.props[5 * i + j];
      const float amplitude = fabsf((g_max - g_min) * ampl);  // mutateHelper, :250-266
      const float semiamplitude = amplitude / 2.0f;
      float tmp_min_max = fminf(g_max, gene + semiamplitude);
      float tmp_max_min = fmaxf(g_min, gene - semiamplitude);
      const float forward_range = tmp_min_max - tmp_max_min;
      tmp_min_max = fminf(g_max, prop + semiamplitude);
      tmp_max_min = fmaxf(g_min, prop - semiamplitude);
      const float backward_range = tmp_min_max - tmp_max_min;
      const float bf = 0.0f < semiamplitude ? backward_range / forward_range : 1.0f;
      const float ng = gene + amplitude * r01 - semiamplitude;
      out[j] = fmaxf(g_min, fminf(g_max, ng));  // constrainDim, :268-270
      fb = j == 0 ? bf : fb * bf;
    }
  } else {  // DE (:700-723, mcmc_DE :309-320)
    const int32_t kk = 8 * i + mv.base;
    const float s1 = sobol(A.dv, kk);
    int32_t k = (int32_t)(s1 * (float)(pop - 1));
    int32_t cand_ub = pop - 1;
    if (k == i) {
      k = pop - 1;
      cand_ub = pop - 2;
    }
    const float s2 = sobol(A.dv, kk + 1);
    int32_t l = (int32_t)(s2 * (float)cand_ub);
    if (l == i || l == k) l = cand_ub;
    const float s3 = sobol(A.dv, kk + 2);
    const float gamma_avg = 2.38f / sqrtf(2.0f * 5.0f);
    const float ampl_ratio = 0.1f * kAmplRatio;
    const float gamma1 = gamma_avg - 0.5f + s3;
    for (int j = 0; j < 5; j++) {
      const float r01 = sobol(A.dv, kk + 3 + j);
      const float mm_diff = gene_max(j) - gene_min(j);
      const float amplitude = fabsf(mm_diff * ampl_ratio);
      const float semiamplitude = amplitude / 2.0f;
      const float perturb = amplitude * r01 - semiamplitude;
      const float ng = A.genomes[5 * i + j] + perturb +
                       gamma1 * (A.genomes[5 * k + j] - A.genomes[5 * l + j]);
      out[j] = fmaxf(gene_min(j), fminf(gene_max(j), ng));
    }
  }
  for (int j = 0; j < 5; j++) A.props[5 * i + j] = out[j];
  A.fb[i] = fb;
  if (A.trace.props && A.it + 1 < A.conv) {
    const int64_t row = (int64_t)(A.it + 1) * pop + i;
    for (int j = 0; j < 5; j++) A.trace.props[5 * row + j] = out[j];
    A.trace.fb[row] = fb;
  }
}

__global__ void ic_winner_kernel(int32_t pop, const float *__restrict__ genomes,
                                 const float *__restrict__ logliks, float *__restrict__ out,
                                 float *__restrict__ winner) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int32_t wi = 0;  // :752-756, the reduction read as a left fold
  float wv = -INFINITY;
  for (int32_t i = 0; i < pop; i++)
    if (wv < logliks[i]) {
      wi = i;
      wv = logliks[i];
    }
  const float *w = genomes + 5 * wi;
  out[0] = w[0];
  out[1] = w[1];
  out[2] = w[4];
  out[3] = w[3];
  out[4] = w[2];
  out[5] = wv;
  for (int j = 0; j < 5; j++) winner[j] = w[j];
}

// ---- launchers ---------------------------------------------------------------------------
extern "C" int ic_launch_setup(void *stream, const float *swaptions, int64_t stride,
                               int32_t nswap, ic_swaption *sw, ic_period *periods) {
  if (nswap <= 0) return 0;
  hipLaunchKernelGGL(ic_setup_kernel, dim3(nswap), dim3(kSetupThreads), 0, (hipStream_t)stream,
                     swaptions, stride, nswap, sw, periods);
  return (int)hipGetLastError();
}

extern "C" int ic_launch_init(void *stream, int32_t pop, const ic_dirvec *dv, float *genomes,
                              float *props) {
  if (pop <= 0) return 0;
  hipLaunchKernelGGL(ic_init_kernel, dim3((pop + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     pop, *dv, genomes, props);
  return (int)hipGetLastError();
}

extern "C" int ic_launch_eval(void *stream, const ic_swaption *sw, const ic_period *periods,
                              const ic_worklists *wl, const float *genomes, int32_t npop,
                              const float *herm_x, const float *herm_w, int32_t nherm,
                              int32_t nswap, int mode, float *out0, float *out1) {
  EvalArgs A;
  A.sw = sw;
  A.per = periods;
  A.genomes = genomes;
  A.npop = npop;
  A.hx = herm_x;
  A.hw = herm_w;
  A.nherm = nherm;
  A.nswap = nswap;
  A.mode = mode;
  A.out0 = out0;
  A.out1 = out1;
  for (int c = IC_LAUNCHES - 1; c >= 0; c--) {
    if (wl->nwork[c] <= 0) continue;
    A.work = wl->work[c];
    A.nwork = wl->nwork[c];
    const dim3 grid((wl->nwork[c] + 3) / 4), block(256);
    if (c == 0)
      hipLaunchKernelGGL(ic_eval_kernel, grid, block, 0, (hipStream_t)stream, A);
    else
      hipLaunchKernelGGL(ic_eval_long_kernel, grid, block, 0, (hipStream_t)stream, A);
  }
  return (int)hipGetLastError();
}

extern "C" int ic_launch_step(void *stream, int32_t pop, int32_t nswap, const float *terms,
                              float *genomes, float *props, float *logliks, float *fb, int init,
                              int32_t acc, const ic_move *next, const ic_dirvec *dv, int32_t it,
                              int32_t conv, const ic_trace *trace) {
  StepArgs A;
  A.pop = pop;
  A.nswap = nswap;
  A.terms = terms;
  A.genomes = genomes;
  A.props = props;
  A.logliks = logliks;
  A.fb = fb;
  A.init = init;
  A.acc = acc;
  A.it = it;
  A.conv = conv;
  A.next = *next;
  A.dv = *dv;
  if (trace) {
    A.trace = *trace;
  } else {
    A.trace.genomes = A.trace.logliks = A.trace.props = A.trace.fb = A.trace.newl = nullptr;
  }
  if (pop <= 0) return 0;
  const int64_t per_block = 256 / kSumLanes;
  hipLaunchKernelGGL(ic_accept_kernel, dim3((unsigned)((pop + per_block - 1) / per_block)),
                     dim3(256), 0, (hipStream_t)stream, A);
  if (A.next.type != 0)
    hipLaunchKernelGGL(ic_propose_kernel, dim3((unsigned)((pop + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, A);
  return (int)hipGetLastError();
}

extern "C" int ic_launch_winner(void *stream, int32_t pop, const float *genomes,
                                const float *logliks, float *out, float *winner) {
  hipLaunchKernelGGL(ic_winner_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, pop, genomes,
                     logliks, out, winner);
  return (int)hipGetLastError();
}
