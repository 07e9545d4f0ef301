/*
 * CPU oracle for finpar/OptionPricing.fut.  TEST INFRASTRUCTURE ONLY: the
 * tests, the smoke check and the benchmark's CPU baseline call it as the
 * checker; the library never does.
 *
 * Every function restates one function of finpar/OptionPricing.fut in f32,
 * in the same evaluation order (C evaluates `a * b + c` as `(a * b) + c`,
 * as Futhark does; built with -ffp-contract=off).  The per-path payoffs are
 * those of the reference; their sum over paths, which the reference reduces
 * in an order its backend chooses (:318), is accumulated here in f64
 * (chunk by chunk of consecutive paths), so that the oracle's mean is the
 * mean of the f32 payoffs up to rounding far below f32's.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Planted-mistake switch: non-zero values make one function of the oracle
 * wrong on purpose, so the tests can show that their checks catch it. */
static int op_oracle_mutation = 0;
void op_oracle_set_mutation(int m) { op_oracle_mutation = m; }
#define OPMUT(k) (op_oracle_mutation == (k))

/* OptionPricing.fut:12 */
static int32_t gray_code(int32_t x) { return OPMUT(1) ? x : (x >> 1) ^ x; }

/* OptionPricing.fut:17-27 (testBit, xorInds) for one direction vector. */
static int32_t xor_inds(int32_t n, const int32_t *dir_v, int64_t num_bits) {
  const int32_t g = gray_code(n);
  int32_t acc = 0;
  for (int64_t i = 0; i < num_bits; i++) {
    const int32_t t = (int32_t)(1u << (uint32_t)i);
    acc ^= ((g & t) == t) ? dir_v[i] : 0;
  }
  return acc;
}

/* OptionPricing.fut:29-35 (sobolIndR): dims D*U values of path n. */
void op_sobol(int32_t n, const int32_t *dir_vs, int64_t dims, int64_t num_bits, float *out) {
  const float divisor = powf(2.0f, (float)num_bits);
  for (int64_t d = 0; d < dims; d++)
    out[d] = (float)xor_inds(n, dir_vs + d * num_bits, num_bits) / divisor;
}

/* OptionPricing.fut:40-56 (polyAppr), Horner as written. */
static float poly_appr(float x, const float a[8], const float b[8]) {
  const float num =
      x * (x * (x * (x * (x * (x * (x * a[7] + a[6]) + a[5]) + a[4]) + a[3]) + a[2]) + a[1]) +
      a[0];
  const float den =
      x * (x * (x * (x * (x * (x * (x * b[7] + b[6]) + b[5]) + b[4]) + b[3]) + b[2]) + b[1]) +
      b[0];
  return num / den;
}

/* OptionPricing.fut:58-79 (smallcase) */
static const float kSmallA[8] = {3.387132872796366608f, 133.14166789178437745f,
                                 1971.5909503065514427f, 13731.693765509461125f,
                                 45921.953931549871457f, 67265.770927008700853f,
                                 33430.575583588128105f, 2509.0809287301226727f};
static const float kSmallB[8] = {1.0f, 42.313330701600911252f, 687.1870074920579083f,
                                 5394.1960214247511077f, 21213.794301586595867f,
                                 39307.89580009271061f, 28729.085735721942674f,
                                 5226.495278852854561f};
/* OptionPricing.fut:81-99 (intermediate) */
static const float kMidA[8] = {1.42343711074968357734f, 4.6303378461565452959f,
                               5.7694972214606914055f, 3.64784832476320460504f,
                               1.27045825245236838258f, 0.24178072517745061177f,
                               0.0227238449892691845833f, 7.7454501427834140764e-4f};
static const float kMidB[8] = {1.0f, 2.05319162663775882187f, 1.6763848301838038494f,
                               0.68976733498510000455f, 0.14810397642748007459f,
                               0.0151986665636164571966f, 5.475938084995344946e-4f,
                               1.05075007164441684324e-9f};
/* OptionPricing.fut:101-119 (tail) */
static const float kTailA[8] = {6.6579046435011037772f, 5.4637849111641143699f,
                                1.7848265399172913358f, 0.29656057182850489123f,
                                0.026532189526576123093f, 0.0012426609473880784386f,
                                2.71155556874348757815e-5f, 2.01033439929228813265e-7f};
static const float kTailB[8] = {1.0f, 0.59983220655588793769f, 0.13692988092273580531f,
                                0.0148753612908506148525f, 7.868691311456132591e-4f,
                                1.8463183175100546818e-5f, 1.4215117583164458887e-7f,
                                2.04426310338993978564e-5f};

/* OptionPricing.fut:121-136 (ugaussianEl) */
float op_ugaussian(float p) {
  const float dp = p - 0.5f;
  if ((dp < 0.0f && 0.0f - dp <= 0.425f) || (0.0f <= dp && dp <= 0.425f))
    return dp * poly_appr(OPMUT(2) ? 0.180625f + dp * dp : 0.180625f - dp * dp, kSmallA,
                          kSmallB);
  const float pp = dp < 0.0f ? dp + 0.5f : 0.5f - dp;
  const float r = sqrtf(-logf(pp));
  const float x = r <= 5.0f ? poly_appr(r - 1.6f, kMidA, kMidB)
                            : poly_appr(r - 5.0f, kTailA, kTailB);
  return dp < 0.0f ? 0.0f - x : x;
}

/* OptionPricing.fut:145-175 (brownianBridgeDates) for one underlying: gauss
 * and out have stride `stride` (the number of underlyings). */
static void bb_dates(int64_t D, const int32_t *bb_inds, const float *bb_data,
                     const float *gauss, int64_t stride, float *row, float *out) {
  const int32_t *bi = bb_inds, *li = bb_inds + D, *ri = bb_inds + 2 * D;
  const float *sd = bb_data, *lw = bb_data + D, *rw = bb_data + 2 * D;
  for (int64_t i = 0; i < D; i++) row[i] = 0.0f;
  row[bi[0] - 1] = sd[0] * gauss[0];
  for (int64_t i = 1; i < D; i++) {
    const int32_t j = li[i] - 1, k = ri[i] - 1, l = bi[i] - 1;
    const float wk = row[k];
    const float zi = gauss[i * stride];
    const float tmp = (OPMUT(3) ? lw[i] : rw[i]) * wk + sd[i] * zi;
    row[l] = j == -1 ? tmp : tmp + lw[i] * row[j];
  }
  for (int64_t ii = 1; ii < D; ii++) {
    const int64_t i = D - ii;
    row[i] = row[i] - row[i - 1];
  }
  for (int64_t i = 0; i < D; i++) out[i * stride] = row[i];
}

/* brownianBridgeDates (:145-183) of one row of D Gaussians, for the tests.
 * Returns 1 if the scratch row cannot be allocated. */
int op_bridge(int64_t D, const int32_t *bb_inds, const float *bb_data, const float *gauss,
              float *out) {
  float *row = malloc((size_t)(D > 0 ? D : 1) * sizeof(float));
  if (!row)
    return 1;
  bb_dates(D, bb_inds, bb_data, gauss, 1, row, out);
  free(row);
  return 0;
}

/* OptionPricing.fut:190-221 (correlateDeltas, combineVs, mkPrices) for one
 * model: bb [D][U] -> prices [D][U]. */
static void black_scholes(int64_t D, int64_t U, const float *bb, const float *md_c,
                          const float *vols, const float *drifts, const float *starts,
                          float *prices) {
  for (int64_t u = 0; u < U; u++) {
    float acc = 1.0f; /* the scan of (map2 (*)) from replicate 1.0 */
    for (int64_t d = 0; d < D; d++) {
      float noise = 0.0f; /* f32.sum of map2 (*) (take (u+1) zi) (take (u+1) md_c[u]) */
      for (int64_t k = 0; k <= (OPMUT(4) ? U - 1 : u); k++) noise += bb[d * U + k] * md_c[u * U + k];
      const float c = drifts[d * U + u] + noise * vols[d * U + u];
      acc = acc * expf(c);
      prices[d * U + u] = starts[u] * acc;
    }
  }
}

/* OptionPricing.fut:230-236 (fminPayoff) */
static float fmin_payoff(const float *xs) {
  const float a = xs[0] / 3758.05f, b = xs[1] / 11840.0f, c = xs[2] / 1200.0f;
  return a < b ? (a < c ? a : c) : (b < c ? b : c);
}

/* OptionPricing.fut:238-293 (trajInner, payoff1-3, genericPayoff) */
static float payoff(int32_t contract, const float *disc, const float *detval,
                    const float *xss) {
  if (contract == 1) {
    const float amount = (xss[0] - 4000.0f) * detval[0];
    const float amount0 = 0.0f < amount ? amount : 0.0f;
    return amount0 * disc[0];
  }
  if (contract == 2) {
    int date;
    float amount;
    if (1.0f <= fmin_payoff(xss + 0)) {
      date = 0;
      amount = 1150.0f;
    } else if (1.0f <= fmin_payoff(xss + 3)) {
      date = 1;
      amount = 1300.0f;
    } else if (1.0f <= fmin_payoff(xss + 6)) {
      date = 2;
      amount = 1450.0f;
    } else if (1.0f <= fmin_payoff(xss + 9)) {
      date = 3;
      amount = 1600.0f;
    } else {
      const float x50 = fmin_payoff(xss + 12);
      date = 4;
      amount = 1.0f <= x50 ? 1750.0f : (0.75f < x50 ? 1000.0f : x50 * 1000.0f);
    }
    return amount * disc[date];
  }
  if (contract == 3) {
    int cond = 0;
    for (int d = 0; d < 367; d++) {
      const float *x = xss + 3 * d;
      cond |= x[0] <= 2630.6349999999998f || x[1] <= 8288.0f || x[2] <= 840.0f;
    }
    const float price1 = 100.0f * disc[0];
    const float *last = xss + 3 * (OPMUT(5) ? 365 : 366);
    const int goto40 = cond && (last[0] < 3758.05f || last[1] < 11840.0f || last[2] < 1200.0f);
    const float amount = goto40 ? 1000.0f * fmin_payoff(last) : 1000.0f;
    const float price2 = amount * disc[1];
    return price1 + price2;
  }
  return 0.0f;
}

/* The payoffs of paths lo .. hi-1 (Sobol indices lo+1 .. hi, :314) summed
 * per model into sums[M] (f64).  D dates, U underlyings, M
 * models, K discounts, B direction-vector bits.  Returns 0, or 1 for a
 * contract whose payoff needs another shape (the reference's size check). */
int op_sum_payoffs(int32_t contract, int64_t lo, int64_t hi, const int32_t *dir_vs,
                   int64_t B, const float *md_cs, const float *md_vols,
                   const float *md_drifts, const float *md_sts, const float *md_detvals,
                   const float *md_discts, int64_t K, const int32_t *bb_inds,
                   const float *bb_data, int64_t D, int64_t U, int64_t M, double *sums,
                   int nthreads) {
  if ((contract == 1 && (D != 1 || U != 1)) || (contract == 2 && (D != 5 || U != 3)) ||
      (contract == 3 && (D != 367 || U != 3)))
    return 1;
  for (int64_t m = 0; m < M; m++) sums[m] = 0.0;
  const int64_t n = hi > lo ? hi - lo : 0;
  const int64_t chunks = nthreads > 1 ? 4 * (int64_t)nthreads : 1;
  double *part = calloc((size_t)(chunks * M), sizeof(double));
  if (!part) return 1;
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
  for (int64_t ch = 0; ch < chunks; ch++) {
    float *g = malloc(sizeof(float) * (size_t)(3 * D * U + D));
    float *s = g, *bb = g + D * U, *px = g + 2 * D * U, *row = g + 3 * D * U;
    const int64_t a = lo + n * ch / chunks, b = lo + n * (ch + 1) / chunks;
    for (int64_t p = a; p < b; p++) {
      op_sobol((int32_t)(p + 1), dir_vs, D * U, B, s);
      for (int64_t i = 0; i < D * U; i++) s[i] = op_ugaussian(s[i]);
      for (int64_t u = 0; u < U; u++) bb_dates(D, bb_inds, bb_data, s + u, U, row, bb + u);
      for (int64_t m = 0; m < M; m++) {
        black_scholes(D, U, bb, md_cs + m * U * U, md_vols + m * D * U,
                      md_drifts + m * D * U, md_sts + m * U, px);
        part[ch * M + m] += (double)payoff(contract, md_discts + m * K, md_detvals + m, px);
      }
    }
    free(g);
  }
  for (int64_t ch = 0; ch < chunks; ch++)
    for (int64_t m = 0; m < M; m++) sums[m] += part[ch * M + m];
  free(part);
  return 0;
}

/* The payoffs themselves, paths lo .. hi-1, model-major: out[m * (hi-lo) + i]. */
int op_payoffs(int32_t contract, int64_t lo, int64_t hi, const int32_t *dir_vs, int64_t B,
               const float *md_cs, const float *md_vols, const float *md_drifts,
               const float *md_sts, const float *md_detvals, const float *md_discts,
               int64_t K, const int32_t *bb_inds, const float *bb_data, int64_t D, int64_t U,
               int64_t M, float *out, int nthreads) {
  if ((contract == 1 && (D != 1 || U != 1)) || (contract == 2 && (D != 5 || U != 3)) ||
      (contract == 3 && (D != 367 || U != 3)))
    return 1;
  const int64_t n = hi > lo ? hi - lo : 0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
  {
    float *g = malloc(sizeof(float) * (size_t)(3 * D * U + D));
    float *s = g, *bb = g + D * U, *px = g + 2 * D * U, *row = g + 3 * D * U;
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      const int64_t p = lo + i;
      op_sobol((int32_t)(p + 1), dir_vs, D * U, B, s);
      for (int64_t q = 0; q < D * U; q++) s[q] = op_ugaussian(s[q]);
      for (int64_t u = 0; u < U; u++) bb_dates(D, bb_inds, bb_data, s + u, U, row, bb + u);
      for (int64_t m = 0; m < M; m++) {
        black_scholes(D, U, bb, md_cs + m * U * U, md_vols + m * D * U,
                      md_drifts + m * D * U, md_sts + m * U, px);
        out[m * n + i] = payoff(contract, md_discts + m * K, md_detvals + m, px);
      }
    }
    free(g);
  }
  return 0;
}
