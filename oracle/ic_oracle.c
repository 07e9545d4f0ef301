/*
 * CPU ORACLE for finpar/InterestCalib.fut -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference program (Futhark source; its
 * compiler is not available here): the G2++ swaption pricer evaluated for a
 * population of genomes on every swaption quote, inside the genetic /
 * differential-evolution MCMC calibration loop, in f32 and in the
 * reference's order of operations (compiled without contraction), with the
 * reference's sequential reductions.  It checks the HIP implementation; only
 * the tests and the benchmark's CPU leg load it.  The product libraries never
 * link or call it.
 *
 * The calendar is a restatement of the reference's vendored date library,
 * lib/github.com/diku-dk/date/date.fut:48-150 (integer minutes since an
 * epoch; Futhark's integer / and % round towards negative infinity).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Planted-mistake switch (the tests show their checks catch each one). */
static int ic_oracle_mutation = 0;
void ic_oracle_set_mutation(int m) { ic_oracle_mutation = m; }
#define ICMUT(k) (ic_oracle_mutation == (k))

/* ---- date.fut ------------------------------------------------------------ */

static int32_t fdiv(int32_t a, int32_t b) { /* Futhark's / on integers */
  int32_t q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
  return q;
}
static int32_t fmod_(int32_t a, int32_t b) { /* Futhark's % on integers */
  int32_t r = a % b;
  if (r != 0 && ((r < 0) != (b < 0))) r += b;
  return r;
}

enum { MINUTES_IN_DAY = 24 * 60, MINUTES_TO_NOON = 12 * 60 };

typedef struct {
  int32_t year, month, day, hour, minute;
} gregorian;

/* date.fut:65-75 (Fliegel and Van Flandern) */
static int32_t date_of_gregorian(gregorian g) {
  const int32_t y = g.year, m = g.month, d = g.day;
  int32_t j;
  if (m == 1 || m == 2)
    j = fdiv(1461 * (y + 4800 - 1), 4) + fdiv(367 * (m + 10), 12) -
        fdiv(3 * fdiv(y + 4900 - 1, 100), 4);
  else
    j = fdiv(1461 * (y + 4800), 4) + fdiv(367 * (m - 2), 12) - fdiv(3 * fdiv(y + 4900, 100), 4);
  return (j + d - 32075 - 2444238) * MINUTES_IN_DAY + g.hour * 60 + g.minute;
}

/* date.fut:77-93 */
static gregorian gregorian_of_date(int32_t t) {
  const int32_t jul = fdiv(t, MINUTES_IN_DAY);
  int32_t l = jul + 68569 + 2444238;
  const int32_t n = fdiv(4 * l, 146097);
  l = l - fdiv(146097 * n + 3, 4);
  const int32_t i = fdiv(4000 * (l + 1), 1461001);
  l = l - fdiv(1461 * i, 4) + 31;
  const int32_t j = fdiv(80 * l, 2447);
  const int32_t d = l - fdiv(2447 * j, 80);
  l = fdiv(j, 11);
  const int32_t m = j + 2 - 12 * l;
  const int32_t y = 100 * (n - 49) + i + l;
  const int32_t daytime = fmod_(t, MINUTES_IN_DAY);
  gregorian g = {y, m, d, 12, 0};
  if (daytime != MINUTES_TO_NOON) {
    g.hour = fdiv(daytime, 60);
    g.minute = fmod_(daytime, 60);
  }
  return g;
}

static int leap(int32_t y) { return fmod_(y, 4) == 0 && (fmod_(y, 400) == 0 || fmod_(y, 100) != 0); }

/* date.fut:98-102 */
static int32_t end_of_month(int32_t y, int32_t m) {
  if (m == 2 && leap(y)) return 29;
  if (m == 2) return 28;
  if (m == 4 || m == 6 || m == 9 || m == 11) return 30;
  return 31;
}

int32_t ic_date_of_triple(int32_t y, int32_t m, int32_t d) {
  gregorian g = {y, m, d, 12, 0};
  return date_of_gregorian(g);
}

void ic_triple_of_date(int32_t t, int32_t *y, int32_t *m, int32_t *d) {
  const gregorian g = gregorian_of_date(t);
  *y = g.year;
  *m = g.month;
  *d = g.day;
}

/* date.fut:126-135 */
int32_t ic_add_months(int32_t date, int32_t nbmonths) {
  gregorian g = gregorian_of_date(date);
  int32_t m = g.month + nbmonths;
  int32_t y = g.year + fdiv(m - 1, 12);
  m = fmod_(m - 1, 12) + 1;
  if (m <= 0) {
    y -= 1;
    m += 12;
  }
  const int32_t eom = end_of_month(y, m);
  gregorian r = {y, m, g.day < eom || ICMUT(6) ? g.day : eom, g.hour, g.minute};
  return date_of_gregorian(r);
}

/* date.fut:137-141: the time from t1 to t2 in 365-day years (f64) */
double ic_diff_dates(int32_t t1, int32_t t2) {
  const double inv_fminutes_in_day = 1.0 / (double)MINUTES_IN_DAY;
  return (double)(t2 - t1) * inv_fminutes_in_day / 365.0;
}

/* ---- InterestCalib.fut:17-48: constants and utilities --------------------- */

#define EPS0 1.0e-3f
#define EPS 1.0e-5f
#define LLHOOD_CAUCHY_OFFS 5.0f
#define R_RATE 0.03f
#define PI_F 3.14159265358979323846f

static int32_t i32_f32(float x) { return (int32_t)x; } /* truncation */

/* date_act_365(t1, t2) = f32.f64 (diff_dates t2 t1) (:31) */
static float date_act_365(int32_t t1, int32_t t2) { return (float)ic_diff_dates(t2, t1); }

static int32_t add_years(int32_t date, float nbyears) { /* :33-34 */
  return ic_add_months(date, i32_f32(nbyears * 12.0f));
}

static int32_t today_date(void) { return ic_date_of_triple(2012, 1, 1); }
static int32_t max_date(void) { return ic_date_of_triple(2229, 12, 31); }
static int32_t min_date(void) { return ic_date_of_triple(1980, 1, 1); }

static float zc(int32_t t) { /* :48 */
  return expf(-R_RATE * date_act_365(t, today_date()));
}

/* ---- math module :56-105 -------------------------------------------------- */

static float erf_poly(float x) {
  const float p = 0.3275911f;
  const float a1 = 0.254829592f, a2 = -0.284496736f, a3 = ICMUT(2) ? -1.421413741f : 1.421413741f, a4 = -1.453152027f,
              a5 = 1.061405429f;
  const float t = 1.0f / (1.0f + p * x);
  const float t2 = t * t;
  const float t3 = t * t2;
  const float t4 = t2 * t2;
  const float t5 = t2 * t3;
  return a1 * t + a2 * t2 + a3 * t3 + a4 * t4 + a5 * t5;
}

static float erf_(float x) { /* :60-69 */
  return 1.0f - erf_poly(x) * expf(-(x * x));
}

float ic_ugaussian_P(float x) { /* :89-93 */
  const float u = x / sqrtf(2.0f);
  const float e = u < 0.0f ? -erf_(-u) : erf_(u);
  return 0.5f * (1.0f + e);
}

static float ugaussian_P_withExpFactor(float x, float exp_factor) { /* :95-101 */
  const float u = fabsf(x / sqrtf(2.0f));
  const float e = erf_poly(u);
  const float res = 0.5f * e * expf(exp_factor - u * u);
  return 0.0f <= x ? expf(exp_factor) - res : res;
}

float ic_ugaussian_P_exp(float x, float exp_factor) { return ugaussian_P_withExpFactor(x, exp_factor); }

/* date.fut:102-105 */
int ic_check_date(int32_t y, int32_t m, int32_t d) {
  return 1 <= d && 1 <= m && m <= 12 && d <= end_of_month(y, m);
}

/* :108-117, fid 1: sum(scales * exp(-bbi*yhat)) - 1 */
static float to_solve(int n, const float *scales, const float *bbi, float yhat) {
  float s = 0.0f;
  for (int i = 0; i < n; i++) s += scales[i] * expf(-bbi[i] * yhat);
  return s - 1.0f;
}

/* :124-185 rootFinding_Brent for fid 1; returns b, *err = fb (or +-inf) */
static float brent(int n, const float *scales, const float *bbi, float lb, float ub, float tol,
                   int32_t iter_max, float *err) {
  tol = tol <= 0.0f ? 1.0e-9f : tol;
  iter_max = iter_max <= 0 ? 10000 : iter_max;
  float a = lb, b = ub;
  float fa = to_solve(n, scales, bbi, a), fb = to_solve(n, scales, bbi, b);
  if (0.0f <= fa * fb) {
    *err = 0.0f <= a ? INFINITY : -INFINITY;
    return 0.0f;
  }
  if (fabsf(fa) < fabsf(fb)) {
    float t = fa; fa = fb; fb = t;
    t = a; a = b; b = t;
  }
  float c = a, fc = fa, d = 0.0f;
  int mflag = 1;
  for (int32_t i = 0; i < iter_max; i++) {
    if (fb == 0.0f || fabsf(b - a) < tol) continue; /* the loop keeps its state */
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
      mflag = 1;
      s = (a + b) / 2.0f;
    } else {
      mflag = 0;
    }
    const float fs = to_solve(n, scales, bbi, s);
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

float ic_brent(int n, const float *scales, const float *bbi, float lb, float ub, float tol,
               int32_t iter_max, float *err) {
  return brent(n, scales, bbi, lb, ub, tol, iter_max, err);
}

/* :190-202 */
float ic_sobol_ind(int m, const int32_t *dir, int32_t n) {
  const float norm_fact = 1.0f / ((float)((int64_t)1 << m) + 1.0f);
  const int32_t n_gray = ICMUT(1) ? n : (n >> 1) ^ n;
  int32_t res = 0;
  for (int i = 0; i < m; i++) {
    const int32_t t = (int32_t)(1u << i);
    if ((n_gray & t) == t) res ^= dir[i];
  }
  return (float)res * norm_fact;
}

/* ---- genomes :220-310 ------------------------------------------------------ */

static const float gb_min[5] = {EPS0, EPS0, EPS0 - 1.0f, EPS0, EPS0};
static const float gb_max[5] = {1.0f - EPS0, 1.0f - EPS0, 1.0f - EPS0, 0.2f, 0.2f};

#define AMPL_RATIO 0.005f /* moves_unif_ampl_ratio, :248 */

static void mutate_helper(float ampl_ratio, float r01, float gene, float prop, int j,
                          float *out_gene, float *out_bf) { /* :250-266 */
  const float g_min = gb_min[j], g_max = gb_max[j];
  const float amplitude = fabsf((g_max - g_min) * ampl_ratio);
  const float semiamplitude = amplitude / 2.0f;
  float tmp_min_max = fminf(g_max, gene + semiamplitude);
  float tmp_max_min = fmaxf(g_min, gene - semiamplitude);
  const float forward_range = tmp_min_max - tmp_max_min;
  tmp_min_max = fminf(g_max, prop + semiamplitude);
  tmp_max_min = fmaxf(g_min, prop - semiamplitude);
  const float backward_range = tmp_min_max - tmp_max_min;
  *out_bf = 0.0f < semiamplitude ? backward_range / forward_range : 1.0f;
  *out_gene = gene + amplitude * r01 - semiamplitude;
}

static float constrain_dim(float gene, int j) { /* :268-270 */
  return fmaxf(gb_min[j], fminf(gb_max[j], gene));
}

/* mutate_dims_all (dim < 0) and mutate_dims_one (dim = dim_j), :278-306 */
static float mutate_dims(int dim, const float *sob_row, const float *orig, const float *muta,
                         float *out) {
  float fb = 1.0f;
  for (int j = 0; j < 5; j++) {
    const float ampl = dim < 0 || j == dim ? AMPL_RATIO : 0.0f;
    float g, r;
    mutate_helper(ampl, sob_row[j], orig[j], muta[j], j, &g, &r);
    out[j] = constrain_dim(g, j);
    fb = j == 0 ? r : fb * r; /* f32.product, left to right */
  }
  return fb;
}

static void mcmc_de(float r01, const float *sob_row, const float *gi, const float *gk,
                    const float *gl, float *out) { /* :309-320 */
  const float gamma_avg = 2.38f / sqrtf(2.0f * 5.0f);
  const float ampl_ratio = 0.1f * AMPL_RATIO;
  const float gamma1 = gamma_avg - 0.5f + r01;
  for (int j = 0; j < 5; j++) {
    const float mm_diff = gb_max[j] - gb_min[j];
    const float amplitude = fabsf(mm_diff * ampl_ratio);
    const float semiamplitude = amplitude / 2.0f;
    const float perturb = amplitude * sob_row[j] - semiamplitude;
    const float g = gi[j] + perturb + gamma1 * (gk[j] - gl[j]);
    out[j] = constrain_dim(g, j);
  }
}

/* ---- G2++ :322-392 ----------------------------------------------------- */

static float b_fun(float z, float tau) { return (1.0f - expf(-z * tau)) / z; }

static float t_fun(float sigma, float x, float tau) {
  const float expxtau = expf(-x * tau);
  const float exp2xtau = expxtau * expxtau;
  return sigma * sigma / (x * x) *
         (tau + 2.0f / x * expxtau - 1.0f / (2.0f * x) * exp2xtau - 3.0f / (2.0f * x));
}

typedef struct {
  float a, b, rho, nu, sigma;
} genome_t;

static void bigv(genome_t g, float tau, float *v, float *ba, float *bb) { /* :336-352 */
  const float sigma = g.sigma == 0.0f ? 1.0e-10f : g.sigma;
  *ba = b_fun(g.a, tau);
  *bb = b_fun(g.b, tau);
  const float t1 = t_fun(sigma, g.a, tau);
  const float t2 = t_fun(g.nu, g.b, tau);
  const float t3 = (ICMUT(3) ? -2.0f : 2.0f) * g.rho * g.nu * sigma / (g.a * g.b) *
                   (tau - *ba - *bb + b_fun(g.a + g.b, tau));
  *v = t1 + t2 + t3;
}

static float bigmx(genome_t g, int32_t today, int32_t tmat, int32_t s, int32_t t) { /* :364-383 */
  const float a = g.a, b = g.b, rho = g.rho, nu = g.nu, sigma = g.sigma;
  const float ts = date_act_365(t, s), tmatt = date_act_365(tmat, t);
  const float tmat0 = date_act_365(tmat, today), tmats = date_act_365(tmat, s);
  const float t0 = date_act_365(t, today), s0 = date_act_365(s, today);
  const float tmp1 = (sigma * sigma) / (a * a) + (sigma * rho * nu) / (a * b);
  const float tmp2 = 1.0f - expf(-a * ts);
  const float tmp3 = sigma * sigma / (2.0f * a * a);
  const float tmp4 = rho * sigma * nu / (b * (a + b));
  const float tmp5 = expf(-a * tmatt) - expf(-a * (tmats + ts));
  const float tmp6 = expf(-b * tmatt) - expf(-b * tmat0 - a * t0 + (a + b) * s0);
  return tmp1 * tmp2 - (tmp3 * tmp5) - (tmp4 * tmp6);
}

static float bigmy(genome_t g, int32_t today, int32_t tmat, int32_t s, int32_t t) { /* :395-413 */
  const float a = g.a, b = g.b, rho = g.rho, nu = g.nu, sigma = g.sigma;
  const float ts = date_act_365(t, s), tmatt = date_act_365(tmat, t);
  const float tmat0 = date_act_365(tmat, today), tmats = date_act_365(tmat, s);
  const float t0 = date_act_365(t, today), s0 = date_act_365(s, today);
  const float tmp1 = nu * nu / (b * b) + sigma * rho * nu / (a * b);
  const float tmp2 = 1.0f - expf(-b * ts);
  const float tmp3 = nu * nu / (2.0f * b * b);
  const float tmp4 = sigma * rho * nu / (a * (a + b));
  const float tmp5 = expf(-b * tmatt) - expf(-b * (tmats + ts));
  const float tmp6 = expf(-a * tmatt) - expf(-a * tmat0 - b * t0 + (a + b) * s0);
  return tmp1 * tmp2 - (tmp3 * tmp5) - (tmp4 * tmp6);
}

#define IC_MAX_SCHED 512

typedef struct {
  float b, sigmax, sigmay, rhoxy, rhoxyc, rhoxycs, mux, muy;
} scals_t;

/* :418-481 exactYhat */
static float exact_yhat(int n, scals_t sc, const float *bai, const float *bbi, const float *aici,
                        const float *log_aici, float x, float eps) {
  const float k = -3.71901648545568f;
  float ups[IC_MAX_SCHED];
  float up = 0.0f, lo = -INFINITY;
  for (int i = 0; i < n; i++) {
    const float baix = bai[i] * x;
    ups[i] = aici[i] * expf(-baix);
    const float l = (log_aici[i] - baix) / bbi[i];
    up = up + ups[i];
    lo = fmaxf(lo, l);
  }
  if (n < 2) return lo;
  const float log_s = logf(up);
  float tmp = log_s / bbi[n - 1];
  if (!(tmp <= 0.0f)) {
    tmp = log_s / bbi[0];
    tmp = 0.0f <= tmp ? tmp : -INFINITY;
  }
  up = tmp;
  const float yl = lo - eps, yu = up + eps;
  const float y0 = sc.sigmay * (sc.rhoxy * (x - sc.mux) / sc.sigmax + k * sc.rhoxycs) -
                   sc.rhoxyc / sc.b + sc.muy;
  const float y1 = sc.sigmay * (sc.rhoxy * (x - sc.mux) / sc.sigmax - k * sc.rhoxycs) + sc.muy;
  if (y1 <= yl) return y1 + 1.0f;
  if (yu <= y0) return y0 - 1.0f;
  const float root_lb = fmaxf(yl, y0), root_ub = fminf(yu, y1);
  float err;
  const float root = brent(n, ups, bbi, root_lb, root_ub, 1.0e-4f, 1000, &err);
  if (err == -INFINITY) return y0 - 1.0f;
  if (err == INFINITY) return y1 + 1.0f;
  return root;
}

/* :486-602 evalGenomeOnSwap: the quote and the model price of one swaption */
void ic_eval_genome_on_swap(const float *genomea, int nherm, const float *herm_x,
                            const float *herm_w, const float *swaption, float *quote,
                            float *price) {
  const genome_t g = {genomea[0], genomea[1], genomea[2], genomea[3], genomea[4]};
  const float a = g.a, b = g.b, rho = g.rho, nu = g.nu, sigma = g.sigma;
  const int32_t today = today_date();
  const float swap_freq = swaption[1];
  const int32_t maturity = add_years(today, swaption[0]);
  const int64_t ns = (int64_t)(12.0f * swaption[2] / swap_freq);
  if (ns < 0 || ns > IC_MAX_SCHED) { /* covered: 0..512 periods (the datasets: 2..60) */
    *quote = NAN;
    *price = NAN;
    return;
  }
  const int n = (int)ns;
  const float tmat0 = date_act_365(maturity, today);

  /* quote (black) price :509-525 */
  float lvl = 0.0f;
  int32_t t0 = max_date(), tn = min_date();
  for (int i = 0; i < n; i++) {
    const int32_t a1 = ic_add_months(maturity, i32_f32(swap_freq * (float)(int64_t)i));
    const int32_t a2 = ic_add_months(a1, i32_f32(swap_freq));
    const float a12 = zc(a2) * date_act_365(a2, a1);
    lvl = lvl + a12;
    t0 = t0 < a1 ? t0 : a1;
    tn = tn > a2 ? tn : a2;
  }
  const float strike = (zc(t0) - zc(tn)) / lvl;
  const float d1 = 0.5f * swaption[3] * tmat0;
  *quote = lvl * strike * (ic_ugaussian_P(d1) - ic_ugaussian_P(-d1));

  /* model price :528-602 */
  float v0_mat, dum1, dum2;
  bigv(g, tmat0, &v0_mat, &dum1, &dum2);
  const float mux = 0.0f - bigmx(g, today, maturity, today, maturity);
  const float muy = 0.0f - bigmy(g, today, maturity, today, maturity);
  const float zc_mat = zc(maturity);
  const float sqrt_bfun_a = sqrtf(b_fun(2.0f * a, tmat0));
  const float sqrt_bfun_b = sqrtf(b_fun(2.0f * b, tmat0));
  const float rhoxy = rho * (ICMUT(4) ? 1.0f : b_fun(a + b, tmat0)) / (sqrt_bfun_a * sqrt_bfun_b);
  const float sigmax = sigma * sqrt_bfun_a;
  const float sigmay = nu * sqrt_bfun_b;
  const float rhoxyc = 1.0f - rhoxy * rhoxy;
  const float rhoxycs = sqrtf(rhoxyc);
  const float sigmay_rhoxycs = sigmay * rhoxycs;
  const float t4 = (rhoxy * sigmay) / sigmax;

  float bas[IC_MAX_SCHED], bbs[IC_MAX_SCHED], aicis[IC_MAX_SCHED], log_aicis[IC_MAX_SCHED];
  float scales[IC_MAX_SCHED], cs[IC_MAX_SCHED], t1_cs[IC_MAX_SCHED];
  for (int i = 0; i < n; i++) {
    const int32_t beg = ic_add_months(maturity, i32_f32(swap_freq * (float)(int64_t)i));
    const int32_t end = ic_add_months(beg, i32_f32(swap_freq));
    const float res = date_act_365(end, beg) * strike;
    const float cii = i == n - 1 && !ICMUT(5) ? 1.0f + res : res;
    const float date_tod1 = date_act_365(end, today);
    float v0_end, vt_end, baii, bbii;
    bigv(g, date_tod1, &v0_end, &dum1, &dum2);
    const float date_tod2 = date_act_365(end, maturity);
    bigv(g, date_tod2, &vt_end, &baii, &bbii);
    const float expo_aici = 0.5f * (vt_end - v0_end + v0_mat);
    const float fact_aici = cii * zc(end) / zc_mat;
    bas[i] = baii;
    bbs[i] = bbii;
    aicis[i] = fact_aici * expf(expo_aici);
    log_aicis[i] = logf(fact_aici) + expo_aici;
    scales[i] = 0.0f - (baii + bbii * t4);
    cs[i] = fact_aici;
    t1_cs[i] = bbii * (mux * t4 - (muy - 0.5f * rhoxyc * sigmay * sigmay * bbii)) + expo_aici;
  }
  const scals_t sc = {b, sigmax, sigmay, rhoxy, rhoxyc, rhoxycs, mux, muy};
  const float eps = 0.5f * sigmax;
  const float f = exact_yhat(n, sc, bas, bbs, aicis, log_aicis, mux, EPS);
  const float gg = exact_yhat(n, sc, bas, bbs, aicis, log_aicis, mux + eps, EPS);
  const float h = exact_yhat(n, sc, bas, bbs, aicis, log_aicis, mux - eps, EPS);
  const float df = 0.5f * (gg - h) / eps;
  const float sqrt2sigmax = sqrtf(2.0f) * sigmax;
  const float t2 = rhoxy / (sigmax * rhoxycs);
  float accum = 0.0f;
  for (int q = 0; q < nherm; q++) {
    const float x_quad = herm_x[q], w_quad = herm_w[q];
    const float x = sqrt2sigmax * x_quad + mux;
    const float yhat_x = f + df * (x - mux);
    const float h1 = ((yhat_x - muy) / sigmay_rhoxycs) - t2 * (x - mux);
    float accum1 = 0.0f;
    for (int i = 0; i < n; i++) {
      const float h2 = h1 + bbs[i] * sigmay_rhoxycs;
      const float expo_aici = t1_cs[i] + scales[i] * x;
      const float expo_part = ugaussian_P_withExpFactor(-h2, expo_aici);
      accum1 = accum1 + cs[i] * expo_part;
    }
    const float tmp = sqrtf(2.0f) * x_quad;
    const float t1 = expf(-0.5f * tmp * tmp);
    accum = accum + w_quad * t1 * (ic_ugaussian_P(-h1) - accum1);
  }
  *price = zc_mat * (accum / sqrtf(PI_F));
}

/* :606-625 log-likelihood (Cauchy) */
float ic_log_likelihood(float y_ref, float y) {
  const float gamma = (fabsf(y_ref) / 50.0f) * LLHOOD_CAUCHY_OFFS + 0.01f;
  const float x = (y - y_ref) / gamma;
  const float pdf = 1.0f / (PI_F * gamma * (1.0f + x * x));
  return logf(pdf + 1.0e-20f);
}

/* sum over the swaptions of the log-likelihoods of one genome (:639-645) */
static float genome_loglik(const float *genome, int nswap, const float *swaptions, int nherm,
                           const float *hx, const float *hw) {
  float s = 0.0f;
  for (int w = 0; w < nswap; w++) {
    float q, p;
    ic_eval_genome_on_swap(genome, nherm, hx, hw, swaptions + 4 * w, &q, &p);
    s = s + ic_log_likelihood(q, p);
  }
  return s;
}

/* Log-likelihoods of n genomes (OpenMP over genomes). */
void ic_logliks(int n, const float *genomes, int nswap, const float *swaptions, int nherm,
                const float *hx, const float *hw, float *out, int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
  for (int i = 0; i < n; i++)
    out[i] = genome_loglik(genomes + 5 * i, nswap, swaptions, nherm, hx, hw);
}

/* One iteration's proposals (:655-724), from the Sobol offset at the top of
 * the iteration: the move type, then new genomes in props (which holds the
 * previous proposals on entry, as mutate_dims_* reads them) and the
 * forward/backward ratios in fb.  Returns the Sobol offset of the acceptance
 * draws (:737). */
int32_t ic_propose(int32_t pop, int32_t nbits, const int32_t *dir, int32_t sob_offs,
                   const float *genomes, float *props, float *fb) {
  const float rand01 = ic_sobol_ind(nbits, dir, sob_offs);
  const int move_type = rand01 <= 0.2f ? 1 : rand01 <= 0.5f ? 2 : 3; /* :240-245 */
  sob_offs = sob_offs + 1;
  if (move_type == 1 || move_type == 2) {
    int dim = -1, base = sob_offs;
    if (move_type == 2) {
      const float s1 = ic_sobol_ind(nbits, dir, sob_offs);
      dim = i32_f32(s1 * (float)5);
      base = sob_offs + 1;
    }
    for (int32_t i = 0; i < pop; i++) {
      float row[5], g[5];
      for (int j = 0; j < 5; j++) row[j] = ic_sobol_ind(nbits, dir, 5 * i + base + j);
      fb[i] = mutate_dims(dim, row, genomes + 5 * i, props + 5 * i, g);
      memcpy(props + 5 * i, g, sizeof(g));
    }
    return move_type == 1 ? sob_offs + 5 * pop : sob_offs + 5 * pop + 1;
  }
  for (int32_t i = 0; i < pop; i++) { /* :700-723 */
    const int32_t kk = 8 * i + sob_offs;
    const float s1 = ic_sobol_ind(nbits, dir, kk);
    int32_t k = i32_f32(s1 * (float)(pop - 1));
    int32_t cand_ub = pop - 1;
    if (k == i) {
      k = pop - 1;
      cand_ub = pop - 2;
    }
    const float s2 = ic_sobol_ind(nbits, dir, kk + 1);
    int32_t l = i32_f32(s2 * (float)cand_ub);
    if (l == i || l == k) l = cand_ub;
    const float s3 = ic_sobol_ind(nbits, dir, kk + 2);
    float row[5];
    for (int j = 0; j < 5; j++) row[j] = ic_sobol_ind(nbits, dir, kk + 3 + j);
    mcmc_de(s3, row, genomes + 5 * i, genomes + 5 * k, genomes + 5 * l, props + 5 * i);
    fb[i] = 1.0f;
  }
  return sob_offs + 8 * pop;
}

/* The accept/reject step (:733-745): genome i takes its proposal when the
 * Sobol number at acc_offs + i is below min(1, exp(new - old) * fb). */
void ic_accept(int32_t pop, int32_t nbits, const int32_t *dir, int32_t acc_offs, float *genomes,
               float *logliks, const float *props, const float *newl, const float *fb) {
  for (int32_t i = 0; i < pop; i++) {
    const float acceptance = fminf(1.0f, expf(newl[i] - logliks[i]) * fb[i]);
    const float r = ic_sobol_ind(nbits, dir, acc_offs + i);
    if (r < acceptance) {
      memcpy(genomes + 5 * i, props + 5 * i, sizeof(float) * 5);
      logliks[i] = newl[i];
    }
  }
}

/* Whether a DE move (:700-724) comes up in the first mcmc_conv iterations:
 * with one genome it indexes genomes[-1], a bounds failure. */
int ic_schedule_has_de(int32_t pop, int32_t mcmc_conv, int32_t nbits, const int32_t *dir) {
  int32_t sob_offs = 5 * pop + 1;
  for (int32_t it = 0; it < mcmc_conv; it++) {
    const float rand01 = ic_sobol_ind(nbits, dir, sob_offs);
    const int move_type = rand01 <= 0.2f ? 1 : rand01 <= 0.5f ? 2 : 3;
    if (move_type == 3) return 1;
    sob_offs = sob_offs + 1 + (move_type == 1 ? 5 * pop : 5 * pop + 1) + pop;
  }
  return 0;
}

/* :627-849 interestCalibKernel + :929-936 main.  out6: a, b, sigma, nu, rho,
 * log-likelihood of the winner; summary [nswap][3].  The genomes and
 * log-likelihoods of the population after the loop go to final_genomes
 * [pop][5] and final_logliks [pop] when those are non-NULL.  Returns 1 where
 * the reference fails (no genome; one genome and a DE move; a direction
 * vector of more than 31 bits, where its i32 shifts wrap). */
int ic_main(int32_t pop, int32_t mcmc_conv, int32_t nswap, const float *swaptions, int32_t nherm,
            const float *hx, const float *hw, int32_t nbits, const int32_t *dir, float *out6,
            float *summary, float *final_genomes, float *final_logliks, int nthreads) {
  if (pop < 1 || nswap < 0 || nbits < 0 || nbits > 31) return 1;
  if (pop == 1 && ic_schedule_has_de(pop, mcmc_conv, nbits, dir)) return 1;
  float *genomes = malloc(sizeof(float) * 5 * (size_t)pop);
  float *props = malloc(sizeof(float) * 5 * (size_t)pop);
  float *logliks = malloc(sizeof(float) * (size_t)pop);
  float *newl = malloc(sizeof(float) * (size_t)pop);
  float *fb = malloc(sizeof(float) * (size_t)pop);
  if (!genomes || !props || !logliks || !newl || !fb) {
    free(genomes); free(props); free(logliks); free(newl); free(fb);
    return 1;
  }
  for (int32_t i = 0; i < pop; i++) { /* :633-638 */
    const int32_t k = 5 * i + 1;
    for (int j = 0; j < 5; j++) {
      const float r01 = ic_sobol_ind(nbits, dir, k + j);
      genomes[5 * i + j] = r01 * (gb_max[j] - gb_min[j]) + gb_min[j];
    }
  }
  ic_logliks(pop, genomes, nswap, swaptions, nherm, hx, hw, logliks, nthreads);
  memcpy(props, genomes, sizeof(float) * 5 * (size_t)pop);
  int32_t sob_offs = 5 * pop + 1;
  for (int32_t it = 0; it < mcmc_conv; it++) {
    const int32_t acc = ic_propose(pop, nbits, dir, sob_offs, genomes, props, fb);
    ic_logliks(pop, props, nswap, swaptions, nherm, hx, hw, newl, nthreads);
    ic_accept(pop, nbits, dir, acc, genomes, logliks, props, newl, fb);
    sob_offs = acc + pop;
  }
  int32_t wi = 0; /* :752-756 */
  float wv = -INFINITY;
  for (int32_t i = 0; i < pop; i++)
    if (wv < logliks[i]) {
      wi = i;
      wv = logliks[i];
    }
  const float *w = genomes + 5 * wi;
  out6[0] = w[0];
  out6[1] = w[1];
  out6[2] = w[4];
  out6[3] = w[3];
  out6[4] = w[2];
  out6[5] = wv;
  for (int32_t s = 0; s < nswap; s++) { /* :205-213 makeSummary */
    float q, p;
    ic_eval_genome_on_swap(w, nherm, hx, hw, swaptions + 4 * s, &q, &p);
    const float black_price = q, calib_price = p;
    const float err_ratio = (calib_price - black_price) / black_price;
    summary[3 * s] = 10000.0f * calib_price;
    summary[3 * s + 1] = 10000.0f * black_price;
    summary[3 * s + 2] = 100.0f * fabsf(err_ratio);
  }
  if (final_genomes) memcpy(final_genomes, genomes, sizeof(float) * 5 * (size_t)pop);
  if (final_logliks) memcpy(final_logliks, logliks, sizeof(float) * (size_t)pop);
  free(genomes); free(props); free(logliks); free(newl); free(fb);
  return 0;
}
