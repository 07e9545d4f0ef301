/*
 * Body of the CPU oracle, instantiated twice by lvc_oracle.c (R = float with
 * suffix f32, R = double with suffix f64).  TEST INFRASTRUCTURE ONLY.
 *
 * Every function restates one function of finpar/LocVolCalib.fut with the
 * same arithmetic in the same evaluation order (C evaluates `a * b * c` as
 * `(a * b) * c`, as Futhark does).  Literals are converted to R so that the
 * f32 instantiation performs every operation in single precision, like the
 * reference (whose real type is f32 throughout).
 */
#ifndef R
#error "define R and SUF before including this file"
#endif

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUF)

/* LocVolCalib.fut:12-23 (initGrid).  x[myXindex] == s0 exactly when the
 * compiler does not contract `i*dx - xi*dx` (the oracle is built with
 * -ffp-contract=off).  *x0log receives the value whose log updateParams takes
 * at column 0 (x[0] itself, but see GRID32 below).
 *
 * GRID32 mode (lvc_oracle_set_grid32): the scalars stdX, dx, myXindex and
 * x[0] of :15-21 are evaluated in f32 from the f32 values of the parameters,
 * as the reference's f32 program does, and the arrays in R from those
 * scalars.  With R = double this is the GPU library's form of the reference's
 * f32 entry (f32 grid index, FP64 grid); with R = float it changes nothing. */
static void FN(init_grid_)(R s0, R alpha, R nu, R t, int64_t numX, int64_t numY,
                           int64_t numT, int32_t *xindex, int32_t *yindex, R *myX,
                           R *myY, R *timeline, R *x0log) {
  R logAlpha = LOG(alpha);
  for (int64_t i = 0; i < numT; i++)
    timeline[i] = t * (R)i / ((R)numT - (R)1.0);
  R stdX = (R)20.0 * alpha * s0 * SQRT(t);
  R stdY = (R)10.0 * nu * SQRT(t);
  R dx = stdX / (R)numX, dy = stdY / (R)numY;
  int32_t xi = (int32_t)(s0 / dx); /* i32.f32 truncates toward zero */
  int32_t yi = (int32_t)numY / 2;
  R x0 = (R)0 * dx - (R)xi * dx + s0;
  if (GRID32) {
    const float s0f = (float)s0, af = (float)alpha, tf = (float)t;
    const float stdXf = 20.0f * af * s0f * sqrtf(tf);
    const float dxf = stdXf / (float)numX;
    xi = (int32_t)(s0f / dxf);
    dx = (R)dxf;
    x0 = (R)(0.0f * dxf - (float)xi * dxf + s0f);
  }
  for (int64_t i = 0; i < numX; i++)
    myX[i] = (R)i * dx - (R)xi * dx + s0;
  for (int64_t i = 0; i < numY; i++)
    myY[i] = (R)i * dy - (R)yi * dy + logAlpha;
  *xindex = xi;
  *yindex = yi;
  *x0log = x0;
}

void FN(lvc_init_grid_)(R s0, R alpha, R nu, R t, int64_t numX, int64_t numY,
                        int64_t numT, int32_t *xindex, int32_t *yindex, R *myX,
                        R *myY, R *timeline) {
  R x0;
  FN(init_grid_)(s0, alpha, nu, t, numX, numY, numT, xindex, yindex, myX, myY,
                 timeline, &x0);
}

/* LocVolCalib.fut:26-44 (initOperator): D and DD are [n][3] row-major. */
void FN(lvc_init_operator_)(const R *x, int64_t n, R *D, R *DD) {
  R dxu = x[1] - x[0];
  D[0] = (R)0.0;
  D[1] = (R)-1.0 / dxu;
  D[2] = (R)1.0 / dxu;
  DD[0] = DD[1] = DD[2] = (R)0.0;
  for (int64_t i = 1; i <= n - 2; i++) {
    R dl = x[i] - x[i - 1];
    R du = x[i + 1] - x[i];
    D[3 * i + 0] = -du / dl / (dl + du);
    D[3 * i + 1] = (du / dl - dl / du) / (dl + du);
    D[3 * i + 2] = dl / du / (dl + du);
    DD[3 * i + 0] = (R)2.0 / dl / (dl + du);
    DD[3 * i + 1] = (R)-2.0 * ((R)1.0 / dl + (R)1.0 / du) / (dl + du);
    if (MUT(1)) /* planted: drop the upper spacing from the centre tap */
      DD[3 * i + 1] = (R)-2.0 * ((R)1.0 / dl) / (dl + du);
    DD[3 * i + 2] = (R)2.0 / du / (dl + du);
  }
  R dxl = x[n - 1] - x[n - 2];
  D[3 * (n - 1) + 0] = (R)-1.0 / dxl;
  D[3 * (n - 1) + 1] = (R)1.0 / dxl;
  D[3 * (n - 1) + 2] = (R)0.0;
  DD[3 * (n - 1) + 0] = DD[3 * (n - 1) + 1] = DD[3 * (n - 1) + 2] = (R)0.0;
}

/* LocVolCalib.fut:46-47 (setPayoff): res is [numY][numX]. */
void FN(lvc_set_payoff_)(R strike, const R *myX, int64_t numX, int64_t numY,
                         R *res) {
  for (int64_t j = 0; j < numY; j++)
    for (int64_t i = 0; i < numX; i++) {
      R v = myX[i] - strike;
      res[j * numX + i] = v > (R)0.0 ? v : (R)0.0; /* f32.max v 0 */
    }
}

/* LocVolCalib.fut:50-66 (updateParams).  Only VarX is non-trivial:
 * MuX = MuY = 0 and VarY = nu*nu everywhere.  varX is [numY][numX]; the log
 * at column 0 takes x0log (init_grid_). */
static void FN(update_params_)(const R *myX, const R *myY, int64_t numX,
                               int64_t numY, R tnow, R beta, R nu, R x0log,
                               R *varX) {
  for (int64_t j = 0; j < numY; j++)
    for (int64_t i = 0; i < numX; i++)
      varX[j * numX + i] =
          EXP((R)2.0 * (beta * LOG(i == 0 ? x0log : myX[i]) + myY[j] -
                        (MUT(4) ? (R)-0.5 : (R)0.5) * nu * nu * tnow));
}

void FN(lvc_update_params_)(const R *myX, const R *myY, int64_t numX,
                            int64_t numY, R tnow, R beta, R nu, R *varX) {
  FN(update_params_)(myX, myY, numX, numY, tnow, beta, nu, myX[0], varX);
}

/* LocVolCalib.fut:68-124 (tridagPar): the three scans evaluated sequentially
 * left to right with the exact operators of the reference.  `out` may alias
 * `y`.  `tmp` holds 5*n scratch values. */
void FN(lvc_tridag_par_)(const R *a, const R *b, const R *c, const R *y,
                         int64_t n, R *out, R *tmp) {
  R *bb = tmp, *yy = tmp + n, *s = tmp + 2 * n;
  /* Recurrence 1 (:74-91): scan of normalised 2x2 matrices. */
  R b0 = b[0];
  R a1 = (R)0.0, a2 = (R)0.0, a3 = (R)1.0; /* scan accumulator (neutral) */
  for (int64_t i = 0; i < n; i++) {
    R m1, m2, m3;
    if (0 < i) {
      R bi = (R)1.0 / b[i];
      m1 = -a[i] * c[MUT(2) ? i : i - 1] * bi; /* planted: wrong index */
      m2 = bi;
      m3 = (R)0.0;
    } else {
      m1 = (R)0.0;
      m2 = (R)0.0;
      m3 = (R)1.0;
    }
    R value = (R)1.0 / ((R)1.0 + m1 * a2);
    R n1 = (a1 + m1 * a3) * value;
    R n2 = (m2 + m3 * a2) * value;
    R n3 = (m2 * a1 + m3 * a3) * value;
    a1 = n1;
    a2 = n2;
    a3 = n3;
    bb[i] = (b0 + a1) / (a2 * b0 + a3);
  }
  /* Recurrence 2 (:96-106): forward substitution as a linear-function scan. */
  R y0 = y[0];
  R f0 = (R)0.0, f1 = (R)1.0;
  for (int64_t i = 0; i < n; i++) {
    R l0, l1;
    if (0 < i) {
      l0 = y[i];
      l1 = (R)0.0 - a[i] / bb[i - 1];
    } else {
      l0 = (R)0.0;
      l1 = (R)1.0;
    }
    R g0 = l0 + l1 * f0;
    R g1 = f1 * l1;
    f0 = g0;
    f1 = g1;
    yy[i] = f0 + f1 * y0;
  }
  /* Recurrence 3 (:111-123): backward substitution on reversed indices. */
  R yn = yy[n - 1] / bb[n - 1];
  f0 = (R)0.0;
  f1 = (R)1.0;
  for (int64_t k = 0; k < n; k++) {
    int64_t i = n - k - 1;
    R l0, l1;
    if (0 < k) {
      l0 = yy[i] / bb[i];
      l1 = (R)0.0 - c[i] / bb[i];
    } else {
      l0 = (R)0.0;
      l1 = (R)1.0;
    }
    R g0 = l0 + l1 * f0;
    R g1 = f1 * l1;
    f0 = g0;
    f1 = g1;
    s[i] = f0 + f1 * yn; /* reverse (:123) folded into the store index */
  }
  for (int64_t i = 0; i < n; i++)
    out[i] = s[i];
}

/* misc/ocean-sim/tridiag.fut:81-95 (tridagSeq): Thomas algorithm.  `out`
 * may alias `y`.  `tmp` holds 2*n scratch values. */
void FN(lvc_tridag_seq_)(const R *a, const R *b, const R *c, const R *y,
                         int64_t n, R *out, R *tmp) {
  R *cp = tmp, *yp = tmp + n;
  cp[0] = c[0] / b[0];
  yp[0] = y[0] / b[0];
  for (int64_t i = 1; i < n; i++) {
    R norm = (R)1.0 / (b[i] - a[i] * cp[i - 1]);
    cp[i] = c[i] * norm;
    yp[i] = (y[i] - a[i] * yp[i - 1]) * norm;
  }
  out[n - 1] = yp[n - 1];
  for (int64_t i = n - 2; i >= 0; i--)
    out[i] = yp[i] - cp[i] * out[i + 1];
}

/* misc/ocean-sim/tridiag.fut:4-79 (the ocean-sim program's tridagPar): the
 * same three recurrences as LocVolCalib.fut:68-124, but recurrence 1 is the
 * scan of un-normalised 2x2 matrices (b_i, -a_i*c_{i-1}, 1, 0) under the
 * product scaled by 1/(a0*b0) (:8-31), a different operator form.  Evaluated
 * left to right.  `out` may alias `y`.  `tmp` holds 5*n scratch values. */
void FN(lvc_tridag_par4_)(const R *a, const R *b, const R *c, const R *y,
                          int64_t n, R *out, R *tmp) {
  R *bb = tmp, *yy = tmp + n, *s = tmp + 2 * n;
  const R b0 = b[0];
  R t0 = (R)1.0, t1 = (R)0.0, t2 = (R)0.0, t3 = (R)1.0; /* identity (:11) */
  for (int64_t i = 0; i < n; i++) {
    if (0 < i) {
      const R m0 = b[i], m1 = (R)0.0 - a[i] * c[i - 1], m2 = (R)1.0, m3 = (R)0.0;
      const R value = (R)1.0 / (t0 * m0);
      const R n0 = (m0 * t0 + m1 * t2) * value;
      const R n1 = (m0 * t1 + m1 * t3) * value;
      const R n2 = (m2 * t0 + m3 * t2) * value;
      const R n3 = (m2 * t1 + m3 * t3) * value;
      t0 = n0;
      t1 = n1;
      t2 = n2;
      t3 = n3;
    }
    bb[i] = (t0 * b0 + t1) / (t2 * b0 + t3);
  }
  /* Recurrences 2 and 3 (:36-78): the linear-function scans of
   * LocVolCalib.fut:96-123, same operators. */
  const R y0 = y[0];
  R f0 = (R)0.0, f1 = (R)1.0;
  for (int64_t i = 0; i < n; i++) {
    if (0 < i) {
      const R l0 = y[i], l1 = (R)0.0 - a[i] / bb[i - 1];
      const R g0 = l0 + l1 * f0, g1 = f1 * l1;
      f0 = g0;
      f1 = g1;
    }
    yy[i] = f0 + f1 * y0;
  }
  const R yn = yy[n - 1] / bb[n - 1];
  f0 = (R)0.0;
  f1 = (R)1.0;
  for (int64_t k = 0; k < n; k++) {
    const int64_t i = n - k - 1;
    if (0 < k) {
      const R l0 = yy[i] / bb[i], l1 = (R)0.0 - c[i] / bb[i];
      const R g0 = l0 + l1 * f0, g1 = f1 * l1;
      f0 = g0;
      f1 = g1;
    }
    s[i] = f0 + f1 * yn;
  }
  for (int64_t i = 0; i < n; i++)
    out[i] = s[i];
}

/* LocVolCalib.fut:126-153 (explicitMethod) on `rows` rows of length m:
 * D, DD are [m][3]; mu, var, res, out are [rows][m]. */
void FN(lvc_explicit_)(const R *D, const R *DD, const R *mu, const R *var,
                       const R *res, int64_t rows, int64_t m, R *out) {
  for (int64_t r = 0; r < rows; r++) {
    const R *mr = mu + r * m, *vr = var + r * m, *rr = res + r * m;
    for (int64_t j = 0; j < m; j++) {
      const R *dx = D + 3 * j, *dxx = DD + 3 * j;
      R c1 = (0 < j) ? (mr[j] * dx[0] + (R)0.5 * vr[j] * dxx[0]) * rr[j - 1]
                     : (R)0.0;
      R c3 = (j < m - 1)
                 ? (mr[j] * dx[2] + (R)0.5 * vr[j] * dxx[2]) * rr[j + 1]
                 : (R)0.0;
      R c2 = (mr[j] * dx[1] + (R)0.5 * vr[j] * dxx[1]) * rr[j];
      out[r * m + j] = c1 + c2 + c3;
    }
  }
}

/* LocVolCalib.fut:156-178 (implicitMethod): solves each of the `rows` rows
 * of u in place.  tmp holds 9*m scratch values. */
void FN(lvc_implicit_)(const R *D, const R *DD, const R *mu, const R *var,
                       R *u, int64_t rows, int64_t m, R dtInv, int thomas,
                       R *tmp) {
  R *a = tmp, *b = tmp + m, *c = tmp + 2 * m, *work = tmp + 3 * m;
  for (int64_t r = 0; r < rows; r++) {
    const R *mr = mu + r * m, *vr = var + r * m;
    for (int64_t j = 0; j < m; j++) {
      const R *d = D + 3 * j, *dd = DD + 3 * j;
      a[j] = (R)0.0 - (R)0.5 * (mr[j] * d[0] + (R)0.5 * vr[j] * dd[0]);
      b[j] = dtInv - (R)0.5 * (mr[j] * d[1] + (R)0.5 * vr[j] * dd[1]);
      c[j] = (R)0.0 - (R)0.5 * (mr[j] * d[2] + (R)0.5 * vr[j] * dd[2]);
    }
    if (thomas)
      FN(lvc_tridag_seq_)(a, b, c, u + r * m, m, u + r * m, work);
    else
      FN(lvc_tridag_par_)(a, b, c, u + r * m, m, u + r * m, work);
  }
}

static void FN(transpose_)(const R *in, int64_t rows, int64_t cols, R *out) {
  for (int64_t r = 0; r < rows; r++)
    for (int64_t c = 0; c < cols; c++)
      out[c * rows + r] = in[r * cols + c];
}

/* Per-option work buffers for value(). */
typedef struct {
  R *myX, *myY, *timeline, *Dx, *Dxx, *Dy, *Dyy;
  R *res, *muX, *varX, *muY, *varY, *u, *v, *resT, *y, *tmp;
} FN(lvc_ws_);

/* LocVolCalib.fut:180-209 (rollback): one ADI time step, res in place. */
static void FN(rollback_)(R tnow, R tnext, int64_t numX, int64_t numY,
                          FN(lvc_ws_) * w, int thomas) {
  int64_t N = numX * numY;
  R dtInv = (R)1.0 / (tnext - tnow);
  /* explicitX (:195-196) */
  FN(lvc_explicit_)(w->Dx, w->Dxx, w->muX, w->varX, w->res, numY, numX, w->u);
  for (int64_t k = 0; k < N; k++)
    w->u[k] = dtInv * w->res[k] + (R)0.5 * w->u[k];
  /* explicitY (:198-200) */
  FN(transpose_)(w->res, numY, numX, w->resT);
  /* planted mistake 5: the untransposed operand (same element count) */
  FN(lvc_explicit_)(w->Dy, w->Dyy, w->muY, w->varY, MUT(5) ? w->res : w->resT,
                    numX, numY, w->v);
  for (int64_t j = 0; j < numY; j++)
    for (int64_t i = 0; i < numX; i++)
      w->u[j * numX + i] = w->u[j * numX + i] + w->v[i * numY + j];
  /* implicitX (:202) */
  FN(lvc_implicit_)(w->Dx, w->Dxx, w->muX, w->varX, w->u, numY, numX, dtInv,
                    thomas, w->tmp);
  /* implicitY (:204-209) */
  for (int64_t i = 0; i < numX; i++)
    for (int64_t j = 0; j < numY; j++)
      w->y[i * numY + j] = dtInv * w->u[j * numX + i] -
                           (MUT(3) ? (R)1.0 : (R)0.5) * w->v[i * numY + j];
  FN(lvc_implicit_)(w->Dy, w->Dyy, w->muY, w->varY, w->y, numX, numY, dtInv,
                    thomas, w->tmp);
  FN(transpose_)(w->y, numX, numY, w->res);
}

/* LocVolCalib.fut:211-236 (value): one option.  Returns NaN (and sets *err
 * when non-NULL) on invalid sizes or allocation failure.  `thomas` selects
 * tridagSeq instead of the reference's scan-form tridagPar. */
R FN(lvc_value_)(int64_t numX, int64_t numY, int64_t numT, R s0, R strike,
                 R t, R alpha, R nu, R beta, int thomas, int *err) {
  if (numX < 3 || numY < 3 || numT < 2) {
    if (err)
      *err = 1;
    return (R)NAN;
  }
  int64_t N = numX * numY, mx = numX > numY ? numX : numY;
  FN(lvc_ws_) w;
  size_t total = (size_t)(7 * (numX + numY) + numT + 9 * N + 9 * mx);
  R *buf = (R *)calloc(total, sizeof(R));
  if (!buf) {
    if (err)
      *err = 2;
    return (R)NAN;
  }
  R *p = buf;
  w.myX = p; p += numX;
  w.myY = p; p += numY;
  w.timeline = p; p += numT;
  w.Dx = p; p += 3 * numX;
  w.Dxx = p; p += 3 * numX;
  w.Dy = p; p += 3 * numY;
  w.Dyy = p; p += 3 * numY;
  w.res = p; p += N;
  w.muX = p; p += N;
  w.varX = p; p += N;
  w.muY = p; p += N;
  w.varY = p; p += N;
  w.u = p; p += N;
  w.v = p; p += N;
  w.resT = p; p += N;
  w.y = p; p += N;
  w.tmp = p; p += 9 * mx;
  if ((size_t)(p - buf) != total)
    abort(); /* layout and allocation out of step */

  int32_t xi, yi;
  R x0log;
  FN(init_grid_)(s0, alpha, nu, t, numX, numY, numT, &xi, &yi, w.myX, w.myY,
                 w.timeline, &x0log);
  FN(lvc_init_operator_)(w.myX, numX, w.Dx, w.Dxx);
  FN(lvc_init_operator_)(w.myY, numY, w.Dy, w.Dyy);

  FN(lvc_set_payoff_)(strike, w.myX, numX, numY, w.res);
  /* updateParams (:58-60): MuX, MuY are zero (calloc), VarY = nu*nu. */
  for (int64_t k = 0; k < N; k++)
    w.varY[k] = nu * nu;
  /* Timeline neighbours, latest interval first (:217-219). */
  for (int64_t k = numT - 2; k >= 0; k--) {
    R tnow = w.timeline[k], tnext = w.timeline[k + 1];
    FN(update_params_)(w.myX, w.myY, numX, numY, tnow, beta, nu, x0log, w.varX);
    FN(rollback_)(tnow, tnext, numX, numY, &w, thomas);
  }
  R out;
  if (xi < 0 || xi >= numX || yi < 0 || yi >= numY) {
    if (err)
      *err = 3; /* the reference's bounds-checked index at :236 fails */
    out = (R)NAN;
  } else {
    out = w.res[(int64_t)yi * numX + xi];
  }
  free(buf);
  return out;
}

/* LocVolCalib.fut:238-252 (main): strikes[i] = 0.001 * i, options
 * independent (OpenMP over options when nthreads > 1).  Returns 0 or an
 * error code. */
int FN(lvc_main_)(int32_t outer, int32_t numX, int32_t numY, int32_t numT,
                  R s0, R t, R alpha, R nu, R beta, int thomas, int nthreads,
                  R *out) {
  if (outer < 0)
    return 4;
  if (numX < 3 || numY < 3 || numT < 2)
    return 1;
  int rc = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
  for (int32_t i = 0; i < outer; i++) {
    int e = 0;
    R strike = (R)0.001 * (R)(int64_t)i;
    out[i] = FN(lvc_value_)(numX, numY, numT, s0, strike, t, alpha, nu, beta,
                            thomas, &e);
    if (e) {
#pragma omp atomic write
      rc = e;
    }
  }
  return rc;
}

/* Per-problem form (the batch entry): each problem has its own sizes and
 * parameters; same arithmetic as value(). */
int FN(lvc_batch_)(int32_t n, const int32_t *numX, const int32_t *numY,
                   const int32_t *numT, const R *s0, const R *strike,
                   const R *t, const R *alpha, const R *nu, const R *beta,
                   int thomas, int nthreads, R *out) {
  int rc = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
  for (int32_t i = 0; i < n; i++) {
    int e = 0;
    out[i] = FN(lvc_value_)(numX[i], numY[i], numT[i], s0[i], strike[i], t[i],
                            alpha[i], nu[i], beta[i], thomas, &e);
    if (e) {
#pragma omp atomic write
      rc = e;
    }
  }
  return rc;
}

#undef FN
#undef CAT
#undef CAT_
