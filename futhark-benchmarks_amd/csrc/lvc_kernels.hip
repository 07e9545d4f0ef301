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
// is cut into segments of M consecutive unknowns, one segment per lane, and
// the `lanes` (a power of two >= ceil(n/M), at most 64) lanes of a system are
// consecutive lanes of one wavefront.  Each lane eliminates its segment
// locally (Thomas on the first M-1 rows, written so that every pivot is a sum
// of positive terms), which leaves one unknown per lane (its last) in a
// tridiagonal reduced system; that is solved by parallel cyclic reduction
// across the lanes, and each lane then back-substitutes.  When a system's
// lanes fit one 16-lane row the PCR exchanges are DPP row shifts (VALU
// operand modifiers, no LDS round trip); wider systems use ds_bpermute.
// Lanes past the end of the system hold identity rows.
//
// Rounding: the file is compiled without floating-point contraction and every
// fused multiply-add is written as fma(), so each operation's rounding is
// fixed by the source, the same in every kernel variant that inlines it
// (single-problem and shared x tiles, the server) whatever the scheduling.
//
// Precision: T = float or double for the state and the sweeps.  REF adds one
// iterative-refinement pass to each implicit solve: the residual of the FP32
// solution is formed in FP64 (with the same taps and the same rounded time
// step dtInv that the explicit part and the combination use, so that the
// scheme stays consistent), solved again in FP32 with the same
// factorisation, and added with one rounding (x + e in FP32 is the correctly
// rounded sum).

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <mutex>

#include "lvc_internal.h"

// Device scratch of the allocation-free entry (futhark_entry_main_noalloc):
// one copy per GPU, held until the process ends.  Host code never uses this
// name as an address (on the host it names a stand-in, not device memory):
// it takes the device address from lvc_scratch_address().
static __device__ __align__(256) unsigned char scratch[LVC_SCRATCH_BYTES];

namespace {

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------

// Timeline point k of initGrid (:14), t * k / (numT - 1): a multiply and a
// divide, bit-identical to a C evaluation of the same expression.
__device__ __forceinline__ double timeline(double t, int k, int nt) {
  return t * (double)k / ((double)nt - 1.0);
}

// Reciprocals: v_rcp_f32 (1 ulp); v_rcp_f64 refined by two Newton steps.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

template <int CTRL>
__device__ __forceinline__ float dpp(float old, float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(
      __float_as_int(old), __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double old, double v) {
  const long long o = __double_as_longlong(old), s = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp((int)o, (int)s, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(s >> 32), CTRL,
                                             0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Lanes of one system: `seg` = this lane's index in its group of `G` lanes.
// Groups of at most 16 lanes lie inside one DPP row and exchange values with
// DPP row shifts; wider groups (32, 64 lanes) use ds_bpermute.  The choice is
// a template parameter (ROW) of every routine that exchanges values, and a
// tile picks its instantiation once, so the exchanges sit in straight-line
// code rather than behind a branch each.
struct Lanes {
  int seg, G;
  int lane4;  // 4 * (lane index in the wavefront): ds_bpermute address base
};

__device__ __forceinline__ Lanes make_lanes(int seg, int G) {
  return Lanes{seg, G, (int)__lane_id() << 2};
}

// ds_bpermute: the value of lane (addr >> 2) mod 64.
__device__ __forceinline__ float bperm(int addr, float v) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}
__device__ __forceinline__ double bperm(int addr, double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)b);
  const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Value of lane seg-st of the group (ident below the group's first lane).
template <bool ROW, typename T>
__device__ __forceinline__ T from_below(const Lanes &L, T v, int st, T ident) {
  T r;
  if constexpr (ROW) {
    switch (st) {
      case 1: r = dpp<0x111>(ident, v); break;  // row_shr:1
      case 2: r = dpp<0x112>(ident, v); break;
      case 4: r = dpp<0x114>(ident, v); break;
      default: r = dpp<0x118>(ident, v); break;
    }
  } else {
    r = bperm(L.lane4 - 4 * st, v);  // outside the group: discarded below
  }
  return L.seg >= st ? r : ident;
}

// Value of lane seg+st of the group (ident past the group's last lane).
template <bool ROW, typename T>
__device__ __forceinline__ T from_above(const Lanes &L, T v, int st, T ident) {
  T r;
  if constexpr (ROW) {
    switch (st) {
      case 1: r = dpp<0x101>(ident, v); break;  // row_shl:1
      case 2: r = dpp<0x102>(ident, v); break;
      case 4: r = dpp<0x104>(ident, v); break;
      default: r = dpp<0x108>(ident, v); break;
    }
  } else {
    r = bperm(L.lane4 + 4 * st, v);  // outside the group: discarded below
  }
  return L.seg + st < L.G ? r : ident;
}

// Whether PCR level with stride st has work: st < G.  Groups exchanging
// through ds_bpermute have at least 32 lanes, so their first five levels
// always do (straight-line code, no branch).
template <bool ROW>
__device__ __forceinline__ bool level_used(const Lanes &L, int st) {
  return (!ROW && st <= 16) || st < L.G;
}

// M consecutive values p[i0 .. i0+M-1] (0 past n); `vec` says p + i0 is 16-byte
// aligned whenever the segment is complete.
template <typename T, int M>
__device__ __forceinline__ void load_seg(const T *__restrict__ p, int i0, int n,
                                         bool vec, T (&o)[M]) {
  if (vec && i0 + M <= n) {
    if constexpr (sizeof(T) == 4) {
#pragma unroll
      for (int q = 0; q < M / 4; ++q) {
        const float4 v = *reinterpret_cast<const float4 *>(p + i0 + 4 * q);
        o[4 * q] = v.x;
        o[4 * q + 1] = v.y;
        o[4 * q + 2] = v.z;
        o[4 * q + 3] = v.w;
      }
    } else {
#pragma unroll
      for (int q = 0; q < M / 2; ++q) {
        const double2 v = *reinterpret_cast<const double2 *>(p + i0 + 2 * q);
        o[2 * q] = v.x;
        o[2 * q + 1] = v.y;
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < M; ++k) o[k] = (i0 + k < n) ? p[i0 + k] : T(0);
  }
}

// s_waitcnt operand for vmcnt(0) alone on gfx9 (expcnt 7, lgkmcnt 15 leave
// those counters unconstrained).
constexpr int kWaitVmcnt0 = 0x0f70;

// The two halves of load_seg, for callers that make the completeness test
// once for several segments (so that their 16-byte loads issue together: a
// test per segment makes each one's loads wait for the ones before).
template <typename T, int M>
__device__ __forceinline__ void load_seg_vec(const T *__restrict__ p, int i0, T (&o)[M]) {
  if constexpr (sizeof(T) == 4) {
#pragma unroll
    for (int q = 0; q < M / 4; ++q) {
      const float4 v = *reinterpret_cast<const float4 *>(p + i0 + 4 * q);
      o[4 * q] = v.x;
      o[4 * q + 1] = v.y;
      o[4 * q + 2] = v.z;
      o[4 * q + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int q = 0; q < M / 2; ++q) {
      const double2 v = *reinterpret_cast<const double2 *>(p + i0 + 2 * q);
      o[2 * q] = v.x;
      o[2 * q + 1] = v.y;
    }
  }
}

template <typename T, int M>
__device__ __forceinline__ void load_seg_tail(const T *__restrict__ p, int i0, int n, T (&o)[M]) {
#pragma unroll
  for (int k = 0; k < M; ++k) o[k] = (i0 + k < n) ? p[i0 + k] : T(0);
}

template <typename T, int M>
__device__ __forceinline__ void store_seg(T *__restrict__ p, int i0, int n,
                                          bool vec, const T (&o)[M]) {
  if (vec && i0 + M <= n) {
    if constexpr (sizeof(T) == 4) {
#pragma unroll
      for (int q = 0; q < M / 4; ++q)
        *reinterpret_cast<float4 *>(p + i0 + 4 * q) =
            make_float4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    } else {
#pragma unroll
      for (int q = 0; q < M / 2; ++q)
        *reinterpret_cast<double2 *>(p + i0 + 2 * q) = make_double2(o[2 * q], o[2 * q + 1]);
    }
  } else {
#pragma unroll
    for (int k = 0; k < M; ++k)
      if (i0 + k < n) p[i0 + k] = o[k];
  }
}

// Buffer resource over one problem's [ny][nx] grid (raw, 32-bit elements
// format word of the gfx9 family); loads and stores through it take a
// per-lane byte offset and a wave-uniform one.
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t grid_rsrc(const T *p, int nx, int ny) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(p), 0, nx * ny * (int)sizeof(T),
                                           0x00020000);
}
template <typename T>
__device__ __forceinline__ T buffer_load(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  }
}
template <typename T>
__device__ __forceinline__ void buffer_store(__amdgpu_buffer_rsrc_t r, T v, int voff, int soff) {
  if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
  } else {
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
  }
}

// A wave-uniform 64-bit value (a descriptor field) in SGPRs: the compiler
// cannot always prove such a load uniform, and a buffer resource built from
// a value it holds in vector registers costs a uniformity loop per access.
__device__ __forceinline__ int64_t wave_uniform(int64_t v) {
  const int lo = __builtin_amdgcn_readfirstlane((int)v);
  const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
  return (int64_t)(((uint64_t)(unsigned)hi << 32) | (unsigned)lo);
}

// Coefficient tables of the geometry buffer in the sweep's real type.
template <typename T>
struct Taps {
  const T *ax, *cx, *ay, *cy;
  __device__ Taps(const double *g, int nx, int ny) {
    if constexpr (sizeof(T) == 4) {
      const float *f = reinterpret_cast<const float *>(g + LVC_GEO_F64(nx, ny));
      ax = f;
      cx = f + LVC_P4(nx);
      ay = cx + LVC_P4(nx);
      cy = ay + LVC_P4(ny);
    } else {
      ax = g + nx;
      cx = g + 2 * nx;
      ay = g + 3 * nx + ny;
      cy = g + 3 * nx + 2 * ny;
    }
  }
};

// ---------------------------------------------------------------------------
// Setup: initGrid's arrays (:12-23), initOperator (:26-44) taps, the VarX
// factor x^(2 beta) of updateParams (:61-65) and setPayoff (:46-47).  The
// scalars (dx, dy, myXindex, myYindex: :15-20) come from the host in the
// descriptor, in the entry's own precision; the arrays are computed here in
// FP64 with multiply-add contraction disabled, so the grid is bit-identical
// to a C restatement of the same FP64 lines.  VarX's log at column 0 takes
// the descriptor's x0log (the reference's own f32 x[0] for the f32 entry):
// it decides whether that log is finite, -inf (x[0] = 0) or NaN (x[0] < 0).
// Grid: (problems, row chunks); chunk 0 also writes the geometry and resets
// the time index.  With `list`, block i sets up problem list[i] from the
// descriptor src[i] (chunk 0 stores it, time index set, to probs[list[i]]).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void lvc_setup_kernel(lvc_problem *probs,
                                                        const lvc_problem *src,
                                                        const int32_t *list,
                                                        double *geo, T *R) {
#pragma clang fp contract(off)
  lvc_problem &dst = probs[list ? list[blockIdx.x] : blockIdx.x];
  const lvc_problem pr = list ? src[blockIdx.x] : dst;
  const int nx = pr.nx, ny = pr.ny;
  const double s0 = pr.s0, alpha = pr.alpha, nu = pr.nu, beta = pr.beta;
  const double dx = pr.dx, dy = pr.dy;
  const int xi = pr.xindex, yi = pr.yindex;

  // payoff rows of this chunk: R[j][i] = max(x_i - K, 0)
  T *res = R + pr.state_off;
  const int rows = (ny + gridDim.y - 1) / gridDim.y;
  const int r0 = blockIdx.y * rows, r1 = min(ny, r0 + rows);
  for (int i = threadIdx.x; i < nx; i += blockDim.x) {
    const double xv = (double)i * dx - (double)xi * dx + s0;
    const double v = xv - pr.strike;
    const T pay = (T)(v > 0.0 ? v : 0.0);
    for (int j = r0; j < r1; ++j) res[(int64_t)j * nx + i] = pay;
  }
  if (blockIdx.y != 0) return;

  const double logAlpha = log(alpha);
  double *x = geo + pr.geo_off;
  double *ax = x + nx, *cx = ax + nx, *y = cx + nx, *ay = y + ny, *cy = ay + ny;
  float *f32 = reinterpret_cast<float *>(geo + pr.geo_off + LVC_GEO_F64(nx, ny));
  float *axf = f32, *cxf = axf + LVC_P4(nx), *ayf = cxf + LVC_P4(nx);
  float *cyf = ayf + LVC_P4(ny);
  for (int i = threadIdx.x; i < nx; i += blockDim.x)
    x[i] = (double)i * dx - (double)xi * dx + s0;
  for (int j = threadIdx.x; j < ny; j += blockDim.x)
    y[j] = (double)j * dy - (double)yi * dy + logAlpha;
  if (threadIdx.x == 0) {
    if (list) {
      lvc_problem w = pr;
      w.k = pr.nt - 2;
      dst = w;
    } else {
      dst.k = pr.nt - 2;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < LVC_P4(nx); i += blockDim.x) {
    double d0 = 0.0, d2 = 0.0;
    if (i >= 1 && i <= nx - 2) {
      const double dl = x[i] - x[i - 1], du = x[i + 1] - x[i];
      d0 = 2.0 / dl / (dl + du);
      d2 = 2.0 / du / (dl + du);
    }
    if (i < nx) {
      const double X = exp(2.0 * (beta * log(i == 0 ? pr.x0log : x[i])));
      ax[i] = -0.25 * X * d0;
      cx[i] = -0.25 * X * d2;
    }
    axf[i] = i < nx ? (float)ax[i] : 0.0f;
    cxf[i] = i < nx ? (float)cx[i] : 0.0f;
  }
  for (int j = threadIdx.x; j < LVC_P4(ny); j += blockDim.x) {
    double d0 = 0.0, d2 = 0.0;
    if (j >= 1 && j <= ny - 2) {
      const double dl = y[j] - y[j - 1], du = y[j + 1] - y[j];
      d0 = 2.0 / dl / (dl + du);
      d2 = 2.0 / du / (dl + du);
    }
    if (j < ny) {
      ay[j] = -0.25 * nu * nu * d0;
      cy[j] = -0.25 * nu * nu * d2;
    }
    ayf[j] = j < ny ? (float)ay[j] : 0.0f;
    cyf[j] = j < ny ? (float)cy[j] : 0.0f;
  }
}

// ---------------------------------------------------------------------------
// Segment solver shared by both sweeps.
// ---------------------------------------------------------------------------
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
  Lanes L;

  // a, c: taps of rows 0..M-1 of this segment; s: diagonal shift per row
  // (dtInv for rows of the system, 1 for padding rows past its end).
  template <bool ROW>
  __device__ __forceinline__ void prepare(const T (&a)[M], const T (&cc)[M],
                                          const T (&s)[M], Lanes lanes) {
    L = lanes;
#pragma unroll
    for (int k = 0; k < M; ++k) c[k] = cc[k];
    aL = a[M - 1];
    cL = c[M - 1];
    // Local factorisation of rows 0..M-2: g = pivot - |c|, every update a sum
    // of positive terms for an M-matrix.
    T g = s[0] - a[0];
    rb[0] = rcp(g - c[0]);
    m[0] = T(0);
#pragma unroll
    for (int k = 1; k < M - 1; ++k) {
      m[k] = a[k] * rb[k - 1];
      g = fma(-m[k], g, s[k]);
      rb[k] = rcp(g - c[k]);
    }
    // Homogeneous responses v (to x[first-1]) and w (to x[last]).
    v[0] = -a[0];
#pragma unroll
    for (int k = 1; k < M - 1; ++k) v[k] = -m[k] * v[k - 1];
    v[M - 2] = v[M - 2] * rb[M - 2];
    w[M - 2] = -c[M - 2] * rb[M - 2];
#pragma unroll
    for (int k = M - 3; k >= 0; --k) {
      v[k] = fma(-c[k], v[k + 1], v[k]) * rb[k];
      w[k] = -c[k] * w[k + 1] * rb[k];
    }
    // Reduced row of this lane: A X[seg-1] + B X[seg] + C X[seg+1] = D.
    const T vn = from_above<ROW>(L, v[0], 1, T(0));
    const T wn = from_above<ROW>(L, w[0], 1, T(0));
    T A = aL * v[M - 2];
    T B = fma(-cL, T(1) - vn, fma(-aL, T(1) - w[M - 2], s[M - 1]));
    T C = cL * wn;
    // PCR on the matrix; the multipliers are kept for the right-hand sides.
#pragma unroll
    for (int l = 0; l < kMaxLevels; ++l) {
      const int st = 1 << l;
      if (level_used<ROW>(L, st)) {
        const T Am = from_below<ROW>(L, A, st, T(0)), Bm = from_below<ROW>(L, B, st, T(1));
        const T Cm = from_below<ROW>(L, C, st, T(0));
        const T Ap = from_above<ROW>(L, A, st, T(0)), Bp = from_above<ROW>(L, B, st, T(1));
        const T Cp = from_above<ROW>(L, C, st, T(0));
        const T q1 = A * rcp(Bm), q2 = C * rcp(Bp);
        k1[l] = q1;
        k2[l] = q2;
        A = -Am * q1;
        C = -Cp * q2;
        B = fma(-Ap, q2, fma(-Cm, q1, B));
      }
    }
    rB = rcp(B);
  }

  // Solves for one right-hand side d (rows 0..M-1 of this segment).
  template <bool ROW>
  __device__ __forceinline__ void solve(const T (&d)[M], T (&x)[M]) const {
    x[0] = d[0];
#pragma unroll
    for (int k = 1; k < M - 1; ++k) x[k] = fma(-m[k], x[k - 1], d[k]);
    x[M - 2] = x[M - 2] * rb[M - 2];
#pragma unroll
    for (int k = M - 3; k >= 0; --k) x[k] = fma(-c[k], x[k + 1], x[k]) * rb[k];
    const T un = from_above<ROW>(L, x[0], 1, T(0));
    T D = fma(-cL, un, fma(-aL, x[M - 2], d[M - 1]));
#pragma unroll
    for (int l = 0; l < kMaxLevels; ++l) {
      const int st = 1 << l;
      if (level_used<ROW>(L, st)) {
        const T Dm = from_below<ROW>(L, D, st, T(0)), Dp = from_above<ROW>(L, D, st, T(0));
        D = fma(-Dp, k2[l], fma(-Dm, k1[l], D));
      }
    }
    const T X = D * rB;
    const T XL = from_below<ROW>(L, X, 1, T(0));
#pragma unroll
    for (int k = 0; k < M - 1; ++k) x[k] = fma(w[k], X, fma(v[k], XL, x[k]));
    x[M - 1] = X;
  }
};

// One refinement pass on the solution x of  s*x + a*(x[-1]-x) + c*(x[+1]-x) = d:
// the residual in FP64, a second solve with the same factorisation, the
// correction added with one rounding.
template <bool ROW, typename T, int M>
__device__ __forceinline__ void refine(const SegSys<T, M> &sys, const T (&d)[M],
                                       const T (&a)[M], const T (&c)[M],
                                       const T (&s)[M], T (&x)[M]) {
  const T xl = from_below<ROW>(sys.L, x[M - 1], 1, T(0));
  const T xr = from_above<ROW>(sys.L, x[0], 1, T(0));
  // One diagonal for every row: s[0] (dt) on a lane with any row of the
  // system, whose rows come first; rows past its end (diagonal 1) have zero
  // taps and right-hand side, so x there is +-0 (or NaN/inf after a
  // non-finite neighbour) and s * x has the same value for any positive s.
  const double s64 = (double)s[0];
  T r[M];
  // One difference per pair of neighbours: x[k-1] - x[k] is the negated
  // x[k] - x[k-1] (round to nearest is symmetric: the same value, only an
  // exact zero may come out with the other sign).
  double dlo = (double)x[0] - (double)xl;  // x[k] - x[k-1]
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const double xk = (double)x[k];
    const double dhi = (double)(k < M - 1 ? x[k + 1] : xr) - xk;  // x[k+1] - x[k]
    const double ax = fma((double)c[k], dhi, fma(-(double)a[k], dlo, s64 * xk));
    r[k] = (T)((double)d[k] - ax);
    dlo = dhi;
  }
  T e[M];
  sys.template solve<ROW>(r, e);
  // x + e rounded once to T: the correctly rounded sum, which is what an FP64
  // addition rounded back to T gives (the exact sum of two floats fits a
  // double whenever their exponents are within 29 of each other).
#pragma unroll
  for (int k = 0; k < M; ++k) x[k] += e[k];
}

// ---------------------------------------------------------------------------
// x sweep: one workgroup = `count` rows of `cpl` consecutive problems that
// share their factorisation (one problem unless they differ only in the
// strike), `lanes` lanes per row.  U (the right-hand side) and V
// (the explicit y term) wait in LDS while the solver's registers are live.
// ---------------------------------------------------------------------------
template <typename T, int M>
constexpr int kXsweepLds = 2 * M * LVC_BLOCK;  // elements of T

// NB = 1: single-problem tiles, the group loop resolved at compile time (fewer
// registers); NB = 0: tile.cpl problems.  Both run the same arithmetic, so a
// problem's result does not depend on which variant its batch used.
template <typename T, int M, bool REF, int NB, bool ROW>
__device__ __forceinline__ void xsweep_tile(const lvc_tile tile,
                                            const lvc_problem *__restrict__ probs,
                                            const double *__restrict__ geo,
                                            const T *__restrict__ R,
                                            T *__restrict__ Y, T *smem) {
  T *lds_u = smem, *lds_v = smem + M * LVC_BLOCK;
  // The tile's problems tile.problem .. + tile.cpl - 1 differ only in their
  // state (the strike): they share sizes, grid, taps, time step and so the
  // factorisation of every row, which is computed once from the first.
  const lvc_problem &pr = probs[tile.problem];
  const int kt = pr.k;
  if (kt < 0) return;  // problems finished: whole workgroup exits
  const int G = tile.lanes;
  const int seg = threadIdx.x & (G - 1);
  const int rloc = threadIdx.x / G;
  if (rloc >= tile.count) return;  // whole lane group exits together
  const int nx = pr.nx, ny = pr.ny;
  const int j = tile.first + rloc;
  const double *g = geo + pr.geo_off;
  const Taps<T> tp(g, nx, ny);

  const double tnow = timeline(pr.t, kt, pr.nt), tnext = timeline(pr.t, kt + 1, pr.nt);
  const T dt = (T)(1.0 / (tnext - tnow));
  const double nu = pr.nu;
  const T E = (T)exp(fma(-nu * nu, tnow, 2.0 * g[3 * nx + j]));  // exp(2y - nu^2 t)
  const T ayj = tp.ay[j], cyj = tp.cy[j];
  const bool vec = (nx & 3) == 0;
  const int i0 = seg * M;

  // Rows j-1, j, j+1 of problem b.  A single problem's rows and explicit
  // terms come before the factorisation (the loads overlap it).  In a group,
  // without refinement each next problem's rows are loaded while the current
  // one is solved; with it, the registers that would take are worth more as
  // occupancy (measured), so they load at the top of each iteration.
  constexpr bool kPrefetch = !REF;
  T rc[M], rm[M], rp[M], rl, rr;
  // FP32 tiles: every lane loads from rows and columns inside the grid (the
  // neighbours of boundary rows and columns clamped to the boundary) and the
  // values outside are replaced by selects, not skipped by branches.  FP64
  // tiles keep the branches (measured faster for them).
  constexpr bool kClampedLoads = sizeof(T) == 4;
  const int64_t dn = j > 0 ? -(int64_t)nx : 0, up = j < ny - 1 ? (int64_t)nx : 0;
  const int il = min(max(i0 - 1, 0), nx - 1), ir = min(i0 + M, nx - 1);
  const bool seg_full = vec && i0 + M <= nx;  // this lane's segment is whole
  auto load_rows = [&](int b) {
    const T *row = R + probs[tile.problem + b].state_off + (int64_t)j * nx;
    if constexpr (kClampedLoads) {
      // One test for the three rows, so that their 16-byte loads issue
      // together.  Both paths write the same registers, so the compiler
      // would wait for every memory operation in flight (the previous
      // problem's stores included) before the 16-byte loads, in case the
      // element loads of lanes past the row end are still landing there;
      // those lanes wait for their own loads instead, and a wave with none
      // skips the wait.
      if (seg_full) {
        load_seg_vec<T, M>(row, i0, rc);
        load_seg_vec<T, M>(row + dn, i0, rm);
        load_seg_vec<T, M>(row + up, i0, rp);
      } else {
        load_seg_tail<T, M>(row, i0, nx, rc);
        load_seg_tail<T, M>(row + dn, i0, nx, rm);
        load_seg_tail<T, M>(row + up, i0, nx, rp);
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
      }
      const T vl = row[il], vr = row[ir];
      rl = (i0 > 0 && i0 - 1 < nx) ? vl : T(0);
      rr = (i0 + M < nx) ? vr : T(0);
    } else {
      load_seg<T, M>(row, i0, nx, vec, rc);
      rl = (i0 > 0 && i0 - 1 < nx) ? row[i0 - 1] : T(0);
      rr = (i0 + M < nx) ? row[i0 + M] : T(0);
      if (j > 0) load_seg<T, M>(row - nx, i0, nx, vec, rm);
      if (j < ny - 1) load_seg<T, M>(row + nx, i0, nx, vec, rp);
    }
  };
  T a[M], c[M], s[M];
  // U (the right-hand side) to registers; V (the explicit y term) and, for
  // the refinement, U to LDS.
  auto explicit_terms = [&](T (&U)[M]) {
    T dlo = rc[0] - rl;  // rc[k] - rc[k-1]; rc[k-1] - rc[k] is its negation (as above)
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const T r0 = rc[k];
      const T dhi = (k < M - 1 ? rc[k + 1] : rr) - r0;
      const T dm = j > 0 ? rm[k] - r0 : T(0);
      const T dp = j < ny - 1 ? rp[k] - r0 : T(0);
      const T ux = fma(c[k], dhi, -(a[k] * dlo));  // 0.5 * explicitX
      dlo = dhi;
      const T vy = T(-2) * fma(cyj, dp, ayj * dm);         // explicitY
      lds_v[k * LVC_BLOCK + threadIdx.x] = vy;
      U[k] = fma(dt, r0, -ux) + vy;
      if (REF) lds_u[k * LVC_BLOCK + threadIdx.x] = U[k];
    }
  };
  if (NB == 1 || kPrefetch) load_rows(0);
  if constexpr (kClampedLoads) {
    if (seg_full) {
      load_seg_vec<T, M>(tp.ax, i0, a);
      load_seg_vec<T, M>(tp.cx, i0, c);
    } else {
      load_seg_tail<T, M>(tp.ax, i0, nx, a);
      load_seg_tail<T, M>(tp.cx, i0, nx, c);
      __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    }
  } else {
    load_seg<T, M>(tp.ax, i0, nx, vec, a);
    load_seg<T, M>(tp.cx, i0, nx, vec, c);
  }
#pragma unroll
  for (int k = 0; k < M; ++k) {
    a[k] *= E;
    c[k] *= E;
    s[k] = i0 + k < nx ? dt : T(1);
  }
  T U[M];
  if (NB == 1) explicit_terms(U);
  SegSys<T, M> sys;
  sys.template prepare<ROW>(a, c, s, make_lanes(seg, G));

  const int nb = NB > 0 ? NB : tile.cpl;
  for (int b = 0; b < nb; ++b) {
    const int64_t off = probs[tile.problem + b].state_off + (int64_t)j * nx;
    if (NB != 1) {
      if (!kPrefetch) load_rows(b);
      explicit_terms(U);
      if (kPrefetch && b + 1 < nb) load_rows(b + 1);
    }
    T x[M];
    sys.template solve<ROW>(U, x);
    if (REF) {
      T d[M];
#pragma unroll
      for (int k = 0; k < M; ++k) d[k] = lds_u[k * LVC_BLOCK + threadIdx.x];
      refine<ROW>(sys, d, a, c, s, x);
    }
    T out[M];
#pragma unroll
    for (int k = 0; k < M; ++k)
      out[k] = fma(dt, x[k], T(-0.5) * lds_v[k * LVC_BLOCK + threadIdx.x]);
    store_seg<T, M>(Y + off, i0, nx, vec, out);
  }
}

constexpr int kFacScratch = 6 * LVC_COL_ROWS;  // elements of the factor scratch
template <typename T>
__device__ void ycol_factor(const lvc_problem &pr, const double *g, T dt, int M,
                            T *__restrict__ f, T *scr);

// yseg > 0: the y sweep runs on column segments of yseg rows, and the first
// nfac workgroups of the grid each write one problem's y factor for this step
// (problem fprobs[blockIdx.x], wave 0 at raised priority: a serial chain that
// should finish long before the row tiles do); the row tiles follow.
// Occupancy floor (waves per SIMD) for the shared FP32+refinement tile with
// 8-unknown segments, which the compiler would otherwise give 198 registers
// (2 waves); 3 waves fit without spills.
template <typename T, int M, bool REF, bool SHARE>
constexpr int kXsweepWaves = (sizeof(T) == 4 && M == 8 && SHARE && REF) ? 3 : 1;

template <typename T, int M, bool REF, bool SHARE>
__global__ __launch_bounds__(LVC_BLOCK, (kXsweepWaves<T, M, REF, SHARE>)) void lvc_xsweep_kernel(
    const lvc_problem *__restrict__ probs, const lvc_tile *__restrict__ tiles,
    const double *__restrict__ geo, const T *__restrict__ R, T *__restrict__ Y,
    T *__restrict__ yfac, int yseg, const int32_t *__restrict__ fprobs, int nfac) {
  // row tiles: U and V panels; factor workgroups: the factor scratch
  __shared__ T smem[kXsweepLds<T, M> > kFacScratch ? kXsweepLds<T, M> : kFacScratch];
  if ((int)blockIdx.x < nfac) {
    if (yseg > 0 && threadIdx.x < 64) {
      const int p = fprobs[blockIdx.x];
      const lvc_problem &pr = probs[p];
      const int kt = pr.k;
      if (kt >= 0) {
        __builtin_amdgcn_s_setprio(3);
        const double tnow = timeline(pr.t, kt, pr.nt), tnext = timeline(pr.t, kt + 1, pr.nt);
        ycol_factor<T>(pr, geo + pr.geo_off, (T)(1.0 / (tnext - tnow)), yseg,
                       yfac + (int64_t)p * LVC_YFAC_LEN, smem);
        __builtin_amdgcn_s_setprio(0);
      }
    }
    return;
  }
  const lvc_tile tile = tiles[blockIdx.x - nfac];
  if (tile.lanes <= 16)
    xsweep_tile<T, M, REF, SHARE ? 0 : 1, true>(tile, probs, geo, R, Y, smem);
  else
    xsweep_tile<T, M, REF, SHARE ? 0 : 1, false>(tile, probs, geo, R, Y, smem);
}

// ---------------------------------------------------------------------------
// y sweep: one workgroup = `count` columns of one problem.  The y matrix is
// the same for every column of a problem (VarY = nu^2, :59), so a lane
// factors its segment once and solves `cpl` columns with it; G lanes per
// column, cpl * 256/G columns per workgroup.  The column panel is staged
// through LDS (coalesced row loads and stores), stored segment-interleaved
// (row r = q*M + k at panel row k*G + q) so the lanes of a column read
// consecutive panel rows.
// ---------------------------------------------------------------------------

// Every thread of the workgroup must call this (it holds barriers).
template <typename T, int M, bool REF, bool ROW>
__device__ __forceinline__ void ysweep_tile(const lvc_tile tile, lvc_problem *probs,
                                            const double *__restrict__ geo,
                                            const T *__restrict__ Y,
                                            T *__restrict__ R, T *lds) {
  lvc_problem &pr = probs[tile.problem];
  const int kt = pr.k;
  if (kt < 0) return;
  const int G = tile.lanes, cpl = tile.cpl;
  const int lc = LVC_BLOCK / G;    // lane-columns per workgroup
  const int cols = cpl * lc;       // columns per workgroup
  const int CP = cols + 1;         // LDS pitch
  const int nx = pr.nx, ny = pr.ny;
  const int ncol = tile.count, col0 = tile.first;
  const int rows_pad = G * M;
  auto lidx = [&](int r, int c) { return ((r % M) * G + r / M) * CP + c; };

  // cols divides LVC_BLOCK: each thread keeps one column and strides rows
  const int pc = threadIdx.x & (cols - 1), pr0 = threadIdx.x / cols;
  const int rstride = LVC_BLOCK / cols;
  const T *src = Y + pr.state_off + col0 + pc;
  for (int r = pr0; r < rows_pad; r += rstride) {
    T val = T(0);
    if (r < ny && pc < ncol) val = src[(int64_t)r * nx];
    lds[lidx(r, pc)] = val;
  }
  __syncthreads();

  const Taps<T> tp(geo + pr.geo_off, nx, ny);
  const double tnow = timeline(pr.t, kt, pr.nt), tnext = timeline(pr.t, kt + 1, pr.nt);
  const T dt = (T)(1.0 / (tnext - tnow));
  const int q = threadIdx.x & (G - 1);
  const int lcol = threadIdx.x / G;
  const int j0 = q * M;
  T a[M], c[M], s[M];
  const bool vec = ((nx | ny) & 3) == 0;
  load_seg<T, M>(tp.ay, j0, ny, vec, a);
  load_seg<T, M>(tp.cy, j0, ny, vec, c);
#pragma unroll
  for (int k = 0; k < M; ++k) s[k] = j0 + k < ny ? dt : T(1);
  SegSys<T, M> sys;
  sys.template prepare<ROW>(a, c, s, make_lanes(q, G));
  for (int ci = 0; ci < cpl; ++ci) {
    const int col = ci * lc + lcol;  // lanes of a column interleave across cpl
    T d[M], x[M];
#pragma unroll
    for (int k = 0; k < M; ++k) d[k] = lds[lidx(j0 + k, col)];
    sys.template solve<ROW>(d, x);
    if (REF) refine<ROW>(sys, d, a, c, s, x);
    // each lane overwrites only its own panel entries, read just above
#pragma unroll
    for (int k = 0; k < M; ++k) lds[lidx(j0 + k, col)] = x[k];
  }
  __syncthreads();
  T *dst = R + pr.state_off + col0 + pc;
  if (pc < ncol)
    for (int r = pr0; r < ny; r += rstride) dst[(int64_t)r * nx] = lds[lidx(r, pc)];
  // The last of the problem's y workgroups to finish advances its time index
  // (every workgroup of the problem has read `k` before it counts itself).
  if (threadIdx.x == 0) {
    const int done = atomicAdd(&pr.ydone, 1) + 1;
    if (done == pr.ytiles) {
      pr.ydone = 0;
      pr.k = kt - 1;
    }
  }
}

template <typename T, int M, bool REF>
__global__ __launch_bounds__(LVC_BLOCK) void lvc_ysweep_kernel(
    lvc_problem *probs, const lvc_tile *__restrict__ tiles,
    const double *__restrict__ geo, const T *__restrict__ Y, T *__restrict__ R) {
  extern __shared__ __align__(16) unsigned char dyn_lds[];
  const lvc_tile tile = tiles[blockIdx.x];
  T *lds = reinterpret_cast<T *>(dyn_lds);
  if (tile.lanes <= 16)
    ysweep_tile<T, M, REF, true>(tile, probs, geo, Y, R, lds);
  else
    ysweep_tile<T, M, REF, false>(tile, probs, geo, Y, R, lds);
}

// ---------------------------------------------------------------------------
// Column y sweep (segment lengths 32 and 64): one lane per column segment,
// Thomas elimination down each segment.  A workgroup is LVC_COL_BLOCK / 64
// waves; wave w holds segment p = w % P of the 64 columns of column group
// w / P (P = ceil(ny / M) segments per column), lane = column.  Every lane of
// a wave works on the same rows, so the factorisation, which is the same for
// every column (VarY = nu^2, :59), is computed once per segment into LDS and
// read at wave-uniform addresses; the column data itself stays in registers
// (coalesced row loads and stores, no transpose).  The P segments of a column
// are coupled through their last unknowns: a tridiagonal reduced system of P
// rows, which every lane solves for its own column by Thomas after one LDS
// exchange.  With P = 1 this is plain thread-per-column Thomas.
//
// Segment algebra (rows k = 0..M-1 of segment p, global row p*M + k): rows
// 0..M-2 give x_k = y_k + v_k X_{p-1} + w_k X_p, where y solves them with the
// couplings to X_{p-1} (the previous segment's last unknown) and X_p (this
// segment's last) set to zero, and v, w are the responses to a unit value of
// each; row M-1 then couples X_{p-1}, X_p and X_{p+1}.  Long segments of a
// weakly shifted system have v, w close to 1, so the reduced diagonal is not
// formed from 1 - w or 1 - v: the rows of the difference form sum to s, which
// gives 1 = v + w + z with z the response to the shift s itself (a sum of
// positive terms), and the reduced system is factored in the same g-form from
// its row sums  s - aL*z_{M-2} - cL*z'_0  (z' of the next segment).
// ---------------------------------------------------------------------------
constexpr int kColWaves = LVC_COL_BLOCK / 64;  // segments per column, at most
static_assert((kColWaves & (kColWaves - 1)) == 0, "wave count must be a power of two");

// The factor table of one problem's y system (elements of T, at problem *
// LVC_YFAC_LEN): per row m, rb, crb, v, w, z (LVC_COL_ROWS rows each), then
// per segment aL, cL, sL (taps of its last row) and the reduced system's
// Thomas factor rm, rrb, rC (kColWaves entries each).
enum { kFm, kFrb, kFcrb, kFv, kFw, kFz };
enum { kSaL, kScL, kSsL, kSrm, kSrrb, kSrC };
static_assert(LVC_YFAC_LEN == 6 * LVC_COL_ROWS + 6 * kColWaves, "factor table layout");
static_assert(LVC_COL_ROWS >= kColWaves * 64, "largest column segment is 64 rows");

template <typename T>
__device__ __forceinline__ T *fac_row(T *f, int which) { return f + which * LVC_COL_ROWS; }
template <typename T>
__device__ __forceinline__ T *fac_seg(T *f, int which) {
  return f + 6 * LVC_COL_ROWS + which * kColWaves;
}

// Factorisation of one problem's y system for segments of M rows (the time
// step's dt), written to its factor table.  Called by all 64 lanes of one
// wavefront; lane q < P factors segment q (Thomas in g-form, g = pivot - |c|,
// every update a sum of positive terms; responses v, w, z), then every lane
// forms the reduced system and lane 0 writes its factor.  `scr` is this
// wavefront's LDS scratch of kFacScratch elements: the taps, staged once with
// coalesced loads so that the serial loops never wait on global memory, and
// the forward values the backward pass reads back (each lane its own rows).

template <typename T>
__device__ void ycol_factor(const lvc_problem &pr, const double *g, T dt, int M,
                            T *__restrict__ f, T *scr) {
  const int ny = pr.ny, P = (ny + M - 1) / M;
  const int q = __lane_id();
  T *srb = scr, *scrb = scr + LVC_COL_ROWS, *sv = scr + 2 * LVC_COL_ROWS;
  T *sz = scr + 3 * LVC_COL_ROWS, *sa = scr + 4 * LVC_COL_ROWS;
  T *sc = scr + 5 * LVC_COL_ROWS;
  {
    const Taps<T> tp(g, pr.nx, ny);
    for (int r = q; r < P * M; r += 64) {
      sa[r] = r < ny ? tp.ay[r] : T(0);
      sc[r] = r < ny ? tp.cy[r] : T(0);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  auto A = [&](int r) { return sa[r]; };
  auto C = [&](int r) { return sc[r]; };
  auto S = [&](int r) { return r < ny ? dt : T(1); };
  T *fm = fac_row(f, kFm), *frb = fac_row(f, kFrb), *fcrb = fac_row(f, kFcrb);
  T *fv = fac_row(f, kFv), *fw = fac_row(f, kFw), *fz = fac_row(f, kFz);
  T vL = T(0), zL = T(0), w0 = T(0), z0 = T(0), aL = T(0), cL = T(0), sL = T(1);
  if (q < P) {
    const int r0 = q * M;
    T a = A(r0), c = C(r0);
    T gg = S(r0) - a;
    T rb = rcp(gg - c);
    T v = -a, z = S(r0);
    fm[r0] = T(0);
    frb[r0] = srb[r0] = rb;
    fcrb[r0] = scrb[r0] = c * rb;
    sv[r0] = v;
    sz[r0] = z;
#pragma unroll 8
    for (int k = 1; k < M - 1; ++k) {
      const int r = r0 + k;
      a = A(r);
      c = C(r);
      const T m = a * rb;
      gg = fma(-m, gg, S(r));
      rb = rcp(gg - c);
      v = -m * v;
      z = fma(-m, z, S(r));
      fm[r] = m;
      frb[r] = srb[r] = rb;
      fcrb[r] = scrb[r] = c * rb;
      sv[r] = v;
      sz[r] = z;
    }
    v *= rb;
    z *= rb;
    T w = -scrb[r0 + M - 2];
    fv[r0 + M - 2] = v;
    fw[r0 + M - 2] = w;
    fz[r0 + M - 2] = z;
    vL = v;
    zL = z;
#pragma unroll 8
    for (int k = M - 3; k >= 0; --k) {
      const int r = r0 + k;
      const T crb = scrb[r], rbk = srb[r];
      v = fma(-crb, v, sv[r] * rbk);
      z = fma(-crb, z, sz[r] * rbk);
      w = -crb * w;
      fv[r] = v;
      fw[r] = w;
      fz[r] = z;
    }
    w0 = w;
    z0 = z;
    aL = A(r0 + M - 1);
    cL = C(r0 + M - 1);
    sL = S(r0 + M - 1);
    fac_seg(f, kSaL)[q] = aL;
    fac_seg(f, kScL)[q] = cL;
    fac_seg(f, kSsL)[q] = sL;
  }
  // Reduced system: sub- and super-diagonal A = aL*v_{M-2} <= 0 and
  // C = cL*w'_0 <= 0 (w' of the next segment), row sum
  // delta = sL - aL*z_{M-2} - cL*z'_0 >= 0; its pivots in g-form.
  const T w0n = __shfl(w0, q + 1), z0n = __shfl(z0, q + 1);
  const bool last = q + 1 >= P;
  const T Aq = aL * vL, Cq = last ? T(0) : cL * w0n;
  const T delta = fma(-cL, last ? T(0) : z0n, fma(-aL, zL, sL));
  T gq = T(0), rrb = T(0);
  for (int j = 0; j < P; ++j) {  // serial over the segments, in every lane
    const T Aj = __shfl(Aq, j), Cj = __shfl(Cq, j), dj = __shfl(delta, j);
    const T m = j == 0 ? T(0) : Aj * rrb;
    gq = j == 0 ? dj - Aj : fma(-m, gq, dj);
    rrb = rcp(gq - Cj);
    if (q == 0) {
      fac_seg(f, kSrm)[j] = m;
      fac_seg(f, kSrrb)[j] = rrb;
      fac_seg(f, kSrC)[j] = Cj;
    }
  }
}

template <typename T, int M>
struct ColLds {
  static constexpr int kRows = kColWaves * M;
  double a64[kRows], c64[kRows];                // the taps in FP64 for the residual
  T y0[2][LVC_COL_BLOCK], Dt[2][LVC_COL_BLOCK];  // exchange, one per solve
};

// Solves in place for the right-hand side held in x (this lane's segment p of
// one column), using exchange slot `sl`.  Returns the neighbouring unknowns
// XL = x_{-1} (previous segment's last) and XN = x_M (next segment's first).
template <typename T, int M, bool HOIST>
__device__ __forceinline__ void col_solve(T (&x)[M], ColLds<T, M> &L, const T *__restrict__ F, int p, int P,
                                          int gbase, int lane, int sl, T &XL, T &XN) {
  const int r0 = p * M;
  const T dlast = x[M - 1];
#pragma unroll
  for (int k = 1; k < M - 1; ++k) x[k] = fma(-F[kFm * LVC_COL_ROWS + r0 + k], x[k - 1], x[k]);
  x[M - 2] *= F[kFrb * LVC_COL_ROWS + r0 + M - 2];
#pragma unroll
  for (int k = M - 3; k >= 0; --k)
    x[k] = fma(-F[kFcrb * LVC_COL_ROWS + r0 + k], x[k + 1], x[k] * F[kFrb * LVC_COL_ROWS + r0 + k]);
  const int me = (gbase + p) * 64 + lane;
  L.y0[sl][me] = x[0];
  L.Dt[sl][me] = fma(-F[6 * LVC_COL_ROWS + kSaL * kColWaves + p], x[M - 2], dlast);
  // HOIST (a factor table in global memory, read by scalar loads): the
  // reduced system's factors of every segment are read here, one round trip
  // that the barrier hides, rather than one read and wait per segment inside
  // the serial sweeps below.  From LDS they are read where used (held in
  // vector registers they would spill the FP64 form).
  const T *Fs = F + 6 * LVC_COL_ROWS;
  T rcL[kColWaves], rm[kColWaves], rC[kColWaves], rrb[kColWaves];
  if constexpr (HOIST) {
#pragma unroll
    for (int q = 0; q < kColWaves; ++q) {
      rcL[q] = Fs[kScL * kColWaves + q];
      rm[q] = Fs[kSrm * kColWaves + q];
      rC[q] = Fs[kSrC * kColWaves + q];
      rrb[q] = Fs[kSrrb * kColWaves + q];
    }
  }
  auto seg = [&](const T(&h)[kColWaves], int which, int q) {
    if constexpr (HOIST)
      return h[q];
    else
      return Fs[which * kColWaves + q];
  };
  __syncthreads();
  // Reduced system of the column: Thomas over its P segments (slots
  // gbase .. gbase+P-1, all below kColWaves).  Its LDS operands are read
  // first, all at once (the slot index clamped inside the workgroup's
  // slots; values past the column's segments are not used), so that the
  // serial sweep does not wait on one LDS round trip per segment.
  T ys[kColWaves], ds[kColWaves];
#pragma unroll
  for (int q = 0; q < kColWaves; ++q) {
    ys[q] = L.y0[sl][min(gbase + q + 1, kColWaves - 1) * 64 + lane];
    ds[q] = L.Dt[sl][min(gbase + q, kColWaves - 1) * 64 + lane];
  }
  T Yr[kColWaves], y0n = T(0);
  T prev = T(0);
#pragma unroll
  for (int q = 0; q < kColWaves; ++q) {
    if (q < P) {
      const T yn = q + 1 < P ? ys[q] : T(0);
      if (q == p) y0n = yn;
      const T D = fma(-seg(rcL, kScL, q), yn, ds[q]);
      prev = q == 0 ? D : fma(-seg(rm, kSrm, q), prev, D);
      Yr[q] = prev;
    }
  }
  T Xm = T(0), Xc = T(0), Xp = T(0), nxt = T(0);
#pragma unroll
  for (int q = kColWaves - 1; q >= 0; --q) {
    if (q < P && q >= p - 1) {  // wave-uniform: the sweep stops at X_{p-1}
      nxt = fma(-seg(rC, kSrC, q), nxt, Yr[q]) * seg(rrb, kSrrb, q);
      if (q == p + 1) Xp = nxt;
      if (q == p) Xc = nxt;
      if (q == p - 1) Xm = nxt;
    }
  }
  // x_k = y_k + v_k X_{p-1} + w_k X_p; the next segment's first unknown is
  // formed by the same fma sequence, so XN is bit-identical to the value that
  // segment holds (the refinement's residual needs the exact neighbour).
#pragma unroll
  for (int k = 0; k < M - 1; ++k)
    x[k] = fma(F[kFw * LVC_COL_ROWS + r0 + k], Xc, fma(F[kFv * LVC_COL_ROWS + r0 + k], Xm, x[k]));
  x[M - 1] = Xc;
  XL = Xm;
  XN = p + 1 < P ? fma(F[kFw * LVC_COL_ROWS + r0 + M], Xp, fma(F[kFv * LVC_COL_ROWS + r0 + M], Xc, y0n)) : T(0);
}

// Segments of 32 rows: at least 4 waves per SIMD (two workgroups per CU), so
// that the 8 segments x 64 columns tiles of 256-row columns run in one round.
template <typename T, int M, bool REF>
__global__ __launch_bounds__(LVC_COL_BLOCK, M <= 32 ? 4 : 2) void lvc_ycol_kernel(
    lvc_problem *probs, const lvc_tile *__restrict__ tiles,
    const double *__restrict__ geo, const T *__restrict__ Y, T *__restrict__ R,
    const T *__restrict__ yfac) {
  __shared__ ColLds<T, M> L;
  const lvc_tile tile = tiles[blockIdx.x];
  lvc_problem &pr = probs[tile.problem];
  const int kt = pr.k;
  if (kt < 0) return;  // problem finished: whole workgroup exits
  const int nx = pr.nx, ny = pr.ny;
  const int P = tile.lanes, groups = tile.cpl;
  if (P < 1 || groups < 1 || P * groups > kColWaves) return;  // not a column tile
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int grp = wave / P, p = wave - grp * P;
  // Waves of column groups past the tile's columns leave before the first
  // barrier (a barrier waits only for the waves still running); the rest are
  // threads 0 .. gused * P * 64 - 1.
  const int gused = min(groups, (tile.count + 63) / 64);
  if (grp >= gused) return;
  const int cidx = grp * 64 + lane;
  const bool active = cidx < tile.count;
  const int col = tile.first + cidx;
  const int r0 = p * M;
  const double *g = geo + pr.geo_off;

  // The step's factor table (written by the x sweep).  Every lane of a wave
  // works on the same rows, so the solve reads it at wave-uniform addresses:
  // with the refinement straight from global memory (scalar loads into
  // SGPRs, operands of the vector arithmetic directly, which keeps VGPRs for
  // the refinement), without it from a copy in LDS in the same layout
  // (measured faster there; with the refinement the two are equal).  The copy takes one row per thread (P * M <=
  // 64 * P, no more than the threads still running) and loads from valid
  // addresses only (clamped indexes; threads past the rows write row 0
  // again), so it stays straight-line and its loads are issued before the
  // column's: the LDS writes and the barrier wait for them alone (loads
  // complete in order).
  const T *__restrict__ f = yfac + (int64_t)pr.fac_src * LVC_YFAC_LEN;
  constexpr bool kFacLds = !REF;
  __shared__ T ftab[kFacLds ? LVC_YFAC_LEN : 1];
  const int fr = threadIdx.x;
  const int fi = fr < P * M ? fr : 0, si = fr < P ? fr : 0;
  constexpr int kRowFields = kFw + 1, kSegFields = kSrC + 1;  // the fields the solve reads
  T fv[kRowFields] = {}, sv[kSegFields] = {};
  if (kFacLds) {
#pragma unroll
    for (int q = 0; q < kRowFields; ++q) fv[q] = f[q * LVC_COL_ROWS + fi];
#pragma unroll
    for (int q = 0; q < kSegFields; ++q) sv[q] = f[6 * LVC_COL_ROWS + q * kColWaves + si];
  }
  // The residual uses the same (T-rounded) taps as the explicit y term of
  // the x sweep, so that the refined solve stays consistent with it: in LDS
  // as FP64, one row per thread, ordered before their first use by the
  // solve's barrier.
  if (REF) {
    const Taps<T> tp(g, nx, ny);
    if (fr < P * M) {
      L.a64[fr] = fr < ny ? (double)tp.ay[fr] : 0.0;
      L.c64[fr] = fr < ny ? (double)tp.cy[fr] : 0.0;
    }
  }
  __builtin_amdgcn_sched_barrier(0);

  // This lane's right-hand side, by buffer loads: the column is the lane's
  // offset and each row a scalar offset (no per-row address arithmetic).
  // Rows past the grid (a last, partial segment: wave-uniform) read its last
  // row and are replaced by zeros.  Idle lanes read a valid column: their
  // (finite) values stay in their own column, which is never stored.
  const int lane_off = (active ? col : tile.first) * (int)sizeof(T);
  const int row_bytes = nx * (int)sizeof(T);
  const int64_t state_off = wave_uniform(pr.state_off);
  const __amdgpu_buffer_rsrc_t in = grid_rsrc(Y + state_off, nx, ny);
  T d[M];
#pragma unroll
  for (int k = 0; k < M; ++k)
    d[k] = buffer_load<T>(in, lane_off, min(r0 + k, ny - 1) * row_bytes);
  if (kFacLds) {
#pragma unroll
    for (int q = 0; q < kRowFields; ++q) ftab[q * LVC_COL_ROWS + fi] = fv[q];
#pragma unroll
    for (int q = 0; q < kSegFields; ++q) ftab[6 * LVC_COL_ROWS + q * kColWaves + si] = sv[q];
    // FP64 columns (twice the bytes) start after their loads have landed:
    // measured faster than overlapping the forward sweep with them.
    if constexpr (sizeof(T) == 8) __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    __syncthreads();
  }
  const T *F = kFacLds ? ftab : f;
  if (r0 + M > ny) {  // wave-uniform
#pragma unroll
    for (int k = 0; k < M; ++k)
      if (r0 + k >= ny) d[k] = T(0);
  }

  const int gbase = grp * P;
  T x[M];
#pragma unroll
  for (int k = 0; k < M; ++k) x[k] = d[k];
  T XL, XN;
  col_solve<T, M, REF>(x, L, F, p, P, gbase, lane, 0, XL, XN);
  if (REF) {
    // Residual in FP64 (T-rounded taps, as above), solved again with the same
    // factorisation; the correction is added in T (one rounding of the
    // exact sum x + e, as an FP64 addition rounded to T would give).
    const double tnow = timeline(pr.t, kt, pr.nt), tnext = timeline(pr.t, kt + 1, pr.nt);
    const double s64 = (double)(T)(1.0 / (tnext - tnow));  // dt as the x sweep rounds it
    // Every row takes the diagonal s64, also the rows past the grid (whose
    // system row is the identity): their taps and right-hand side are zero,
    // so the solve gave them +-0 (NaN only after a non-finite neighbour) and
    // the residual is the same for any positive diagonal.  No per-row test
    // (32 row masks held until the stores spilled the scalar registers).
    // One difference per pair of neighbouring rows (negated for the row
    // below, as in the x sweep's residual).
    double dlo = (double)x[0] - (double)XL;
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const double xk = (double)x[k];
      const double dhi = (double)(k < M - 1 ? x[k + 1] : XN) - xk;
      const double ax = fma(L.c64[r0 + k], dhi, fma(-L.a64[r0 + k], dlo, s64 * xk));
      d[k] = (T)((double)d[k] - ax);
      dlo = dhi;
    }
    // The second solve forms its segment masks again (scalar compares)
    // rather than holding the first one's across the residual: the empty asm
    // hides that p, P and gbase are unchanged.
    int p2 = p, P2 = P, gbase2 = gbase;
    asm volatile("" : "+s"(p2), "+s"(P2), "+s"(gbase2));
    T EL, EN;
    col_solve<T, M, REF>(d, L, F, p2, P2, gbase2, lane, 1, EL, EN);
#pragma unroll
    for (int k = 0; k < M; ++k) x[k] += d[k];
  }
  const __amdgpu_buffer_rsrc_t out = grid_rsrc(R + state_off, nx, ny);
  // With the refinement, the stores' row offsets are formed again here (a
  // scalar multiply each) rather than shared with the loads': the empty asm
  // hides that the stride is the loads' one, since 32 offsets held across
  // the two solves spilled the scalar registers.  (Without it the form
  // holds no such offsets; the asm would cost vector registers there.)
  int st_row = row_bytes, st_r0 = r0;
  if constexpr (REF) asm volatile("" : "+s"(st_row), "+s"(st_r0));
  if (active) {
    if (st_r0 + M <= ny) {  // wave-uniform
#pragma unroll
      for (int k = 0; k < M; ++k) buffer_store<T>(out, x[k], lane_off, (st_r0 + k) * st_row);
    } else {
#pragma unroll
      for (int k = 0; k < M; ++k)
        if (st_r0 + k < ny) buffer_store<T>(out, x[k], lane_off, (st_r0 + k) * st_row);
    }
  }
  // The last of the problem's workgroups to finish advances its time index
  // (every workgroup of the problem read `k` before its first barrier).
  if (threadIdx.x == 0) {
    const int done = atomicAdd(&pr.ydone, 1) + 1;
    if (done == pr.ytiles) {
      pr.ydone = 0;
      pr.k = kt - 1;
    }
  }
}

// value (:236): the price at (yindex, xindex), to the problem's output slot
// (with `list`: of problem list[i], to out[i]).
template <typename T, typename O>
__global__ void lvc_gather_kernel(const lvc_problem *probs, const int32_t *list, int n,
                                  const T *R, O *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const lvc_problem &pr = probs[list ? list[i] : i];
  const int o = list ? i : pr.out_index;
  if (pr.status != 0) {
    out[o] = (O)NAN;
    return;
  }
  out[o] = (O)R[pr.state_off + (int64_t)pr.yindex * pr.nx + pr.xindex];
}

// main's descriptors, tiles and factor leaders (lvc_main_layout): thread i
// writes descriptor i, x tile i, y tile i and leader i where they exist.
__global__ __launch_bounds__(256) void lvc_main_init_kernel(const lvc_problem tmpl,
                                                            const lvc_main_layout lay,
                                                            lvc_problem *probs, lvc_tile *tiles,
                                                            int32_t *fprobs) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nxt = (int64_t)lay.groups * lay.xtpg, nyt = (int64_t)lay.outer * lay.ytpo;
  if (i < lay.outer) {
    const int p = (int)i;
    lvc_problem q = tmpl;
    // LocVolCalib.fut:247, strike_i = 0.001 * i (in f32 for the f32 entry)
    q.strike = lay.strike_f32 ? (double)(0.001f * (float)p) : 0.001 * (double)p;
    q.geo_off = (int64_t)p * lay.geo_stride;
    q.state_off = (int64_t)p * lay.state_stride;
    q.k = -1;
    q.ydone = 0;
    q.ytiles = lay.ytpo;
    q.fac_src = (p / lay.xshare) * lay.xshare;
    q.out_index = p;
    probs[p] = q;
  }
  if (i < lay.groups) fprobs[i] = (int32_t)i * lay.xshare;
  if (i < nxt) {
    const int g = (int)(i / lay.xtpg), first = (int)(i % lay.xtpg) * lay.rows;
    const lvc_tile t = {g * lay.xshare, first, min(lay.rows, tmpl.ny - first), lay.gx,
                        min(lay.xshare, lay.outer - g * lay.xshare), 0};
    tiles[i] = t;
  }
  if (i < nyt) {
    const int p = (int)(i / lay.ytpo), first = (int)(i % lay.ytpo) * lay.cols;
    const lvc_tile t = {p, first, min(lay.cols, tmpl.nx - first), lay.gy, lay.ycpl, 0};
    tiles[nxt + i] = t;
  }
}

template <typename T, int M, bool REF>
int launch_sweeps(hipStream_t st, bool xs, const lvc_problem *probs,
                  const lvc_tile *tiles, int ntiles, const double *geo,
                  const void *in, void *out, int lds_bytes, void *yfac, int yseg,
                  const int32_t *fprobs, int nfac, bool xshared) {
  if (ntiles <= 0) return 0;
  if constexpr (M >= 32) {
    if (xs || !yfac) return (int)hipErrorInvalidValue;  // column segments: y sweep only
    lvc_ycol_kernel<T, M, REF><<<ntiles, LVC_COL_BLOCK, 0, st>>>(
        const_cast<lvc_problem *>(probs), tiles, geo, (const T *)in, (T *)out,
        (const T *)yfac);
    return (int)hipGetLastError();
  } else {
    if (xs) {
      if (yseg > 0 && (!yfac || !fprobs || !lvc_seg_is_col(yseg))) return (int)hipErrorInvalidValue;
      if (yseg <= 0 || nfac < 0) nfac = 0;
      if (xshared)
        lvc_xsweep_kernel<T, M, REF, true><<<ntiles + nfac, LVC_BLOCK, 0, st>>>(
            probs, tiles, geo, (const T *)in, (T *)out, (T *)yfac, yseg, fprobs, nfac);
      else
        lvc_xsweep_kernel<T, M, REF, false><<<ntiles + nfac, LVC_BLOCK, 0, st>>>(
            probs, tiles, geo, (const T *)in, (T *)out, (T *)yfac, yseg, fprobs, nfac);
    } else {
      if (lds_bytes <= 0 || lds_bytes > 65536) return (int)hipErrorInvalidValue;
      lvc_ysweep_kernel<T, M, REF><<<ntiles, LVC_BLOCK, lds_bytes, st>>>(
          const_cast<lvc_problem *>(probs), tiles, geo, (const T *)in, (T *)out);
    }
    return (int)hipGetLastError();
  }
}

template <typename T, bool REF>
int dispatch_seg(hipStream_t st, bool xs, int seg, const lvc_problem *probs,
                 const lvc_tile *tiles, int ntiles, const double *geo,
                 const void *in, void *out, int lds, void *yfac, int yseg,
                 const int32_t *fprobs, int nfac, bool xshared) {
#define LVC_SEG_CASE(S)                                                         \
  case S:                                                                      \
    return launch_sweeps<T, S, REF>(st, xs, probs, tiles, ntiles, geo, in, out, \
                                    lds, yfac, yseg, fprobs, nfac, xshared);
  switch (seg) {
    LVC_SEG_CASE(4)
    LVC_SEG_CASE(8)
    LVC_SEG_CASE(16)
    LVC_SEG_CASE(32)
    LVC_SEG_CASE(64)
    default: return (int)hipErrorInvalidValue;
  }
#undef LVC_SEG_CASE
}

int dispatch(hipStream_t st, bool xs, int precision, int seg,
             const lvc_problem *probs, const lvc_tile *tiles, int ntiles,
             const double *geo, const void *in, void *out, int lds, void *yfac,
             int yseg, const int32_t *fprobs, int nfac, bool xshared) {
  switch (precision) {
    case LVC_FP32:
      return dispatch_seg<float, false>(st, xs, seg, probs, tiles, ntiles, geo, in, out,
                                        lds, yfac, yseg, fprobs, nfac, xshared);
    case LVC_FP32_REFINE:
      return dispatch_seg<float, true>(st, xs, seg, probs, tiles, ntiles, geo, in, out,
                                       lds, yfac, yseg, fprobs, nfac, xshared);
    case LVC_FP64:
      return dispatch_seg<double, false>(st, xs, seg, probs, tiles, ntiles, geo, in, out,
                                         lds, yfac, yseg, fprobs, nfac, xshared);
    default: return (int)hipErrorInvalidValue;
  }
}

}  // namespace

extern "C" {

int lvc_scratch_address(int device, void **ptr) {
  // one lookup per device, kept for the life of the process: every GPU holds
  // its own copy of the array at its own address
  constexpr int kMaxDevices = 64;
  static void *addr[kMaxDevices];
  static std::mutex mu;
  *ptr = nullptr;
  if (device < 0 || device >= kMaxDevices) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> guard(mu);
  if (!addr[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return (int)e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return (int)e;
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return (int)e;
    if (!p) return (int)hipErrorInvalidSymbol;
    addr[device] = p;
  }
  *ptr = addr[device];
  return 0;
}

int lvc_launch_main_init(void *stream, const lvc_problem *tmpl,
                         const lvc_main_layout *lay, lvc_problem *probs,
                         lvc_tile *tiles, int32_t *fprobs) {
  const int64_t n = max(max((int64_t)lay->outer, (int64_t)lay->groups * lay->xtpg),
                        (int64_t)lay->outer * lay->ytpo);
  if (n <= 0) return 0;
  if (lay->xshare < 1 || lay->xtpg < 1 || lay->ytpo < 1) return (int)hipErrorInvalidValue;
  lvc_main_init_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      *tmpl, *lay, probs, tiles, fprobs);
  return (int)hipGetLastError();
}

int lvc_seg_supported(int seg, int ysweep) {
  return seg == 4 || seg == 8 || seg == 16 || (ysweep && (seg == 32 || seg == 64));
}

int lvc_launch_setup_list(void *stream, int precision, lvc_problem *probs,
                          const lvc_problem *src, const int32_t *list, int nprob,
                          double *geo, void *R) {
  if (nprob <= 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid(nprob, 8);
  if (precision == LVC_FP64)
    lvc_setup_kernel<double><<<grid, 256, 0, st>>>(probs, src, list, geo, (double *)R);
  else
    lvc_setup_kernel<float><<<grid, 256, 0, st>>>(probs, src, list, geo, (float *)R);
  return (int)hipGetLastError();
}

int lvc_launch_setup(void *stream, int precision, lvc_problem *probs, int nprob,
                     double *geo, void *R) {
  return lvc_launch_setup_list(stream, precision, probs, nullptr, nullptr, nprob, geo, R);
}

int lvc_launch_xsweep(void *stream, int precision, int seg,
                      const lvc_problem *probs, const lvc_tile *tiles,
                      int ntiles, const double *geo, const void *R, void *Y,
                      void *yfac, int yseg, const int32_t *fprobs, int nfac,
                      int xshared) {
  return dispatch((hipStream_t)stream, true, precision, seg, probs, tiles,
                  ntiles, geo, R, Y, 0, yfac, yseg, fprobs, nfac, xshared != 0);
}

int lvc_launch_ysweep(void *stream, int precision, int seg,
                      const lvc_problem *probs, const lvc_tile *tiles,
                      int ntiles, const double *geo, const void *Y, void *R,
                      int lds_bytes, const void *yfac) {
  return dispatch((hipStream_t)stream, false, precision, seg, probs, tiles,
                  ntiles, geo, Y, R, lds_bytes, const_cast<void *>(yfac), 0, nullptr, 0,
                  false);
}

int lvc_launch_gather_list(void *stream, int precision, lvc_problem *probs,
                           const int32_t *list, int nprob, const void *R, void *out,
                           int out_f64) {
  if (nprob <= 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  const int nb = (nprob + 255) / 256;
  if (precision == LVC_FP64) {
    if (out_f64)
      lvc_gather_kernel<double, double><<<nb, 256, 0, st>>>(probs, list, nprob, (const double *)R, (double *)out);
    else
      lvc_gather_kernel<double, float><<<nb, 256, 0, st>>>(probs, list, nprob, (const double *)R, (float *)out);
  } else {
    if (out_f64)
      lvc_gather_kernel<float, double><<<nb, 256, 0, st>>>(probs, list, nprob, (const float *)R, (double *)out);
    else
      lvc_gather_kernel<float, float><<<nb, 256, 0, st>>>(probs, list, nprob, (const float *)R, (float *)out);
  }
  return (int)hipGetLastError();
}

int lvc_launch_gather(void *stream, int precision, lvc_problem *probs,
                      int nprob, const void *R, void *out, int out_f64) {
  return lvc_launch_gather_list(stream, precision, probs, nullptr, nprob, R, out, out_f64);
}

}  // extern "C"
