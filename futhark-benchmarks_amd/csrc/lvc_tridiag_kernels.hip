// HIP kernels for gfx950: the batched tridiagonal solves of the ocean-sim
// test program (misc/ocean-sim/tridiag-test.fut:14-26), [n][m] f32 systems,
// one system per row.
//
//   tridag_seq_kernel  tridagSeq (misc/ocean-sim/tridiag.fut:81-95): Thomas.
//   tridag_par_kernel  tridagPar (tridiag.fut:4-79): the three recurrences
//                      in the form of the reference's scans, evaluated left to
//                      right.
//   tridag_seq_aligned_kernel, tridag_par_staged_kernel: the same two for
//                      systems of up to 128 unknowns, staged through LDS.
//
// One thread per system, in the reference's order of operations with IEEE
// division and no contraction (the file is built with -ffp-contract=off), so
// the results are the bits of a sequential f32 evaluation: these systems are
// not diagonally dominant, and a reordered (partitioned or cyclic-reduction)
// solve would meet pivots the reference's never does.  Systems of up to 128
// unknowns go through the staged kernels below, which keep the forward
// pass's per-row values in registers.  For longer ones the two kernels above
// keep them in LDS (c' or the modified diagonal, one column per lane) when
// m <= kLdsRows, otherwise in a global scratch array; y' goes to the output
// array in place.

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int kThreads = 64;   // systems per workgroup (one wavefront)
constexpr int kLdsRows = 256;  // longest system kept in LDS: 64 KiB

// Per-row storage of the forward pass: LDS column `lane` (stride kThreads,
// conflict-free) or the system's own row of a global scratch array.
struct RowStore {
  float *p;
  int stride;
  __device__ float &operator[](int i) const { return p[i * stride]; }
};

// tridiag.fut:81-95
template <bool LDS>
__global__ __launch_bounds__(kThreads) void tridag_seq_kernel(
    const float *__restrict__ a, const float *__restrict__ b,
    const float *__restrict__ c, const float *__restrict__ y,
    float *__restrict__ out, float *__restrict__ scratch, int64_t n, int m) {
  extern __shared__ float lds[];  // m * kThreads floats when LDS
  const int64_t sys = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (sys >= n) return;
  const int64_t off = sys * m;
  const float *A = a + off, *B = b + off, *C = c + off, *Yv = y + off;
  float *X = out + off;
  const RowStore cp = LDS ? RowStore{lds + threadIdx.x, kThreads} : RowStore{scratch + off, 1};
  float cprev = C[0] / B[0];
  float yprev = Yv[0] / B[0];
  cp[0] = cprev;
  X[0] = yprev;
  for (int i = 1; i < m; ++i) {
    const float ai = A[i];
    const float norm = 1.0f / (B[i] - ai * cprev);
    cprev = C[i] * norm;
    yprev = (Yv[i] - ai * yprev) * norm;
    cp[i] = cprev;
    X[i] = yprev;
  }
  float next = yprev;  // solution[m-1] = yp[m-1]
  for (int i = m - 2; i >= 0; --i) {
    next = X[i] - cp[i] * next;
    X[i] = next;
  }
}

// tridiag.fut:4-79: recurrence 1 as the scan of normalised 2x2 matrices,
// recurrences 2 and 3 as scans of linear-function composition, in the
// operator forms of the reference, accumulated left to right.
template <bool LDS>
__global__ __launch_bounds__(kThreads) void tridag_par_kernel(
    const float *__restrict__ a, const float *__restrict__ b,
    const float *__restrict__ c, const float *__restrict__ y,
    float *__restrict__ out, float *__restrict__ scratch, int64_t n, int m) {
  extern __shared__ float lds[];  // m * kThreads floats when LDS
  const int64_t sys = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (sys >= n) return;
  const int64_t off = sys * m;
  const float *A = a + off, *B = b + off, *C = c + off, *Yv = y + off;
  float *X = out + off;
  const RowStore bb = LDS ? RowStore{lds + threadIdx.x, kThreads} : RowStore{scratch + off, 1};
  // Recurrence 1 (:8-31): b'_i = (t0*b0 + t1) / (t2*b0 + t3) from the scan
  // of (b_i, -a_i*c_{i-1}, 1, 0) under the normalised product.
  const float b0 = B[0];
  float t0 = 1.0f, t1 = 0.0f, t2 = 0.0f, t3 = 1.0f;  // scan of the identity
  bb[0] = (t0 * b0 + t1) / (t2 * b0 + t3);
  for (int i = 1; i < m; ++i) {
    const float m0 = B[i], m1 = 0.0f - A[i] * C[i - 1], m2 = 1.0f, m3 = 0.0f;
    const float value = 1.0f / (t0 * m0);
    const float n0 = (m0 * t0 + m1 * t2) * value;
    const float n1 = (m0 * t1 + m1 * t3) * value;
    const float n2 = (m2 * t0 + m3 * t2) * value;
    const float n3 = (m2 * t1 + m3 * t3) * value;
    t0 = n0;
    t1 = n1;
    t2 = n2;
    t3 = n3;
    bb[i] = (t0 * b0 + t1) / (t2 * b0 + t3);
  }
  // Recurrence 2 (:36-53): y'_i = f0 + f1*y0 from the scan of
  // (y_i, -a_i/b'_{i-1}) under (b0 + b1*a0, a1*b1).
  const float y0 = Yv[0];
  float f0 = 0.0f, f1 = 1.0f;
  X[0] = f0 + f1 * y0;
  for (int i = 1; i < m; ++i) {
    const float l0 = Yv[i], l1 = 0.0f - A[i] / bb[i - 1];
    const float g0 = l0 + l1 * f0, g1 = f1 * l1;
    f0 = g0;
    f1 = g1;
    X[i] = f0 + f1 * y0;
  }
  // Recurrence 3 (:57-78): from yn = y'_{m-1}/b'_{m-1}, the reversed scan of
  // (y'_i/b'_i, -c_i/b'_i); the final reverse is the store index.
  const float yn = X[m - 1] / bb[m - 1];
  f0 = 0.0f;
  f1 = 1.0f;
  X[m - 1] = f0 + f1 * yn;
  for (int k = 1; k < m; ++k) {
    const int i = m - k - 1;
    const float bi = bb[i];
    const float l0 = X[i] / bi, l1 = 0.0f - C[i] / bi;
    const float g0 = l0 + l1 * f0, g1 = f1 * l1;
    f0 = g0;
    f1 = g1;
    X[i] = f0 + f1 * yn;
  }
}

// Staged kernels for systems of up to kStagedRows unknowns: one wavefront
// per 64 systems.  Chunks of CH columns of a, b, c, y are loaded into LDS
// with coalesced row reads (each system's row is contiguous: CH * 4 bytes of
// a row per chunk, 16 bytes a lane) and read back one system per lane.  The
// forward pass's two per-row values for the backward pass (c' and y' of
// Thomas, l1 and l0 of the scan form's third recurrence) stay in this lane's
// registers: every loop over rows is unrolled to its full length, so each
// row has a register of its own, and with one wavefront per SIMD (launches
// of these kernels never hold more) they fit in the unified register file.
// The backward pass therefore reads nothing from memory; its solution goes
// out through the same LDS transpose, coalesced.  The arithmetic and its
// order per system are those of the kernels above (recurrences 1-3 of the
// scan form advance together: each step uses only earlier steps' values, so
// interleaving them changes no result).
//
// The scan form (tridag_par_staged_kernel) takes chunks of 16 columns
// aligned to the systems' columns: with its more numerous carries and five
// divisions a row it is bound by its division chain rather than its reads,
// and 16 measured best (32, or 32-column chunks aligned to the arrays'
// lines as Thomas takes them, were slower).
// Thomas (tridag_seq_aligned_kernel, below) takes 32-column chunks aligned
// to the arrays' 128-byte lines instead.
constexpr int kStagedRows = 128;
constexpr int kVec = 4;  // floats per global load and store of the staging
// dword-aligned float4 (a row of m floats starts at any multiple of 4 bytes)
typedef float float4u __attribute__((ext_vector_type(kVec), aligned(4)));

// A chunk of one array, CH columns of the 64 rows, is moved in kQ
// instructions of 16 bytes a lane: kLanesPerRow lanes cover one row's CH
// columns, kRowStep rows per instruction.  Whole instructions, not bytes,
// are what bound a single wavefront here: gfx9 tracks at most 63 vector
// memory instructions in flight per wave, so 4-float moves quarter the
// round trips of 1-float ones.
template <int CH>
struct Stage {
  static constexpr int kLanesPerRow = CH / kVec;
  static constexpr int kRowStep = kThreads / kLanesPerRow;
  static constexpr int kQ = kThreads / kRowStep;
  float v[4][kThreads][CH + 1];  // a, b, c, y; the solution through v[3]
};

// The same rows and columns loaded into registers (this lane's slots), to be
// written to LDS later: the next chunk's loads fly while this one computes.
// Slots outside the systems (rows >= rows, columns >= m) hold 0.  A chunk
// wholly inside them (every chunk of a full workgroup but a ragged last one)
// takes unguarded 16-byte loads; the test is made once for the four arrays,
// so that their loads issue back to back.
template <int CH>
using Chunk = float[Stage<CH>::kQ][kVec];

template <int CH>
__device__ __forceinline__ void fetch(Chunk<CH> &pa, Chunk<CH> &pb, Chunk<CH> &pc,
                                      Chunk<CH> &py, const float *__restrict__ a,
                                      const float *__restrict__ b, const float *__restrict__ c,
                                      const float *__restrict__ y, int64_t s0, int rows, int m,
                                      int i0, int lane) {
  using St = Stage<CH>;
  const int sr = lane / St::kLanesPerRow, sc = (lane % St::kLanesPerRow) * kVec;
  const int i = i0 + sc;
  const float *__restrict__ src[4] = {a, b, c, y};
  float(*const dst[4])[kVec] = {pa, pb, pc, py};
  if (rows == kThreads && i0 + CH <= m) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int q = 0; q < St::kQ; ++q) {
        const float4u v = *(const float4u *)(src[k] + (s0 + q * St::kRowStep + sr) * m + i);
        dst[k][q][0] = v.x;
        dst[k][q][1] = v.y;
        dst[k][q][2] = v.z;
        dst[k][q][3] = v.w;
      }
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int q = 0; q < St::kQ; ++q) {
        const int r = q * St::kRowStep + sr;
#pragma unroll
        for (int e = 0; e < kVec; ++e)
          dst[k][q][e] = (r < rows && i + e < m) ? src[k][(s0 + r) * m + i + e] : 0.0f;
      }
  }
}

template <int CH>
__device__ __forceinline__ void deposit(float (&st)[kThreads][CH + 1], const Chunk<CH> &reg,
                                        int lane) {
  using St = Stage<CH>;
  const int sr = lane / St::kLanesPerRow, sc = (lane % St::kLanesPerRow) * kVec;
#pragma unroll
  for (int q = 0; q < St::kQ; ++q)
#pragma unroll
    for (int e = 0; e < kVec; ++e) st[q * St::kRowStep + sr][sc + e] = reg[q][e];
}

template <int CH>
__device__ __forceinline__ void stage_out(const float (&st)[kThreads][CH + 1],
                                          float *__restrict__ dst, int64_t s0, int rows,
                                          int m, int i0, int lane) {
  using St = Stage<CH>;
  const int sr = lane / St::kLanesPerRow, sc = (lane % St::kLanesPerRow) * kVec;
  const int i = i0 + sc;
  float v[St::kQ][kVec];
#pragma unroll
  for (int q = 0; q < St::kQ; ++q)
#pragma unroll
    for (int e = 0; e < kVec; ++e) v[q][e] = st[q * St::kRowStep + sr][sc + e];
  if (rows == kThreads && i0 + CH <= m) {
#pragma unroll
    for (int q = 0; q < St::kQ; ++q) {
      float4u w;
      w.x = v[q][0];
      w.y = v[q][1];
      w.z = v[q][2];
      w.w = v[q][3];
      *(float4u *)(dst + (s0 + q * St::kRowStep + sr) * m + i) = w;
    }
  } else {
#pragma unroll
    for (int q = 0; q < St::kQ; ++q) {
      const int r = q * St::kRowStep + sr;
#pragma unroll
      for (int e = 0; e < kVec; ++e)
        if (r < rows && i + e < m) dst[(s0 + r) * m + i + e] = v[q][e];
    }
  }
}

template <int CH>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, 1)))
void tridag_par_staged_kernel(const float *__restrict__ a, const float *__restrict__ b,
                              const float *__restrict__ c, const float *__restrict__ y,
                              float *__restrict__ out, int64_t n, int m) {
  constexpr int kChunks = kStagedRows / CH;
  __shared__ Stage<CH> S;
  const int lane = threadIdx.x;
  const int64_t s0 = (int64_t)blockIdx.x * kThreads;
  const int rows = (int)min((int64_t)kThreads, n - s0);
  const int nch = (m + CH - 1) / CH;
  float fw0[kStagedRows], fw1[kStagedRows];  // recurrence 3's l0 and l1 per row
  float cprev = 0.0f;  // c_{i-1}
  float t0 = 1.0f, t1 = 0.0f, t2 = 0.0f, t3 = 1.0f, b0 = 0.0f, bprev = 0.0f;
  float f0 = 0.0f, f1 = 1.0f, y0 = 0.0f;
  float yn = 0.0f;  // l0 of row m-1, recurrence 3's start

  // The row loops have no per-row branch, so that the scheduler can overlap
  // one row's LDS reads with the previous row's arithmetic: rows m..(the end
  // of the last chunk) run on the zeros that fetch() leaves there, and what
  // they produce is never stored nor carried into rows below m (the forward
  // pass ends with them; the backward pass restarts at row m-1 by select).
  Chunk<CH> pa, pb, pc, py;  // the next chunk, in flight
  fetch<CH>(pa, pb, pc, py, a, b, c, y, s0, rows, m, 0, lane);
#pragma unroll
  for (int ch = 0; ch < kChunks; ++ch) {
    if (ch >= nch) break;
    const int i0 = ch * CH;
    deposit<CH>(S.v[0], pa, lane);
    deposit<CH>(S.v[1], pb, lane);
    deposit<CH>(S.v[2], pc, lane);
    deposit<CH>(S.v[3], py, lane);
    __syncthreads();
    if (ch + 1 < nch) {
      fetch<CH>(pa, pb, pc, py, a, b, c, y, s0, rows, m, i0 + CH, lane);
    }
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int i = i0 + j;
      const float ai = S.v[0][lane][j], bi = S.v[1][lane][j];
      const float ci = S.v[2][lane][j], yi = S.v[3][lane][j];
      // recurrence 1 (c_{i-1} carried in cprev)
      if (i == 0) {
        b0 = bi;
      } else {
        const float m0 = bi, m1 = 0.0f - ai * cprev, m2 = 1.0f, m3 = 0.0f;
        const float value = 1.0f / (t0 * m0);
        const float n0 = (m0 * t0 + m1 * t2) * value;
        const float n1 = (m0 * t1 + m1 * t3) * value;
        const float n2 = (m2 * t0 + m3 * t2) * value;
        const float n3 = (m2 * t1 + m3 * t3) * value;
        t0 = n0;
        t1 = n1;
        t2 = n2;
        t3 = n3;
      }
      const float bb = (t0 * b0 + t1) / (t2 * b0 + t3);
      // recurrence 2
      float yy;
      if (i == 0) {
        y0 = yi;
        yy = f0 + f1 * y0;
      } else {
        const float l0 = yi, l1 = 0.0f - ai / bprev;
        const float g0 = l0 + l1 * f0, g1 = f1 * l1;
        f0 = g0;
        f1 = g1;
        yy = f0 + f1 * y0;
      }
      // this row's element of recurrence 3
      fw0[i] = yy / bb;
      fw1[i] = 0.0f - ci / bb;
      bprev = bb;
      cprev = ci;
      yn = i < m ? fw0[i] : yn;  // ends as y'_{m-1}/b'_{m-1}
    }
    __syncthreads();  // every lane has read the chunk before the next deposit
  }

  // backward pass (recurrence 3) over the chunks in reverse, from registers
  float g0 = 0.0f, g1 = 1.0f;
#pragma unroll
  for (int ch = kChunks - 1; ch >= 0; --ch) {
    if (ch < nch) {
      const int i0 = ch * CH;
#pragma unroll
      for (int j = CH - 1; j >= 0; --j) {
        const int i = i0 + j;
        // recurrence 3 restarts at row m-1 from (g0, g1) = (0, 1)
        const float h0 = fw0[i] + fw1[i] * g0, h1 = g1 * fw1[i];
        const bool below = i < m - 1;
        g0 = below ? h0 : 0.0f;
        g1 = below ? h1 : 1.0f;
        S.v[3][lane][j] = g0 + g1 * yn;
      }
      __syncthreads();
      stage_out<CH>(S.v[3], out, s0, rows, m, i0, lane);
      __syncthreads();
    }
  }
}

// Thomas, staged with line-aligned chunks.  Row r of the workgroup starts
// at float g_r = (s0 + r) * m of each array, sh_r = g_r mod CH floats into
// an aligned CH-float unit (a 128-byte line at CH = 32); chunk k of row r is
// that row's aligned unit k, floats [g_r - sh_r + k*CH, +CH), so every load
// is a whole unit and each unit is read once per row (a unit that holds the
// end of one row and the start of the next is read by both).  In chunk k,
// step j is slot s = k*CH + j, which is column s - sh_r of row r: columns
// differ between lanes, slots do not, so the per-row registers are indexed by
// slot.  Slots outside [sh_r, sh_r + m) compute on the neighbours' values (or
// zeros) and are discarded: the forward carries restart at column 0 (slot
// sh_r, in chunk 0) and the backward pass at column m-1.  The solution goes
// out by columns, as in the scan form's kernel (16-byte stores of whole
// column chunks, each row's own columns only).  Against chunks aligned to
// the columns, each row's line is fetched once rather than twice (a 128-byte
// chunk at any offset spans two lines, and the L2 does not keep the second
// until the row's next chunk asks for it), at the price of one more chunk
// per system (160 slots for 128 unknowns).
template <int CH>
constexpr int aligned_chunks() { return (kStagedRows + 2 * CH - 2) / CH; }

template <int CH>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, 1)))
void tridag_seq_aligned_kernel(const float *__restrict__ a, const float *__restrict__ b,
                               const float *__restrict__ c, const float *__restrict__ y,
                               float *__restrict__ out, int64_t n, int m) {
  using St = Stage<CH>;
  constexpr int kChunks = aligned_chunks<CH>();
  constexpr int kSlots = kChunks * CH;
  __shared__ St S;
  const int lane = threadIdx.x;
  const int64_t s0 = (int64_t)blockIdx.x * kThreads;
  const int rows = (int)min((int64_t)kThreads, n - s0);
  const int nch = (m + 2 * CH - 2) / CH;  // chunks of the row that ends latest
  // 16-byte pieces sit at global indices that are multiples of 4 (units are
  // CH-aligned), so when n*m is one too a piece is wholly inside the arrays
  // or wholly past their end; pieces past the end only feed slots past the
  // last row's end, which are discarded, and load the arrays' last piece
  // instead.  Otherwise (and for a partial workgroup) guarded loads.
  const bool whole = rows == kThreads && (n * m) % kVec == 0;
  const int tail = (int)min(n * m - s0 * m, (int64_t)kThreads * m + 2 * CH);  // floats from src to the end, capped
  // offsets from the workgroup's first row, 32-bit (64 rows of <= 128)
  const int base_mod = (int)((s0 * m) % CH);
  // this lane's system (compute) and its slots
  const int sh = (base_mod + lane * m) % CH;
  const int last = sh + m - 1;
  // this lane's staging row and unit offset (loads, stores)
  const int sr = lane / St::kLanesPerRow, sc = (lane % St::kLanesPerRow) * kVec;
  // start of row q*kRowStep + sr, and its shift into its first unit
  auto row_start = [&](int q) { return (q * St::kRowStep + sr) * m; };
  auto row_shift = [&](int q) { return (base_mod + (q * St::kRowStep + sr) * m) % CH; };

  float fw0[kSlots], fw1[kSlots];  // y' and c' by slot
  float cprev = 0.0f, yprev = 0.0f;

  const float *__restrict__ src[4] = {a + s0 * m, b + s0 * m, c + s0 * m, y + s0 * m};
  float *__restrict__ dst = out + s0 * m;
  Chunk<CH> pf[4];  // the next chunk, in flight
  auto fetch_chunk = [&](int k) {
    if (whole) {
#pragma unroll
      for (int q = 0; q < St::kQ; ++q) {
        // may be negative for row 0 (its unit starts before the workgroup's
        // first row), never before the array
        const int u = min(row_start(q) - row_shift(q) + k * CH + sc, tail - kVec);
#pragma unroll
        for (int arr = 0; arr < 4; ++arr) {
          const float4u v = *(const float4u *)(src[arr] + u);
          pf[arr][q][0] = v.x;
          pf[arr][q][1] = v.y;
          pf[arr][q][2] = v.z;
          pf[arr][q][3] = v.w;
        }
      }
    } else {
      // the last workgroup: rows past n hold zeros; units stop at the end of
      // the arrays (floats from src to there: tail), elementwise only in the
      // one 16-byte piece that crosses it
#pragma unroll
      for (int q = 0; q < St::kQ; ++q) {
        const int r = q * St::kRowStep + sr;
        const int u = row_start(q) - row_shift(q) + k * CH + sc;
#pragma unroll
        for (int arr = 0; arr < 4; ++arr) {
          float4u v = {0.0f, 0.0f, 0.0f, 0.0f};
          if (r < rows && u + kVec <= tail) {
            v = *(const float4u *)(src[arr] + u);
          } else if (r < rows) {
#pragma unroll
            for (int e = 0; e < kVec; ++e)
              if (u + e < tail) v[e] = src[arr][u + e];
          }
          pf[arr][q][0] = v.x;
          pf[arr][q][1] = v.y;
          pf[arr][q][2] = v.z;
          pf[arr][q][3] = v.w;
        }
      }
    }
  };

  fetch_chunk(0);
#pragma unroll
  for (int k = 0; k < kChunks; ++k) {
    if (k >= nch) break;
#pragma unroll
    for (int arr = 0; arr < 4; ++arr) deposit<CH>(S.v[arr], pf[arr], lane);
    __syncthreads();
    if (k + 1 < nch) fetch_chunk(k + 1);
    // Thomas' first row (tridiag.fut:82-83), c_0/b_0 and y_0/b_0, from this
    // lane's column 0 (slot sh, in chunk 0), divided once here rather than
    // at every step of the chunk
    float c_first = 0.0f, y_first = 0.0f;
    if (k == 0) {
      const float b_0 = S.v[1][lane][sh];
      c_first = S.v[2][lane][sh] / b_0;
      y_first = S.v[3][lane][sh] / b_0;
    }
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int s = k * CH + j;
      const float ai = S.v[0][lane][j], bi = S.v[1][lane][j];
      const float ci = S.v[2][lane][j], yi = S.v[3][lane][j];
      // column 0 falls in chunk 0 (sh < CH); before it the carries hold
      // whatever the slots before produced, and they restart there (:82-88)
      const float norm = 1.0f / (bi - ai * cprev);
      float cn = ci * norm, yv = (yi - ai * yprev) * norm;
      if (k == 0) {
        const bool first = j == sh;
        cn = first ? c_first : cn;
        yv = first ? y_first : yv;
      }
      cprev = cn;
      yprev = yv;
      fw1[s] = cprev;
      fw0[s] = yprev;
    }
    __syncthreads();  // every lane has read the chunk before the next deposit
  }

  // backward pass over the chunks in reverse, from registers.  Slot s's
  // solution goes to an LDS ring of two chunks per row (the staging buffer,
  // free by now); once slot chunk k is done, every row's columns
  // [k*CH, k*CH + CH) are (their slots lie in chunks k and k+1), and they go
  // out as the scan form kernel's column chunks.
  float(*ring)[2 * CH + 1] = reinterpret_cast<float(*)[2 * CH + 1]>(&S.v[0][0][0]);
  static_assert(sizeof(float) * kThreads * (2 * CH + 1) <= sizeof(St), "ring fits the stage");
  const int ncol = (m + CH - 1) / CH;  // column chunks
  float xnext = 0.0f;
#pragma unroll
  for (int k = kChunks - 1; k >= 0; --k) {
    if (k < nch) {
#pragma unroll
      for (int j = CH - 1; j >= 0; --j) {
        const int s = k * CH + j;
        // x_{m-1} = y'_{m-1}; below it x_i = y'_i - c'_i x_{i+1} (:90-95)
        float x = fw0[s] - fw1[s] * xnext;
        x = s == last ? fw0[s] : x;
        xnext = x;
        ring[lane][s % (2 * CH)] = x;
      }
      __syncthreads();
      if (k < ncol) {
        const int i0 = k * CH, i = i0 + sc;
        float v[St::kQ][kVec];
#pragma unroll
        for (int q = 0; q < St::kQ; ++q)
#pragma unroll
          for (int e = 0; e < kVec; ++e)
            v[q][e] = ring[q * St::kRowStep + sr][(i + e + row_shift(q)) % (2 * CH)];
        if (rows == kThreads && i0 + CH <= m) {
#pragma unroll
          for (int q = 0; q < St::kQ; ++q) {
            float4u w;
            w.x = v[q][0];
            w.y = v[q][1];
            w.z = v[q][2];
            w.w = v[q][3];
            *(float4u *)(dst + (row_start(q) + i)) = w;
          }
        } else {
#pragma unroll
          for (int q = 0; q < St::kQ; ++q) {
            const int r = q * St::kRowStep + sr;
#pragma unroll
            for (int e = 0; e < kVec; ++e)
              if (r < rows && i + e < m) dst[row_start(q) + i + e] = v[q][e];
          }
        }
      }
      __syncthreads();
    }
  }
}

}  // namespace

extern "C" {

// form: 0 = tridagPar (scan form), 1 = tridagSeq (Thomas).  scratch: n*m
// floats, used when m > lvc_tridag_lds_rows() (the forward pass's per-row
// values); may be NULL otherwise.
int lvc_launch_tridag(void *stream, int form, const float *a, const float *b,
                      const float *c, const float *y, float *out, float *scratch,
                      int64_t n, int m) {
  if (n <= 0) return 0;
  if (m <= 0) return (int)hipErrorInvalidValue;
  hipStream_t st = (hipStream_t)stream;
  const int64_t blocks = (n + kThreads - 1) / kThreads;
  if (blocks > 0x7fffffff) return (int)hipErrorInvalidValue;
  if (form != 0 && form != 1) return (int)hipErrorInvalidValue;
  if (m <= kStagedRows) {
    if (form == 1)
      tridag_seq_aligned_kernel<32><<<(unsigned)blocks, kThreads, 0, st>>>(a, b, c, y, out, n, m);
    else
      tridag_par_staged_kernel<16><<<(unsigned)blocks, kThreads, 0, st>>>(a, b, c, y, out, n, m);
    return (int)hipGetLastError();
  }
  const bool lds = m <= kLdsRows;
  if (!lds && !scratch) return (int)hipErrorInvalidValue;
  const size_t shm = lds ? (size_t)m * kThreads * sizeof(float) : 0;
  if (form == 1) {
    if (lds)
      tridag_seq_kernel<true><<<(unsigned)blocks, kThreads, shm, st>>>(a, b, c, y, out, scratch, n, m);
    else
      tridag_seq_kernel<false><<<(unsigned)blocks, kThreads, 0, st>>>(a, b, c, y, out, scratch, n, m);
  } else if (form == 0) {
    if (lds)
      tridag_par_kernel<true><<<(unsigned)blocks, kThreads, shm, st>>>(a, b, c, y, out, scratch, n, m);
    else
      tridag_par_kernel<false><<<(unsigned)blocks, kThreads, 0, st>>>(a, b, c, y, out, scratch, n, m);
  } else {
    return (int)hipErrorInvalidValue;
  }
  return (int)hipGetLastError();
}

int lvc_tridag_lds_rows(void) { return kLdsRows; }
int lvc_tridag_staged_rows(void) { return kStagedRows; }

}  // extern "C"
