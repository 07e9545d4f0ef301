// HIP kernels for gfx950: the batched tridiagonal solves of the ocean-sim
// test program (misc/ocean-sim/tridiag-test.fut:14-26), [n][m] f32 systems,
// one system per row.
//
//   tridag_seq_kernel  tridagSeq (misc/ocean-sim/tridiag.fut:81-95): Thomas.
//   tridag_par_kernel  tridagPar (tridiag.fut:4-79): the three recurrences
//                      in the form of the reference's scans, evaluated left to
//                      right.
//
// One thread per system (systems of up to 128 unknowns through the staged
// kernel below), in the reference's order of operations with IEEE
// division and no contraction (the file is built with -ffp-contract=off), so
// the results are the bits of a sequential f32 evaluation: these systems are
// not diagonally dominant, and a reordered (partitioned or cyclic-reduction)
// solve would meet pivots the reference's never does.  The forward pass keeps
// its per-row values for the backward pass in LDS (c' or the modified
// diagonal, one column per lane) when m <= kLdsRows, otherwise in a global
// scratch array; y' goes to the output array in place.

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

// Staged variant for systems of up to kStagedRows unknowns: one wavefront per
// 64 systems.  Chunks of kChunk columns of a, b, c, y are loaded into LDS
// with coalesced row reads (each system's row is contiguous; kChunk * 4 =
// 128 bytes per system row) and read back one system per lane.  The forward
// pass's per-row values for the backward pass go back through the same
// transpose to global memory: y' (or l0) into the output array in place, c'
// (or the scan form's l1) into a scratch array of the same shape, so that
// the workgroup's LDS is the staging buffer alone and every workgroup of a
// launch is resident at once.  The backward pass reads both back chunk by
// chunk in reverse.  The arithmetic and its order per system are those of
// the kernels above (recurrences 1-3 of the scan form advance together: each
// step uses only earlier steps' values, so interleaving them changes no
// result).
constexpr int kStagedRows = 128;
constexpr int kChunk = 32;
constexpr int kRowStep = kThreads / kChunk;  // rows per staging instruction
constexpr int kQ = kThreads / kRowStep;      // staging instructions per array

struct Stage {
  float v[4][kThreads][kChunk + 1];  // a, b, c, y; after use c' (l1) and y' (l0)
};

// The same rows and columns loaded into registers (this lane's slots), to be
// written to LDS later: the next chunk's loads fly while this one computes.
__device__ __forceinline__ void fetch(float (&reg)[kQ], const float *__restrict__ src,
                                      int64_t s0, int rows, int m, int i0, int lane) {
  const int sr = lane / kChunk, sc = lane % kChunk, i = i0 + sc;
#pragma unroll
  for (int q = 0; q < kQ; ++q) {
    const int r = q * kRowStep + sr;
    reg[q] = (r < rows && i < m) ? src[(s0 + r) * m + i] : 0.0f;
  }
}

__device__ __forceinline__ void deposit(float (&st)[kThreads][kChunk + 1],
                                        const float (&reg)[kQ], int lane) {
  const int sr = lane / kChunk, sc = lane % kChunk;
#pragma unroll
  for (int q = 0; q < kQ; ++q) st[q * kRowStep + sr][sc] = reg[q];
}

__device__ __forceinline__ void stage_out(const float (&st)[kThreads][kChunk + 1],
                                          float *__restrict__ dst, int64_t s0, int rows,
                                          int m, int i0, int lane) {
  const int sr = lane / kChunk, sc = lane % kChunk, i = i0 + sc;
#pragma unroll
  for (int q = 0; q < kQ; ++q) {
    const int r = q * kRowStep + sr;
    if (r < rows && i < m) dst[(s0 + r) * m + i] = st[r][sc];
  }
}

template <bool SEQ>
__global__ __launch_bounds__(kThreads) void tridag_staged_kernel(
    const float *__restrict__ a, const float *__restrict__ b,
    const float *__restrict__ c, const float *__restrict__ y,
    float *__restrict__ out, float *__restrict__ cw, int64_t n, int m) {
  __shared__ Stage S;
  const int lane = threadIdx.x;
  const int64_t s0 = (int64_t)blockIdx.x * kThreads;
  const int rows = (int)min((int64_t)kThreads, n - s0);
  const int nch = (m + kChunk - 1) / kChunk;
  // Thomas carries
  float cprev = 0.0f, yprev = 0.0f;
  // scan-form carries
  float t0 = 1.0f, t1 = 0.0f, t2 = 0.0f, t3 = 1.0f, b0 = 0.0f, bprev = 0.0f;
  float f0 = 0.0f, f1 = 1.0f, y0 = 0.0f, last_bb = 1.0f, last_yy = 0.0f;

  float pa[kQ], pb[kQ], pc[kQ], py[kQ];  // the next chunk, in flight
  fetch(pa, a, s0, rows, m, 0, lane);
  fetch(pb, b, s0, rows, m, 0, lane);
  fetch(pc, c, s0, rows, m, 0, lane);
  fetch(py, y, s0, rows, m, 0, lane);
  for (int ch = 0; ch < nch; ++ch) {
    const int i0 = ch * kChunk;
    deposit(S.v[0], pa, lane);
    deposit(S.v[1], pb, lane);
    deposit(S.v[2], pc, lane);
    deposit(S.v[3], py, lane);
    __syncthreads();
    if (ch + 1 < nch) {
      fetch(pa, a, s0, rows, m, i0 + kChunk, lane);
      fetch(pb, b, s0, rows, m, i0 + kChunk, lane);
      fetch(pc, c, s0, rows, m, i0 + kChunk, lane);
      fetch(py, y, s0, rows, m, i0 + kChunk, lane);
    }
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      const int i = i0 + j;
      if (i < m) {
        const float ai = S.v[0][lane][j], bi = S.v[1][lane][j];
        const float ci = S.v[2][lane][j], yi = S.v[3][lane][j];
        if (SEQ) {
          if (i == 0) {
            cprev = ci / bi;
            yprev = yi / bi;
          } else {
            const float norm = 1.0f / (bi - ai * cprev);
            cprev = ci * norm;
            yprev = (yi - ai * yprev) * norm;
          }
          S.v[2][lane][j] = cprev;
          S.v[3][lane][j] = yprev;
        } else {
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
          S.v[3][lane][j] = yy / bb;
          S.v[2][lane][j] = 0.0f - ci / bb;
          bprev = bb;
          cprev = ci;
          last_bb = bb;
          last_yy = yy;
        }
      }
    }
    __syncthreads();
    stage_out(S.v[2], cw, s0, rows, m, i0, lane);
    stage_out(S.v[3], out, s0, rows, m, i0, lane);
    __syncthreads();
  }

  // backward pass over the chunks in reverse; the solution replaces y' / l0
  float xnext = yprev;                                  // Thomas: x_{m-1} = y'_{m-1}
  const float yn = SEQ ? 0.0f : last_yy / last_bb;      // scan form
  float g0 = 0.0f, g1 = 1.0f;
  fetch(pc, cw, s0, rows, m, (nch - 1) * kChunk, lane);
  fetch(py, out, s0, rows, m, (nch - 1) * kChunk, lane);
  for (int ch = nch - 1; ch >= 0; --ch) {
    const int i0 = ch * kChunk;
    deposit(S.v[2], pc, lane);
    deposit(S.v[3], py, lane);
    __syncthreads();
    if (ch > 0) {
      fetch(pc, cw, s0, rows, m, i0 - kChunk, lane);
      fetch(py, out, s0, rows, m, i0 - kChunk, lane);
    }
#pragma unroll
    for (int j = kChunk - 1; j >= 0; --j) {
      const int i = i0 + j;
      if (i < m) {
        const float fwi = S.v[2][lane][j];
        if (i == m - 1) {
          S.v[3][lane][j] = SEQ ? yprev : 0.0f + 1.0f * yn;
        } else if (SEQ) {
          xnext = S.v[3][lane][j] - fwi * xnext;
          S.v[3][lane][j] = xnext;
        } else {
          const float h0 = S.v[3][lane][j] + fwi * g0, h1 = g1 * fwi;
          g0 = h0;
          g1 = h1;
          S.v[3][lane][j] = g0 + g1 * yn;
        }
      }
    }
    __syncthreads();
    stage_out(S.v[3], out, s0, rows, m, i0, lane);
    __syncthreads();
  }
}

}  // namespace

extern "C" {

// form: 0 = tridagPar (scan form), 1 = tridagSeq (Thomas).  scratch: n*m
// floats, used when m <= lvc_tridag_staged_rows() (the staged kernel's c')
// or m > lvc_tridag_lds_rows(); may be NULL otherwise.
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
    if (!scratch) return (int)hipErrorInvalidValue;
    if (form == 1)
      tridag_staged_kernel<true><<<(unsigned)blocks, kThreads, 0, st>>>(a, b, c, y, out, scratch, n, m);
    else
      tridag_staged_kernel<false><<<(unsigned)blocks, kThreads, 0, st>>>(a, b, c, y, out, scratch, n, m);
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
