// Batched Householder QR of many independent column-major L x k matrices
// (k <= L <= kQrRows), panel by panel over all of them at once, and the
// application of their Q factors: the preconditioner of the batched one-sided Jacobi
// (tn_batched.hip).  One-sided Jacobi on the rows of R instead of the
// columns of X = Q R converges in far fewer sweeps on graded and
// rank-deficient matrices (Drmac & Veselic, "New fast and accurate Jacobi
// SVD algorithm I", SIAM J. Matrix Anal. Appl. 29 (2008)): the k - rank
// rows of R that carry only rounding are negligible from the start.
//
// Per panel of QNB = 16 columns (LAPACK dgeqrf's steps with dlarfg's
// reflector convention, beta = -sign(alpha) ||x||, and dlarft's forward
// column-wise T), every job's panel in the same two launches:
//   * bqr_panel, one workgroup per job: the panel's rows are held in
//     registers, 4 rows per thread (rows tid, tid + 256, ...), and each
//     reflector costs one block reduction of 16 dot products (the update of
//     the panel's other columns and the new column of T from the same dots);
//   * bqr_trail, one workgroup per 64 trailing columns of a job: C -= V T^T
//     V^T C by v_mfma_f64_16x16x4, each wave a 16-column tile, V read back
//     from the stored panel through LDS.  Few registers, so many waves per
//     CU hide the loads' latency.
#pragma once
#include "tn_common.h"

#include <utility>

namespace tn {

constexpr int QNB = 16;         // reflectors per panel (T is QNB x QNB)
constexpr int kQrRows = 1024;   // rows a workgroup holds in registers (4 per thread)

typedef double f64x4q __attribute__((ext_vector_type(4)));

struct QJob {
  int64_t x;    // the job's L x k matrix (column-major, ld = L): on exit R on and
                // above the diagonal, the reflectors below it
  int64_t r;    // the rows of R as k vectors of length k
  int64_t qz;   // Q [Z; 0]: L x ncols, column-major
  int64_t tf;   // the panels' T factors, QNB x QNB row-major each
  int64_t z;    // the job's rotation accumulator Z (k x k, columns contiguous)
  int32_t L, k;
  int32_t job;  // index of the job in the batch
  int32_t pad;
};

// v[0..N) summed over the block's 256 threads; every thread gets the sums
// (bit-identical in all of them): each wave's values by DPP, the four waves'
// partials one per lane from LDS, summed by two shuffles and read back by
// readlane.
template <int N>
__device__ __forceinline__ void qr_block_sum(double (&v)[N], double (*red)[QNB]) {
  static_assert(4 * N <= 64, "one partial per lane");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < N; ++c) v[c] = wave_sum_dpp(v[c]);
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < N; ++c) red[wave][c] = v[c];
  }
  __syncthreads();
  double p = lane < 4 * N ? red[lane / N][lane % N] : 0.0;
#pragma unroll
  for (int sh = N; sh < 4 * N; sh <<= 1) p += __shfl_xor(p, sh, 64);
  const long long u = __double_as_longlong(p);
#pragma unroll
  for (int c = 0; c < N; ++c) {
    const int lo = __builtin_amdgcn_readlane((int)u, c);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), c);
    v[c] = __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
  }
  __syncthreads();
}

// C[p0:L, c0:c1) -= V op(T) V^T C[p0:L, c0:c1) for c1 - c0 <= 64 (column-
// major C, ld = ldc; called by the whole block: wave w takes the 16-column
// tile at c0 + 16 w), V the nb reflectors of the panel at column p0 of X
// (ld = L; unit diagonal, zero above it), op(T) = T^T if trans (applying
// Q^T) else T.  V is staged through LDS (Vs) 64 rows at a time with
// coalesced loads shared by the four waves; each wave's C loads of a chunk
// are issued together.  MFMA operand layouts (16x16x4 f64): A[m][k] in lane
// m + 16 k, B[k][n] in lane n + 16 k, D[k + 4 e][n] in element e of lane
// n + 16 k.
__device__ __forceinline__ void qr_block_reflect(const double* __restrict__ X, int L, int p0,
                                                 int nb, const double (*Ts)[QNB + 1], bool trans,
                                                 double* __restrict__ C, int64_t ldc, int c0,
                                                 int c1, double (*Vs)[QNB + 1]) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const int cc = c0 + 16 * wave + lr;
  const bool cok = cc < c1;
  double* Cc = C + (int64_t)(cok ? cc : c0) * ldc;
  // V rows r0 .. r0 + 63 into Vs (zero past L, above the diagonal and past nb)
  auto stage = [&](int r0) {
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u, i = e & 63, c = e >> 6, row = r0 + i, dr = p0 + c;
      double v = 0.0;
      if (c < nb && row < L && row >= dr) v = row == dr ? 1.0 : X[(int64_t)dr * L + row];
      Vs[i][c] = v;
    }
    __syncthreads();
  };
  // Y = V^T C_tile (16 x 16), summed over the rows p0 .. L-1
  f64x4q y = {0.0, 0.0, 0.0, 0.0};
  for (int r0 = p0; r0 < L; r0 += 64) {
    stage(r0);
    double b[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int row = r0 + 4 * i + lk;
      b[i] = (cok && row < L) ? Cc[row] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i)
      y = __builtin_amdgcn_mfma_f64_16x16x4f64(Vs[4 * i + lk][lr], b[i], y, 0, 0, 0);
  }
  // Y2 = op(T) Y: element u of y holds Y[lk + 4u][lr], the B operand of
  // the k-step whose summation indices are lk + 4u
  f64x4q y2 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int kk = lk + 4 * u;
    const double a = trans ? Ts[kk][lr] : Ts[lr][kk];
    y2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, y[u], y2, 0, 0, 0);
  }
  // C_tile -= V Y2, 16 rows at a time
  for (int r0 = p0; r0 < L; r0 += 64) {
    stage(r0);
#pragma unroll
    for (int s16 = 0; s16 < 4; ++s16) {
      const int R0 = r0 + 16 * s16;
      f64x4q d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int u = 0; u < 4; ++u)
        d = __builtin_amdgcn_mfma_f64_16x16x4f64(Vs[16 * s16 + lr][lk + 4 * u], y2[u], d, 0, 0, 0);
      if (cok) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = R0 + lk + 4 * e;
          if (row < L) Cc[row] -= d[e];
        }
      }
    }
  }
}

// Reflector JJ of the panel at column p0 (nb columns; JJ a compile-time
// index, so that the panel stays in registers): forms it from column JJ,
// applies it to the panel's later columns and adds column JJ of T.
template <int JJ>
__device__ __forceinline__ void qr_panel_step(double (&P)[4][QNB], int p0, int nb,
                                              double (*red)[QNB], double (*Ts)[QNB + 1]) {
  if (JJ >= nb) return;  // block-uniform
  const int tid = threadIdx.x, col = p0 + JJ;
  double sa[2] = {0.0, 0.0};  // squared norm below the diagonal, alpha
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = tid + 256 * q;
    if (r > col) sa[0] = fma(P[q][JJ], P[q][JJ], sa[0]);
    if (r == col) sa[1] = P[q][JJ];
  }
  qr_block_sum<2>(sa, red);
  const double s = sa[0], alpha = sa[1];
  double beta = alpha, tau = 0.0, scal = 1.0;
  if (s != 0.0) {
    beta = -copysign(sqrt(fma(alpha, alpha, s)), alpha);
    tau = (beta - alpha) / beta;
    scal = 1.0 / (alpha - beta);
  }
  double v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = tid + 256 * q;
    v[q] = r > col ? P[q][JJ] * scal : (r == col ? 1.0 : 0.0);
    if (r > col) P[q][JJ] = v[q];
  }
  // d_c = v^T P[:, c]: for c > JJ the update of column c, for c < JJ (the
  // reflectors already formed) the new column of T
  double d[QNB];
#pragma unroll
  for (int c = 0; c < QNB; ++c) {
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) acc = fma(v[q], P[q][c], acc);
    d[c] = acc;
  }
  qr_block_sum<QNB>(d, red);
#pragma unroll
  for (int c = JJ + 1; c < QNB; ++c) {
    const double f = -tau * d[c];
#pragma unroll
    for (int q = 0; q < 4; ++q) P[q][c] = fma(f, v[q], P[q][c]);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (tid + 256 * q == col) P[q][JJ] = beta;
  // T[0:JJ, JJ] = -tau T[0:JJ, 0:JJ] d[0:JJ], T[JJ][JJ] = tau: thread i owns
  // row i of T (its earlier entries are its own writes)
  if (tid < JJ) {
    double t = 0.0;
#pragma unroll
    for (int l = 0; l < JJ; ++l) t = fma(Ts[tid][l], d[l], t);
    Ts[tid][JJ] = -tau * t;
  }
  if (tid == JJ) Ts[JJ][JJ] = tau;
}

template <int... J>
__device__ __forceinline__ void qr_panel(std::integer_sequence<int, J...>, double (&P)[4][QNB],
                                         int p0, int nb, double (*red)[QNB],
                                         double (*Ts)[QNB + 1]) {
  (qr_panel_step<J>(P, p0, nb, red, Ts), ...);
}

// Panel p of every job (one workgroup per job that has it; items: the
// indices into qjobs): the panel's QNB columns factored in registers, the
// reflectors and R's entries stored back, T into Tf.
__global__ void __launch_bounds__(256) bqr_panel(double* __restrict__ Wv, double* __restrict__ Tf,
                                                 const QJob* __restrict__ qjobs,
                                                 const int* __restrict__ items, int p) {
  __shared__ double red[4][QNB];
  __shared__ double Ts[QNB][QNB + 1];
  const QJob Q = qjobs[items[blockIdx.x]];
  double* X = Wv + Q.x;
  const int L = Q.L, k = Q.k, tid = threadIdx.x;
  const int p0 = p * QNB, nb = min(QNB, k - p0);
  double P[4][QNB];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = tid + 256 * q;
#pragma unroll
    for (int c = 0; c < QNB; ++c)
      P[q][c] = (r >= p0 && r < L && c < nb) ? X[(int64_t)(p0 + c) * L + r] : 0.0;
  }
  Ts[tid >> 4][tid & 15] = 0.0;
  // (the first reduction's barriers order the zeroing before any T write)
  qr_panel(std::make_integer_sequence<int, QNB>{}, P, p0, nb, red, Ts);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = tid + 256 * q;
    if (r >= p0 && r < L) {
#pragma unroll
      for (int c = 0; c < QNB; ++c)
        if (c < nb) X[(int64_t)(p0 + c) * L + r] = P[q][c];
    }
  }
  __syncthreads();
  Tf[Q.tf + (int64_t)p * (QNB * QNB) + tid] = Ts[tid >> 4][tid & 15];
}

// Columns of a job in groups of 64 (one workgroup, a 16-column tile per
// wave): item = (index into qjobs, group).
struct QItem {
  int32_t q, g;
};

// Trailing update after panel p: the group's columns (after the panel) of
// X get C -= V T^T V^T C.
__global__ void __launch_bounds__(256) bqr_trail(double* __restrict__ Wv,
                                                 const double* __restrict__ Tf,
                                                 const QJob* __restrict__ qjobs,
                                                 const QItem* __restrict__ items, int p) {
  __shared__ double Ts[QNB][QNB + 1];
  __shared__ double Vs[64][QNB + 1];
  const QItem it = items[blockIdx.x];
  const QJob Q = qjobs[it.q];
  const int tid = threadIdx.x;
  const int p0 = p * QNB, nb = min(QNB, Q.k - p0);
  const int c0 = p0 + nb + 64 * it.g, c1 = min(Q.k, c0 + 64);
  Ts[tid >> 4][tid & 15] = Tf[Q.tf + (int64_t)p * (QNB * QNB) + tid];
  // (the first staging barrier orders Ts before its use)
  double* X = Wv + Q.x;
  qr_block_reflect(X, Q.L, p0, nb, Ts, true, X, Q.L, c0, c1, Vs);
}

// The rows of R as vectors: Rv[c * k + i] = R[c][i] (i >= c), 0 below.
__global__ void __launch_bounds__(256) bqr_rvec(double* __restrict__ Wv,
                                                const QJob* __restrict__ qjobs) {
  const QJob Q = qjobs[blockIdx.y];
  const int64_t kk = (int64_t)Q.k * Q.k;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < kk; e += (int64_t)gridDim.x * 256) {
    const int64_t c = e / Q.k, i = e - c * Q.k;
    Wv[Q.r + e] = i >= c ? Wv[Q.x + i * Q.L + c] : 0.0;
  }
}

// Q [Z; 0] for the columns c < ncols of Z in the order perm (column-major,
// ld = L): first the copy (grid-wide), then bqr_apply per panel.
__global__ void __launch_bounds__(256) bqr_qz_init(double* __restrict__ QZ,
                                                   const double* __restrict__ Zb,
                                                   const QJob* __restrict__ qjobs,
                                                   const int* __restrict__ perm,
                                                   const int64_t* __restrict__ sbase,
                                                   const int* __restrict__ ncols) {
  const QJob Q = qjobs[blockIdx.y];
  const int nc = ncols[Q.job];
  const int64_t tot = (int64_t)Q.L * nc;
  const int* pm = perm + sbase[Q.job];
  const double* Z = Zb + Q.z;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < tot; e += (int64_t)gridDim.x * 256) {
    const int64_t c = e / Q.L, i = e - c * Q.L;
    QZ[Q.qz + e] = i < Q.k ? Z[(int64_t)pm[c] * Q.k + i] : 0.0;
  }
}

// Panel p's block reflector applied to the group's columns of Q Z (panels in
// reverse order make Q [Z; 0]); groups at or past the job's ncols exit.
__global__ void __launch_bounds__(256) bqr_apply(const double* __restrict__ Wv,
                                                 const double* __restrict__ Tf,
                                                 const QJob* __restrict__ qjobs,
                                                 const QItem* __restrict__ items,
                                                 const int* __restrict__ ncols, int p,
                                                 double* __restrict__ QZ) {
  __shared__ double Ts[QNB][QNB + 1];
  __shared__ double Vs[64][QNB + 1];
  const QItem it = items[blockIdx.x];
  const QJob Q = qjobs[it.q];
  const int nc = ncols[Q.job];
  const int c0 = 64 * it.g, c1 = min(nc, c0 + 64);
  if (c0 >= c1) return;  // block-uniform
  const int tid = threadIdx.x;
  const int p0 = p * QNB, nb = min(QNB, Q.k - p0);
  Ts[tid >> 4][tid & 15] = Tf[Q.tf + (int64_t)p * (QNB * QNB) + tid];
  qr_block_reflect(Wv + Q.x, Q.L, p0, nb, Ts, false, QZ + Q.qz, Q.L, c0, c1, Vs);
}

}  // namespace tn
