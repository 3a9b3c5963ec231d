// Householder QR with FP64 reflectors (replaces np.linalg.qr(mode="reduced"),
// i.e. LAPACK geqrf + orgqr, called by tensor_qr at tensor.py:1044).
//
// Layout: the m x n input is gathered once into a column-major workspace W
// (ld = m) so every reflector is a contiguous column.  Two-level blocking:
//  * inner panels of w <= 8 columns are factored by ONE 1024-thread
//    workgroup with the whole panel held in registers (R rows per thread),
//    so the two reductions per column are LDS/shuffle reductions, not grid
//    synchronisations;  the panel also forms its w x w compact-WY factor T;
//  * the remaining columns of the 64-wide outer block are updated by a
//    column-parallel kernel, and the trailing matrix by three MFMA GEMMs
//    (Y = V^T A, Y = T^T Y, A -= V Y) with split-K for the first.
// Q is formed explicitly (orgqr order: blocks applied last to first to
// [I; 0]).  Sign convention as LAPACK dlarfg: beta = -sign(alpha)*norm.
#include "tn_internal.h"

#include <vector>

namespace tn {

constexpr int QR_NB = 64;
constexpr int PT = 512;  // panel threads (2 waves per SIMD: 256 VGPRs per thread)
constexpr int NWAVE = PT / 64;

// Block-wide sum of nv (<= NV) values; every thread receives the totals.
// One barrier per call: each wave's partials (DPP-reduced) go to one of two
// LDS buffers (alternating with `par`); then lane l of every wave reads the
// one partial (wave l / NV, value l % NV), three xor-shuffles sum each
// value over the waves, and every lane takes the NV totals from lanes
// 0 .. NV-1 (bit-identical in all threads).  A buffer is written again only
// two calls later, after the intervening call's barrier, which every thread
// reaches after its read.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], int nv, double (*lds)[NWAVE * NV],
                                          int& par) {
  static_assert(NWAVE * NV <= 64, "one partial per lane");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double* buf = lds[par];
  par ^= 1;
#pragma unroll
  for (int q = 0; q < NV; ++q)
    if (q < nv) v[q] = wave_sum_dpp(v[q]);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q)
      if (q < nv) buf[wave * NV + q] = v[q];
  }
  __syncthreads();
  double p = lane < NWAVE * NV ? buf[lane] : 0.0;
#pragma unroll
  for (int sh = NV; sh < NWAVE * NV; sh <<= 1) p += __shfl_xor(p, sh, 64);
  const long long u = __double_as_longlong(p);
#pragma unroll
  for (int q = 0; q < NV; ++q)
    if (q < nv) {
      const int lo = __builtin_amdgcn_readlane((int)u, q);
      const int hi = __builtin_amdgcn_readlane((int)(u >> 32), q);
      v[q] = __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
    }
}

// Factor columns [c0, c0+w) (w <= W) over rows [c0, m) of A (column-major,
// ld), one workgroup, the panel in registers: thread t owns rows
// c0 + t + PT*i.  Per column one block reduction, carrying ||x_j|| below the
// pivot, the Gram entries v_a . v_{j-1} (a < j-1) of the compact-WY factor
// and the unscaled dot products of x_j with the later columns below the
// pivot, from which the update's dots v . x_q follow without a second
// reduction.
template <int R, int W>
__global__ void __launch_bounds__(PT) qr_panel_kernel(double* __restrict__ A, int64_t m,
                                                      int64_t ld, int64_t c0, int w,
                                                      double* __restrict__ tau,
                                                      double* __restrict__ Tp) {
  __shared__ double lds[2][NWAVE * W];
  int par = 0;
  __shared__ double piv[W];       // pivot value of each column (one slot each)
  __shared__ double prow[W][W];   // prow[j][q]: pivot row of column j, entry q > j (one row each)
  __shared__ double gsh[W][W];    // v_a . v_b, a < b
  __shared__ double taus[W];
  const int t = threadIdx.x;
  double x[R][W];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    int64_t r = c0 + t + (int64_t)PT * i;
#pragma unroll
    for (int j = 0; j < W; ++j) x[i][j] = (r < m && j < w) ? A[r + (c0 + j) * ld] : 0.0;
  }
  // every index below is a compile-time constant after unrolling (no
  // early exit), so x[][] stays in registers
#pragma unroll
  for (int j = 0; j < W; ++j) {
    if (j < w) {
      // One reduction per column.  vals[0] = ||x_j below the pivot||^2;
      // vals[1 + a] = v_a . v_{j-1} (a < j-1, the compact-WY Gram); vals[q] =
      // S_q = x_j . x_q below the pivot (q > j, x_j not yet scaled).  The
      // update's dots follow from them: with v = [1; x_j,below * scale],
      // v . x_q = x_q[p] + scale * S_q, the pivot row x_q[p] broadcast with it.
      double vals[W];
#pragma unroll
      for (int q = 0; q < W; ++q) vals[q] = 0.0;
      const int64_t p = c0 + j;
#pragma unroll
      for (int i = 0; i < R; ++i) {
        int64_t r = c0 + t + (int64_t)PT * i;
        if (r < m) {
          if (r > p) {
            vals[0] = fma(x[i][j], x[i][j], vals[0]);
#pragma unroll
            for (int q = j + 1; q < W; ++q) vals[q] = fma(x[i][j], x[i][q], vals[q]);
          }
          if (j >= 2 && r > c0 + j - 1) {
#pragma unroll
            for (int a = 0; a < j - 1; ++a) vals[1 + a] = fma(x[i][a], x[i][j - 1], vals[1 + a]);
          }
        }
      }
      if (j >= 2 && t == j - 1) {
        // row c0 + (j-1) belongs to thread j-1 (i = 0): v_{j-1} = 1 there
#pragma unroll
        for (int a = 0; a < j - 1; ++a) vals[1 + a] += x[0][a];
      }
      if (t == j) {
        piv[j] = x[0][j];
#pragma unroll
        for (int q = j + 1; q < W; ++q) prow[j][q] = x[0][q];
      }
      block_sum<W>(vals, W, lds, par);
      if (j >= 2 && t == 0) {
#pragma unroll
        for (int a = 0; a < j - 1; ++a) gsh[a][j - 1] = vals[1 + a];
      }
      const double alpha = piv[j];
      const double xnorm2 = vals[0];
      double tj = 0.0, scale = 0.0, beta = alpha;
      if (xnorm2 > 0.0) {
        double nrm = sqrt(alpha * alpha + xnorm2);
        beta = alpha >= 0.0 ? -nrm : nrm;
        tj = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
      }
#pragma unroll
      for (int i = 0; i < R; ++i) {
        int64_t r = c0 + t + (int64_t)PT * i;
        if (r > p && r < m) x[i][j] *= scale;
      }
      if (t == 0) taus[j] = tj;
      if (j + 1 < w && tj != 0.0) {
#pragma unroll
        for (int q = j + 1; q < W; ++q) vals[q] = tj * fma(scale, vals[q], prow[j][q]);
#pragma unroll
        for (int i = 0; i < R; ++i) {
          int64_t r = c0 + t + (int64_t)PT * i;
          if (r >= p && r < m) {
            const double v = (r == p) ? 1.0 : x[i][j];
#pragma unroll
            for (int q = j + 1; q < W; ++q)
              if (q < w) x[i][q] = fma(-vals[q], v, x[i][q]);
          }
        }
      }
      if (t == j) x[0][j] = beta;
    }
  }
  // Gram entries v_a . v_{w-1} of the last column (a < w-1)
#pragma unroll
  for (int b = 1; b < W; ++b) {
    if (b == w - 1) {
      double vals[W];
#pragma unroll
      for (int q = 0; q < W; ++q) vals[q] = 0.0;
#pragma unroll
      for (int i = 0; i < R; ++i) {
        int64_t r = c0 + t + (int64_t)PT * i;
        if (r < m && r > c0 + b) {
#pragma unroll
          for (int a = 0; a < b; ++a) vals[a] = fma(x[i][a], x[i][b], vals[a]);
        }
      }
      if (t == b) {
#pragma unroll
        for (int a = 0; a < b; ++a) vals[a] += x[0][a];
      }
      block_sum<W>(vals, b, lds, par);
      if (t == 0) {
#pragma unroll
        for (int a = 0; a < b; ++a) gsh[a][b] = vals[a];
      }
    }
  }
  __syncthreads();
  if (t < w) {
    // T (upper triangular, w x w): T[b][b] = tau_b, T[a][b] = -tau_b
    // sum_{a<=c<b} T[a][c] G[c][b]; thread a builds row a in registers (a row
    // needs only its own earlier entries), so the rows run in parallel.
    const int a = t;
    double Tr[W];
#pragma unroll
    for (int b = 0; b < W; ++b) {
      double v = 0.0;
      if (b < w) {
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < b; ++c)
          if (c >= a) acc += Tr[c] * gsh[c][b];
        v = a == b ? taus[b] : (a < b ? -taus[b] * acc : 0.0);
      }
      Tr[b] = v;
    }
    tau[c0 + a] = taus[a];
#pragma unroll
    for (int b = 0; b < W; ++b)
      if (b < w) Tp[a * w + b] = Tr[b];
  }
#pragma unroll
  for (int i = 0; i < R; ++i) {
    int64_t r = c0 + t + (int64_t)PT * i;
#pragma unroll
    for (int j = 0; j < W; ++j)
      if (r < m && j < w) A[r + (c0 + j) * ld] = x[i][j];
  }
}

// Inner-block update W[c0:m, cb:ce] = (I - V T^T V^T) W[c0:m, cb:ce], split
// over row chunks so the whole chip participates:
//   qr_panel_dots:  Ypart[chunk][a][c] = sum_{r in chunk} V[r,a] W[r,c]
//   qr_panel_update: y = sum_chunks Ypart, z = T^T y, W[r,c] -= sum_a V[r,a] z[a]
// V[r,a] = 1 at r = c0+a, W[r, c0+a] below it, 0 above (implicit reflectors).
constexpr int QCH = 256;  // rows per chunk
constexpr int QCOL = 8;   // columns per block

__device__ __forceinline__ double vref(const double* A, int64_t ld, int64_t c0, int a, int64_t r) {
  return (r == c0 + a) ? 1.0 : (r > c0 + a ? A[r + (c0 + a) * ld] : 0.0);
}

template <int W>
__global__ void __launch_bounds__(256) qr_panel_dots(const double* __restrict__ A, int64_t m,
                                                     int64_t ld, int64_t c0, int w, int64_t cb,
                                                     int64_t ncol, double* __restrict__ Ypart) {
  __shared__ double Vs[QCH][W];
  __shared__ double red[4][W][QCOL];
  const int64_t r0 = c0 + (int64_t)blockIdx.y * QCH;
  const int64_t cbase = cb + (int64_t)blockIdx.x * QCOL;
  const int tid = threadIdx.x;
  {
    // all of a thread's reflector loads issued before its LDS stores
    double vv[QCH * W / 256];
#pragma unroll
    for (int i = 0; i < QCH * W / 256; ++i) {
      const int e = tid + 256 * i, rr = e / W, a = e % W;
      const int64_t r = r0 + rr;
      vv[i] = (a < w && r < m) ? vref(A, ld, c0, a, r) : 0.0;
    }
#pragma unroll
    for (int i = 0; i < QCH * W / 256; ++i) {
      const int e = tid + 256 * i;
      Vs[e / W][e % W] = vv[i];
    }
  }
  __syncthreads();
  double y[W][QCOL];
#pragma unroll
  for (int a = 0; a < W; ++a)
#pragma unroll
    for (int j = 0; j < QCOL; ++j) y[a][j] = 0.0;
  const int64_t r = r0 + tid;
  if (r < m) {
#pragma unroll
    for (int j = 0; j < QCOL; ++j) {
      int64_t c = cbase + j;
      double v = (c < cb + ncol) ? A[r + c * ld] : 0.0;
#pragma unroll
      for (int a = 0; a < W; ++a) y[a][j] = fma(Vs[tid][a], v, y[a][j]);
    }
  }
#pragma unroll
  for (int a = 0; a < W; ++a)
#pragma unroll
    for (int j = 0; j < QCOL; ++j) y[a][j] = wave_sum_dpp(y[a][j]);
  if ((tid & 63) == 0) {
#pragma unroll
    for (int a = 0; a < W; ++a)
#pragma unroll
      for (int j = 0; j < QCOL; ++j) red[tid >> 6][a][j] = y[a][j];
  }
  __syncthreads();
  if (tid < W * QCOL) {
    int a = tid / QCOL, j = tid % QCOL;
    double acc = red[0][a][j] + red[1][a][j] + red[2][a][j] + red[3][a][j];
    int64_t c = cbase + j;
    if (c < cb + ncol)
      Ypart[((int64_t)blockIdx.y * W + a) * ncol + (c - cb)] = acc;
  }
}

template <int W>
__global__ void __launch_bounds__(256) qr_panel_update(double* __restrict__ A, int64_t m,
                                                       int64_t ld, int64_t c0, int w, int64_t cb,
                                                       int64_t ncol, int nchunk,
                                                       const double* __restrict__ Ypart,
                                                       const double* __restrict__ Tp) {
  __shared__ double ys[W][QCOL];
  __shared__ double z[W][QCOL];
  const int64_t r0 = c0 + (int64_t)blockIdx.y * QCH;
  const int64_t cbase = cb + (int64_t)blockIdx.x * QCOL;
  const int tid = threadIdx.x;
  if (tid < W * QCOL) {
    // y[a][j] = sum over the chunks' partials (in chunk order), four loads
    // in flight at a time
    const int a = tid / QCOL, j = tid % QCOL;
    const int64_t c = cbase + j;
    double ya = 0.0;
    if (a < w && c < cb + ncol) {
      const double* yp = Ypart + (int64_t)a * ncol + (c - cb);
      const int64_t cs = (int64_t)W * ncol;
      int ch = 0;
      for (; ch + 4 <= nchunk; ch += 4) {
        const double v0 = yp[ch * cs], v1 = yp[(ch + 1) * cs];
        const double v2 = yp[(ch + 2) * cs], v3 = yp[(ch + 3) * cs];
        ya += v0;
        ya += v1;
        ya += v2;
        ya += v3;
      }
      for (; ch < nchunk; ++ch) ya += yp[ch * cs];
    }
    ys[a][j] = ya;
  }
  __syncthreads();
  if (tid < W * QCOL) {
    // z[b][j] = sum_{a <= b} T[a][b] y[a][j]
    const int b = tid / QCOL, j = tid % QCOL;
    const int64_t c = cbase + j;
    double acc = 0.0;
    if (b < w && c < cb + ncol)
      for (int a = 0; a <= b; ++a) acc += Tp[a * w + b] * ys[a][j];
    z[b][j] = acc;
  }
  __syncthreads();
  const int64_t r = r0 + tid;
  if (r >= m) return;
  double v[W];
#pragma unroll
  for (int a = 0; a < W; ++a) v[a] = a < w ? vref(A, ld, c0, a, r) : 0.0;
#pragma unroll
  for (int j = 0; j < QCOL; ++j) {
    int64_t c = cbase + j;
    if (c >= cb + ncol) break;
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < W; ++a) acc = fma(v[a], z[a][j], acc);
    A[r + c * ld] -= acc;
  }
}

// Explicit V (mr x nb, column-major, ld = mr) of the outer block at j0.
__global__ void qr_build_v_kernel(const double* __restrict__ A, int64_t m, int64_t ld,
                                  int64_t j0, int nb, double* __restrict__ V) {
  const int64_t mr = m - j0;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= mr * nb) return;
  int64_t a = i / mr, r = i - a * mr;
  V[i] = (r == a) ? 1.0 : (r > a ? A[(j0 + r) + (j0 + a) * ld] : 0.0);
}

// T (nb x nb, row-major) from G = V^T V and tau (LAPACK dlarft,
// forward/columnwise: H_1 ... H_nb = I - V T V^T), recursively: with the
// reflectors split into blocks A | B,
//   T = [[T_A, -T_A G_AB T_B], [0, T_B]].
// Wave w builds the 16 x 16 diagonal block w by the column recurrence
// T[a][b] = -tau_b sum_{a<=c<b} T[a][c] G[c][b] (lane a owns row a; only
// 16 dependent columns), then all 256 threads form the off-diagonal blocks
// of the two 32 x 32 halves and finally the 32 x 32 block between them,
// each as two small products through LDS.  Indices >= nb carry tau = 0 and
// G = 0, so their rows and columns of T are zero.
__global__ void __launch_bounds__(256) qr_larft_kernel(const double* __restrict__ G, int nb,
                                                       const double* __restrict__ tau,
                                                       double* __restrict__ T) {
  __shared__ double Gs[QR_NB][QR_NB + 1];
  __shared__ double Ts[QR_NB][QR_NB + 1];
  __shared__ double Ms[QR_NB][QR_NB + 1];
  __shared__ double ts[QR_NB];
  const int tid = threadIdx.x;
  {
    double g[QR_NB * QR_NB / 256];
#pragma unroll
    for (int k = 0; k < QR_NB * QR_NB / 256; ++k) {
      const int i = tid + 256 * k, r = i / QR_NB, c = i % QR_NB;
      g[k] = (r < nb && c < nb) ? G[r * nb + c] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < QR_NB * QR_NB / 256; ++k) {
      const int i = tid + 256 * k;
      Gs[i / QR_NB][i % QR_NB] = g[k];
      Ts[i / QR_NB][i % QR_NB] = 0.0;
    }
  }
  if (tid < QR_NB) ts[tid] = tid < nb ? tau[tid] : 0.0;
  __syncthreads();
  // diagonal 16 x 16 blocks: wave w, lane a < 16 owns row 16 w + a
  {
    const int w = tid >> 6, a = tid & 63, o = 16 * w;
    if (a < 16) {
      double Tr[16];
#pragma unroll
      for (int bb = 0; bb < 16; ++bb) {
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < bb; ++c)
          if (c >= a) acc = fma(Tr[c], Gs[o + c][o + bb], acc);
        Tr[bb] = a == bb ? ts[o + bb] : (a < bb ? -ts[o + bb] * acc : 0.0);
      }
#pragma unroll
      for (int bb = 0; bb < 16; ++bb) Ts[o + a][o + bb] = Tr[bb];
    }
  }
  __syncthreads();
  // off-diagonal block of the half at h0 (size 2 hs, hs = 16 then 32):
  // T[h0 + i][h0 + hs + j] = -sum_p T_A[i][p] (sum_q G_AB[p][q] T_B[q][j])
  auto combine = [&](int hs, int nh) {
    // M = G_AB T_B for nh halves at once: entries (half, p, j)
    const int per = hs * hs;
    for (int e = tid; e < nh * per; e += 256) {
      const int hh = e / per, p = (e % per) / hs, j = e % hs, h0 = 2 * hs * hh;
      double acc = 0.0;
      for (int q = 0; q <= j; ++q) acc = fma(Gs[h0 + p][h0 + hs + q], Ts[h0 + hs + q][h0 + hs + j], acc);
      Ms[h0 + p][h0 + hs + j] = acc;
    }
    __syncthreads();
    for (int e = tid; e < nh * per; e += 256) {
      const int hh = e / per, i = (e % per) / hs, j = e % hs, h0 = 2 * hs * hh;
      double acc = 0.0;
      for (int p = i; p < hs; ++p) acc = fma(Ts[h0 + i][h0 + p], Ms[h0 + p][h0 + hs + j], acc);
      Ts[h0 + i][h0 + hs + j] = -acc;
    }
    __syncthreads();
  };
  combine(16, 2);
  combine(32, 1);
  for (int e = tid; e < nb * nb; e += 256) T[e] = Ts[e / nb][e % nb];
}

__global__ void qr_init_from_kernel(const double* __restrict__ X, double* __restrict__ Q, int64_t m,
                                    int64_t k) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m * k) return;
  int64_t c = i / m, r = i - c * m;
  Q[i] = r < k ? X[c * k + r] : 0.0;
}

__global__ void qr_init_q_kernel(double* __restrict__ Q, int64_t m, int64_t k) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m * k) return;
  int64_t c = i / m, r = i - c * m;
  Q[i] = (r == c) ? 1.0 : 0.0;
}

// R (k x n row-major) = upper triangle of W.
template <typename T>
__global__ void qr_extract_r_kernel(const double* __restrict__ A, int64_t m, int64_t k, int64_t n,
                                    T* __restrict__ R) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k * n) return;
  int64_t r = i / n, c = i - r * n;
  R[i] = (T)(c >= r ? A[r + c * m] : 0.0);
}

template <typename T>
__global__ void colmajor_to_rowmajor_kernel(const double* __restrict__ A, int64_t m, int64_t k,
                                            T* __restrict__ Q) {
  __shared__ double tile[32][33];
  int64_t bx = (int64_t)blockIdx.x * 32, by = (int64_t)blockIdx.y * 32;
  int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    int64_t r = bx + tx, c = by + j;
    if (r < m && c < k) tile[j][tx] = A[r + c * m];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    int64_t r = bx + j, c = by + tx;
    if (r < m && c < k) Q[r * k + c] = (T)tile[tx][j];
  }
}

// gather: W[i + j*m] = A(row i, col j) for any strided input (T -> double)
template <typename T>
__global__ void gather_colmajor_kernel(const T* __restrict__ src, Modes rows, Modes cols,
                                       int64_t m, int64_t n, double* __restrict__ W) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < m * n; i += stride) {
    int64_t c = i / m, r = i - c * m;
    W[i] = (double)src[mode_offset(rows, r) + mode_offset(cols, c)];
  }
}

size_t qr_layout(int64_t m, int64_t n, char* base, size_t cap, QrWs* ws) {
  const int64_t k = std::min(m, n);
  const int64_t nblk = ceil_div(std::max<int64_t>(k, 1), QR_NB);
  const int64_t wn = std::max(n, k);
  Arena a{base, cap, 0};
  QrWs w{};
  w.W = a.take<double>((size_t)(m * n));
  w.tau = a.take<double>((size_t)std::max<int64_t>(k, 1));
  w.V = a.take<double>((size_t)(m * QR_NB));
  w.Tall = a.take<double>((size_t)(nblk * QR_NB * QR_NB));
  w.G = a.take<double>((size_t)(QR_NB * QR_NB));
  w.Y = a.take<double>((size_t)(QR_NB * wn));
  w.Y2 = a.take<double>((size_t)(QR_NB * wn));
  w.Tp = a.take<double>(64);
  w.Ypart = a.take<double>((size_t)ceil_div(m, QCH) * 8 * QR_NB);
  w.Qw = a.take<double>((size_t)(m * k));
  w.part_bytes = (size_t)16 * QR_NB * wn * sizeof(double);
  w.part = a.take<double>((size_t)16 * QR_NB * wn);
  if (ws) *ws = w;
  return a.used;
}

// Panels taller than the register panel holds (w = 1): one column per step,
// three grid launches -- partial sums of squares below the pivot, the
// reflector (beta, tau, scale; same convention as the register panel) from
// one block, then the scaling of the reflector below the pivot.  The
// partials and the scale live in ws.Ypart, which is free between a panel and
// its update.
constexpr int QTALL = 4096;  // rows per block of the tall-column kernels

__device__ __forceinline__ double block_sum256(double s, double* red) {
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void __launch_bounds__(256) qr_tall_sumsq(const double* __restrict__ A, int64_t m,
                                                      int64_t ld, int64_t c0,
                                                      double* __restrict__ part) {
  __shared__ double red[4];
  const int64_t r0 = c0 + 1 + (int64_t)blockIdx.x * QTALL;
  const int64_t r1 = r0 + QTALL < m ? r0 + QTALL : m;
  const double* col = A + c0 * ld;
  double s = 0.0;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += 256) s = fma(col[r], col[r], s);
  s = block_sum256(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(256) qr_tall_reflector(double* __restrict__ A, int64_t ld,
                                                          int64_t c0, double* __restrict__ part,
                                                          int nchunk, double* __restrict__ tau,
                                                          double* __restrict__ Tp) {
  __shared__ double red[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < nchunk; i += 256) s += part[i];
  const double xnorm2 = block_sum256(s, red);
  if (threadIdx.x == 0) {
    double* pv = A + c0 + c0 * ld;
    const double alpha = *pv;
    double tj = 0.0, scale = 0.0, beta = alpha;
    if (xnorm2 > 0.0) {
      const double nrm = sqrt(alpha * alpha + xnorm2);
      beta = alpha >= 0.0 ? -nrm : nrm;
      tj = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    *pv = beta;
    tau[c0] = tj;
    Tp[0] = tj;
    part[nchunk] = scale;
  }
}

__global__ void __launch_bounds__(256) qr_tall_scale(double* __restrict__ A, int64_t m, int64_t ld,
                                                      int64_t c0, const double* __restrict__ part,
                                                      int nchunk) {
  const double scale = part[nchunk];
  const int64_t r0 = c0 + 1 + (int64_t)blockIdx.x * QTALL;
  const int64_t r1 = r0 + QTALL < m ? r0 + QTALL : m;
  double* col = A + c0 * ld;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += 256) col[r] *= scale;
}

static int run_tall_column(double* A, int64_t m, int64_t ld, int64_t c0, double* tau, double* Tp,
                           double* part, hipStream_t st) {
  const int nchunk = (int)ceil_div(std::max<int64_t>(m - c0 - 1, 0), QTALL);
  if (nchunk > 0) {
    hipLaunchKernelGGL(qr_tall_sumsq, dim3((unsigned)nchunk), dim3(256), 0, st, A, m, ld, c0, part);
    TN_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(qr_tall_reflector, dim3(1), dim3(256), 0, st, A, ld, c0, part, nchunk, tau, Tp);
  TN_CHECK_LAUNCH();
  if (nchunk > 0) {
    hipLaunchKernelGGL(qr_tall_scale, dim3((unsigned)nchunk), dim3(256), 0, st, A, m, ld, c0, part,
                       nchunk);
    TN_CHECK_LAUNCH();
  }
  return TN_OK;
}

typedef void (*PanelFn)(double*, int64_t, int64_t, int64_t, int, double*, double*);

struct PanelChoice {
  int R, W;
  const void* fn;
};

template <int R, int W>
static int launch_panel(double* A, int64_t m, int64_t ld, int64_t c0, int w, double* tau,
                        double* Tp, hipStream_t st) {
  hipLaunchKernelGGL((qr_panel_kernel<R, W>), dim3(1), dim3(PT), 0, st, A, m, ld, c0, w, tau, Tp);
  TN_CHECK_LAUNCH();
  return TN_OK;
}

// Pick the register-panel instantiation for m_rem rows; returns panel width.
static int panel_width(int64_t mrem) {
  int64_t R = ceil_div(mrem, PT);
  if (R <= 10) return 8;
  if (R <= 14) return 4;
  if (R <= 24) return 2;
  return 1;
}

static int run_panel(double* A, int64_t m, int64_t ld, int64_t c0, int w, double* tau, double* Tp,
                     double* part, hipStream_t st) {
  int64_t R = ceil_div(m - c0, PT);
  if (R <= 1) return launch_panel<1, 8>(A, m, ld, c0, w, tau, Tp, st);
  if (R <= 2) return launch_panel<2, 8>(A, m, ld, c0, w, tau, Tp, st);
  if (R <= 4) return launch_panel<4, 8>(A, m, ld, c0, w, tau, Tp, st);
  if (R <= 6) return launch_panel<6, 8>(A, m, ld, c0, w, tau, Tp, st);
  if (R <= 8) return launch_panel<8, 8>(A, m, ld, c0, w, tau, Tp, st);
  if (R <= 10) return launch_panel<10, 8>(A, m, ld, c0, w, tau, Tp, st);
  if (R <= 14) return launch_panel<14, 4>(A, m, ld, c0, w, tau, Tp, st);
  if (R <= 24) return launch_panel<24, 2>(A, m, ld, c0, w, tau, Tp, st);
  if (R <= 48) return launch_panel<48, 1>(A, m, ld, c0, w, tau, Tp, st);
  TN_REQUIRE(w == 1, "tn_qr: tall panel of width %d", w);
  return run_tall_column(A, m, ld, c0, tau, Tp, part, st);
}

static int run_apply_panel(double* A, int64_t m, int64_t ld, int64_t c0, int w, const double* Tp,
                           int64_t cb, int64_t ce, double* Ypart, hipStream_t st) {
  if (ce <= cb) return TN_OK;
  const int64_t ncol = ce - cb;
  const int nchunk = (int)ceil_div(m - c0, QCH);
  dim3 grid((unsigned)ceil_div(ncol, QCOL), (unsigned)nchunk);
  if (w > 4) {
    hipLaunchKernelGGL(qr_panel_dots<8>, grid, dim3(256), 0, st, A, m, ld, c0, w, cb, ncol, Ypart);
    hipLaunchKernelGGL(qr_panel_update<8>, grid, dim3(256), 0, st, A, m, ld, c0, w, cb, ncol,
                       nchunk, Ypart, Tp);
  } else if (w > 2) {
    hipLaunchKernelGGL(qr_panel_dots<4>, grid, dim3(256), 0, st, A, m, ld, c0, w, cb, ncol, Ypart);
    hipLaunchKernelGGL(qr_panel_update<4>, grid, dim3(256), 0, st, A, m, ld, c0, w, cb, ncol,
                       nchunk, Ypart, Tp);
  } else {
    hipLaunchKernelGGL(qr_panel_dots<2>, grid, dim3(256), 0, st, A, m, ld, c0, w, cb, ncol, Ypart);
    hipLaunchKernelGGL(qr_panel_update<2>, grid, dim3(256), 0, st, A, m, ld, c0, w, cb, ncol,
                       nchunk, Ypart, Tp);
  }
  TN_CHECK_LAUNCH();
  return TN_OK;
}

// Factor W (m x n column-major, ld=m) in place; fills tau and the per-block T.
int qr_factor(double* W, int64_t m, int64_t n, const QrWs& ws, hipStream_t st) {
  const int64_t k = std::min(m, n);
  for (int64_t j0 = 0; j0 < k; j0 += QR_NB) {
    const int nbo = (int)std::min<int64_t>(QR_NB, k - j0);
    // inner panels inside the outer block
    for (int64_t jj = j0; jj < j0 + nbo;) {
      int wmax = panel_width(m - jj);
      int w = (int)std::min<int64_t>(wmax, j0 + nbo - jj);
      TN_TRY(run_panel(W, m, m, jj, w, ws.tau, ws.Tp, ws.Ypart, st));
      TN_TRY(run_apply_panel(W, m, m, jj, w, ws.Tp, jj + w, j0 + nbo, ws.Ypart, st));
      jj += w;
    }
    // compact-WY for the whole outer block (also used to form Q)
    const int64_t mr = m - j0;
    hipLaunchKernelGGL(qr_build_v_kernel, dim3((unsigned)ceil_div(mr * nbo, 256)), dim3(256), 0,
                       st, W, m, m, j0, nbo, ws.V);
    TN_CHECK_LAUNCH();
    // G = V^T V
    TN_TRY(contract_launch(ws.V, mk({nbo}, {mr}), mk({mr}, {1}), ws.V, mk({mr}, {1}),
                           mk({nbo}, {mr}), ws.G, mk({nbo}, {nbo}), mk({nbo}, {1}), TN_F64,
                           TN_PREC_NATIVE, 1.0, 0.0, ws.part, ws.part_bytes, st));
    double* T = ws.Tall + (j0 / QR_NB) * QR_NB * QR_NB;
    hipLaunchKernelGGL(qr_larft_kernel, dim3(1), dim3(256), 0, st, ws.G, nbo, ws.tau + j0, T);
    TN_CHECK_LAUNCH();
    const int64_t nt = n - (j0 + nbo);
    if (nt > 0) {
      double* At = W + j0 + (j0 + nbo) * m;
      // Y = V^T At   (nbo x nt)
      TN_TRY(contract_launch(ws.V, mk({nbo}, {mr}), mk({mr}, {1}), At, mk({mr}, {1}),
                             mk({nt}, {m}), ws.Y, mk({nbo}, {nt}), mk({nt}, {1}), TN_F64,
                             TN_PREC_NATIVE, 1.0, 0.0, ws.part, ws.part_bytes, st));
      // Y2 = T^T Y
      TN_TRY(contract_launch(T, mk({nbo}, {1}), mk({nbo}, {nbo}), ws.Y, mk({nbo}, {nt}),
                             mk({nt}, {1}), ws.Y2, mk({nbo}, {nt}), mk({nt}, {1}), TN_F64,
                             TN_PREC_NATIVE, 1.0, 0.0, nullptr, 0, st));
      // At -= V Y2
      TN_TRY(contract_launch(ws.V, mk({mr}, {1}), mk({nbo}, {mr}), ws.Y2, mk({nbo}, {nt}),
                             mk({nt}, {1}), At, mk({mr}, {1}), mk({nt}, {m}), TN_F64,
                             TN_PREC_NATIVE, -1.0, 1.0, nullptr, 0, st));
    }
  }
  return TN_OK;
}

// Qw (m x k column-major) = H_1 ... H_k [X0; 0]; X0 = I when init == nullptr,
// else the k x k column-major matrix init (ld = k).
int qr_form_q(const double* W, int64_t m, int64_t n, const QrWs& ws, double* Qw, hipStream_t st,
              const double* init) {
  const int64_t k = std::min(m, n);
  if (init == nullptr)
    hipLaunchKernelGGL(qr_init_q_kernel, dim3((unsigned)ceil_div(m * k, 256)), dim3(256), 0, st,
                       Qw, m, k);
  else
    hipLaunchKernelGGL(qr_init_from_kernel, dim3((unsigned)ceil_div(m * k, 256)), dim3(256), 0,
                       st, init, Qw, m, k);
  TN_CHECK_LAUNCH();
  const int64_t nblk = ceil_div(k, QR_NB);
  for (int64_t b = nblk - 1; b >= 0; --b) {
    const int64_t j0 = b * QR_NB;
    const int nbo = (int)std::min<int64_t>(QR_NB, k - j0);
    // with [I; 0] the columns left of j0 are still unit vectors that this
    // block's reflectors (rows >= j0) leave alone; a general start matrix
    // needs every column updated.
    const int64_t c0 = init == nullptr ? j0 : 0;
    const int64_t mr = m - j0, nc = k - c0;
    hipLaunchKernelGGL(qr_build_v_kernel, dim3((unsigned)ceil_div(mr * nbo, 256)), dim3(256), 0,
                       st, W, m, m, j0, nbo, ws.V);
    TN_CHECK_LAUNCH();
    const double* T = ws.Tall + b * QR_NB * QR_NB;
    double* Qs = Qw + j0 + c0 * m;
    TN_TRY(contract_launch(ws.V, mk({nbo}, {mr}), mk({mr}, {1}), Qs, mk({mr}, {1}),
                           mk({nc}, {m}), ws.Y, mk({nbo}, {nc}), mk({nc}, {1}), TN_F64,
                           TN_PREC_NATIVE, 1.0, 0.0, ws.part, ws.part_bytes, st));
    TN_TRY(contract_launch(T, mk({nbo}, {nbo}), mk({nbo}, {1}), ws.Y, mk({nbo}, {nc}),
                           mk({nc}, {1}), ws.Y2, mk({nbo}, {nc}), mk({nc}, {1}), TN_F64,
                           TN_PREC_NATIVE, 1.0, 0.0, nullptr, 0, st));
    TN_TRY(contract_launch(ws.V, mk({mr}, {1}), mk({nbo}, {mr}), ws.Y2, mk({nbo}, {nc}),
                           mk({nc}, {1}), Qs, mk({mr}, {1}), mk({nc}, {m}), TN_F64,
                           TN_PREC_NATIVE, -1.0, 1.0, nullptr, 0, st));
  }
  return TN_OK;
}

size_t qr_workspace_bytes(int64_t m, int64_t n) { return qr_layout(m, n, nullptr, 0, nullptr) + 256; }

}  // namespace tn

using namespace tn;

extern "C" size_t tn_qr_workspace(int64_t m, int64_t n, int dtype) {
  (void)dtype;
  if (m <= 0 || n <= 0) return 256;
  return tn::qr_workspace_bytes(m, n);
}

extern "C" int tn_qr(const void* A, const tn_modes* a_rows, const tn_modes* a_cols, int64_t m,
                     int64_t n, void* Q, void* R, int dtype, void* ws, size_t ws_bytes,
                     void* stream) {
  TN_REQUIRE(m >= 0 && n >= 0, "tn_qr: negative size");
  TN_REQUIRE(dtype == TN_F64 || dtype == TN_F32, "tn_qr: unsupported dtype %d", dtype);
  Modes rows = to_modes(a_rows), cols = to_modes(a_cols);
  TN_REQUIRE(modes_size(rows) == m && modes_size(cols) == n, "tn_qr: modes do not match m x n");
  const int64_t k = std::min(m, n);
  if (k == 0) return TN_OK;
  TN_REQUIRE(A && Q && R, "tn_qr: null pointer");
  TN_REQUIRE(ws && ws_bytes >= tn_qr_workspace(m, n, dtype), "tn_qr: workspace too small");
  hipStream_t st = as_stream(stream);
  QrWs w{};
  qr_layout(m, n, (char*)ws, ws_bytes, &w);
  unsigned gb = (unsigned)std::min<int64_t>(ceil_div(m * n, 256), 65536);
  if (dtype == TN_F64)
    hipLaunchKernelGGL(gather_colmajor_kernel<double>, dim3(gb), dim3(256), 0, st,
                       (const double*)A, rows, cols, m, n, w.W);
  else
    hipLaunchKernelGGL(gather_colmajor_kernel<float>, dim3(gb), dim3(256), 0, st,
                       (const float*)A, rows, cols, m, n, w.W);
  TN_CHECK_LAUNCH();
  TN_TRY(qr_factor(w.W, m, n, w, st));
  TN_TRY(qr_form_q(w.W, m, n, w, w.Qw, st, nullptr));
  dim3 tg((unsigned)ceil_div(m, 32), (unsigned)ceil_div(k, 32));
  unsigned rb = (unsigned)ceil_div(k * n, 256);
  if (dtype == TN_F64) {
    hipLaunchKernelGGL(colmajor_to_rowmajor_kernel<double>, tg, dim3(256), 0, st, w.Qw, m, k,
                       (double*)Q);
    hipLaunchKernelGGL(qr_extract_r_kernel<double>, dim3(rb), dim3(256), 0, st, w.W, m, k, n,
                       (double*)R);
  } else {
    hipLaunchKernelGGL(colmajor_to_rowmajor_kernel<float>, tg, dim3(256), 0, st, w.Qw, m, k,
                       (float*)Q);
    hipLaunchKernelGGL(qr_extract_r_kernel<float>, dim3(rb), dim3(256), 0, st, w.W, m, k, n,
                       (float*)R);
  }
  TN_CHECK_LAUNCH();
  return TN_OK;
}
