// Thin SVD by one-sided block Jacobi (Hestenes) with FP64 rotations
// (replaces np.linalg.svd(full_matrices=False), LAPACK gesdd, called by
// tensor_svd at tensor.py:915).
//
// The k = min(m,n) vectors to orthogonalise (columns of A if m >= n, rows of
// A otherwise) are gathered into a workspace Wv with each vector contiguous
// (length L = max(m,n)); Z (k x k) accumulates the rotations.  Vectors are
// grouped in blocks of 16; each outer step pairs the blocks round-robin and,
// for every pair (I,J) at once:
//   1. jac_gram:  partial Gram G = P^T P of P = [W_I W_J] over a row chunk,
//                 v_mfma_f64_16x16x4 from an LDS tile (grid: pairs x chunks);
//   2. jac_eig:   sums the partials, records the pair's largest normalised
//                 off-diagonal |G_pq|/sqrt(G_pp G_qq), and diagonalises the
//                 32x32 G with a parallel cyclic two-sided Jacobi in LDS
//                 (G = Q diag Q^T);
//   3. jac_apply: [W_I W_J] <- [W_I W_J] Q and [Z_I Z_J] <- [Z_I Z_J] Q,
//                 row-parallel (grid: pairs x row chunks).
// A sweep is nblk-1 steps; sweeps repeat until the largest normalised
// off-diagonal of a whole sweep is <= tol.  Singular values are the vector
// norms; they are sorted descending on the device.
#include "tn_internal.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace tn {

constexpr int JB = 16;       // vectors per block
constexpr int J2 = 2 * JB;   // pair width
constexpr int JCH = 256;     // rows per chunk (jac_gram / jac_apply)

__device__ __forceinline__ int rr_block(int pos, int step, int nblk) {
  return pos == 0 ? 0 : 1 + (pos - 1 + step) % (nblk - 1);
}

// pair pi at step s -> blocks (I, J)
__device__ __forceinline__ void pair_blocks(int pi, int s, int nblk, int& I, int& J) {
  I = rr_block(pi, s, nblk);
  J = rr_block(nblk - 1 - pi, s, nblk);
}

__device__ __forceinline__ int64_t pair_vec(int c, int I, int J) {
  return c < JB ? (int64_t)I * JB + c : (int64_t)J * JB + (c - JB);
}

typedef double f64x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) jac_gram(const double* __restrict__ Wv, int64_t L,
                                                int64_t nv, int s, int nblk, int nchunk,
                                                double* __restrict__ Gpart) {
  __shared__ double Ps[2][64][J2 + 1];
  const int pi = blockIdx.x, ch = blockIdx.y;
  int I, J;
  pair_blocks(pi, s, nblk, I, J);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int64_t r0 = (int64_t)ch * JCH;
  const int64_t r1 = min(L, r0 + JCH);
  const int ti = wave >> 1, tj = wave & 1;
  f64x4 acc = {0, 0, 0, 0};
  double reg[8];
  auto load = [&](int64_t rb) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      int e = tid + 256 * j;
      int row = e & 63, col = e >> 6;
      int64_t v = pair_vec(col, I, J);
      int64_t r = rb + row;
      reg[j] = (v < nv && r < r1) ? Wv[v * L + r] : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      int e = tid + 256 * j;
      Ps[buf][e & 63][e >> 6] = reg[j];
    }
  };
  const int nsub = (int)((r1 - r0 + 63) / 64);
  if (nsub > 0) {
    load(r0);
    store(0);
  }
  __syncthreads();
  for (int sb = 0; sb < nsub; ++sb) {
    const int buf = sb & 1;
    if (sb + 1 < nsub) load(r0 + (int64_t)(sb + 1) * 64);
#pragma unroll
    for (int k0 = 0; k0 < 64; k0 += 4) {
      double a = Ps[buf][k0 + (lane >> 4)][16 * ti + (lane & 15)];
      double b = Ps[buf][k0 + (lane >> 4)][16 * tj + (lane & 15)];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    if (sb + 1 < nsub) store(buf ^ 1);
    __syncthreads();
  }
  double* out = Gpart + ((int64_t)pi * nchunk + ch) * (J2 * J2);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    int r = 16 * ti + (lane >> 4) + 4 * e, c = 16 * tj + (lane & 15);
    out[r * J2 + c] = acc[e];
  }
}

// Largest normalised off-diagonal over the pair, encoded for atomicMax.
__device__ __forceinline__ unsigned long long dbits(double x) {
  return (unsigned long long)__double_as_longlong(x);
}

__global__ void __launch_bounds__(256) jac_eig(const double* __restrict__ Gpart, int nchunk,
                                               int64_t nv, int s, int nblk, double tol,
                                               double* __restrict__ Qall, int* __restrict__ skip,
                                               unsigned long long* __restrict__ offmax) {
  __shared__ double G[J2][J2 + 1];
  __shared__ double Q[J2][J2 + 1];
  __shared__ double cs[J2 / 2][2];
  __shared__ int pp[J2 / 2][2];
  __shared__ double red[4];
  __shared__ int rotated;
  const int pi = blockIdx.x, tid = threadIdx.x;
  int I, J;
  pair_blocks(pi, s, nblk, I, J);
  for (int e = tid; e < J2 * J2; e += 256) {
    double acc = 0.0;
    for (int c = 0; c < nchunk; ++c) acc += Gpart[((int64_t)pi * nchunk + c) * (J2 * J2) + e];
    G[e / J2][e % J2] = acc;
    Q[e / J2][e % J2] = (e / J2 == e % J2) ? 1.0 : 0.0;
  }
  __syncthreads();
  // off-diagonal measure (all pairs p < q of the 32 vectors)
  double off = 0.0;
  for (int e = tid; e < J2 * J2; e += 256) {
    int p = e / J2, q = e % J2;
    if (p < q) {
      double d = G[p][p] * G[q][q];
      if (d > 0.0) off = fmax(off, fabs(G[p][q]) / sqrt(d));
    }
  }
  off = wave_max(off);
  if ((tid & 63) == 0) red[tid >> 6] = off;
  __syncthreads();
  off = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  if (tid == 0) atomicMax(offmax, dbits(off));
  if (off <= tol) {
    if (tid == 0) skip[pi] = 1;
    return;
  }
  if (tid == 0) skip[pi] = 0;
  // parallel cyclic Jacobi on the 32 x 32 Gram matrix
  const double inner_tol = 1e-15;
  for (int sweep = 0; sweep < 30; ++sweep) {
    if (tid == 0) rotated = 0;
    __syncthreads();
    for (int r = 0; r < J2 - 1; ++r) {
      if (tid < J2 / 2) {
        int p = tid == 0 ? 0 : 1 + (tid - 1 + r) % (J2 - 1);
        int q = 1 + (J2 - 2 - tid + r) % (J2 - 1);
        if (p > q) { int t = p; p = q; q = t; }
        double gpq = G[p][q], gpp = G[p][p], gqq = G[q][q];
        double c = 1.0, sn = 0.0;
        if (fabs(gpq) > inner_tol * sqrt(fmax(gpp * gqq, 0.0)) && gpq != 0.0) {
          double th = (gqq - gpp) / (2.0 * gpq);
          double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(1.0 + th * th));
          c = 1.0 / sqrt(1.0 + t * t);
          sn = t * c;
          rotated = 1;
        }
        cs[tid][0] = c; cs[tid][1] = sn;
        pp[tid][0] = p; pp[tid][1] = q;
      }
      __syncthreads();
      // rows: G <- J^T G
      for (int e = tid; e < (J2 / 2) * J2; e += 256) {
        int k = e / J2, col = e % J2;
        int p = pp[k][0], q = pp[k][1];
        double c = cs[k][0], sn = cs[k][1];
        double gp = G[p][col], gq = G[q][col];
        G[p][col] = c * gp - sn * gq;
        G[q][col] = sn * gp + c * gq;
      }
      __syncthreads();
      // columns: G <- G J, Q <- Q J
      for (int e = tid; e < (J2 / 2) * J2; e += 256) {
        int k = e / J2, row = e % J2;
        int p = pp[k][0], q = pp[k][1];
        double c = cs[k][0], sn = cs[k][1];
        double gp = G[row][p], gq = G[row][q];
        G[row][p] = c * gp - sn * gq;
        G[row][q] = sn * gp + c * gq;
        double qp = Q[row][p], qq = Q[row][q];
        Q[row][p] = c * qp - sn * qq;
        Q[row][q] = sn * qp + c * qq;
      }
      __syncthreads();
    }
    const int any = rotated;
    __syncthreads();  // everyone has read `rotated` before thread 0 resets it
    if (!any) break;
  }
  double* out = Qall + (int64_t)pi * (J2 * J2);
  for (int e = tid; e < J2 * J2; e += 256) out[e] = Q[e / J2][e % J2];
}

// rows [ch*JCH, ...) of the stacked [Wv (L rows); Z (nv rows)] for pair pi
__global__ void __launch_bounds__(256) jac_apply(double* __restrict__ Wv, int64_t L,
                                                 double* __restrict__ Z, int64_t nv, int s,
                                                 int nblk, int nchW, const double* __restrict__ Qall,
                                                 const int* __restrict__ skip) {
  __shared__ double Q[J2][J2];
  const int pi = blockIdx.x, ch = blockIdx.y;
  if (skip[pi]) return;
  int I, J;
  pair_blocks(pi, s, nblk, I, J);
  for (int e = threadIdx.x; e < J2 * J2; e += 256) Q[e / J2][e % J2] = Qall[(int64_t)pi * J2 * J2 + e];
  __syncthreads();
  double* base;
  int64_t ld, r;
  if (ch < nchW) {
    base = Wv; ld = L; r = (int64_t)ch * JCH + threadIdx.x;
    if (r >= L) return;
  } else {
    base = Z; ld = nv; r = (int64_t)(ch - nchW) * JCH + threadIdx.x;
    if (r >= nv) return;
  }
  double x[J2];
#pragma unroll
  for (int c = 0; c < J2; ++c) {
    int64_t v = pair_vec(c, I, J);
    x[c] = v < nv ? base[v * ld + r] : 0.0;
  }
#pragma unroll
  for (int c = 0; c < J2; ++c) {
    int64_t v = pair_vec(c, I, J);
    if (v >= nv) continue;
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < J2; ++i) acc = fma(x[i], Q[i][c], acc);
    base[v * ld + r] = acc;
  }
}

// sigma_j = ||Wv_j|| (one block per vector)
__global__ void __launch_bounds__(256) vec_norms(const double* __restrict__ Wv, int64_t L,
                                                 double* __restrict__ sig) {
  __shared__ double red[4];
  const int64_t v = blockIdx.x;
  double acc = 0.0;
  for (int64_t r = threadIdx.x; r < L; r += 256) {
    double x = Wv[v * L + r];
    acc = fma(x, x, acc);
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) sig[v] = sqrt(red[0] + red[1] + red[2] + red[3]);
}

// perm[rank] = j with rank = #{i : s_i > s_j or (s_i == s_j and i < j)}
__global__ void __launch_bounds__(1024) rank_sort(const double* __restrict__ sig, int64_t k,
                                                  int* __restrict__ perm, double* __restrict__ S) {
  for (int64_t j = threadIdx.x; j < k; j += 1024) {
    double sj = sig[j];
    int64_t rank = 0;
    for (int64_t i = 0; i < k; ++i) {
      double si = sig[i];
      rank += (si > sj || (si == sj && i < j)) ? 1 : 0;
    }
    perm[rank] = (int)j;
    S[rank] = sj;
  }
}

// U (m x k row-major) and Vt (k x n row-major) from Wv / Z.
template <typename T>
__global__ void svd_out_vec(const double* __restrict__ Wv, int64_t L, const double* __restrict__ Z,
                            int64_t k, const int* __restrict__ perm, const double* __restrict__ S,
                            int tall, T* __restrict__ U, T* __restrict__ Vt, int64_t m,
                            int64_t n) {
  // tall (m >= n): U[i, j] = Wv[perm j][i] / S[j];  Vt[j, c] = Z[perm j][c]
  // wide (m <  n): U[i, j] = Z[perm j][i];           Vt[j, c] = Wv[perm j][c] / S[j]
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = m * k + k * n;
  if (idx >= total) return;
  if (idx < m * k) {
    int64_t i = idx / k, j = idx - i * k;
    int64_t src = perm[j];
    double v;
    if (tall) {
      double s = S[j];
      v = s > 0.0 ? Wv[src * L + i] / s : 0.0;
    } else {
      v = Z[src * k + i];
    }
    U[idx] = (T)v;
  } else {
    int64_t e = idx - m * k;
    int64_t j = e / n, c = e - j * n;
    int64_t src = perm[j];
    double v;
    if (tall) {
      v = Z[src * k + c];
    } else {
      double s = S[j];
      v = s > 0.0 ? Wv[src * L + c] / s : 0.0;
    }
    Vt[e] = (T)v;
  }
}

template <typename T>
__global__ void gather_vectors(const T* __restrict__ src, Modes rows, Modes cols, int64_t m,
                               int64_t n, int tall, double* __restrict__ Wv) {
  // tall: Wv[j*m + i] = A(i,j)   wide: Wv[i*n + j] = A(i,j)
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; idx < m * n; idx += stride) {
    int64_t i, j;
    if (tall) { j = idx / m; i = idx - j * m; } else { i = idx / n; j = idx - i * n; }
    Wv[idx] = (double)src[mode_offset(rows, i) + mode_offset(cols, j)];
  }
}

__global__ void init_identity(double* __restrict__ Z, int64_t k) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= k * k) return;
  Z[idx] = (idx / k == idx % k) ? 1.0 : 0.0;
}

struct SvdWs {
  double *Wv, *Z, *Gpart, *Q, *sig, *S;
  int *skip, *perm;
  unsigned long long* off;
  int nblk, npairs, nchW, nchZ, nch;
};

static size_t svd_layout(int64_t m, int64_t n, char* base, size_t cap, SvdWs* out) {
  const int64_t k = std::min(m, n), L = std::max(m, n);
  SvdWs w{};
  w.nblk = (int)ceil_div(std::max<int64_t>(k, 1), JB);
  if (w.nblk % 2) w.nblk += 1;
  w.npairs = w.nblk / 2;
  w.nchW = (int)ceil_div(L, JCH);
  w.nchZ = (int)ceil_div(k, JCH);
  w.nch = w.nchW;
  Arena a{base, cap, 0};
  w.Wv = a.take<double>((size_t)(L * k));
  w.Z = a.take<double>((size_t)(k * k));
  w.Gpart = a.take<double>((size_t)w.npairs * w.nchW * J2 * J2);
  w.Q = a.take<double>((size_t)w.npairs * J2 * J2);
  w.sig = a.take<double>((size_t)k);
  w.S = a.take<double>((size_t)k);
  w.skip = a.take<int>((size_t)w.npairs);
  w.perm = a.take<int>((size_t)k);
  w.off = a.take<unsigned long long>(4);
  if (out) *out = w;
  return a.used;
}

}  // namespace tn

using namespace tn;

extern "C" size_t tn_svd_workspace(int64_t m, int64_t n, int dtype) {
  (void)dtype;
  if (m <= 0 || n <= 0) return 256;
  return svd_layout(m, n, nullptr, 0, nullptr) + 256;
}

extern "C" int tn_svd(const void* A, const tn_modes* a_rows, const tn_modes* a_cols, int64_t m,
                      int64_t n, void* U, double* S, void* Vt, int dtype, double tol,
                      int max_sweeps, int* sweeps_out, void* ws, size_t ws_bytes, void* stream) {
  TN_REQUIRE(m >= 0 && n >= 0, "tn_svd: negative size");
  TN_REQUIRE(dtype == TN_F64 || dtype == TN_F32, "tn_svd: unsupported dtype %d", dtype);
  Modes rows = to_modes(a_rows), cols = to_modes(a_cols);
  TN_REQUIRE(modes_size(rows) == m && modes_size(cols) == n, "tn_svd: modes do not match m x n");
  const int64_t k = std::min(m, n), L = std::max(m, n);
  if (sweeps_out) *sweeps_out = 0;
  if (k == 0) return TN_OK;
  TN_REQUIRE(A && U && S && Vt, "tn_svd: null pointer");
  TN_REQUIRE(ws && ws_bytes >= tn_svd_workspace(m, n, dtype), "tn_svd: workspace too small");
  TN_REQUIRE(k <= (int64_t)1 << 30, "tn_svd: too large");
  if (tol <= 0.0) tol = std::max(1e-15, std::sqrt((double)L) * 2.220446049250313e-16);
  if (max_sweeps <= 0) max_sweeps = 40;
  hipStream_t st = as_stream(stream);
  SvdWs w{};
  svd_layout(m, n, (char*)ws, ws_bytes, &w);
  const int tall = m >= n ? 1 : 0;
  unsigned gb = (unsigned)std::min<int64_t>(ceil_div(m * n, 256), 65536);
  if (dtype == TN_F64)
    hipLaunchKernelGGL(gather_vectors<double>, dim3(gb), dim3(256), 0, st, (const double*)A, rows,
                       cols, m, n, tall, w.Wv);
  else
    hipLaunchKernelGGL(gather_vectors<float>, dim3(gb), dim3(256), 0, st, (const float*)A, rows,
                       cols, m, n, tall, w.Wv);
  TN_CHECK_LAUNCH();
  hipLaunchKernelGGL(init_identity, dim3((unsigned)ceil_div(k * k, 256)), dim3(256), 0, st, w.Z, k);
  TN_CHECK_LAUNCH();
  int sweep = 0;
  bool converged = false;
  unsigned long long host_off = 0;
  for (; sweep < max_sweeps; ++sweep) {
    TN_HIP(hipMemsetAsync(w.off, 0, sizeof(unsigned long long), st));
    for (int s = 0; s < w.nblk - 1; ++s) {
      hipLaunchKernelGGL(jac_gram, dim3(w.npairs, w.nchW), dim3(256), 0, st, w.Wv, L, k, s,
                         w.nblk, w.nchW, w.Gpart);
      hipLaunchKernelGGL(jac_eig, dim3(w.npairs), dim3(256), 0, st, w.Gpart, w.nchW, k, s, w.nblk,
                         tol, w.Q, w.skip, w.off);
      hipLaunchKernelGGL(jac_apply, dim3(w.npairs, w.nchW + w.nchZ), dim3(256), 0, st, w.Wv, L,
                         w.Z, k, s, w.nblk, w.nchW, w.Q, w.skip);
      TN_CHECK_LAUNCH();
    }
    TN_HIP(hipMemcpyAsync(&host_off, w.off, sizeof(host_off), hipMemcpyDeviceToHost, st));
    TN_HIP(hipStreamSynchronize(st));
    double off;
    memcpy(&off, &host_off, sizeof(off));
    if (off <= tol) {
      converged = true;
      ++sweep;
      break;
    }
  }
  if (sweeps_out) *sweeps_out = sweep;
  hipLaunchKernelGGL(vec_norms, dim3((unsigned)k), dim3(256), 0, st, w.Wv, L, w.sig);
  hipLaunchKernelGGL(rank_sort, dim3(1), dim3(1024), 0, st, w.sig, k, w.perm, S);
  TN_CHECK_LAUNCH();
  unsigned ob = (unsigned)ceil_div(m * k + k * n, 256);
  if (dtype == TN_F64)
    hipLaunchKernelGGL(svd_out_vec<double>, dim3(ob), dim3(256), 0, st, w.Wv, L, w.Z, k, w.perm, S,
                       tall, (double*)U, (double*)Vt, m, n);
  else
    hipLaunchKernelGGL(svd_out_vec<float>, dim3(ob), dim3(256), 0, st, w.Wv, L, w.Z, k, w.perm, S,
                       tall, (float*)U, (float*)Vt, m, n);
  TN_CHECK_LAUNCH();
  if (!converged) {
    set_error(TN_ERR_NOCONV, "tn_svd: Jacobi did not converge in %d sweeps", max_sweeps);
    return TN_ERR_NOCONV;
  }
  return TN_OK;
}
