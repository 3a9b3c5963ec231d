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
//   2. jac_eig_apply: every block of a pair sums the partials, records the
//                 largest normalised off-diagonal |G_pq|/sqrt(G_pp G_qq),
//                 runs one cyclic sweep of two-sided Jacobi on the 32x32 G
//                 in LDS accumulating Q (the same Q in every block), and
//                 applies [W_I W_J] <- [W_I W_J] Q (or [Z_I Z_J] Q) to its
//                 row chunk by MFMA (grid: pairs x row chunks).
// The two kernels of every step of a sweep are captured once as a HIP graph
// and replayed.
// A sweep is nblk-1 steps; sweeps repeat until the largest normalised
// off-diagonal of a whole sweep is <= tol.  Singular values are the vector
// norms; they are sorted descending on the device.
#include "tn_internal.h"
#include "tn_jacobi.h"
#include "tn_prof.h"


#include <cmath>
#include <mutex>
#include <cstring>
#include <vector>

namespace tn {

// Rows per chunk of jac_eig_apply: 64 NT with NT = 2 (128 rows), or NT = 4
// (256 rows) when 128-row chunks would put more than two blocks on a CU:
// the blocks' rotation rounds then share the SIMDs' issue slots and stop
// being latency-bound (n = 2048: 64 pairs x 8 chunks of 256 instead of 16
// chunks of 128).
constexpr int kEigResident = 512;
static int eig_nt(int npairs, int64_t L, int64_t kz) {
  return (int64_t)npairs * (ceil_div(L, 128) + ceil_div(kz, 128)) <= kEigResident ? 2 : 4;
}
constexpr int GCH = 128;     // rows per Gram partial of jac_gram (twice the blocks)

// Partial Gram P^T P of the pair's 32 vectors over rows [r0, r1) into out
// (32 x 32, row stride ostride; global or LDS), v_mfma_f64_16x16x4 from a
// double-buffered LDS tile.  256 threads.
__device__ void gram_partial(double (*Ps)[64][J2 + 1], const double* __restrict__ Wv, int64_t L,
                             int64_t nv, int I, int J, int64_t r0, int64_t r1,
                             double* __restrict__ out, int ostride = J2) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
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
  lds_barrier();
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
    lds_barrier();
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    int r = 16 * ti + (lane >> 4) + 4 * e, c = 16 * tj + (lane & 15);
    out[r * ostride + c] = acc[e];
  }
}

// G (sm.G[0]) = sum of the pair's nchunk partials (fixed order, so every
// block that sums them gets the same bits); Q (sm.Q[0]) = I.  The partials
// are read four chunks at a time with every load issued before the first
// add: a load-add chain would pay one memory latency per partial element.
__device__ void eig_load(EigLds& sm, const double* __restrict__ Gp, int nchunk) {
  constexpr int PER = J2 * J2 / 256;  // elements per thread
  double acc[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) acc[i] = 0.0;
  int c = 0;
  for (; c + 4 <= nchunk; c += 4) {
    double v[4][PER];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int i = 0; i < PER; ++i)
        v[u][i] = Gp[(int64_t)(c + u) * (J2 * J2) + threadIdx.x + 256 * i];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int i = 0; i < PER; ++i) acc[i] += v[u][i];
  }
  for (; c < nchunk; ++c) {
    double v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) v[i] = Gp[(int64_t)c * (J2 * J2) + threadIdx.x + 256 * i];
#pragma unroll
    for (int i = 0; i < PER; ++i) acc[i] += v[i];
  }
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = threadIdx.x + 256 * i;
    sm.G[0][e / J2][e % J2] = acc[i];
    sm.Q[0][e / J2][e % J2] = (e / J2 == e % J2) ? 1.0 : 0.0;
  }
  lds_barrier();
}

__global__ void __launch_bounds__(256) jac_gram(const double* __restrict__ Wv, int64_t L,
                                                int64_t nv, int s, int nblk, int nchunk,
                                                double* __restrict__ Gpart) {
  __shared__ double Ps[2][64][J2 + 1];
  const int pi = blockIdx.x, ch = blockIdx.y;
  int I, J;
  pair_blocks(pi, s, nblk, I, J);
  const int64_t r0 = (int64_t)ch * GCH;
  const int64_t r1 = min(L, r0 + GCH);
  gram_partial(Ps, Wv, L, nv, I, J, r0, r1, Gpart + ((int64_t)pi * nchunk + ch) * (J2 * J2));
}

// Pair eigensolve + apply, grid pairs x (row chunks of Wv, then of Z):
// every block of a pair sums the pair's Gram partials in the same order,
// records the normalised off-diagonal (chunk 0 only), runs the same inner
// Jacobi sweep in LDS (bit-identical Q in all of them) and applies Q to its
// own row chunk by MFMA.  Computing the 32 x 32 eigensolve redundantly per
// chunk costs nothing in latency and saves a launch and the Q round trip
// through memory.
template <int NT>
__global__ void __launch_bounds__(256) jac_eig_apply(double* __restrict__ Wv, int64_t L,
                                                     double* __restrict__ Z, int64_t nv, int s,
                                                     int nblk, int nchW, int nchG,
                                                     const double* __restrict__ Gpart, double tol,
                                                     unsigned long long* __restrict__ offmax) {
  constexpr int CH = 64 * NT;
  __shared__ EigLds sm;
  const int pi = blockIdx.x, ch = blockIdx.y;
  int I, J;
  pair_blocks(pi, s, nblk, I, J);
  double* base;
  int64_t ld, rs, re;
  if (ch < nchW) {
    base = Wv; ld = L; rs = (int64_t)ch * CH; re = min(L, rs + CH);
  } else {
    base = Z; ld = nv; rs = (int64_t)(ch - nchW) * CH; re = min(nv, rs + CH);
  }
  eig_load(sm, Gpart + (int64_t)pi * nchG * (J2 * J2), nchG);
  double x[8][NT];
  apply_rows_fetch<NT>(base, ld, nv, rs, re, I, J, x);   // in flight during the rounds
  const double off = eig_off(sm, tol);
  if (ch == 0 && threadIdx.x == 0) atomicMax(offmax, dbits(off));
  if (off <= tol) return;
  const int cur = eig_rotate(sm, tol);
  apply_rows_finish<NT>(base, ld, nv, rs, re, I, J, x, sm.Q[cur]);
}

// Convergence of the whole vector set from its Gram matrix G = Wv^T Wv
// (k x k, either triangle): eig_off's measure over every pair (p < q), so
// "off <= tol" holds exactly when a sweep would rotate no pair of G.
__global__ void __launch_bounds__(256) gram_off(const double* __restrict__ G, int64_t k,
                                                double tol, unsigned long long* __restrict__ offmax) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const float above = (float)tol * (1.0f + 0x1p-20f) + 0x1p-126f;
  float off = 0.0f;
  if (e < k * k) {
    const int64_t p = e / k, q = e - p * k;
    if (p < q) {
      double spq, spp, sqq;
      const bool rot = pair_rotates(G[e], G[p * k + p], G[q * k + q], tol, 0.0, spq, spp, sqq);
      const float v = fminf(fabsf((float)spq) * __builtin_amdgcn_rsqf((float)(spp * sqq)), 1e30f);
      if (rot) off = fmaxf(v, above);
    }
  }
  off = wave_max(off);
  if ((threadIdx.x & 63) == 0 && off > 0.0f) atomicMax(offmax, dbits((double)off));
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

// perm[rank] = j with rank = #{i : s_i > s_j or (s_i == s_j and i < j)};
// NaN ranks as -inf so that the ranks stay a permutation of 0..k-1.
__device__ __forceinline__ double rank_key(double x) { return x == x ? x : -INFINITY; }

// 64 elements j per block of 256 threads (k = 1024 spreads over 16 CUs;
// 1024-thread blocks of 256 elements kept it on 4 and took 15.5 us): the
// four lanes l, l + 16, l + 32, l + 48 of a wave count over every fourth key
// of an LDS tile of 2048 (interleaved, so each reads from LDS in step with
// the others) and their counts are summed by shuffles; the comparisons are
// unrolled by eight so that the LDS reads of a group are in flight together.
constexpr int RANK_TILE = 2048;
constexpr int RANK_NT = 256;
__global__ void __launch_bounds__(RANK_NT) rank_sort(const double* __restrict__ sig, int64_t k,
                                                     int* __restrict__ perm,
                                                     double* __restrict__ S) {
  __shared__ double tile[RANK_TILE];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane >> 4;
  const int64_t j = (int64_t)blockIdx.x * (RANK_NT / 4) + wave * 16 + (lane & 15);
  const double sj = j < k ? rank_key(sig[j]) : 0.0;
  int cnt = 0;
  for (int64_t t0 = 0; t0 < k; t0 += RANK_TILE) {
    const int tn = (int)min<int64_t>(RANK_TILE, k - t0);
    __syncthreads();
    for (int i = threadIdx.x; i < tn; i += RANK_NT) tile[i] = rank_key(sig[t0 + i]);
    __syncthreads();
    // this lane's keys: i = q + 4 m
    const int64_t jr = j - t0;  // key i precedes j on ties when i < jr
    int i = q;
    for (; i + 28 < tn; i += 32) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = tile[i + 4 * u];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        cnt += (v[u] > sj || (v[u] == sj && i + 4 * u < jr)) ? 1 : 0;
    }
    for (; i < tn; i += 4) {
      const double si = tile[i];
      cnt += (si > sj || (si == sj && i < jr)) ? 1 : 0;
    }
  }
  cnt += __shfl_xor(cnt, 16, 64);
  cnt += __shfl_xor(cnt, 32, 64);
  if (j < k && q == 0) {
    perm[cnt] = (int)j;
    S[cnt] = sig[j];
  }
}

static void launch_rank_sort(const double* sig, int64_t k, int* perm, double* S, hipStream_t st) {
  hipLaunchKernelGGL(rank_sort, dim3((unsigned)ceil_div(std::max<int64_t>(k, 1), RANK_NT / 4)),
                     dim3(RANK_NT), 0, st, sig, k, perm, S);
}

// U (m x k row-major) and Vt (k x n row-major).  Column j of U is vector
// perm[j] of Uvec (ld ldu), row j of Vt is vector perm[j] of Vvec (ld ldv);
// a "scaled" source holds sigma_j * (unit vector) and is divided by S[j].
// upm / vpm (nullable): component i of a U source vector is row upm[i] of U,
// component c of a V source vector is column vpm[c] of Vt (the pivot order
// of the preconditioned path).
template <typename T>
__global__ void svd_out_kernel(const double* __restrict__ Uvec, int64_t ldu, int u_scaled,
                               const double* __restrict__ Vvec, int64_t ldv, int v_scaled,
                               int64_t k, const int* __restrict__ perm,
                               const double* __restrict__ S, const int* __restrict__ upm,
                               const int* __restrict__ vpm, T* __restrict__ U,
                               T* __restrict__ Vt, int64_t m, int64_t n) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= m * k + k * n) return;
  if (idx < m * k) {
    int64_t i = idx / k, j = idx - i * k;
    double v = Uvec[(int64_t)perm[j] * ldu + i];
    if (u_scaled) v = S[j] > 0.0 ? v / S[j] : 0.0;
    const int64_t row = upm ? (int64_t)upm[i] : i;
    U[row * k + j] = (T)v;
  } else {
    int64_t e = idx - m * k;
    int64_t j = e / n, c = e - j * n;
    double v = Vvec[(int64_t)perm[j] * ldv + c];
    if (v_scaled) v = S[j] > 0.0 ? v / S[j] : 0.0;
    const int64_t col = vpm ? (int64_t)vpm[c] : c;
    Vt[j * n + col] = (T)v;
  }
}

// Xp[c*L + i] = X[p[c]*L + i]  (k vectors of length L)
__global__ void permute_vectors(const double* __restrict__ X, int64_t L, int64_t k,
                                const int* __restrict__ p, double* __restrict__ Xp) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= L * k) return;
  const int64_t c = idx / L, i = idx - c * L;
  Xp[idx] = X[(int64_t)p[c] * L + i];
}

// Rows of the upper-triangular R (in the column-major QR workspace, ld = L)
// as contiguous vectors: V[j*k + c] = R[j][c].
__global__ void r_rows_kernel(const double* __restrict__ W, int64_t L, int64_t k,
                              double* __restrict__ V) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= k * k) return;
  int64_t j = idx / k, c = idx - j * k;
  V[idx] = c >= j ? W[j + c * L] : 0.0;
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

__global__ void reset_u64(unsigned long long* p) { *p = 0ull; }

__global__ void init_identity(double* __restrict__ Z, int64_t k) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= k * k) return;
  Z[idx] = (idx / k == idx % k) ? 1.0 : 0.0;
}

struct SvdWs {
  double *Wv, *Z, *Gpart, *sig, *S, *R, *QZ, *Rtmp;
  int *skip, *perm, *vp;
  unsigned long long* off;
  QrWs qr;
  char* cq;  // Cholesky QR workspace
  size_t cq_bytes;
  int precond;
};

// Precondition with a QR (static norm pivot, then Jacobi on the rows of R)
// whenever there are at least 32 vectors: besides shortening long vectors, it
// grades the problem, so square matrices with spread spectra (bond matrices)
// need far fewer sweeps (22 -> 8 for a 128 x 128 matrix with singular values
// from 1 to 1e-6, 10 -> 9 for a Gaussian one, in a simulation of this block
// Jacobi).
static bool use_precond(int64_t L, int64_t k) { return k >= 32 && L >= k; }

static size_t svd_layout(int64_t m, int64_t n, char* base, size_t cap, SvdWs* out) {
  const int64_t k = std::min(m, n), L = std::max(m, n);
  SvdWs w{};
  w.precond = use_precond(L, k) ? 1 : 0;
  const int64_t Lj = w.precond ? k : L;  // vector length inside the Jacobi sweeps
  const int nblk = (int)ceil_div(std::max<int64_t>(k, 1), JB) + ((ceil_div(std::max<int64_t>(k, 1), JB)) % 2);
  const int npairs = nblk / 2;
  Arena a{base, cap, 0};
  w.Wv = a.take<double>((size_t)(L * k));
  w.Z = a.take<double>((size_t)(k * k));
  w.Gpart = a.take<double>((size_t)npairs * ceil_div(Lj, GCH) * J2 * J2);
  w.sig = a.take<double>((size_t)k);
  w.S = a.take<double>((size_t)k);
  w.skip = a.take<int>((size_t)npairs);
  w.perm = a.take<int>((size_t)k);
  w.off = a.take<unsigned long long>(4);
  if (w.precond) {
    w.R = a.take<double>((size_t)(k * k));
    w.Rtmp = a.take<double>((size_t)(k * k));
    w.QZ = a.take<double>((size_t)(L * k));
    w.vp = a.take<int>((size_t)k);
    size_t qbytes = qr_layout(L, k, nullptr, 0, nullptr) + 256;
    char* qbase = a.take<char>(qbytes);
    if (base) qr_layout(L, k, qbase, qbytes, &w.qr);
    w.cq_bytes = cholqr_ws_bytes(L, k);
    w.cq = a.take<char>(w.cq_bytes);
  }
  if (out) *out = w;
  return a.used;
}

__global__ void copy_f64(const double* __restrict__ src, double* __restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// One-sided block Jacobi on k vectors of length L (Wv, vector-contiguous),
// accumulating the rotations into Z (k x k, vector-contiguous).  sort_tmp
// (max(L, k) x k doubles, or nullptr): the vectors, and Z's columns with
// them, are put in descending norm order at the start of every sweep, so that
// the blocks pair vectors of similar length (one sweep fewer at n = 256 ..
// 1024, simulated on the Cholesky factors of random PSD matrices: 13 -> 12
// at n = 1024).
static int jacobi_run(double* Wv, int64_t L, double* Z, int64_t k, const SvdWs& w, double tol,
                      int max_sweeps, int* sweeps_out, bool* converged, hipStream_t st,
                      bool allow_graph = true, double* sort_tmp = nullptr) {
  int nblk = (int)ceil_div(std::max<int64_t>(k, 1), JB);
  if (nblk % 2) nblk += 1;
  const int npairs = nblk / 2;
  // Z == nullptr: the rotations are not accumulated (eigensolver use)
  const int64_t kz0 = Z ? k : 0;
  const int ent = eig_nt(npairs, L, kz0);
  const int nchW = (int)ceil_div(L, 64 * ent), nchZ = Z ? (int)ceil_div(k, 64 * ent) : 0;
  const int nchG = (int)ceil_div(L, GCH);
  const int64_t kz = Z ? k : 0;
  if (Z)
    hipLaunchKernelGGL(init_identity, dim3((unsigned)ceil_div(k * k, 256)), dim3(256), 0, st, Z, k);
  TN_CHECK_LAUNCH();
  int sweep = 0;
  *converged = false;
  // algorithmic work per step: Gram 2 L (2b)^2 per pair, apply 2 (L + k) (2b)^2 per pair
  const double gram_fl = 2.0 * (double)L * J2 * J2 * npairs;
  const double apply_fl = 2.0 * (double)(L + kz) * J2 * J2 * npairs;
  unsigned long long host_off = 0;
  // the sort's six launches per sweep pay off only where a sweep is long
  // (and graph-replayed): at least 16 steps, k >= 256
  if (nblk - 1 < 16) sort_tmp = nullptr;
  auto enqueue_sweep = [&]() -> int {
    hipLaunchKernelGGL(reset_u64, dim3(1), dim3(1), 0, st, w.off);
    if (sort_tmp) {
      const unsigned ge = (unsigned)ceil_div(L * k, 256);
      hipLaunchKernelGGL(vec_norms, dim3((unsigned)k), dim3(256), 0, st, Wv, L, w.sig);
      launch_rank_sort(w.sig, k, w.perm, w.S, st);
      hipLaunchKernelGGL(permute_vectors, dim3(ge), dim3(256), 0, st, Wv, L, k, w.perm, sort_tmp);
      hipLaunchKernelGGL(copy_f64, dim3(ge), dim3(256), 0, st, sort_tmp, Wv, L * k);
      if (Z) {
        const unsigned gz = (unsigned)ceil_div(k * k, 256);
        hipLaunchKernelGGL(permute_vectors, dim3(gz), dim3(256), 0, st, Z, k, k, w.perm, sort_tmp);
        hipLaunchKernelGGL(copy_f64, dim3(gz), dim3(256), 0, st, sort_tmp, Z, k * k);
      }
    }
    for (int s = 0; s < nblk - 1; ++s) {
      {
        ProfScope p(PROF_JAC_GRAM, st, gram_fl, 8.0 * (double)L * J2 * npairs);
        hipLaunchKernelGGL(jac_gram, dim3(npairs, nchG), dim3(256), 0, st, Wv, L, k, s, nblk,
                           nchG, w.Gpart);
      }
      {
        // a kernel boundary between the partials and the pair eigensolve is
        // cheaper than a device-scope fence + arrival counter in the Gram
        // kernel (measured).  Work: the inner sweep (31 rounds x 9216 flops)
        // once per pair plus the apply; HBM bytes: each pair's partials once
        // (the other row blocks of the pair re-read them from L2) + the
        // chunks read and written
        ProfScope p(PROF_JAC_EIG, st, 31.0 * 9216.0 * npairs + apply_fl,
                    8.0 * (double)nchG * J2 * J2 * npairs + 16.0 * (double)(L + kz) * J2 * npairs);
        if (ent == 2)
          hipLaunchKernelGGL(jac_eig_apply<2>, dim3(npairs, nchW + nchZ), dim3(256), 0, st, Wv,
                             L, Z, k, s, nblk, nchW, nchG, w.Gpart, tol, w.off);
        else
          hipLaunchKernelGGL(jac_eig_apply<4>, dim3(npairs, nchW + nchZ), dim3(256), 0, st, Wv,
                             L, Z, k, s, nblk, nchW, nchG, w.Gpart, tol, w.off);
      }
    }
    TN_CHECK_LAUNCH();
    return TN_OK;
  };
  // A sweep is the same launch sequence every time: capture it once as a HIP
  // graph and replay it (long sweeps only; not while launch timing is on).
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  if (allow_graph && nblk - 1 >= 16 && !prof_on()) {
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
      int rc = enqueue_sweep();
      hipError_t e = hipStreamEndCapture(st, &graph);
      if (rc != TN_OK || e != hipSuccess ||
          hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        graph = nullptr;
        exec = nullptr;
        (void)hipGetLastError();
      }
    } else {
      (void)hipGetLastError();
    }
  }
  int rc_loop = TN_OK;
  for (; sweep < max_sweeps; ++sweep) {
    if (exec) {
      if (hipGraphLaunch(exec, st) != hipSuccess) {
        set_error(TN_ERR_HIP, "tn_svd: graph launch failed");
        rc_loop = TN_ERR_HIP;
        break;
      }
    } else {
      rc_loop = enqueue_sweep();
      if (rc_loop != TN_OK) break;
    }
    if (hipMemcpyAsync(&host_off, w.off, sizeof(host_off), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      set_error(TN_ERR_HIP, "tn_svd: convergence readback failed");
      rc_loop = TN_ERR_HIP;
      break;
    }
    double off;
    memcpy(&off, &host_off, sizeof(off));
    if (off <= tol) {
      *converged = true;
      ++sweep;
      break;
    }
    // Once the measure is near sqrt(tol), this sweep has most likely
    // converged the set (the off-diagonal is about squared per sweep there,
    // at most ~10 off^2 in simulations): check that on the Gram matrix of the
    // rotated vectors (one GEMM and one reduction, a few percent of a sweep)
    // instead of running a confirming sweep that rotates nothing.
    if (sort_tmp && off * off <= 10.0 * tol) {
      rc_loop = contract_launch(Wv, mk({k}, {L}), mk({L}, {1}), Wv, mk({L}, {1}), mk({k}, {L}),
                                sort_tmp, mk({k}, {k}), mk({k}, {1}), TN_F64, TN_PREC_NATIVE, 1.0,
                                0.0, nullptr, 0, st);
      if (rc_loop != TN_OK) break;
      hipLaunchKernelGGL(reset_u64, dim3(1), dim3(1), 0, st, w.off + 1);
      hipLaunchKernelGGL(gram_off, dim3((unsigned)ceil_div(k * k, 256)), dim3(256), 0, st,
                         sort_tmp, k, tol, w.off + 1);
      if (hipGetLastError() != hipSuccess ||
          hipMemcpyAsync(&host_off, w.off + 1, sizeof(host_off), hipMemcpyDeviceToHost, st) !=
              hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess) {
        set_error(TN_ERR_HIP, "tn_svd: convergence readback failed");
        rc_loop = TN_ERR_HIP;
        break;
      }
      memcpy(&off, &host_off, sizeof(off));
      if (off <= tol) {
        *converged = true;
        ++sweep;
        break;
      }
    }
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (rc_loop != TN_OK) return rc_loop;
  if (sweeps_out) *sweeps_out = sweep;
  return TN_OK;
}

}  // namespace tn

using namespace tn;

extern "C" size_t tn_svd_workspace(int64_t m, int64_t n, int dtype) {
  (void)dtype;
  if (m <= 0 || n <= 0) return 256;
  return svd_layout(m, n, nullptr, 0, nullptr) + 256;
}

namespace tn {

// tn_svd with the graph replay of the Jacobi sweeps optional (tn_svd_noalloc
// leaves it off: instantiating a graph is not free of allocations).
static int svd_impl(const void* A, const tn_modes* a_rows, const tn_modes* a_cols, int64_t m,
                    int64_t n, void* U, double* S, void* Vt, int dtype, double tol,
                    int max_sweeps, int* sweeps_out, void* ws, size_t ws_bytes, void* stream,
                    bool allow_graph) {
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
  if (max_sweeps <= 0) max_sweeps = 40;
  hipStream_t st = as_stream(stream);
  SvdWs w{};
  svd_layout(m, n, (char*)ws, ws_bytes, &w);
  const int tall = m >= n ? 1 : 0;
  // X (L x k, column-major) holds the vectors: columns of A (tall) or rows (wide)
  unsigned gb = (unsigned)std::min<int64_t>(ceil_div(m * n, 256), 65536);
  if (dtype == TN_F64)
    hipLaunchKernelGGL(gather_vectors<double>, dim3(gb), dim3(256), 0, st, (const double*)A, rows,
                       cols, m, n, tall, w.Wv);
  else
    hipLaunchKernelGGL(gather_vectors<float>, dim3(gb), dim3(256), 0, st, (const float*)A, rows,
                       cols, m, n, tall, w.Wv);
  TN_CHECK_LAUNCH();
  bool converged = false;
  const double* Uvec;
  const double* Vvec;
  int64_t ldu, ldv;
  int u_scaled, v_scaled;
  const double* orth_src;  // vectors whose norms are the singular values
  int64_t orth_len;
  const int* upm = nullptr;
  const int* vpm = nullptr;
  if (w.precond) {
    // Static pivot: X' = X P with the vectors by descending norm (R R^T is
    // then closer to diagonal and the Jacobi needs fewer sweeps).
    // X' = Q R; one-sided Jacobi on the rows of R (R^T Z = W''), then
    // X' = (Q Z) Sigma (W''/Sigma)^T and X = X' P^T.
    hipLaunchKernelGGL(vec_norms, dim3((unsigned)k), dim3(256), 0, st, w.Wv, L, w.sig);
    launch_rank_sort(w.sig, k, w.vp, w.S, st);
    hipLaunchKernelGGL(permute_vectors, dim3((unsigned)ceil_div(L * k, 256)), dim3(256), 0, st,
                       w.Wv, L, k, w.vp, w.QZ);
    TN_CHECK_LAUNCH();
    // The factorisation only has to give an orthonormal Q with X' = Q R
    // (any such R has the singular values and right vectors of X'): Cholesky
    // QR with a Newton-Schulz correction (a few GEMMs; Q into Wv, R row-major,
    // i.e. as row vectors) when X' is well enough conditioned for it to reach
    // orthogonality 1e-13, blocked Householder otherwise.
    const int cq = cholqr_run(w.QZ, mk({L}, {1}), mk({k}, {L}), L, k, w.Wv, 1, L, w.R, 1e-13,
                              nullptr, w.cq, w.cq_bytes, st);
    if (cq != TN_OK && cq != TN_ERR_NOCONV) return cq;
    const bool householder = cq != TN_OK;
    if (householder) {
      TN_TRY(qr_factor(w.QZ, L, k, w.qr, st));
      hipLaunchKernelGGL(r_rows_kernel, dim3((unsigned)ceil_div(k * k, 256)), dim3(256), 0, st,
                         w.QZ, L, k, w.R);
      TN_CHECK_LAUNCH();
    }
    if (tol <= 0.0) tol = std::max(1e-15, (double)k * 2.220446049250313e-16);
    TN_TRY(jacobi_run(w.R, k, w.Z, k, w, tol, max_sweeps, sweeps_out, &converged, st,
                      allow_graph, w.Rtmp));
    double* QZv;  // Q Z (L x k, vector-contiguous)
    if (householder) {
      TN_TRY(qr_form_q(w.QZ, L, k, w.qr, w.Wv, st, w.Z));
      QZv = w.Wv;
    } else {
      TN_TRY(contract_launch(w.Wv, mk({L}, {1}), mk({k}, {L}), w.Z, mk({k}, {1}), mk({k}, {k}),
                             w.QZ, mk({L}, {1}), mk({k}, {L}), TN_F64, TN_PREC_NATIVE, 1.0, 0.0,
                             nullptr, 0, st));
      QZv = w.QZ;
    }
    orth_src = w.R;
    orth_len = k;
    // left vectors of X: Q Z (orthonormal); right vectors of X: rows of R
    // rotated, whose components are in pivot order
    if (tall) { Uvec = QZv; ldu = L; u_scaled = 0; Vvec = w.R; ldv = k; v_scaled = 1; vpm = w.vp; }
    else { Uvec = w.R; ldu = k; u_scaled = 1; Vvec = QZv; ldv = L; v_scaled = 0; upm = w.vp; }
  } else {
    // LAPACK dgesvj convergence threshold: (vector length) * eps
    if (tol <= 0.0) tol = std::max(1e-15, (double)L * 2.220446049250313e-16);
    TN_TRY(jacobi_run(w.Wv, L, w.Z, k, w, tol, max_sweeps, sweeps_out, &converged, st,
                      allow_graph));
    orth_src = w.Wv;
    orth_len = L;
    if (tall) { Uvec = w.Wv; ldu = L; u_scaled = 1; Vvec = w.Z; ldv = k; v_scaled = 0; }
    else { Uvec = w.Z; ldu = k; u_scaled = 0; Vvec = w.Wv; ldv = L; v_scaled = 1; }
  }
  hipLaunchKernelGGL(vec_norms, dim3((unsigned)k), dim3(256), 0, st, orth_src, orth_len, w.sig);
  launch_rank_sort(w.sig, k, w.perm, S, st);
  TN_CHECK_LAUNCH();
  unsigned ob = (unsigned)ceil_div(m * k + k * n, 256);
  if (dtype == TN_F64)
    hipLaunchKernelGGL(svd_out_kernel<double>, dim3(ob), dim3(256), 0, st, Uvec, ldu, u_scaled,
                       Vvec, ldv, v_scaled, k, w.perm, S, upm, vpm, (double*)U, (double*)Vt, m,
                       n);
  else
    hipLaunchKernelGGL(svd_out_kernel<float>, dim3(ob), dim3(256), 0, st, Uvec, ldu, u_scaled,
                       Vvec, ldv, v_scaled, k, w.perm, S, upm, vpm, (float*)U, (float*)Vt, m,
                       n);
  TN_CHECK_LAUNCH();
  if (!converged) {
    set_error(TN_ERR_NOCONV, "tn_svd: Jacobi did not converge in %d sweeps", max_sweeps);
    return TN_ERR_NOCONV;
  }
  return TN_OK;
}

// ---- scratch of the allocation-free SVD entry point (tn_svd_noalloc) ----
// One array per GPU, held by the loaded code object until the process ends.
// In host code `scratch` names a host-side stand-in, not the device array:
// its device address is fetched per device with hipGetSymbolAddress and
// only that address is handed to kernels.
constexpr size_t kSvdScratchBytes = (size_t)TN_SVD_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kSvdScratchBytes];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_dev[kMaxScratchDevices] = {};

// Device address of the current device's scratch array, looked up once per
// device (with that device current) and kept.
static int scratch_address(void** out) {
  int dev = -1;
  TN_HIP(hipGetDevice(&dev));
  TN_REQUIRE(dev >= 0 && dev < kMaxScratchDevices, "tn_svd_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error(TN_ERR_HIP, "tn_svd_noalloc: hipGetSymbolAddress failed on device %d: %s", dev,
                hipGetErrorString(e));
      return TN_ERR_HIP;
    }
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return TN_OK;
}

}  // namespace tn

extern "C" int tn_svd(const void* A, const tn_modes* a_rows, const tn_modes* a_cols, int64_t m,
                      int64_t n, void* U, double* S, void* Vt, int dtype, double tol,
                      int max_sweeps, int* sweeps_out, void* ws, size_t ws_bytes, void* stream) {
  return svd_impl(A, a_rows, a_cols, m, n, U, S, Vt, dtype, tol, max_sweeps, sweeps_out, ws,
                  ws_bytes, stream, true);
}

extern "C" int tn_svd_noalloc(const void* A, const tn_modes* a_rows, const tn_modes* a_cols,
                              int64_t m, int64_t n, void* U, double* S, void* Vt, int dtype,
                              double tol, int max_sweeps, int* sweeps_out, void* stream) {
  TN_REQUIRE(m >= 0 && n >= 0, "tn_svd_noalloc: negative size");
  const size_t need = tn_svd_workspace(m, n, dtype);
  if (need > kSvdScratchBytes) {
    set_error(TN_ERR_WORKSPACE,
              "tn_svd_noalloc: a %lld x %lld SVD needs %zu bytes of scratch, more than the "
              "%zu bytes built in",
              (long long)m, (long long)n, need, kSvdScratchBytes);
    return TN_ERR_WORKSPACE;
  }
  void* ws = nullptr;
  TN_TRY(scratch_address(&ws));
  return svd_impl(A, a_rows, a_cols, m, n, U, S, Vt, dtype, tol, max_sweeps, sweeps_out, ws,
                  kSvdScratchBytes, stream, false);
}

// ---------------------------------------------------------------------------
// Eigendecomposition of a symmetric positive semi-definite matrix,
// K = U diag(lam) U^T, for the truncation step of the structured
// apply+compress sweep, where only the bond's Gram matrix K = C^T E C is
// formed (the reference takes the SVD of C's left-isometric completion,
// onedim_core.py:317 -> tensor.py:915; K has the same left vectors and
// lam = sigma^2).
//
// Method:
//   1. P orders K by descending diagonal (a static pivot); shift
//      delta = 8 sqrt(n) eps_in tr(K), eps_in the input's unit roundoff;
//   2. P K P^T + delta I = L L^T, blocked right-looking Cholesky (64-column
//      blocks: the diagonal block and its inverse by one kernel, the panel
//      L21 = A21 L11^{-T} and the trailing update by MFMA GEMMs; a column
//      whose pivot falls below delta/2 is dropped);
//   3. one-sided block Jacobi on the columns of L (no rotation accumulation):
//      L Z = U' Sigma, hence P K P^T + delta I = U' Sigma^2 U'^T, so
//      U = P^T U' and lam = sigma^2 - delta.
// Jacobi on the Cholesky factor works on a matrix with the singular values
// sqrt(lam); Jacobi on K itself would work on K^2, whose condition number
// reaches the FP64 floor for truncated-MPS Gram matrices and needs about
// three times as many sweeps.  The shift leaves the eigenvectors unchanged.
// ---------------------------------------------------------------------------
namespace tn {

constexpr int CB = 64;  // Cholesky block

// diag[j] = A[j][j]; delta = 8 sqrt(n) eps_in sum_j |A[j][j]|
__global__ void __launch_bounds__(256) psd_diag(const double* __restrict__ A, int64_t n,
                                                double eps_in, double* __restrict__ diag,
                                                double* __restrict__ delta) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int64_t j = threadIdx.x; j < n; j += 256) {
    const double v = A[j * n + j];
    diag[j] = v;
    acc += fabs(v);
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
    *delta = 8.0 * sqrt((double)n) * eps_in * (red[0] + red[1] + red[2] + red[3]);
}

// B[c][r] = sym(A)[p[c]][p[r]] + delta [r == c]   (column-major, ld n);
// the symmetric part, since a K formed by rounded products is symmetric only
// to rounding
__global__ void psd_permute(const double* __restrict__ A, int64_t n, const int* __restrict__ p,
                            const double* __restrict__ delta, double* __restrict__ B) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const int64_t c = idx / n, r = idx - c * n;
  const int64_t pc = p[c], pr = p[r];
  double v = 0.5 * (A[pc * n + pr] + A[pr * n + pc]);
  if (r == c) v += *delta;
  B[idx] = v;
}

// U[p[i]][j] = Wv[perm[j]][i] / S[j];  lam[j] = S[j]^2 - delta
__global__ void eig_out_kernel(const double* __restrict__ Wv, int64_t n, const int* __restrict__ p,
                               const int* __restrict__ perm, const double* __restrict__ S,
                               const double* __restrict__ delta, double* __restrict__ U,
                               double* __restrict__ lam) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const int64_t i = idx / n, j = idx - i * n;
  const double s = S[j];
  U[(int64_t)p[i] * n + j] = s > 0.0 ? Wv[(int64_t)perm[j] * n + i] / s : 0.0;
  if (i == 0) lam[j] = s * s - *delta;
}

struct EigWs {
  SvdWs jw;  // Jacobi buffers (Wv holds L)
  double *B, *diag, *Sd, *delta, *X;
  int* p;
};

static size_t eig_layout(int64_t n, char* base, size_t cap, EigWs* out) {
  EigWs e{};
  const int64_t nb0 = ceil_div(std::max<int64_t>(n, 1), JB);
  const int nblk = (int)(nb0 + nb0 % 2);
  const int npairs = nblk / 2;
  Arena a{base, cap, 0};
  e.jw.Wv = a.take<double>((size_t)(n * n));
  e.B = a.take<double>((size_t)(n * n));
  e.diag = a.take<double>((size_t)n);
  e.Sd = a.take<double>((size_t)n);
  e.delta = a.take<double>(4);
  e.X = a.take<double>((size_t)CB * CB);
  e.p = a.take<int>((size_t)n);
  e.jw.Gpart = a.take<double>((size_t)npairs * ceil_div(n, GCH) * J2 * J2);
  e.jw.sig = a.take<double>((size_t)n);
  e.jw.S = a.take<double>((size_t)n);
  e.jw.skip = a.take<int>((size_t)npairs);
  e.jw.perm = a.take<int>((size_t)n);
  e.jw.off = a.take<unsigned long long>(4);
  if (out) *out = e;
  return a.used;
}

}  // namespace tn

extern "C" size_t tn_eigh_psd_workspace(int64_t n) {
  if (n <= 0) return 256;
  return eig_layout(n, nullptr, 0, nullptr) + 256;
}

extern "C" int tn_eigh_psd(const void* K, const tn_modes* k_rows, const tn_modes* k_cols, int64_t n,
                           double* U, double* lam, int dtype, double tol, int max_sweeps,
                           int* sweeps_out, void* ws, size_t ws_bytes, void* stream) {
  TN_REQUIRE(n >= 0, "tn_eigh_psd: negative size");
  TN_REQUIRE(dtype == TN_F64 || dtype == TN_F32, "tn_eigh_psd: unsupported dtype %d", dtype);
  Modes rows = to_modes(k_rows), cols = to_modes(k_cols);
  TN_REQUIRE(modes_size(rows) == n && modes_size(cols) == n,
             "tn_eigh_psd: modes do not match n x n");
  if (sweeps_out) *sweeps_out = 0;
  if (n == 0) return TN_OK;
  TN_REQUIRE(K && U && lam, "tn_eigh_psd: null pointer");
  TN_REQUIRE(ws && ws_bytes >= tn_eigh_psd_workspace(n), "tn_eigh_psd: workspace too small");
  TN_REQUIRE(n <= (int64_t)1 << 15, "tn_eigh_psd: too large");
  if (max_sweeps <= 0) max_sweeps = 40;
  hipStream_t st = as_stream(stream);
  EigWs e{};
  eig_layout(n, (char*)ws, ws_bytes, &e);
  double* Wv = e.jw.Wv;
  // 1. gather K column-major into Wv (scratch), diagonal, shift, static pivot
  unsigned gb = (unsigned)std::min<int64_t>(ceil_div(n * n, 256), 65536);
  if (dtype == TN_F64)
    hipLaunchKernelGGL(gather_vectors<double>, dim3(gb), dim3(256), 0, st, (const double*)K, rows,
                       cols, n, n, 1, Wv);
  else
    hipLaunchKernelGGL(gather_vectors<float>, dim3(gb), dim3(256), 0, st, (const float*)K, rows,
                       cols, n, n, 1, Wv);
  // shift allowance from the input's precision (K formed in FP32 is
  // indefinite at ~eps_32 |K|)
  const double eps_in = dtype == TN_F64 ? 2.220446049250313e-16 : 1.1920928955078125e-07;
  hipLaunchKernelGGL(psd_diag, dim3(1), dim3(256), 0, st, Wv, n, eps_in, e.diag, e.delta);
  launch_rank_sort(e.diag, n, e.p, e.Sd, st);
  const unsigned nb = (unsigned)ceil_div(n * n, 256);
  hipLaunchKernelGGL(psd_permute, dim3(nb), dim3(256), 0, st, Wv, n, e.p, e.delta, e.B);
  TN_CHECK_LAUNCH();
  TN_HIP(hipMemsetAsync(Wv, 0, sizeof(double) * (size_t)(n * n), st));
  // 2. blocked Cholesky: B (working copy) -> L in Wv
  for (int64_t j0 = 0; j0 < n; j0 += CB) {
    const int w = (int)std::min<int64_t>(CB, n - j0);
    // L11 and its inverse X in one kernel (a dropped column has a zero X row)
    TN_TRY(chol_block(e.B, n, j0, w, e.delta, Wv, e.X, st));
    const int64_t m2 = n - j0 - w;
    if (m2 <= 0) break;
    // L21 = A21 L11^{-T} = A21 X^T   (m2 x w):  B[k][c] = X[c][k] = X[k * CB + c]
    TN_TRY(contract_launch(e.B + j0 * n + j0 + w, mk({m2}, {1}), mk({w}, {n}), e.X, mk({w}, {CB}),
                           mk({w}, {1}), Wv + j0 * n + j0 + w, mk({m2}, {1}), mk({w}, {n}), TN_F64,
                           TN_PREC_NATIVE, 1.0, 0.0, nullptr, 0, st));
    // A22 -= L21 L21^T
    const double* L21 = Wv + j0 * n + j0 + w;
    TN_TRY(contract_launch(L21, mk({m2}, {1}), mk({w}, {n}), L21, mk({w}, {n}), mk({m2}, {1}),
                           e.B + (j0 + w) * n + j0 + w, mk({m2}, {1}), mk({m2}, {n}), TN_F64,
                           TN_PREC_NATIVE, -1.0, 1.0, nullptr, 0, st));
  }
  // 3. one-sided Jacobi on the columns of L (length n)
  if (tol <= 0.0) tol = std::max(1e-15, (double)n * 2.220446049250313e-16);
  bool converged = false;
  // (B, the Cholesky's working copy, is free now: the sort's scratch)
  TN_TRY(jacobi_run(Wv, n, nullptr, n, e.jw, tol, max_sweeps, sweeps_out, &converged, st, true,
                    e.B));
  hipLaunchKernelGGL(vec_norms, dim3((unsigned)n), dim3(256), 0, st, Wv, n, e.jw.sig);
  launch_rank_sort(e.jw.sig, n, e.jw.perm, e.jw.S, st);
  hipLaunchKernelGGL(eig_out_kernel, dim3(nb), dim3(256), 0, st, Wv, n, e.p, e.jw.perm, e.jw.S,
                     e.delta, U, lam);
  TN_CHECK_LAUNCH();
  if (!converged) {
    set_error(TN_ERR_NOCONV, "tn_eigh_psd: Jacobi did not converge in %d sweeps", max_sweeps);
    return TN_ERR_NOCONV;
  }
  return TN_OK;
}
