// Cholesky QR with a Newton-Schulz correction for a tall FP64 matrix:
// A = Q R with orthonormal Q.
//
// The variational sweep (onedim_core.py:560-631) re-orthogonalises one site
// per update through tensor_lq (tensor.py:1058-1109), whose only use there
// is the row space: Q becomes the site, L is absorbed into the neighbour, so
// any A = L Q with orthonormal Q gives the same state.  Householder QR of
// these 1024 x 512 blocks is a chain of latency-bound panel kernels; here:
//   G = A^T A = L L^T (blocked Cholesky),  Q1 = A L^{-T}     (Cholesky QR)
//   Q = Q1 (3 I - Q1^T Q1) / 2                              (Newton-Schulz)
//   R = Q^T A.
// Q1^T Q1 = I + E with |E| ~ cond(A)^2 eps, and the correction leaves
// O(|E|^2): orthogonal to rounding for cond(A) up to ~1e4-1e5.  R is upper
// triangular up to O(|E|) (R = (I + E/2) L^T + O(|E|^2)).  The driver
// checks ||Q^T Q - I||_max at the end and reports inputs outside that range
// (TN_ERR_NOCONV) so the caller can fall back to Householder.
#include "tn_internal.h"
#include "tn_prof.h"

#include <algorithm>
#include <cstring>

namespace tn {

int splitk_factor(int64_t M, int64_t N, int64_t K);

namespace {

constexpr int QB = 64;  // Cholesky block
constexpr int QC = 8;   // columns of the block per wave (eight waves)

// 1/sqrt(x) for FP64: the hardware estimate (v_rsq_f64) refined by two
// Newton steps, y <- y + y (1/2 - x y^2 / 2), each in fused form.
__device__ __forceinline__ double rsq_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  y = fma(y, fma(-hx * y, y, 0.5), y);
  return fma(y, fma(-hx * y, y, 0.5), y);
}

// S[n + i, j] = A(i, j): the m x n matrix under the n x n Gram block of the
// stacked (n + m) x n column-major buffer S (ld = n + m).
__global__ void cq_gather(const double* __restrict__ A, Modes rows, Modes cols, int64_t m,
                          int64_t n, double* __restrict__ S) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= m * n) return;
  const int64_t j = idx / m, i = idx - j * m;
  S[j * (n + m) + n + i] = A[mode_offset(rows, i) + mode_offset(cols, j)];
}

// pivmin = thr * max_i S[i, i] over the leading n x n block; one block of 256.
__global__ void __launch_bounds__(256) cq_pivmin(const double* __restrict__ S, int64_t ld,
                                                 int64_t n, double thr,
                                                 double* __restrict__ pivmin) {
  __shared__ double part[4];
  double v = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) v = fmax(v, S[i * ld + i]);
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) *pivmin = thr * fmax(fmax(part[0], part[1]), fmax(part[2], part[3]));
}

// Factor the w x w diagonal block at (j0, j0) of S (column-major, ld):
// L11 (lower, zeros above) into F at the same position and X = L11^{-1}
// (QB x QB, column-major).  512 threads: thread (q, t) keeps row t of the
// block, columns QC q .. QC q + QC - 1 (wave q's panel), in registers, for
// both the Cholesky factor and X.  Blocked right-looking, one barrier per
// panel:
//   * wave b factors panel b alone: per column c the pivot and the panel's
//     entries of column c come from other lanes by readlane, and the
//     panel's later columns take the rank-1 update (no barrier); it
//     publishes the panel's L columns and pivot reciprocals to LDS;
//   * after the barrier, every later wave applies the panel's rank-8 update
//     A[t][k] -= sum_c L[t][c] L[k][c] to its own columns (wave b + 1 then
//     factors panel b + 1 at once), and every wave carries X's rows through
//     the panel: the panel's rows of X are finalised by forward
//     substitution with the panel's diagonal block (X_p <- L_pp^{-1} X_p,
//     lanes 0..7, one per column of the wave), then the rows below take
//     X[t] -= sum_c L[t][c] X_p[c].  Wave b + 1 does its X step for panel b
//     one iteration later, after factoring panel b + 1, so that the chain of
//     panel factorisations does not wait for it (every panel keeps its own
//     LDS slot).
// This is the elimination X <- I; X[c,:] /= L[c,c]; X[t,:] -= L[t,c] X[c,:]
// (t > c) column by column, grouped by panel.  A pivot not above *pivmin
// sets *fail and is replaced by *pivmin (or 1); the caller discards the
// result.  Padding rows/columns (>= w) are zero with unit pivots.
// DROP (the shifted PSD factorisation of tn_eigh_psd): *pivmin is the
// shift delta; a pivot below delta / 2 (or not positive) drops its column
// (L column and X row zero) instead of failing.
template <bool DROP>
__global__ void __launch_bounds__(512) cq_diag(const double* __restrict__ S, int64_t ld,
                                               int64_t j0, int w,
                                               const double* __restrict__ pivmin,
                                               double* __restrict__ F, double* __restrict__ X,
                                               int* __restrict__ fail) {
  // panel columns of L (Ls[b][kc][t] = L[t][QC b + kc], zero above the
  // diagonal) and 1 / L[c][c] (0 for a dropped column); XP: per wave, the
  // panel's rows of X at its columns
  __shared__ double Ls[QB / QC][QC][QB];
  __shared__ double invs[QB / QC][QC];
  __shared__ double XP[QB / QC][QC][QC + 1];
  const int t = threadIdx.x & 63;
  const int q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave index: uniform
  const double pm = DROP ? 0.5 * *pivmin : *pivmin;
  bool bad = false;
  double r[QC], x[QC];
#pragma unroll
  for (int k = 0; k < QC; ++k) {
    const int c = QC * q + k;
    r[k] = (t < w && c < w) ? S[(j0 + c) * ld + j0 + t] : 0.0;
    x[k] = c == t ? 1.0 : 0.0;
  }
  // wave b: factor panel b (its own columns) and publish it into slot pb
  auto factor = [&](int b, int pb) {
    const int c0 = QC * b;
#pragma unroll
    for (int kc = 0; kc < QC; ++kc) {
      const int c = c0 + kc;
      const double a = r[kc];
      const double dv = readlane_f64(a, c);
      double d, inv;
      if (DROP) {
        const bool keep = c >= w || (dv >= pm && dv > 0.0);
        d = c < w ? (keep ? dv : 1.0) : 1.0;
        inv = keep ? rsq_nr(d) : 0.0;
      } else {
        const bool ok = dv > pm;  // false for NaN
        bad = bad || (c < w && !ok);
        d = c < w ? (ok ? dv : (pm > 0.0 ? pm : 1.0)) : 1.0;
        inv = rsq_nr(d);
      }
      const double l = t > c ? a * inv : 0.0;
      const double co = l * inv;
#pragma unroll
      for (int k = kc + 1; k < QC; ++k) r[k] -= co * readlane_f64(a, c0 + k);
      r[kc] = t == c ? d * inv : l;
      Ls[pb][kc][t] = r[kc];
      if (t == 0) invs[pb][kc] = inv;
    }
  };
  // X through panel b at this wave's columns: the panel's rows finalised,
  // then the rank-8 update of the rows below
  auto xstep = [&](int b) {
    const int c0 = QC * b;
    double lt[QC];
#pragma unroll
    for (int kc = 0; kc < QC; ++kc) lt[kc] = Ls[b][kc][t];
    if (t >= c0 && t < c0 + QC) {
#pragma unroll
      for (int k = 0; k < QC; ++k) XP[q][t - c0][k] = x[k];
    }
    wave_lds_sync();
    if (t < QC) {
      double s[QC];
#pragma unroll
      for (int i = 0; i < QC; ++i) s[i] = XP[q][i][t];
#pragma unroll
      for (int i = 0; i < QC; ++i) {
#pragma unroll
        for (int kk = 0; kk < i; ++kk) s[i] -= Ls[b][kk][c0 + i] * s[kk];
        s[i] *= invs[b][i];
      }
#pragma unroll
      for (int i = 0; i < QC; ++i) XP[q][i][t] = s[i];
    }
    wave_lds_sync();
    const int ip = min(max(t - c0, 0), QC - 1);
#pragma unroll
    for (int k = 0; k < QC; ++k) {
      double acc = x[k];
#pragma unroll
      for (int i = 0; i < QC; ++i) acc -= lt[i] * XP[q][i][k];
      x[k] = t >= c0 + QC ? acc : (t >= c0 ? XP[q][ip][k] : x[k]);
    }
    wave_lds_sync();
  };
  if (q == 0) factor(0, 0);
#pragma unroll 1
  for (int b = 0; b < QB / QC; ++b) {
    __syncthreads();  // panel b published
    if (q > b) {
#pragma unroll
      for (int kc = 0; kc < QC; ++kc) {
        const double l = Ls[b][kc][t];
#pragma unroll
        for (int k = 0; k < QC; ++k) r[k] -= l * Ls[b][kc][QC * q + k];
      }
    }
    if (q == b + 1) {
      factor(b + 1, b + 1);
    } else {
      if (q == b && b > 0) xstep(b - 1);
      xstep(b);
    }
  }
#pragma unroll
  for (int k = 0; k < QC; ++k) {
    const int c = QC * q + k;
    if (t < w && c < w) F[(j0 + c) * ld + j0 + t] = c <= t ? r[k] : 0.0;
    X[c * QB + t] = (t < w && c < w) ? x[k] : 0.0;
  }
  if (!DROP && bad && t == 0) fail[0] = 1;
}

// orth = max |G - I| over the n x n block G (column-major, ld); block c
// scans column c; orth must start at zero (non-negative doubles order as
// their bit patterns).
__global__ void __launch_bounds__(256) cq_orth(const double* __restrict__ G, int64_t ld,
                                               int64_t n, unsigned long long* __restrict__ orth) {
  __shared__ double part[4];
  const int64_t c = blockIdx.x;
  double v = 0.0;
  for (int64_t r = threadIdx.x; r < n; r += 256)
    v = fmax(v, fabs(G[c * ld + r] - (r == c ? 1.0 : 0.0)));
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double b = fmax(fmax(part[0], part[1]), fmax(part[2], part[3]));
    atomicMax(orth, (unsigned long long)__double_as_longlong(b));
  }
}

// M = (3 I - G) / 2 in place (n x n, column-major).
__global__ void cq_ns(double* __restrict__ G, int64_t n) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const int64_t c = idx / n, r = idx - c * n;
  G[idx] = (r == c ? 1.5 : 0.0) - 0.5 * G[idx];
}

// The dims of `shape` with the strides of a row-major layout whose last
// dimension has stride `unit`.
Modes row_major_like(const Modes& shape, int64_t unit) {
  Modes r = shape;
  int64_t st = unit;
  for (int i = shape.n - 1; i >= 0; --i) {
    r.strides[i] = st;
    st *= shape.dims[i];
  }
  return r;
}

struct CqWs {
  double *B0, *B1, *M, *X, *pivmin, *part;
  unsigned long long* orth;
  int* fail;
  size_t part_bytes;
};

size_t cq_layout(int64_t m, int64_t n, char* base, size_t cap, CqWs* ws) {
  Arena a{base, cap, 0};
  CqWs w{};
  w.B0 = a.take<double>((size_t)((n + m) * n));
  w.B1 = a.take<double>((size_t)((n + m) * n));
  w.M = a.take<double>((size_t)(n * n));
  w.X = a.take<double>((size_t)QB * QB);
  w.pivmin = a.take<double>(2);
  w.orth = a.take<unsigned long long>(1);
  w.fail = a.take<int>(1);
  // split-K partials for the n x n Grams (K = m) and the m x n correction (K = n)
  const size_t pe = std::max((size_t)splitk_factor(n, n, m) * n * n,
                             (size_t)splitk_factor(m, n, n) * m * n);
  w.part_bytes = pe > (size_t)(n * n) ? pe * sizeof(double) : 0;
  w.part = w.part_bytes ? a.take<double>(pe) : nullptr;
  if (ws) *ws = w;
  return a.used;
}

// One Cholesky QR pass.  S (ld = n + m): W = S[n:, :] holds the m x n input;
// its Gram W^T W is formed into S[:n, :] and the stacked matrix is
// eliminated block by block: F[:n, :] receives L (lower, W^T W = L L^T) and
// F[n:, :] receives Q = W L^{-T}, because the rows of W go through the same
// column eliminations as the rows of the Gram.  S is destroyed; F is
// cleared first.
int cq_pass(double* S, double* F, int64_t m, int64_t n, double thr, const CqWs& ws,
            hipStream_t st) {
  const int64_t ld = n + m;
  TN_TRY(contract_launch(S + n, mk({n}, {ld}), mk({m}, {1}), S + n, mk({m}, {1}), mk({n}, {ld}),
                         S, mk({n}, {1}), mk({n}, {ld}), TN_F64, TN_PREC_NATIVE, 1.0, 0.0, ws.part,
                         ws.part_bytes, st));
  hipLaunchKernelGGL(cq_pivmin, dim3(1), dim3(256), 0, st, S, ld, n, thr, ws.pivmin);
  TN_HIP(hipMemsetAsync(F, 0, sizeof(double) * (size_t)(ld * n), st));
  for (int64_t j0 = 0; j0 < n; j0 += QB) {
    const int w = (int)std::min<int64_t>(QB, n - j0);
    hipLaunchKernelGGL(cq_diag<false>, dim3(1), dim3(512), 0, st, S, ld, j0, w, ws.pivmin, F, ws.X,
                       ws.fail);
    TN_CHECK_LAUNCH();
    // [L21; Q_j] = [G21; W_j] L11^{-T}:  B[k][c] = X[c][k] = X[k * QB + c]
    const int64_t rows = ld - j0 - w;
    TN_TRY(contract_launch(S + j0 * ld + j0 + w, mk({rows}, {1}), mk({w}, {ld}), ws.X,
                           mk({w}, {QB}), mk({w}, {1}), F + j0 * ld + j0 + w, mk({rows}, {1}),
                           mk({w}, {ld}), TN_F64, TN_PREC_NATIVE, 1.0, 0.0, nullptr, 0, st));
    const int64_t m2 = n - j0 - w;
    if (m2 <= 0) break;
    // [G22; W_rest] -= [L21; Q_j] L21^T
    const double* P = F + j0 * ld + j0 + w;
    TN_TRY(contract_launch(P, mk({rows}, {1}), mk({w}, {ld}), P, mk({w}, {ld}), mk({m2}, {1}),
                           S + (j0 + w) * ld + j0 + w, mk({rows}, {1}), mk({m2}, {ld}), TN_F64,
                           TN_PREC_NATIVE, -1.0, 1.0, nullptr, 0, st));
  }
  return TN_OK;
}

}  // namespace

int chol_block(const double* S, int64_t ld, int64_t j0, int w, const double* delta, double* F,
               double* X, hipStream_t st) {
  hipLaunchKernelGGL(cq_diag<true>, dim3(1), dim3(512), 0, st, S, ld, j0, w, delta, F, X,
                     (int*)nullptr);
  TN_CHECK_LAUNCH();
  return TN_OK;
}

size_t cholqr_ws_bytes(int64_t m, int64_t n) {
  if (m <= 0 || n <= 0) return 256;
  return cq_layout(m, n, nullptr, 0, nullptr) + 256;
}

int cholqr_run(const void* A, const Modes& rows, const Modes& cols, int64_t m, int64_t n,
               double* Q, int64_t q_rs, int64_t q_cs, double* R, double orth_tol,
               double* orth_out, void* ws, size_t ws_bytes, hipStream_t st) {
  CqWs w{};
  cq_layout(m, n, (char*)ws, ws_bytes, &w);
  const int64_t ld = n + m;
  TN_HIP(hipMemsetAsync(w.fail, 0, sizeof(int), st));
  TN_HIP(hipMemsetAsync(w.orth, 0, sizeof(unsigned long long), st));
  hipLaunchKernelGGL(cq_gather, dim3((unsigned)ceil_div(m * n, 256)), dim3(256), 0, st,
                     (const double*)A, rows, cols, m, n, w.B0);
  TN_CHECK_LAUNCH();
  // Cholesky QR (B0 -> B1): pivots must stay above 1e-14 of the largest
  // diagonal of A^T A; Q1 = B1[n:, :]
  TN_TRY(cq_pass(w.B0, w.B1, m, n, 1e-14, w, st));
  // M = (3 I - Q1^T Q1) / 2
  TN_TRY(contract_launch(w.B1 + n, mk({n}, {ld}), mk({m}, {1}), w.B1 + n, mk({m}, {1}),
                         mk({n}, {ld}), w.M, mk({n}, {1}), mk({n}, {n}), TN_F64, TN_PREC_NATIVE,
                         1.0, 0.0, w.part, w.part_bytes, st));
  hipLaunchKernelGGL(cq_ns, dim3((unsigned)ceil_div(n * n, 256)), dim3(256), 0, st, w.M, n);
  TN_CHECK_LAUNCH();
  // Q = Q1 M  (element (i, j) at i * q_rs + j * q_cs)
  TN_TRY(contract_launch(w.B1 + n, mk({m}, {1}), mk({n}, {ld}), w.M, mk({n}, {1}), mk({n}, {n}),
                         Q, mk({m}, {q_rs}), mk({n}, {q_cs}), TN_F64, TN_PREC_NATIVE, 1.0, 0.0,
                         w.part, w.part_bytes, st));
  // R = Q^T A  (row-major n x n); the contraction launcher pairs the K modes
  // of both operands and the N modes of B and C mode by mode, so Q's row
  // index and R's column index are split like A's row and column modes
  TN_TRY(contract_launch(Q, mk({n}, {q_cs}), row_major_like(rows, q_rs), A, rows, cols, R,
                         mk({n}, {n}), row_major_like(cols, 1), TN_F64, TN_PREC_NATIVE, 1.0, 0.0,
                         w.part, w.part_bytes, st));
  // orthogonality of the result: max |Q^T Q - I|, Gram formed in B0
  TN_TRY(contract_launch(Q, mk({n}, {q_cs}), mk({m}, {q_rs}), Q, mk({m}, {q_rs}),
                         mk({n}, {q_cs}), w.B0, mk({n}, {1}), mk({n}, {ld}), TN_F64,
                         TN_PREC_NATIVE, 1.0, 0.0, w.part, w.part_bytes, st));
  hipLaunchKernelGGL(cq_orth, dim3((unsigned)n), dim3(256), 0, st, w.B0, ld, n, w.orth);
  TN_CHECK_LAUNCH();
  int fail = 0;
  unsigned long long ob = 0;
  TN_HIP(hipMemcpyAsync(&fail, w.fail, sizeof(int), hipMemcpyDeviceToHost, st));
  TN_HIP(hipMemcpyAsync(&ob, w.orth, sizeof(ob), hipMemcpyDeviceToHost, st));
  TN_HIP(hipStreamSynchronize(st));
  double orth;
  std::memcpy(&orth, &ob, sizeof(orth));
  if (orth_out) *orth_out = orth;
  if (fail || !(orth <= orth_tol)) {
    set_error(TN_ERR_NOCONV,
              "tn_cholqr: %lld x %lld matrix too ill-conditioned for Cholesky QR "
              "(pivot failure %d, ||Q^T Q - I||_max = %.3g)",
              (long long)m, (long long)n, fail, orth);
    return TN_ERR_NOCONV;
  }
  return TN_OK;
}

}  // namespace tn

using namespace tn;

extern "C" size_t tn_cholqr_workspace(int64_t m, int64_t n) { return cholqr_ws_bytes(m, n); }

extern "C" int tn_cholqr(const void* A, const tn_modes* a_rows, const tn_modes* a_cols, int64_t m,
                         int64_t n, double* Q, double* R, double orth_tol, double* orth_out,
                         void* ws, size_t ws_bytes, void* stream) {
  Modes rows = to_modes(a_rows), cols = to_modes(a_cols);
  TN_REQUIRE(m >= 0 && n >= 0 && m >= n, "tn_cholqr: needs m >= n >= 0 (got %lld x %lld)",
             (long long)m, (long long)n);
  TN_REQUIRE(modes_size(rows) == m && modes_size(cols) == n, "tn_cholqr: modes do not match m x n");
  if (orth_out) *orth_out = 0.0;
  if (n == 0) return TN_OK;
  TN_REQUIRE(A && Q && R, "tn_cholqr: null pointer");
  TN_REQUIRE(ws && ws_bytes >= tn_cholqr_workspace(m, n), "tn_cholqr: workspace too small");
  if (orth_tol <= 0.0) orth_tol = 1e-12;
  return cholqr_run(A, rows, cols, m, n, Q, n, 1, R, orth_tol, orth_out, ws, ws_bytes,
                    as_stream(stream));
}
