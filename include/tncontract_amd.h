/*
 * tncontract_amd -- C ABI of the MI355X (gfx950) contraction-plus-truncation
 * library.  Every entry point takes raw device pointers, plain integer shapes
 * and strides, and a hipStream_t passed as void*.  Buffers are owned by the
 * caller (the Python layer allocates them with PyTorch-ROCm); work is
 * stream-ordered; no C++ exception crosses this boundary.  Each function
 * returns TN_OK (0) or a negative status; tn_last_error() gives the message.
 *
 * The reference (andrewdarmawan/tncontract) is pure Python over NumPy/LAPACK
 * and has no FFI of its own.  Each entry point below names the reference call
 * site whose arithmetic it replaces; the Python layer (tncontract_amd) binds
 * them with ctypes behind the reference's own Python API.
 */
#ifndef TNCONTRACT_AMD_H
#define TNCONTRACT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TN_OK 0
#define TN_ERR_INVALID -1     /* bad shapes / arguments        -> ValueError */
#define TN_ERR_HIP -2         /* HIP runtime failure           -> RuntimeError */
#define TN_ERR_NOCONV -3      /* factorisation did not converge -> LinAlgError */
#define TN_ERR_UNSUPPORTED -4 /* dtype / size not supported    -> NotImplementedError */
#define TN_ERR_WORKSPACE -5   /* workspace too small           -> ValueError */

#define TN_MAX_MODES 8

/* Element types. */
#define TN_F64 0
#define TN_F32 1

/* Arithmetic of a contraction (tn_contract). */
#define TN_PREC_NATIVE 0 /* same as the element type (FP64 MFMA for TN_F64) */
#define TN_PREC_F32 1    /* FP32 MFMA, FP32 accumulate                       */
#define TN_PREC_BF16X3 2 /* split FP32 = hi+lo bf16, 3 bf16 MFMAs, FP32 acc  */

/* A group of tensor modes: element offset of the linear (row-major, slowest
 * mode first) index i is sum_j idx_j(i) * strides[j].  Strides are in
 * elements and may be any non-negative value (a permuted view). */
typedef struct tn_modes {
  int32_t nmodes;
  int64_t dims[TN_MAX_MODES];
  int64_t strides[TN_MAX_MODES];
} tn_modes;

int tn_version(void);
const char* tn_last_error(void);
int tn_last_error_code(void);
/* Number of HIP devices visible to the runtime this library is bound to. */
int tn_device_count(int* count);

/* ---- data movement (tensor.py:354-394, 482: np.rollaxis + np.reshape) ----
 * dst (dense, row-major over src->dims) = src gathered through src->strides. */
int tn_permute(const void* src, const tn_modes* src_modes, void* dst, int dtype, void* stream);

/* dst = (dst_dtype) src, n elements. */
int tn_convert(const void* src, int src_dtype, void* dst, int dst_dtype, int64_t n, void* stream);

/* x *= alpha (tensor.py:95-123 scalar multiply; onedim_core.py:358, 482). */
int tn_scale(void* x, int dtype, int64_t n, double alpha, void* stream);

/* x = sqrt(x) elementwise (np.sqrt of the singular values when tensor_svd
 * absorbs them into both factors, tensor.py:950-956). */
int tn_sqrt(void* x, int dtype, int64_t n, void* stream);

/* out = a*x + b*y elementwise (Tensor.__add__, tensor.py:125-140; distance,
 * tensor.py:782-802). */
int tn_axpby(const void* x, const void* y, void* out, int dtype, int64_t n, double a, double b,
             void* stream);

/* out (k x k, row-major) = diag(s) (S = np.diag(s) in tensor_svd, tensor.py:940). */
int tn_diag_embed(const double* s, int64_t k, void* out, int dtype, void* stream);

/* Inverse of a diagonal matrix (Tensor.inv = np.linalg.inv, tensor.py:487-488,
 * applied to the singular-value matrices of MatrixProductStateCanonical,
 * onedim_core.py:1079-1082, 1160-1163, 1211-1214, 1876).  out (n x n,
 * row-major) = diag(1 / A_ii), A read as A[i*rs + j*cs]; info (device, 2
 * doubles) = {number of nonzero off-diagonal entries, number of zero diagonal
 * entries} so the caller can detect a non-diagonal or singular A.
 * ws: >= tn_diag_inverse_workspace(n) bytes. */
size_t tn_diag_inverse_workspace(int64_t n);
int tn_diag_inverse(const void* A, int64_t n, int64_t rs, int64_t cs, void* out, int dtype,
                    double* info, void* ws, size_t ws_bytes, void* stream);

/* *out (device double) = sum x_i^2 in FP64 (np.linalg.norm, tensor.py:590;
 * onedim_core.py:272, 305, 656).  ws: >= tn_sumsq_workspace(n) bytes. */
size_t tn_sumsq_workspace(int64_t n);
int tn_sumsq(const void* x, int dtype, int64_t n, double* out, void* ws, size_t ws_bytes,
             void* stream);

/* out[0] (device) = number of infinite entries of x, out[1] = number of NaN
 * entries: the diagnostic tensor_svd prints before re-raising when an SVD
 * fails (tensor.py:922-929).  One workgroup; for error paths only. */
int tn_count_nonfinite(const void* x, int dtype, int64_t n, unsigned long long* out,
                       void* stream);

/* x /= ||x||_F in place, ||x|| computed in FP64 and stored at *norm_out
 * (device double, may be NULL); x untouched when the norm is 0.  Keeps the
 * left environments of the structured sweep O(1) without a host round trip. */
size_t tn_normalize_workspace(int64_t n);
int tn_normalize(void* x, int dtype, int64_t n, double* norm_out, void* ws, size_t ws_bytes,
                 void* stream);

/* Internal contraction (np.trace, tensor.py:318-338):
 * dst[o] = sum_{t<diag_dim} src[off(out_modes, o) + t * diag_stride]. */
int tn_trace(const void* src, const tn_modes* out_modes, int64_t diag_dim, int64_t diag_stride,
             void* dst, int dtype, void* stream);

/* ---- pairwise contraction (np.tensordot in contract(), tensor.py:735) ----
 * C[m,n] = alpha * sum_k A[m,k] B[k,n] + beta * C[m,n], where the row/col
 * index of each operand is decoded through its mode group, so any label
 * permutation of the inputs and of the output is fused into the loads and the
 * stores.  a_k and b_k must have the same total size.  precision: TN_PREC_*.
 * beta == 0 never reads C. */
int tn_contract(const void* A, const tn_modes* a_m, const tn_modes* a_k,
                const void* B, const tn_modes* b_k, const tn_modes* b_n,
                void* C, const tn_modes* c_m, const tn_modes* c_n,
                int dtype, int precision, double alpha, double beta, void* stream);

/* tn_contract with a split-K workspace (same contraction; no counterpart in
 * the reference, whose tensordot has no device scratch).  When the output has
 * fewer 128 x 128 tiles than two per CU, K is split over workgroups whose
 * partial sums go to ws and are reduced into C by a second launch.
 * tn_contract_workspace gives the bytes for the full split of an M x N x K
 * contraction (0: no split); a smaller ws_bytes only limits the split, and
 * ws == NULL runs tn_contract.  Calls sharing one ws must be stream-ordered. */
size_t tn_contract_workspace(int64_t M, int64_t N, int64_t K, int dtype);
int tn_contract_ws(const void* A, const tn_modes* a_m, const tn_modes* a_k,
                   const void* B, const tn_modes* b_k, const tn_modes* b_n,
                   void* C, const tn_modes* c_m, const tn_modes* c_n,
                   int dtype, int precision, double alpha, double beta,
                   void* ws, size_t ws_bytes, void* stream);

/* ---- QR (np.linalg.qr(mode="reduced") in tensor_qr, tensor.py:1044) ------
 * A is the m x n matrix whose row index is decoded through a_rows and column
 * index through a_cols.  Q: m x k row-major, R: k x n row-major, k=min(m,n).
 * Blocked Householder with FP64 reflectors; R's diagonal follows the LAPACK
 * dgeqrf sign convention (beta = -sign(alpha) * norm).  dtype TN_F64. */
size_t tn_qr_workspace(int64_t m, int64_t n, int dtype);
int tn_qr(const void* A, const tn_modes* a_rows, const tn_modes* a_cols, int64_t m, int64_t n,
          void* Q, void* R, int dtype, void* ws, size_t ws_bytes, void* stream);

/* ---- Cholesky QR (row space of a variational site update) ---------------
 * A = Q R for the m x n FP64 matrix A, m >= n (decoded as in tn_qr): Q m x n
 * row-major with orthonormal columns, R = Q^T A n x n row-major.  Replaces
 * tensor_lq (tensor.py:1058-1109) where variational_compress calls it
 * (onedim_core.py:560-631): there only the row space matters (Q becomes the
 * site, L = R^T is absorbed into the neighbour), so Q may differ from
 * Householder's by an orthogonal n x n factor, and R is upper triangular
 * only up to O(cond(A)^2 eps).  Method: G = A^T A = L L^T (blocked
 * Cholesky), Q1 = A L^{-T}, one Newton-Schulz step Q = Q1 (3I - Q1^T Q1)/2,
 * R = Q^T A.  *orth_out (host, may be NULL) receives max |Q^T Q - I|.
 * Returns TN_ERR_NOCONV (outputs unusable) when a Cholesky pivot falls below
 * 1e-14 of the largest diagonal of A^T A or the orthogonality exceeds
 * orth_tol (<= 0 -> 1e-12): the caller then uses tn_qr. */
size_t tn_cholqr_workspace(int64_t m, int64_t n);
int tn_cholqr(const void* A, const tn_modes* a_rows, const tn_modes* a_cols, int64_t m, int64_t n,
              double* Q, double* R, double orth_tol, double* orth_out, void* ws, size_t ws_bytes,
              void* stream);

/* ---- SVD (np.linalg.svd(full_matrices=False), tensor.py:915) ------------
 * Thin SVD of the m x n matrix A (decoded as in tn_qr): U m x k row-major,
 * S k (FP64, descending), Vt k x n row-major, k = min(m,n).  One-sided
 * block Jacobi with FP64 rotations; with k >= 32 the vectors are first
 * ordered by norm and QR-factored (Cholesky QR, Householder when that does
 * not reach orthogonality 1e-13) and the Jacobi runs on the rows of R.
 * tol: orthogonality tolerance (<= 0 -> (vector length) * eps, as LAPACK
 * dgesvj); max_sweeps <= 0 -> 40.  *sweeps_out (host, may be NULL) receives
 * the number of sweeps used.  Returns TN_ERR_NOCONV if not converged.
 * A singular value that is exactly zero gets a zero column of U or a zero
 * row of Vt (in whichever factor is formed by dividing rotated vectors by
 * their norms), where LAPACK returns an orthonormal completion; the
 * reference's callers truncate such values away or multiply them by zero, so
 * the product U S Vt is the same. */
size_t tn_svd_workspace(int64_t m, int64_t n, int dtype);
int tn_svd(const void* A, const tn_modes* a_rows, const tn_modes* a_cols, int64_t m, int64_t n,
           void* U, double* S, void* Vt, int dtype, double tol, int max_sweeps, int* sweeps_out,
           void* ws, size_t ws_bytes, void* stream);

/* ---- allocation-free SVD ------------------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (serving loops and harnesses that forbid hipMalloc).  The same
 * computation and results as tn_svd, with the workspace taken from a
 * library-owned scratch array of TN_SVD_NOALLOC_SCRATCH_BYTES per GPU (held
 * until the process ends) instead of a caller's buffer; it calls no
 * allocating routine and does not replay its sweeps as a HIP graph.  An SVD
 * whose tn_svd_workspace exceeds the scratch size is refused with
 * TN_ERR_WORKSPACE before any device work.  Every call on one GPU uses that
 * GPU's one array: calls that overlap on the same GPU (other streams or host
 * threads) race and must be serialised by the caller. */
#define TN_SVD_NOALLOC_SCRATCH_BYTES (64u << 20)
int tn_svd_noalloc(const void* A, const tn_modes* a_rows, const tn_modes* a_cols, int64_t m,
                   int64_t n, void* U, double* S, void* Vt, int dtype, double tol,
                   int max_sweeps, int* sweeps_out, void* stream);

/* ---- PSD eigendecomposition (truncation from a bond Gram matrix) ---------
 * K = U diag(lam) U^T for a symmetric positive semi-definite n x n matrix K
 * (decoded as in tn_qr; FP64 or FP32 input).  U: n x n row-major FP64
 * (column j = eigenvector j), lam: n FP64, descending.  Used where only the
 * Gram matrix K = C^T E C of a bond is formed: its eigenvectors are the left
 * singular vectors the reference's truncating SVD computes
 * (onedim_core.py:317 -> tensor.py:915) and lam = sigma^2.
 * Method: static diagonal pivot, shift delta = 8 sqrt(n) eps_in tr(K) with
 * eps_in the unit roundoff of dtype (eigenvectors unchanged; pass a K formed
 * in FP32 as TN_F32 so the shift covers its rounding), blocked Cholesky
 * (a pivot below delta/2 drops its column), one-sided block Jacobi on the
 * Cholesky factor's columns (FP64 rotations); lam = sigma^2 - delta.  tol <= 0 ->
 * n * eps; max_sweeps <= 0 -> 40; *sweeps_out (host, may be NULL).
 * Returns TN_ERR_NOCONV if not converged. */
size_t tn_eigh_psd_workspace(int64_t n);
int tn_eigh_psd(const void* K, const tn_modes* k_rows, const tn_modes* k_cols, int64_t n,
                double* U, double* lam, int dtype, double tol, int max_sweeps, int* sweeps_out,
                void* ws, size_t ws_bytes, void* stream);

/* ---- batched uneven contraction + truncation --------------------------
 * No single reference call: each job is what the reference does with one
 * tensordot (tensor.py:735) followed by truncated_svd (tensor.py:1112) with a
 * relative threshold -- here for many independent jobs of different sizes in
 * shared launches (grouped GEMM of the FP32 factors on FP64 MFMA with FP64
 * accumulation and autotuned tile shapes, then a batched one-sided block
 * Jacobi with FP64 rotations; a job with 32 < min(m, n) and max(m, n) <=
 * 1024 is first factored C_j = Q R by a batched Householder QR, the Jacobi
 * runs on the rows of R and the left vectors are Q Z).
 * C_j = L_j R_j (m x n); C_j = U S V^T (thin, sorted descending).  kept_out[j]
 * (host) = #{sigma/sigma_0 > rel_threshold}, capped at chi_max (>0): C_j is
 * FP64-accurate, so this is the rank of the exact product.  The first kept
 * columns of U / rows of Vt are the truncated factors.  The Jacobi stops at
 * orthogonality max(1e-8, L eps) (L the Jacobi vector length: max(m, n),
 * or min(m, n) when QR-preconditioned); vectors whose squared norm is
 * below (rel_threshold / 2k)^2 ||C_j||_F^2 are not rotated against each
 * other (their span lies below the threshold), so the singular values and
 * vectors past the kept ones are not those of an SVD. */
typedef struct tn_trunc_job {
  const void* L; /* m x k FP32 row-major */
  const void* R; /* k x n FP32 row-major */
  int64_t m, k, n;
  void* U;       /* m x min(m,n) FP32 row-major */
  double* S;     /* min(m,n), FP64, descending */
  void* Vt;      /* min(m,n) x n FP32 row-major */
} tn_trunc_job;
size_t tn_batched_truncate_workspace(const tn_trunc_job* jobs, int njobs);
int tn_batched_truncate(const tn_trunc_job* jobs, int njobs, double rel_threshold, int chi_max,
                        int max_sweeps, int* kept_out, int* sweeps_out, void* ws, size_t ws_bytes,
                        void* stream);

/* Many independent FP64 thin SVDs in shared launches: np.linalg.svd
 * (tensor.py:915) of each job's A (m x n, FP64, row-major), A = U S Vt with S
 * descending; the same batched one-sided block Jacobi (FP64 rotations,
 * QR-preconditioned as above), orthogonality to L eps (L the Jacobi vector
 * length).  All min(m, n) singular triplets are returned, an exactly zero
 * singular value with zero vectors; the caller truncates.  No counterpart as a batch in the
 * reference: it serves a layer of TEBD gates on disjoint bonds
 * (onedim_core.py:1100-1186), whose SVDs are independent. */
typedef struct tn_svd_job {
  const double* A; /* m x n FP64 row-major */
  int64_t m, n;
  double* U;       /* m x min(m,n) FP64 row-major */
  double* S;       /* min(m,n), descending */
  double* Vt;      /* min(m,n) x n FP64 row-major */
} tn_svd_job;
size_t tn_batched_svd_workspace(const tn_svd_job* jobs, int njobs);
int tn_batched_svd(const tn_svd_job* jobs, int njobs, int max_sweeps, int* sweeps_out, void* ws,
                   size_t ws_bytes, void* stream);

/* ---- learned-rank truncation (no reference counterpart: the reference
 * always runs the full SVD at tensor.py:915) -------------------------------
 * Rank-prediction MLP 34 -> 64 -> 64 -> 1 (ReLU, FP32).  Parameters are one
 * FP32 vector of tn_mlp_param_count() floats: W1[34][64], b1[64],
 * W2[64][64], b2[64], w3[64], b3.  X: B x 34 features, Y: B outputs; H1/H2
 * (B x 64, may be NULL in inference) keep the activations for the backward
 * pass.  tn_mlp_backward writes d/dP of (1/B) sum (y - t)^2 into grad (P
 * floats) using per_sample (B x P floats) as deterministic scratch. */
int tn_mlp_param_count(void);
int tn_mlp_forward(const float* X, int B, const float* P, float* Y, float* H1, float* H2,
                   void* stream);
int tn_mlp_backward(const float* X, const float* T, int B, const float* P, const float* Y,
                    const float* H1, const float* H2, float* grad, float* per_sample,
                    void* stream);
int tn_adam_step(float* P, const float* grad, float* m, float* v, int n, float lr, float beta1,
                 float beta2, float eps, int step, void* stream);
/* n standard-normal samples (counter-based, deterministic per seed): the
 * Gaussian test matrix of the randomized range finder. */
int tn_randn(void* out, int dtype, int64_t n, unsigned long long seed, void* stream);

/* ---- launch timing (no reference counterpart: measurement support) ------
 * tn_prof_enable(1) makes the library record a HIP event pair, on the launch
 * stream, around every launch of the instrumented kernels; tn_prof_collect
 * synchronises those events and returns, for one tag, the summed elapsed
 * time, launch count and the algorithmic flops / bytes the launcher assigned
 * to each launch.  Tags: 0 FP64 GEMM, 1 FP32 GEMM, 2 BF16x3 GEMM, 3 Jacobi
 * Gram, 4 Jacobi pair eigensolve (+ apply in tn_svd), 5 Jacobi apply, 6 QR panel, 7 direct
 * (small-K) contraction. */
int tn_prof_enable(int on);
int tn_prof_collect(int tag, double* total_ms, int64_t* launches, double* flops, double* bytes,
                    int reset);

#ifdef __cplusplus
}
#endif
#endif /* TNCONTRACT_AMD_H */
