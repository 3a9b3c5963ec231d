// Internal (non-ABI) entry points shared between the library's .hip files.
#pragma once
#include "tn_common.h"

#include <initializer_list>

namespace tn {

// GEMM-view contraction launcher used by tn_contract and by the QR / SVD
// drivers.  splitk_ws (optional) lets it split K across blocks when the
// output has few tiles; it must hold S*M*N elements of the data type.
int contract_launch(const void* A, const Modes& a_m, const Modes& a_k, const void* B,
                    const Modes& b_k, const Modes& b_n, void* C, const Modes& c_m,
                    const Modes& c_n, int dtype, int prec, double alpha, double beta,
                    void* splitk_ws, size_t splitk_bytes, hipStream_t st);

// Mode group from literal dims/strides (slowest first).
Modes mk(std::initializer_list<int64_t> dims, std::initializer_list<int64_t> strides);

// Householder QR internals (tn_qr.hip), reused by the QR-preconditioned SVD.
struct QrWs {
  double *W, *tau, *V, *Tall, *G, *Y, *Y2, *Tp, *Qw, *part, *Ypart;
  size_t part_bytes;
};
// Carve the QR workspace for an m x n factorisation (base == nullptr: size query).
size_t qr_layout(int64_t m, int64_t n, char* base, size_t cap, QrWs* ws);
// Factor the column-major m x n matrix W (ld = m) in place: R above the
// diagonal, reflectors below, tau and per-block T factors in ws.
int qr_factor(double* W, int64_t m, int64_t n, const QrWs& ws, hipStream_t st);
// Qw (m x k, column-major) = H_1...H_k [X0; 0], X0 = I (init == nullptr) or init (k x k).
int qr_form_q(const double* W, int64_t m, int64_t n, const QrWs& ws, double* Qw, hipStream_t st,
              const double* init);

// Cholesky QR with a Newton-Schulz correction (tn_cholqr.hip): A (m x n,
// m >= n, FP64, decoded through rows/cols) = Q R, Q element (i, j) at
// i * q_rs + j * q_cs, R = Q^T A row-major n x n.  Synchronises the stream;
// TN_ERR_NOCONV when a pivot fails or max |Q^T Q - I| > orth_tol.
// Shifted Cholesky of the w x w diagonal block at (j0, j0) of the
// column-major S (ld): L11 into F at the same position, L11^{-1} into X
// (64 x 64, column-major); a pivot below delta/2 (delta on the device)
// drops its column (zero L column, zero X row).  w <= 64.
int chol_block(const double* S, int64_t ld, int64_t j0, int w, const double* delta, double* F,
               double* X, hipStream_t st);
size_t cholqr_ws_bytes(int64_t m, int64_t n);
int cholqr_run(const void* A, const Modes& rows, const Modes& cols, int64_t m, int64_t n,
               double* Q, int64_t q_rs, int64_t q_cs, double* R, double orth_tol,
               double* orth_out, void* ws, size_t ws_bytes, hipStream_t st);

}  // namespace tn
