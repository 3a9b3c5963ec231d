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

}  // namespace tn
