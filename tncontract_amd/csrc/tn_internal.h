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

}  // namespace tn
