// Pairwise tensor contraction (np.tensordot in contract(), tensor.py:735).
//
// Two kernels:
//  * gemm_mfma<Prec>: C[m,n] = alpha*sum_k A[m,k]B[k,n] + beta*C, 128x128
//    block tile, 4 waves (2x2) each owning a 64x64 tile of 16x16 MFMA
//    fragments.  The row/col/k index of every operand is decoded through its
//    mode group, so a label permutation of A, B or C is fused into the
//    global->LDS tile loads and into the epilogue stores (no transpose copy).
//    Prec = FP64 (v_mfma_f64_16x16x4), FP32 (v_mfma_f32_16x16x4) or BF16X3
//    (each FP32 operand split into hi+lo bf16, three v_mfma_f32_16x16x32_bf16
//    per tile product: hi*hi + hi*lo + lo*hi, FP32 accumulation).
//  * contract_direct: one thread per output element, iterated in the
//    output's memory order; used when K is small (e.g. the K = d = 2 MPO
//    application), where the work is HBM-bound on the stores.  With 32-bit
//    offsets and K <= 8 (the usual case) contract_direct32 decodes the
//    output index by magic-number division and reads the contracted offsets
//    from a host-built table.
#include "tn_internal.h"
#include "tn_prof.h"

#include <algorithm>
#include <type_traits>
#include <vector>

namespace tn {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

enum { P_F64 = 0, P_F32 = 1, P_BF16X3 = 2 };

constexpr int BM = 128, BN = 128;
constexpr int NT = 256;

struct GemmArgs {
  const void* A;
  const void* B;
  void* C;
  Modes am, ak, bk, bn, cm, cn;
  int64_t M, N, K;
  double alpha, beta;
  int a_kfast;  // A tile loads iterate k fastest (A unit stride along k)
  int b_nfast;  // B tile loads iterate n fastest
  int64_t kchunk;  // split-K: block z handles k in [z*kchunk, (z+1)*kchunk)
  void* part;      // split-K: dense M x N partial sums per z (nullptr: write C)
};

// Offset of index i in a mode group; single-mode groups avoid division and
// indices below 2^32 (every k, and every row or column index in practice)
// are decoded with 32-bit divisions, several times cheaper on the GPU than
// 64-bit ones (the tile loads decode one k per element and k-tile).
__device__ __forceinline__ int64_t goff(const Modes& m, int64_t i) {
  if (m.n == 1) return i * m.strides[0];
  if (m.n == 0) return 0;
  if ((uint64_t)i <= 0xffffffffull) return mode_offset32(m, (uint32_t)i);
  return mode_offset(m, i);
}

template <int P> struct PrecTraits;
template <> struct PrecTraits<P_F64> { using T = double; static constexpr int BK = 16; };
template <> struct PrecTraits<P_F32> { using T = float; static constexpr int BK = 16; };
template <> struct PrecTraits<P_BF16X3> { using T = float; static constexpr int BK = 32; };

// ------------------------------------------------------------ FP64 / FP32
// LDS layout [BK][BM + 2]: fragment reads by 16 lanes hit 16 consecutive rows.
// Two LDS buffers: while the MFMAs read the current tile, the next tile (loaded
// into registers before the MFMAs) is stored into the other buffer, so each
// k-tile needs one barrier.  SIMPLE: every mode group has at most one mode
// (the common case after merging), so operand offsets are index * stride
// computed in registers instead of read from the per-tile row/column tables.
template <int P, bool SIMPLE, int AKF, int BNF, int TM>
__global__ void __launch_bounds__(NT, 2) gemm_mfma_kernel(GemmArgs g) {
  constexpr int BM = TM, BN = TM;   // tile edge (128, or 64 for few-tile shapes)
  constexpr int FRG = TM / 32;      // 16 x 16 fragments per wave and dimension
  using T = typename PrecTraits<P>::T;
  constexpr int BK = PrecTraits<P>::BK;
  constexpr int LDA = BM + 2, LDB = BN + 2;
  constexpr int EPT_A = BM * BK / NT, EPT_B = BN * BK / NT;  // 8 each (4 at TM = 64)
  __shared__ T As[2][BK * LDA];
  __shared__ T Bs[2][BK * LDB];
  __shared__ int64_t rowA[SIMPLE ? 1 : BM], colB[SIMPLE ? 1 : BN], rowC[SIMPLE ? 1 : BM],
      colC[SIMPLE ? 1 : BN];
  // multi-mode K groups: the BK offsets of a k-tile, decoded once per tile
  // (by 2 BK threads, two tiles ahead) instead of once per element
  __shared__ int64_t kA[SIMPLE ? 1 : 2][SIMPLE ? 1 : BK], kB[SIMPLE ? 1 : 2][SIMPLE ? 1 : BK];

  const T* __restrict__ A = (const T*)g.A;
  const T* __restrict__ B = (const T*)g.B;
  T* __restrict__ C = (T*)g.C;
  const int tid = threadIdx.x;
  const int64_t m0 = (int64_t)blockIdx.y * BM, n0 = (int64_t)blockIdx.x * BN;
  const int64_t kbeg = (int64_t)blockIdx.z * g.kchunk;
  const int64_t kend = min(g.K, kbeg + g.kchunk);
  // single-mode strides (SIMPLE)
  const int64_t sam = g.am.n ? g.am.strides[0] : 0, sak = g.ak.n ? g.ak.strides[0] : 0;
  const int64_t sbk = g.bk.n ? g.bk.strides[0] : 0, sbn = g.bn.n ? g.bn.strides[0] : 0;
  const int64_t scm = g.cm.n ? g.cm.strides[0] : 0, scn = g.cn.n ? g.cn.strides[0] : 0;

  auto koffs = [&](int slot, int64_t k0) {
    if (tid < BK) {
      const int64_t k = k0 + tid;
      kA[slot][tid] = k < kend ? goff(g.ak, k) : 0;
    } else if (tid < 2 * BK) {
      const int64_t k = k0 + tid - BK;
      kB[slot][tid - BK] = k < kend ? goff(g.bk, k) : 0;
    }
  };
  if (!SIMPLE) {
    koffs(0, kbeg);
    koffs(1, kbeg + BK);
    for (int i = tid; i < BM; i += NT) {
      int64_t m = m0 + i;
      rowA[i] = m < g.M ? goff(g.am, m) : 0;
      rowC[i] = m < g.M ? goff(g.cm, m) : 0;
    }
    for (int i = tid; i < BN; i += NT) {
      int64_t n = n0 + i;
      colB[i] = n < g.N ? goff(g.bn, n) : 0;
      colC[i] = n < g.N ? goff(g.cn, n) : 0;
    }
    __syncthreads();
  }

  T ra[EPT_A], rb[EPT_B];
  auto load = [&](int64_t k0, int slot) {
#pragma unroll
    for (int j = 0; j < EPT_A; ++j) {
      int e = tid + NT * j;
      int r, kk;
      if (AKF >= 0 ? AKF == 1 : g.a_kfast != 0) { kk = e % BK; r = e / BK; } else { r = e % BM; kk = e / BM; }
      int64_t m = m0 + r, k = k0 + kk;
      // SIMPLE operands are known (host check) to have every offset of the
      // tile, masked rows included, below 2^31: 32-bit index arithmetic
      const int64_t off = SIMPLE ? (int64_t)((uint32_t)m * (uint32_t)sam + (uint32_t)k * (uint32_t)sak)
                                 : rowA[r] + kA[slot][kk];
      ra[j] = (m < g.M && k < kend) ? A[off] : (T)0;
    }
#pragma unroll
    for (int j = 0; j < EPT_B; ++j) {
      int e = tid + NT * j;
      int c, kk;
      if (BNF >= 0 ? BNF == 1 : g.b_nfast != 0) { c = e % BN; kk = e / BN; } else { kk = e % BK; c = e / BK; }
      int64_t n = n0 + c, k = k0 + kk;
      const int64_t off = SIMPLE ? (int64_t)((uint32_t)k * (uint32_t)sbk + (uint32_t)n * (uint32_t)sbn)
                                 : kB[slot][kk] + colB[c];
      rb[j] = (n < g.N && k < kend) ? B[off] : (T)0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < EPT_A; ++j) {
      int e = tid + NT * j;
      int r, kk;
      if (AKF >= 0 ? AKF == 1 : g.a_kfast != 0) { kk = e % BK; r = e / BK; } else { r = e % BM; kk = e / BM; }
      As[buf][kk * LDA + r] = ra[j];
    }
#pragma unroll
    for (int j = 0; j < EPT_B; ++j) {
      int e = tid + NT * j;
      int c, kk;
      if (BNF >= 0 ? BNF == 1 : g.b_nfast != 0) { c = e % BN; kk = e / BN; } else { kk = e % BK; c = e / BK; }
      Bs[buf][kk * LDB + c] = rb[j];
    }
  };

  const int wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * (TM / 2), wc = (wave & 1) * (TM / 2);
  const int lr = lane & 15, lk = lane >> 4;

  using Acc = typename std::conditional<P == P_F64, f64x4, f32x4>::type;
  Acc acc[FRG][FRG];
#pragma unroll
  for (int i = 0; i < FRG; ++i)
#pragma unroll
    for (int j = 0; j < FRG; ++j) acc[i][j] = Acc{0, 0, 0, 0};

  const int64_t nk = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;
  if (nk > 0) {
    load(kbeg, 0);
    store(0);
    __syncthreads();
  }
  int cur = 0;
  for (int64_t kt = 0; kt < nk; ++kt) {
    if (kt + 1 < nk) load(kbeg + (kt + 1) * BK, (int)((kt + 1) & 1));
    // tile kt + 2 reuses the slot tile kt's loads read (one barrier ago)
    if (!SIMPLE && kt + 2 < nk) koffs((int)(kt & 1), kbeg + (kt + 2) * BK);
    const T* as = As[cur];
    const T* bs = Bs[cur];
    // two k-steps of fragments in flight at a time (a full unroll hoists all
    // four and pushes the FP64 kernel past the two-wave register budget)
#pragma unroll 2
    for (int ks = 0; ks < BK; ks += 4) {
      T af[FRG], bf[FRG];
#pragma unroll
      for (int i = 0; i < FRG; ++i) af[i] = as[(ks + lk) * LDA + wr + 16 * i + lr];
#pragma unroll
      for (int j = 0; j < FRG; ++j) bf[j] = bs[(ks + lk) * LDB + wc + 16 * j + lr];
#pragma unroll
      for (int i = 0; i < FRG; ++i)
#pragma unroll
        for (int j = 0; j < FRG; ++j) {
          if constexpr (P == P_F64)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
          else
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
    }
    // the other buffer was last read before the previous barrier
    if (kt + 1 < nk) store(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }

  // epilogue: FP64 fragment element e of lane l sits at row (l/16)+4e,
  // FP32 at row 4*(l/16)+e; column l%16 in both.
  const T alpha = (T)g.alpha, beta = (T)g.beta;
#pragma unroll
  for (int i = 0; i < FRG; ++i)
#pragma unroll
    for (int j = 0; j < FRG; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        int r = wr + 16 * i + (P == P_F64 ? (lk + 4 * e) : (4 * lk + e));
        int c = wc + 16 * j + lr;
        int64_t m = m0 + r, n = n0 + c;
        if (m < g.M && n < g.N) {
          if (g.part) {
            ((T*)g.part)[((int64_t)blockIdx.z * g.M + m) * g.N + n] = acc[i][j][e];
          } else {
            T* p = SIMPLE ? C + (int64_t)((uint32_t)m * (uint32_t)scm + (uint32_t)n * (uint32_t)scn)
                          : C + rowC[r] + colC[c];
            T v = alpha * acc[i][j][e];
            if (g.beta != 0.0) v += beta * *p;
            *p = v;
          }
        }
      }
}

// ------------------------------------------------------------ BF16x3
// LDS layout [row][BK + 8] bf16 (k contiguous): lane l of a 16x16x32 MFMA
// reads 8 consecutive k of row l%16 starting at k = 8*(l/16).
__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

__global__ void __launch_bounds__(NT) gemm_bf16x3_kernel(GemmArgs g) {
  constexpr int BK = 32;
  constexpr int LD = BK + 8;
  constexpr int EPT = BM * BK / NT;  // 16
  __shared__ __attribute__((aligned(16))) __bf16 Ah[BM * LD], Al[BM * LD];
  __shared__ __attribute__((aligned(16))) __bf16 Bh[BN * LD], Bl[BN * LD];
  __shared__ int64_t rowA[BM], colB[BN], rowC[BM], colC[BN];
  // the BK offsets of a k-tile, decoded once per tile, two tiles ahead
  __shared__ int64_t kA[2][BK], kB[2][BK];

  const float* __restrict__ A = (const float*)g.A;
  const float* __restrict__ B = (const float*)g.B;
  float* __restrict__ C = (float*)g.C;
  const int tid = threadIdx.x;
  const int64_t m0 = (int64_t)blockIdx.y * BM, n0 = (int64_t)blockIdx.x * BN;
  const int64_t kbeg = (int64_t)blockIdx.z * g.kchunk;
  const int64_t kend = min(g.K, kbeg + g.kchunk);
  auto koffs = [&](int slot, int64_t k0) {
    if (tid < BK) {
      const int64_t k = k0 + tid;
      kA[slot][tid] = k < kend ? goff(g.ak, k) : 0;
    } else if (tid < 2 * BK) {
      const int64_t k = k0 + tid - BK;
      kB[slot][tid - BK] = k < kend ? goff(g.bk, k) : 0;
    }
  };
  koffs(0, kbeg);
  koffs(1, kbeg + BK);
  for (int i = tid; i < BM; i += NT) {
    int64_t m = m0 + i;
    rowA[i] = m < g.M ? goff(g.am, m) : 0;
    rowC[i] = m < g.M ? goff(g.cm, m) : 0;
  }
  for (int i = tid; i < BN; i += NT) {
    int64_t n = n0 + i;
    colB[i] = n < g.N ? goff(g.bn, n) : 0;
    colC[i] = n < g.N ? goff(g.cn, n) : 0;
  }
  __syncthreads();

  float ra[EPT], rb[EPT];
  auto load = [&](int64_t k0, int slot) {
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      int e = tid + NT * j;
      int r, kk;
      if (g.a_kfast) { kk = e % BK; r = e / BK; } else { r = e % BM; kk = e / BM; }
      int64_t m = m0 + r, k = k0 + kk;
      ra[j] = (m < g.M && k < kend) ? A[rowA[r] + kA[slot][kk]] : 0.f;
    }
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      int e = tid + NT * j;
      int c, kk;
      if (g.b_nfast) { c = e % BN; kk = e / BN; } else { kk = e % BK; c = e / BK; }
      int64_t n = n0 + c, k = k0 + kk;
      rb[j] = (n < g.N && k < kend) ? B[kB[slot][kk] + colB[c]] : 0.f;
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      int e = tid + NT * j;
      int r, kk;
      if (g.a_kfast) { kk = e % BK; r = e / BK; } else { r = e % BM; kk = e / BM; }
      split_bf16(ra[j], Ah[r * LD + kk], Al[r * LD + kk]);
    }
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      int e = tid + NT * j;
      int c, kk;
      if (g.b_nfast) { c = e % BN; kk = e / BN; } else { kk = e % BK; c = e / BK; }
      split_bf16(rb[j], Bh[c * LD + kk], Bl[c * LD + kk]);
    }
  };

  const int wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0, 0, 0, 0};

  const int64_t nk = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;
  if (nk > 0) {
    load(kbeg, 0);
    store();
    __syncthreads();
  }
  for (int64_t kt = 0; kt < nk; ++kt) {
    if (kt + 1 < nk) load(kbeg + (kt + 1) * BK, (int)((kt + 1) & 1));
    // tile kt + 2 reuses the slot tile kt's loads read (before the last barrier)
    if (kt + 2 < nk) koffs((int)(kt & 1), kbeg + (kt + 2) * BK);
    bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      ah[i] = *(const bf16x8*)&Ah[(wr + 16 * i + lr) * LD + lk];
      al[i] = *(const bf16x8*)&Al[(wr + 16 * i + lr) * LD + lk];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      bh[j] = *(const bf16x8*)&Bh[(wc + 16 * j + lr) * LD + lk];
      bl[j] = *(const bf16x8*)&Bl[(wc + 16 * j + lr) * LD + lk];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
      }
    __syncthreads();
    if (kt + 1 < nk) {
      store();
      __syncthreads();
    }
  }
  const float alpha = (float)g.alpha, beta = (float)g.beta;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        int r = wr + 16 * i + 4 * (lane >> 4) + e;
        int c = wc + 16 * j + lr;
        int64_t m = m0 + r, n = n0 + c;
        if (m < g.M && n < g.N) {
          if (g.part) {
            ((float*)g.part)[((int64_t)blockIdx.z * g.M + m) * g.N + n] = acc[i][j][e];
          } else {
            float* p = C + rowC[r] + colC[c];
            float v = alpha * acc[i][j][e];
            if (g.beta != 0.0) v += beta * *p;
            *p = v;
          }
        }
      }
}

// ------------------------------------------------------------ direct
struct DirectArgs {
  const void* A;
  const void* B;
  void* C;
  int no, nk;
  int64_t odim[2 * TN_MAX_MODES], oa[2 * TN_MAX_MODES], ob[2 * TN_MAX_MODES],
      oc[2 * TN_MAX_MODES];
  int64_t kdim[TN_MAX_MODES], ka[TN_MAX_MODES], kb[TN_MAX_MODES];
  int64_t nout, K;
  double alpha, beta;
};

template <typename T, typename Acc>
__global__ void __launch_bounds__(256) contract_direct(DirectArgs g) {
  const T* __restrict__ A = (const T*)g.A;
  const T* __restrict__ B = (const T*)g.B;
  T* __restrict__ C = (T*)g.C;
  int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; o < g.nout; o += stride) {
    int64_t idx = o, pa = 0, pb = 0, pc = 0;
    for (int i = g.no - 1; i >= 0; --i) {
      int64_t d = g.odim[i], q = idx / d, r = idx - q * d;
      pa += r * g.oa[i];
      pb += r * g.ob[i];
      pc += r * g.oc[i];
      idx = q;
    }
    Acc acc = 0;
    for (int64_t k = 0; k < g.K; ++k) {
      int64_t kidx = k, qa = 0, qb = 0;
      for (int i = g.nk - 1; i >= 0; --i) {
        int64_t d = g.kdim[i], q = kidx / d, r = kidx - q * d;
        qa += r * g.ka[i];
        qb += r * g.kb[i];
        kidx = q;
      }
      acc += (Acc)A[pa + qa] * (Acc)B[pb + qb];
    }
    T v = (T)(g.alpha * (double)acc);
    if (g.beta != 0.0) v += (T)g.beta * C[pc];
    C[pc] = v;
  }
}

// Division by a runtime constant d >= 1 for 0 <= n < 2^31: q = (mulhi(n, m) +
// n) >> l with l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1 (the
// round-up multiplier method of Granlund and Montgomery, PLDI 1994; the sum
// stays below 2^32 because n < 2^31).  Exact for every such n (checked on the
// host against integer division over the whole range of d).
struct FastDiv {
  uint32_t d, m, l;
};

static FastDiv make_fastdiv(uint32_t d) {
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  const uint64_t m = (((1ull << 32) * ((1ull << l) - d)) / d) + 1;
  return FastDiv{d, (uint32_t)m, l};
}

__device__ __forceinline__ uint32_t fast_div(uint32_t n, const FastDiv& f) {
  return (__umulhi(n, f.m) + n) >> f.l;
}

// contract_direct with 32-bit offsets (every offset of A, B and C and the
// output count below 2^31), magic-number index decoding and the K <= 8
// contracted offsets precomputed on the host: the MPO application's
// per-element work is then a few integer multiply-adds and K fused
// multiply-adds, and the kernel is bound by its stores.
struct Direct32Args {
  const void* A;
  const void* B;
  void* C;
  int no, K;
  FastDiv od[2 * TN_MAX_MODES];
  int32_t oa[2 * TN_MAX_MODES], ob[2 * TN_MAX_MODES], oc[2 * TN_MAX_MODES];
  int32_t ka[8], kb[8];
  uint32_t nout;
  double alpha, beta;
};

template <typename T, typename Acc>
__global__ void __launch_bounds__(256) contract_direct32(Direct32Args g) {
  const T* __restrict__ A = (const T*)g.A;
  const T* __restrict__ B = (const T*)g.B;
  T* __restrict__ C = (T*)g.C;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t o = blockIdx.x * blockDim.x + threadIdx.x; o < g.nout; o += stride) {
    uint32_t idx = o;
    int32_t pa = 0, pb = 0, pc = 0;
    for (int i = g.no - 1; i >= 0; --i) {
      const uint32_t q = fast_div(idx, g.od[i]);
      const int32_t r = (int32_t)(idx - q * g.od[i].d);
      pa += r * g.oa[i];
      pb += r * g.ob[i];
      pc += r * g.oc[i];
      idx = q;
    }
    Acc acc = 0;
    for (int k = 0; k < g.K; ++k) acc += (Acc)A[pa + g.ka[k]] * (Acc)B[pb + g.kb[k]];
    T v = (T)(g.alpha * (double)acc);
    if (g.beta != 0.0) v += (T)g.beta * C[pc];
    C[pc] = v;
  }
}

// Merge adjacent modes of a group when the merge is valid for every operand
// that shares the group (strides of the slower mode = stride*dim of the faster).
static void merge_group(std::vector<int64_t>& dims, std::vector<std::vector<int64_t>*> strides) {
  size_t i = 0;
  // drop unit dims first
  for (size_t j = 0; j < dims.size();) {
    if (dims[j] == 1 && dims.size() > 1) {
      dims.erase(dims.begin() + j);
      for (auto* s : strides) s->erase(s->begin() + j);
    } else {
      ++j;
    }
  }
  while (i + 1 < dims.size()) {
    bool ok = true;
    for (auto* s : strides)
      if ((*s)[i] != (*s)[i + 1] * dims[i + 1]) ok = false;
    if (ok) {
      dims[i] *= dims[i + 1];
      for (auto* s : strides) {
        (*s)[i] = (*s)[i + 1];
        s->erase(s->begin() + i + 1);
      }
      dims.erase(dims.begin() + i + 1);
    } else {
      ++i;
    }
  }
}

static Modes make_modes(const std::vector<int64_t>& d, const std::vector<int64_t>& s) {
  Modes m{};
  m.n = (int)d.size();
  for (int i = 0; i < TN_MAX_MODES; ++i) {
    m.dims[i] = i < m.n ? d[i] : 1;
    m.strides[i] = i < m.n ? s[i] : 0;
  }
  return m;
}

static void unpack(const tn_modes* m, std::vector<int64_t>& d, std::vector<int64_t>& s) {
  d.clear();
  s.clear();
  if (!m) return;
  for (int i = 0; i < m->nmodes; ++i) {
    d.push_back(m->dims[i]);
    s.push_back(m->strides[i]);
  }
}

static int64_t prod(const std::vector<int64_t>& d) {
  int64_t p = 1;
  for (auto x : d) p *= x;
  return p;
}

static bool dims_equal(const tn_modes* a, const tn_modes* b) {
  int na = a ? a->nmodes : 0, nb = b ? b->nmodes : 0;
  if (na != nb) return false;
  for (int i = 0; i < na; ++i)
    if (a->dims[i] != b->dims[i]) return false;
  return true;
}


// Split-K reduction: C = alpha * sum_z part[z] + beta * C.
template <typename T>
__global__ void __launch_bounds__(256) splitk_reduce(const T* __restrict__ part, int S, int64_t M,
                                                     int64_t N, Modes cm, Modes cn, T* C,
                                                     double alpha, double beta) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t MN = M * N;
  if (i >= MN) return;
  double acc = 0.0;
  for (int z = 0; z < S; ++z) acc += (double)part[(int64_t)z * MN + i];
  int64_t m = i / N, n = i - m * N;
  T* p = C + goff(cm, m) + goff(cn, n);
  double v = alpha * acc;
  if (beta != 0.0) v += beta * (double)*p;
  *p = (T)v;
}

// Split-K factor: with fewer output tiles than two per CU, split K so that
// about 512 workgroups (two resident per CU) run, keeping >= 64 of K (four
// k-tiles) each.  A single FP64 tile runs at one CU's MFMA rate (~0.3
// TFLOP/s), so few-tile, long-K products (V^T V, V^T A in the blocked QR)
// are bound by how many CUs share their K.
int splitk_factor(int64_t M, int64_t N, int64_t K) {
  const int64_t tiles = ceil_div(M, BM) * ceil_div(N, BN);
  if (tiles >= 256 || K < 128) return 1;
  const int64_t want = (512 + tiles / 2) / tiles;
  return (int)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(want, K / 64), 64));
}

int contract_launch(const void* A, const Modes& a_m, const Modes& a_k, const void* B,
                    const Modes& b_k, const Modes& b_n, void* C, const Modes& c_m,
                    const Modes& c_n, int dtype, int prec, double alpha, double beta,
                    void* splitk_ws, size_t splitk_bytes, hipStream_t st) {
  if (prec == TN_PREC_NATIVE) prec = dtype == TN_F64 ? P_F64 : P_F32;
  TN_REQUIRE(dtype == TN_F64 || dtype == TN_F32, "contract: unsupported dtype %d", dtype);
  TN_REQUIRE(!(dtype == TN_F64 && prec != P_F64),
             "contract: FP64 data supports only native precision (convert to FP32 first)");
  TN_REQUIRE(!(dtype == TN_F32 && prec == P_F64), "contract: FP32 data cannot use FP64 MFMA");
  std::vector<int64_t> md, mas, mcs, nd, nbs, ncs, kd, kas, kbs;
  for (int i = 0; i < a_m.n; ++i) { md.push_back(a_m.dims[i]); mas.push_back(a_m.strides[i]); mcs.push_back(c_m.strides[i]); }
  for (int i = 0; i < b_n.n; ++i) { nd.push_back(b_n.dims[i]); nbs.push_back(b_n.strides[i]); ncs.push_back(c_n.strides[i]); }
  for (int i = 0; i < a_k.n; ++i) { kd.push_back(a_k.dims[i]); kas.push_back(a_k.strides[i]); kbs.push_back(b_k.strides[i]); }
  const int64_t M = prod(md), N = prod(nd), K = prod(kd);
  if (M == 0 || N == 0) return TN_OK;
  TN_REQUIRE(A && B && C, "contract: null pointer");

  bool small_k = K <= 8;
  bool tiny = (double)M * (double)N * (double)std::max<int64_t>(K, 1) < 64.0 * 64.0 * 64.0;
  if (K == 0 || small_k || tiny || (M < 16 && N < 16)) {
    // direct kernel: output modes sorted by C stride (descending) so that
    // consecutive threads store consecutive addresses.
    struct OM { int64_t d, a, b, c; };
    std::vector<OM> om;
    for (size_t i = 0; i < md.size(); ++i) om.push_back({md[i], mas[i], 0, mcs[i]});
    for (size_t i = 0; i < nd.size(); ++i) om.push_back({nd[i], 0, nbs[i], ncs[i]});
    std::stable_sort(om.begin(), om.end(), [](const OM& x, const OM& y) { return x.c > y.c; });
    std::vector<OM> om2;
    for (auto& x : om)
      if (x.d != 1) om2.push_back(x);
    TN_REQUIRE(om2.size() <= 2 * TN_MAX_MODES, "contract: too many output modes");
    merge_group(kd, {&kas, &kbs});
    TN_REQUIRE(kd.size() <= TN_MAX_MODES, "contract: too many contracted modes");
    DirectArgs g{};
    g.A = A; g.B = B; g.C = C;
    g.no = (int)om2.size();
    for (int i = 0; i < g.no; ++i) {
      g.odim[i] = om2[i].d; g.oa[i] = om2[i].a; g.ob[i] = om2[i].b; g.oc[i] = om2[i].c;
    }
    g.nk = K == 0 ? 0 : (int)kd.size();
    for (int i = 0; i < g.nk; ++i) { g.kdim[i] = kd[i]; g.ka[i] = kas[i]; g.kb[i] = kbs[i]; }
    g.nout = M * N;
    g.K = K;
    g.alpha = alpha;
    g.beta = beta;
    unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(g.nout, 256), 65536 * 4);
    ProfScope prof(PROF_DIRECT, st, 2.0 * (double)g.nout * (double)K,
                   (double)(dtype == TN_F64 ? 8 : 4) * ((double)g.nout * 2.0));
    // 32-bit variant: every offset below 2^31 and K <= 8
    auto span = [](const int64_t* d, const int64_t* st_, int n) {
      int64_t m = 0;
      for (int i = 0; i < n; ++i) m += (d[i] - 1) * (st_[i] < 0 ? -st_[i] : st_[i]);
      return m;
    };
    const int64_t lim = (int64_t)1 << 31;
    bool fits = g.nout < lim && K >= 1 && K <= 8;
    for (int i = 0; i < g.no && fits; ++i)
      fits = g.oa[i] >= 0 && g.ob[i] >= 0 && g.oc[i] >= 0;
    for (int i = 0; i < g.nk && fits; ++i) fits = g.ka[i] >= 0 && g.kb[i] >= 0;
    fits = fits && span(g.odim, g.oa, g.no) + span(g.kdim, g.ka, g.nk) < lim &&
           span(g.odim, g.ob, g.no) + span(g.kdim, g.kb, g.nk) < lim &&
           span(g.odim, g.oc, g.no) < lim;
    if (fits) {
      Direct32Args h{};
      h.A = A; h.B = B; h.C = C;
      h.no = g.no;
      h.K = (int)K;
      for (int i = 0; i < g.no; ++i) {
        h.od[i] = make_fastdiv((uint32_t)g.odim[i]);
        h.oa[i] = (int32_t)g.oa[i]; h.ob[i] = (int32_t)g.ob[i]; h.oc[i] = (int32_t)g.oc[i];
      }
      for (int64_t k = 0; k < K; ++k) {   // contracted offsets, mode 0 slowest
        int64_t kidx = k, qa = 0, qb = 0;
        for (int i = g.nk - 1; i >= 0; --i) {
          const int64_t r = kidx % g.kdim[i];
          qa += r * g.ka[i];
          qb += r * g.kb[i];
          kidx /= g.kdim[i];
        }
        h.ka[k] = (int32_t)qa;
        h.kb[k] = (int32_t)qb;
      }
      h.nout = (uint32_t)g.nout;
      h.alpha = alpha;
      h.beta = beta;
      if (dtype == TN_F64)
        hipLaunchKernelGGL((contract_direct32<double, double>), dim3(blocks), dim3(256), 0, st, h);
      else
        hipLaunchKernelGGL((contract_direct32<float, float>), dim3(blocks), dim3(256), 0, st, h);
      TN_CHECK_LAUNCH();
      return TN_OK;
    }
    if (dtype == TN_F64)
      hipLaunchKernelGGL((contract_direct<double, double>), dim3(blocks), dim3(256), 0, st, g);
    else
      hipLaunchKernelGGL((contract_direct<float, float>), dim3(blocks), dim3(256), 0, st, g);
    TN_CHECK_LAUNCH();
    return TN_OK;
  }

  merge_group(md, {&mas, &mcs});
  merge_group(nd, {&nbs, &ncs});
  merge_group(kd, {&kas, &kbs});
  GemmArgs g{};
  g.A = A; g.B = B; g.C = C;
  g.am = make_modes(md, mas);
  g.cm = make_modes(md, mcs);
  g.bn = make_modes(nd, nbs);
  g.cn = make_modes(nd, ncs);
  g.ak = make_modes(kd, kas);
  g.bk = make_modes(kd, kbs);
  g.M = M; g.N = N; g.K = K;
  g.alpha = alpha;
  g.beta = beta;
  // Lanes of a tile load walk the operand's unit-stride dimension.
  const bool a_k_unit = !kas.empty() && kas.back() == 1;
  const bool b_k_unit = !kbs.empty() && kbs.back() == 1;
  const bool b_n_unit = !nbs.empty() && nbs.back() == 1;
  g.a_kfast = a_k_unit ? 1 : 0;
  g.b_nfast = (b_k_unit && !b_n_unit) ? 0 : 1;
  TN_REQUIRE(ceil_div(M, BM) <= 65535, "contract: M too large");
  const int64_t tiles = ceil_div(M, BM) * ceil_div(N, BN);
  // Split K when few output tiles would leave most CUs idle.
  int S = 1;
  const size_t esz = dtype == TN_F64 ? 8 : 4;
  if (splitk_ws != nullptr) {
    S = splitk_factor(M, N, K);
    while (S > 1 && (size_t)S * M * N * esz > splitk_bytes) --S;
  }
  g.kchunk = S > 1 ? ceil_div(ceil_div(K, S), 32) * 32 : K;
  if (S > 1) S = (int)ceil_div(K, g.kchunk);
  g.part = S > 1 ? splitk_ws : nullptr;
  // FP64 products with few 128 x 128 tiles and no K split (short K): 64 x 64
  // tiles, four times the workgroups
  const bool t64 = prec == P_F64 && S == 1 && tiles < 128;
  const int tm = t64 ? 64 : BM;
  dim3 grid((unsigned)ceil_div(N, tm), (unsigned)ceil_div(M, tm), (unsigned)S);
  const size_t esz_in = dtype == TN_F64 ? 8 : 4;
  ProfScope prof(prec == P_F64 ? PROF_GEMM_F64 : (prec == P_F32 ? PROF_GEMM_F32 : PROF_GEMM_BF16X3),
                 st, 2.0 * (double)M * (double)N * (double)K,
                 (double)esz_in * ((double)M * K + (double)K * N + (double)M * N));
  // SIMPLE: single-mode groups whose offsets (tile overshoot included) fit
  // in 31 bits, so the kernel indexes in 32-bit arithmetic
  auto st1 = [](const Modes& x) { return x.n ? x.strides[0] : (int64_t)0; };
  const double lim = 2147483647.0;
  const bool simple =
      g.am.n <= 1 && g.ak.n <= 1 && g.bk.n <= 1 && g.bn.n <= 1 && g.cm.n <= 1 && g.cn.n <= 1 &&
      (double)(M + BM) * st1(g.am) + (double)(K + 64) * st1(g.ak) < lim &&
      (double)(K + 64) * st1(g.bk) + (double)(N + BN) * st1(g.bn) < lim &&
      (double)(M + BM) * st1(g.cm) + (double)(N + BN) * st1(g.cn) < lim;
  // single-mode FP64 operands (the hot case): the tile-load mapping of each
  // operand is a template flag, so only one index mapping is live
  if (prec == P_F64 && simple && t64) {
    if (g.a_kfast && g.b_nfast)
      hipLaunchKernelGGL((gemm_mfma_kernel<P_F64, true, 1, 1, 64>), grid, dim3(NT), 0, st, g);
    else if (g.a_kfast)
      hipLaunchKernelGGL((gemm_mfma_kernel<P_F64, true, 1, 0, 64>), grid, dim3(NT), 0, st, g);
    else if (g.b_nfast)
      hipLaunchKernelGGL((gemm_mfma_kernel<P_F64, true, 0, 1, 64>), grid, dim3(NT), 0, st, g);
    else
      hipLaunchKernelGGL((gemm_mfma_kernel<P_F64, true, 0, 0, 64>), grid, dim3(NT), 0, st, g);
  } else if (prec == P_F64 && simple) {
    if (g.a_kfast && g.b_nfast)
      hipLaunchKernelGGL((gemm_mfma_kernel<P_F64, true, 1, 1, 128>), grid, dim3(NT), 0, st, g);
    else if (g.a_kfast)
      hipLaunchKernelGGL((gemm_mfma_kernel<P_F64, true, 1, 0, 128>), grid, dim3(NT), 0, st, g);
    else if (g.b_nfast)
      hipLaunchKernelGGL((gemm_mfma_kernel<P_F64, true, 0, 1, 128>), grid, dim3(NT), 0, st, g);
    else
      hipLaunchKernelGGL((gemm_mfma_kernel<P_F64, true, 0, 0, 128>), grid, dim3(NT), 0, st, g);
  } else if (prec == P_F64 && t64)
    hipLaunchKernelGGL((gemm_mfma_kernel<P_F64, false, -1, -1, 64>), grid, dim3(NT), 0, st, g);
  else if (prec == P_F64)
    hipLaunchKernelGGL((gemm_mfma_kernel<P_F64, false, -1, -1, 128>), grid, dim3(NT), 0, st, g);
  else if (prec == P_F32 && simple)
    hipLaunchKernelGGL((gemm_mfma_kernel<P_F32, true, -1, -1, 128>), grid, dim3(NT), 0, st, g);
  else if (prec == P_F32)
    hipLaunchKernelGGL((gemm_mfma_kernel<P_F32, false, -1, -1, 128>), grid, dim3(NT), 0, st, g);
  else if (prec == P_BF16X3)
    hipLaunchKernelGGL(gemm_bf16x3_kernel, grid, dim3(NT), 0, st, g);
  else {
    set_error(TN_ERR_UNSUPPORTED, "contract: unknown precision %d", prec);
    return TN_ERR_UNSUPPORTED;
  }
  TN_CHECK_LAUNCH();
  if (S > 1) {
    unsigned blocks = (unsigned)ceil_div(M * N, 256);
    if (dtype == TN_F64)
      hipLaunchKernelGGL(splitk_reduce<double>, dim3(blocks), dim3(256), 0, st,
                         (const double*)splitk_ws, S, M, N, g.cm, g.cn, (double*)C, alpha, beta);
    else
      hipLaunchKernelGGL(splitk_reduce<float>, dim3(blocks), dim3(256), 0, st,
                         (const float*)splitk_ws, S, M, N, g.cm, g.cn, (float*)C, alpha, beta);
    TN_CHECK_LAUNCH();
  }
  return TN_OK;
}

Modes mk(std::initializer_list<int64_t> dims, std::initializer_list<int64_t> strides) {
  Modes m{};
  m.n = (int)dims.size();
  int i = 0;
  for (auto d : dims) m.dims[i++] = d;
  i = 0;
  for (auto s : strides) m.strides[i++] = s;
  for (int j = m.n; j < TN_MAX_MODES; ++j) { m.dims[j] = 1; m.strides[j] = 0; }
  return m;
}

}  // namespace tn

using namespace tn;

extern "C" int tn_contract(const void* A, const tn_modes* a_m, const tn_modes* a_k,
                           const void* B, const tn_modes* b_k, const tn_modes* b_n, void* C,
                           const tn_modes* c_m, const tn_modes* c_n, int dtype, int precision,
                           double alpha, double beta, void* stream) {
  for (const tn_modes* m : {a_m, a_k, b_k, b_n, c_m, c_n})
    TN_REQUIRE(m == nullptr || (m->nmodes >= 0 && m->nmodes <= TN_MAX_MODES),
               "tn_contract: mode group has too many modes");
  TN_REQUIRE(dims_equal(a_m, c_m), "tn_contract: A row modes and C row modes differ");
  TN_REQUIRE(dims_equal(b_n, c_n), "tn_contract: B col modes and C col modes differ");
  TN_REQUIRE(dims_equal(a_k, b_k), "tn_contract: contracted modes of A and B differ");
  Modes am = to_modes(a_m), ak = to_modes(a_k), bk = to_modes(b_k), bn = to_modes(b_n),
        cm = to_modes(c_m), cn = to_modes(c_n);
  for (const Modes* m : {&am, &ak, &bn})
    for (int i = 0; i < m->n; ++i) TN_REQUIRE(m->dims[i] >= 0, "tn_contract: negative dimension");
  return contract_launch(A, am, ak, B, bk, bn, C, cm, cn, dtype, precision, alpha, beta, nullptr,
                         0, as_stream(stream));
}

extern "C" size_t tn_contract_workspace(int64_t M, int64_t N, int64_t K, int dtype) {
  if (M <= 0 || N <= 0 || K <= 0) return 0;
  const int S = splitk_factor(M, N, K);
  return S > 1 ? (size_t)S * (size_t)M * (size_t)N * (dtype == TN_F64 ? 8 : 4) : 0;
}

extern "C" int tn_contract_ws(const void* A, const tn_modes* a_m, const tn_modes* a_k,
                              const void* B, const tn_modes* b_k, const tn_modes* b_n, void* C,
                              const tn_modes* c_m, const tn_modes* c_n, int dtype, int precision,
                              double alpha, double beta, void* ws, size_t ws_bytes, void* stream) {
  for (const tn_modes* m : {a_m, a_k, b_k, b_n, c_m, c_n})
    TN_REQUIRE(m == nullptr || (m->nmodes >= 0 && m->nmodes <= TN_MAX_MODES),
               "tn_contract_ws: mode group has too many modes");
  TN_REQUIRE(dims_equal(a_m, c_m), "tn_contract_ws: A row modes and C row modes differ");
  TN_REQUIRE(dims_equal(b_n, c_n), "tn_contract_ws: B col modes and C col modes differ");
  TN_REQUIRE(dims_equal(a_k, b_k), "tn_contract_ws: contracted modes of A and B differ");
  Modes am = to_modes(a_m), ak = to_modes(a_k), bk = to_modes(b_k), bn = to_modes(b_n),
        cm = to_modes(c_m), cn = to_modes(c_n);
  for (const Modes* m : {&am, &ak, &bn})
    for (int i = 0; i < m->n; ++i) TN_REQUIRE(m->dims[i] >= 0, "tn_contract_ws: negative dimension");
  return contract_launch(A, am, ak, B, bk, bn, C, cm, cn, dtype, precision, alpha, beta,
                         ws_bytes ? ws : nullptr, ws ? ws_bytes : 0, as_stream(stream));
}
