// Batched uneven contraction + truncated SVD (many independent jobs of
// different sizes, e.g. two-site MPS updates: C_j = L_j R_j, then a truncated
// SVD of C_j).  The reference has no batched path: each job is one
// tensordot (tensor.py:735) + one tensor_svd/truncated_svd (tensor.py:830,
// 1112).  Everything below runs every job in shared launches:
//
//  * grouped GEMM (FP32 factors, FP64 MFMA v_mfma_f64_16x16x4 with FP64
//    accumulation, so the rank decided at a relative threshold is the exact
//    product's): one launch covers the tiles of all jobs (block -> (job,
//    tile) by binary search over a tile prefix); the epilogue writes each C_j
//    directly into the Jacobi vector layout (columns of C_j if tall, rows if
//    wide).  Two tile shapes (64x64, 128x128); a small autotuner times both
//    per size bucket on first use and caches the choice.
//  * batched one-sided block Jacobi: every job advances its own round-robin
//    schedule (step g mod (nblk_j - 1)) in the same gram / eig / apply
//    launches; a per-job sweep-end kernel retires converged jobs on the
//    device.  A set of steps (at least one sweep of the largest job) is
//    captured once as a HIP graph and replayed, the device advancing its own
//    step counter; the host reads each set's active-job count while the next
//    set runs and stops replaying when none is left.
//    Vectors whose squared norm is below (thr / 2k)^2 ||C_j||_F^2 are not
//    rotated against each other (bjob_neg2): the rounding-level vectors of a
//    rank-deficient product lie below the keep threshold whatever their
//    orthogonality, and orthogonalising that cluster was most of the sweeps
//    of such jobs.
//    Jobs with 32 < k and L <= 1024 are preconditioned by a
//    Householder QR (tn_bqr.h: per panel, every job's panel in one launch
//    and every job's trailing update in another): the Jacobi runs on the k
//    rows of R (vectors of length k) and the left vectors are Q Z.  On the
//    graded, rank-deficient products of the truncation workload (1024 jobs)
//    this takes at most 12 sweeps where the columns of the products took up
//    to 41.
//  * per-job sort, rank selection (sigma/sigma_0 > threshold, <= chi) and
//    output of U, S, V^T.
#include "tn_bqr.h"
#include "tn_internal.h"
#include "tn_jacobi.h"
#include "tn_prof.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

namespace tn {

typedef float f32x4b __attribute__((ext_vector_type(4)));
typedef double f64x4b __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------- grouped GEMM
struct GJob {
  const float* A;  // M x K row-major
  const float* B;  // K x N row-major
  double* C;       // output: element (r, c) at C[r * crs + c * ccs]
  int32_t M, N, K;
  int64_t crs, ccs;
};

// FP64 = true: the FP32 operands are converted to FP64 (exactly) in the LDS
// store and multiplied by v_mfma_f64_16x16x4, so C carries FP64 rounding
// (~1e-16 relative) instead of FP32's (~1e-7): the rank a relative threshold
// of 1e-8 selects is then the rank of the exact product, as in the
// reference's FP64 tensordot of the same data.  FP64 = false: FP32 MFMA.
template <int FR, bool FP64>  // FR: MFMA fragments per wave along each dimension (tile 32*FR)
__global__ void __launch_bounds__(256) gemm_grouped(const GJob* __restrict__ jobs,
                                                   const int* __restrict__ tile_base,
                                                   const int* __restrict__ job_ids, int njobs) {
  typedef typename std::conditional<FP64, double, float>::type T;
  typedef typename std::conditional<FP64, f64x4b, f32x4b>::type V4;
  constexpr int BT = 32 * FR;  // block tile (square)
  constexpr int BK = 16;
  constexpr int LD = BT + 2;
  constexpr int PER = BT * BK / 256;  // elements of each operand tile per thread
  // double-buffered: the next k-tile's global loads are in flight in
  // registers during the current tile's MFMAs; one barrier per k-tile
  __shared__ T As[2][BK * LD];
  __shared__ T Bs[2][BK * LD];
  // locate job: largest j with tile_base[j] <= blockIdx.x
  int lo = 0, hi = njobs - 1;
  const int b = blockIdx.x;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tile_base[mid] <= b) lo = mid; else hi = mid - 1;
  }
  const GJob g = jobs[job_ids[lo]];
  const int t = b - tile_base[lo];
  const int ntn = (g.N + BT - 1) / BT;
  const int m0 = (t / ntn) * BT, n0 = (t % ntn) * BT;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * (16 * FR), wc = (wave & 1) * (16 * FR);
  const int lr = lane & 15, lk = lane >> 4;
  V4 acc[FR][FR];
#pragma unroll
  for (int i = 0; i < FR; ++i)
#pragma unroll
    for (int j = 0; j < FR; ++j) acc[i][j] = V4{0, 0, 0, 0};
  float ra[PER], rb[PER];
  // A tile: lanes walk k (unit stride); B tile: lanes walk n (unit stride)
  auto fetch = [&](int k0) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = tid + 256 * i;
      const int kk = e % BK, r = e / BK;
      const int m = m0 + r, k = k0 + kk;
      ra[i] = (m < g.M && k < g.K) ? g.A[(int64_t)m * g.K + k] : 0.f;
      const int c = e % BT, kb = e / BT;
      const int n = n0 + c, k2 = k0 + kb;
      rb[i] = (n < g.N && k2 < g.K) ? g.B[(int64_t)k2 * g.N + n] : 0.f;
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = tid + 256 * i;
      As[buf][(e % BK) * LD + e / BK] = (T)ra[i];
      Bs[buf][(e / BT) * LD + e % BT] = (T)rb[i];
    }
  };
  const int nkt = (g.K + BK - 1) / BK;
  if (nkt > 0) {
    fetch(0);
    stash(0);
  }
  __syncthreads();
  for (int kt = 0; kt < nkt; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nkt) fetch((kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      T af[FR], bf[FR];
#pragma unroll
      for (int i = 0; i < FR; ++i) af[i] = As[buf][(ks + lk) * LD + wr + 16 * i + lr];
#pragma unroll
      for (int j = 0; j < FR; ++j) bf[j] = Bs[buf][(ks + lk) * LD + wc + 16 * j + lr];
#pragma unroll
      for (int i = 0; i < FR; ++i)
#pragma unroll
        for (int j = 0; j < FR; ++j) {
          if constexpr (FP64)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
          else
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
    }
    if (kt + 1 < nkt) stash(buf ^ 1);
    __syncthreads();
  }
  // accumulator element e of lane (lr, lk): row 4 lk + e (FP32) or lk + 4 e
  // (FP64) of the 16 x 16 fragment, column lr
#pragma unroll
  for (int i = 0; i < FR; ++i)
#pragma unroll
    for (int j = 0; j < FR; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        int r = m0 + wr + 16 * i + (FP64 ? lk + 4 * e : 4 * lk + e), c = n0 + wc + 16 * j + lr;
        if (r < g.M && c < g.N) g.C[(int64_t)r * g.crs + (int64_t)c * g.ccs] = (double)acc[i][j][e];
      }
}

// ---------------------------------------------------------------- batched Jacobi
constexpr int BJB = 16, BJ2 = 32, BCH = 256;
constexpr int BANT = 2, BACH = 64 * BANT;   // row tiles per wave / rows per chunk of bjac_apply
static_assert(BJB == JB && BJ2 == J2, "batched pairs use the shared pair layout");

struct BJob {
  int64_t wv;     // offset of the job's Jacobi vectors (k vectors of length L: the
                  // columns or rows of its matrix, or the rows of R)
  int64_t z;      // offset of the job's rotation accumulator (k x k)
  int32_t L, k, nblk, nch;
  int32_t ncha, nchaz;   // row chunks of W and Z in bjac_apply (BACH rows)
  int32_t gram_base, eig_base;
  double tol;
  // squared norm below which a vector is negligible (pairs of two such
  // vectors are not rotated; 0: none is)
  double neg2;
};

// acc += the 32 x 32 Gram of the pair's vectors (blocks I and K2 of the
// job's k vectors of length L at W) over rows [r0, r1): 64-row tiles staged
// through LDS, v_mfma_f64_16x16x4 by four waves (one 16 x 16 quarter each);
// the next tile's loads are issued into registers before the current
// tile's MFMAs, so their latency overlaps the MFMAs.
__device__ __forceinline__ void bgram_rows(const double* __restrict__ W, int64_t L, int k, int I,
                                           int K2, int64_t r0, int64_t r1,
                                           double (*Ps)[BJ2 + 1], f64x4b& acc) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int ti = wave >> 1, tj = wave & 1;
  constexpr int PE = 64 * BJ2 / 256;
  double pv[PE];
  auto fetch = [&](int64_t rb) {
#pragma unroll
    for (int i = 0; i < PE; ++i) {
      const int e = tid + 256 * i, row = e & 63, col = e >> 6;
      const int64_t v = col < BJB ? (int64_t)I * BJB + col : (int64_t)K2 * BJB + (col - BJB);
      const int64_t r = rb + row;
      pv[i] = (v < k && r < r1) ? W[v * L + r] : 0.0;
    }
  };
  if (r0 < r1) fetch(r0);
  for (int64_t rb = r0; rb < r1; rb += 64) {
#pragma unroll
    for (int i = 0; i < PE; ++i) {
      const int e = tid + 256 * i;
      Ps[e & 63][e >> 6] = pv[i];
    }
    __syncthreads();
    if (rb + 64 < r1) fetch(rb + 64);
#pragma unroll
    for (int k0 = 0; k0 < 64; k0 += 4) {
      double a = Ps[k0 + (lane >> 4)][16 * ti + (lane & 15)];
      double b = Ps[k0 + (lane >> 4)][16 * tj + (lane & 15)];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
}

__device__ __forceinline__ int brr(int pos, int step, int nblk) {
  return pos == 0 ? 0 : 1 + (pos - 1 + step) % (nblk - 1);
}

__global__ void __launch_bounds__(256) bjac_gram(double* __restrict__ Wv, const BJob* __restrict__ jobs,
                                                 const int* __restrict__ it_job,
                                                 const int* __restrict__ it_pair,
                                                 const int* __restrict__ it_chunk,
                                                 const int* __restrict__ done,
                                                 const int* __restrict__ gbase, int goff,
                                                 double* __restrict__ Gpart,
                                                 unsigned long long* __restrict__ work) {
  __shared__ double Ps[64][BJ2 + 1];
  const int item = blockIdx.x;
  const int j = it_job[item];
  if (done[j]) return;
  const BJob J = jobs[j];
  const int s = (*gbase + goff) % (J.nblk - 1);
  const int pi = it_pair[item], ch = it_chunk[item];
  const int I = brr(pi, s, J.nblk), K2 = brr(J.nblk - 1 - pi, s, J.nblk);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int ti = wave >> 1, tj = wave & 1;
  const int64_t r0 = (int64_t)ch * BCH, r1 = min((int64_t)J.L, r0 + BCH);
  if (work && tid == 0) atomicAdd(work, (unsigned long long)(r1 - r0));
  const double* base = Wv + J.wv;
  f64x4b acc = {0, 0, 0, 0};
  bgram_rows(base, J.L, J.k, I, K2, r0, r1, Ps, acc);
  double* out = Gpart + (int64_t)item * (BJ2 * BJ2);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    int r = 16 * ti + (lane >> 4) + 4 * e, c = 16 * tj + (lane & 15);
    out[r * BJ2 + c] = acc[e];
  }
}

// eig items are (job, pair) in job-major order; Gpart items for (job, pair,
// chunk) start at gram_base + pair * nch.
__global__ void __launch_bounds__(256) bjac_eig(const double* __restrict__ Gpart,
                                                const BJob* __restrict__ jobs,
                                                const int* __restrict__ e_job,
                                                const int* __restrict__ e_pair,
                                                const int* __restrict__ done,
                                                double* __restrict__ Qall, int* __restrict__ skip,
                                                unsigned long long* __restrict__ offmax,
                                                unsigned long long* __restrict__ work) {
  __shared__ EigLds sm;
  const int item = blockIdx.x, tid = threadIdx.x;
  const int j = e_job[item];
  if (done[j]) return;
  const BJob J = jobs[j];
  const int pi = e_pair[item];
  {
    // every partial of a group of four chunks loaded before the first add
    constexpr int PER = BJ2 * BJ2 / 256;
    const double* Gp = Gpart + ((int64_t)J.gram_base + (int64_t)pi * J.nch) * (BJ2 * BJ2);
    double acc[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) acc[i] = 0.0;
    int c = 0;
    for (; c + 4 <= J.nch; c += 4) {
      double v[4][PER];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int i = 0; i < PER; ++i) v[u][i] = Gp[(int64_t)(c + u) * (BJ2 * BJ2) + tid + 256 * i];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int i = 0; i < PER; ++i) acc[i] += v[u][i];
    }
    for (; c < J.nch; ++c) {
      double v[PER];
#pragma unroll
      for (int i = 0; i < PER; ++i) v[i] = Gp[(int64_t)c * (BJ2 * BJ2) + tid + 256 * i];
#pragma unroll
      for (int i = 0; i < PER; ++i) acc[i] += v[i];
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = tid + 256 * i;
      sm.G[0][e / BJ2][e % BJ2] = acc[i];
      sm.Q[0][e / BJ2][e % BJ2] = (e / BJ2 == e % BJ2) ? 1.0 : 0.0;
    }
  }
  __syncthreads();
  const double off = eig_off(sm, J.tol, J.neg2);
  if (tid == 0) atomicMax(offmax + j, dbits(off));
  if (off <= J.tol) {
    if (tid == 0) skip[item] = 1;
    return;
  }
  if (tid == 0) skip[item] = 0;
  if (work && tid == 0) atomicAdd(work + 1, 1ull);
  const int cur = eig_rotate(sm, J.tol, J.neg2);
  double* out = Qall + (int64_t)item * (BJ2 * BJ2);
  for (int e = tid; e < BJ2 * BJ2; e += 256) out[e] = sm.Q[cur][e / BJ2][e % BJ2];
}

// apply items are (job, pair, chunk) over the L + k rows of [W; Z]
__global__ void __launch_bounds__(256) bjac_apply(double* __restrict__ Wv, double* __restrict__ Zb,
                                                  const BJob* __restrict__ jobs,
                                                  const int* __restrict__ a_job,
                                                  const int* __restrict__ a_pair,
                                                  const int* __restrict__ a_chunk,
                                                  const int* __restrict__ done,
                                                  const int* __restrict__ gbase, int goff,
                                                  const double* __restrict__ Qall,
                                                  const int* __restrict__ skip,
                                                  unsigned long long* __restrict__ work) {
  __shared__ double Q[BJ2][BJ2 + 1];
  const int item = blockIdx.x;
  const int j = a_job[item];
  if (done[j]) return;
  const BJob J = jobs[j];
  const int pi = a_pair[item], ch = a_chunk[item];
  const int eidx = J.eig_base + pi;
  if (skip[eidx]) return;
  const int s = (*gbase + goff) % (J.nblk - 1);
  const int I = brr(pi, s, J.nblk), K2 = brr(J.nblk - 1 - pi, s, J.nblk);
  const double* Qg = Qall + (int64_t)eidx * BJ2 * BJ2;
  const bool onW = ch < J.ncha;
  double* base = onW ? Wv + J.wv : Zb + J.z;
  const int64_t ld = onW ? J.L : J.k;
  const int64_t r = (int64_t)(onW ? ch : ch - J.ncha) * BACH;
  const int64_t re = min(ld, r + BACH);
  if (work && threadIdx.x == 0) atomicAdd(work + 2, (unsigned long long)(re - r));
  // Q's loads and the chunk's loads are all issued before Q goes into LDS
  constexpr int PER = BJ2 * BJ2 / 256;
  double qv[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) qv[i] = Qg[threadIdx.x + 256 * i];
  double x[8][BANT];
  apply_rows_fetch<BANT>(base, ld, J.k, r, re, I, K2, x);
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = threadIdx.x + 256 * i;
    Q[e / BJ2][e % BJ2] = qv[i];
  }
  lds_barrier();
  apply_rows_finish<BANT>(base, ld, J.k, r, re, I, K2, x, Q);
}

// One Jacobi step for batches whose jobs fit one row chunk (L <= BCH and
// k <= BCH): the block of each (job, pair) eig item forms the pair's
// 32 x 32 Gram over all L rows of W in LDS, runs the pair eigensolve and
// rotates the pair's rows of W and then of Z.  One block owns the pair's
// vectors for the whole step (a second block rotating them while this one
// still reads them for its Gram would race).  The same arithmetic as
// bjac_gram + bjac_eig + bjac_apply, in one launch and without the partial
// Grams and rotations passing through global memory.
__global__ void __launch_bounds__(256) bjac_fused(double* __restrict__ Wv, double* __restrict__ Zb,
                                                  const BJob* __restrict__ jobs,
                                                  const int* __restrict__ e_job,
                                                  const int* __restrict__ e_pair,
                                                  const int* __restrict__ done,
                                                  const int* __restrict__ gbase, int goff,
                                                  unsigned long long* __restrict__ offmax,
                                                  unsigned long long* __restrict__ work) {
  __shared__ double Ps[64][BJ2 + 1];
  __shared__ EigLds sm;
  const int item = blockIdx.x, tid = threadIdx.x;
  const int j = e_job[item];
  if (done[j]) return;
  const BJob J = jobs[j];
  const int s = (*gbase + goff) % (J.nblk - 1);
  const int pi = e_pair[item];
  const int I = brr(pi, s, J.nblk), K2 = brr(J.nblk - 1 - pi, s, J.nblk);
  const int wave = tid >> 6, lane = tid & 63;
  const int ti = wave >> 1, tj = wave & 1;
  double* W = Wv + J.wv;
  double* Z = Zb + J.z;
  double x[8][4], z[8][4];
  apply_rows_fetch<4>(W, J.L, J.k, 0, J.L, I, K2, x);  // in flight during the Gram and rounds
  apply_rows_fetch<4>(Z, J.k, J.k, 0, J.k, I, K2, z);
  f64x4b acc = {0, 0, 0, 0};
  bgram_rows(W, J.L, J.k, I, K2, 0, J.L, Ps, acc);
#pragma unroll
  for (int e = 0; e < 4; ++e) sm.G[0][16 * ti + (lane >> 4) + 4 * e][16 * tj + (lane & 15)] = acc[e];
  for (int e = tid; e < BJ2 * BJ2; e += 256) sm.Q[0][e / BJ2][e % BJ2] = (e / BJ2 == e % BJ2) ? 1.0 : 0.0;
  __syncthreads();
  const double off = eig_off(sm, J.tol, J.neg2);
  if (tid == 0) atomicMax(offmax + j, dbits(off));
  if (off <= J.tol) return;
  if (work && tid == 0) {
    atomicAdd(work, (unsigned long long)J.L);
    atomicAdd(work + 1, 1ull);
    atomicAdd(work + 2, (unsigned long long)(J.L + J.k));
  }
  const int cur = eig_rotate(sm, J.tol, J.neg2);
  apply_rows_finish<4>(W, J.L, J.k, 0, J.L, I, K2, x, sm.Q[cur]);
  apply_rows_finish<4>(Z, J.k, J.k, 0, J.k, I, K2, z, sm.Q[cur]);
}

// After global step g: retire jobs whose sweep just ended with off <= tol.
__global__ void bjac_sweep_end(const BJob* __restrict__ jobs, int njobs,
                               const int* __restrict__ gbase, int goff, int max_sweeps,
                               int* __restrict__ done, int* __restrict__ sweeps,
                               unsigned long long* __restrict__ offmax, int* __restrict__ active) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= njobs || done[j]) return;
  const BJob J = jobs[j];
  if ((*gbase + goff + 1) % (J.nblk - 1) != 0) return;
  double off = __longlong_as_double((long long)offmax[j]);
  offmax[j] = 0ull;
  int sw = ++sweeps[j];
  if (off <= J.tol) {
    done[j] = 1;
    atomicSub(active, 1);
  } else if (sw >= max_sweeps) {
    done[j] = 2;
    atomicSub(active, 1);
  }
}

// The global step of a launch is *gbase + goff: a captured set of steps
// (goff = 0 .. n-1) is replayed with the base advanced by n after each set.
__global__ void badvance(int* __restrict__ gbase, int n) {
  if (threadIdx.x == 0) *gbase += n;
}

__global__ void binit_identity(double* __restrict__ Zb, const BJob* __restrict__ jobs) {
  const BJob J = jobs[blockIdx.y];
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)J.k * J.k) return;
  Zb[J.z + idx] = (idx / J.k == idx % J.k) ? 1.0 : 0.0;
}

// neg2 of every job from its Frobenius norm F (the sum of its vectors'
// squared norms, which the rotations keep): with every vector of squared norm
// <= neg2 = (thr / (2 k))^2 F^2 left unrotated against the others like it,
// the space those vectors span has singular values <= sqrt(k neg2) <=
// thr F / (2 sqrt(k)) <= thr sigma_0 / 2 (F <= sqrt(k) sigma_0), so it lies
// below the relative threshold thr whatever its internal orthogonality.
__global__ void __launch_bounds__(256) bjob_neg2(const double* __restrict__ Wv,
                                                 BJob* __restrict__ jobs, double thr) {
  __shared__ double red[4];
  BJob& J = jobs[blockIdx.x];
  const int64_t tot = (int64_t)J.L * J.k;
  double acc = 0.0;
  for (int64_t e = threadIdx.x; e < tot; e += 256) {
    const double x = Wv[J.wv + e];
    acc = fma(x, x, acc);
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double f2 = red[0] + red[1] + red[2] + red[3];
    const double q = thr / (2.0 * (double)J.k);
    J.neg2 = q * q * f2;
  }
}

__global__ void __launch_bounds__(256) bvec_norms(const double* __restrict__ Wv,
                                                  const BJob* __restrict__ jobs,
                                                  double* __restrict__ sig, const int64_t* sbase) {
  __shared__ double red[4];
  const BJob J = jobs[blockIdx.y];
  const int v = blockIdx.x;
  if (v >= J.k) return;
  double acc = 0.0;
  for (int64_t r = threadIdx.x; r < J.L; r += 256) {
    double x = Wv[J.wv + (int64_t)v * J.L + r];
    acc = fma(x, x, acc);
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) sig[sbase[blockIdx.y] + v] = sqrt(red[0] + red[1] + red[2] + red[3]);
}

// per job: sort descending, count kept = #{sigma/sigma_0 > thr}, <= chi
__global__ void __launch_bounds__(1024) bsort_rank(const double* __restrict__ sig,
                                                   const BJob* __restrict__ jobs,
                                                   const int64_t* __restrict__ sbase,
                                                   int* __restrict__ perm,
                                                   double* __restrict__ Ssorted, double thr,
                                                   int chi, int* __restrict__ kept) {
  __shared__ int cnt;
  const int j = blockIdx.x;
  const BJob J = jobs[j];
  const double* s = sig + sbase[j];
  double smax = 0.0;
  for (int i = 0; i < J.k; ++i) smax = fmax(smax, s[i]);
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int local = 0;
  // NaN ranks as -inf so that the ranks stay a permutation of 0..k-1
  for (int v = threadIdx.x; v < J.k; v += 1024) {
    const double sv = s[v];
    const double kv = sv == sv ? sv : -INFINITY;
    int rank = 0;
    for (int i = 0; i < J.k; ++i) {
      const double ki = s[i] == s[i] ? s[i] : -INFINITY;
      rank += (ki > kv || (ki == kv && i < v)) ? 1 : 0;
    }
    perm[sbase[j] + rank] = v;
    Ssorted[sbase[j] + rank] = sv;
    if (smax > 0.0 && sv / smax > thr) ++local;
  }
  atomicAdd(&cnt, local);
  __syncthreads();
  if (threadIdx.x == 0) {
    int k = cnt;
    if (chi > 0 && k > chi) k = chi;
    kept[j] = k;
  }
}

struct OJob {
  void* U;     // m x k row-major (FP32 or FP64 output)
  void* Vt;    // k x n row-major
  double* S;   // k
  const double* qz;  // QR-preconditioned jobs: Q Z, the left vectors of the
                     // job's vector matrix in sorted order (L x kept); else null
  int32_t m, n;
  int32_t tall;
  int32_t f64;  // outputs are FP64
};

// FP64 input path: Wv vectors of job j from its row-major m x n matrix A
// (columns when m >= n, rows otherwise).
struct AJob {
  const double* A;
  int64_t x;  // offset of the job's vector matrix (L x k) in the vector buffer
};

__global__ void bgather_f64(double* __restrict__ Wv, const AJob* __restrict__ ajobs,
                            const OJob* __restrict__ outs) {
  const int j = blockIdx.y;
  const OJob O = outs[j];
  const double* A = ajobs[j].A;
  const int64_t L = O.tall ? O.m : O.n, k = O.tall ? O.n : O.m;
  const int64_t tot = L * k;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = e / L, i = e - v * L;
    Wv[ajobs[j].x + e] = O.tall ? A[i * O.n + v] : A[v * O.n + i];
  }
}

// Vector matrix X (L x k: the columns of the job's matrix if tall, its rows
// otherwise) = U_X S V_X^T.  Plain jobs: U_X = the rotated vectors / S, V_X =
// Z.  QR-preconditioned jobs (X = Q R, Jacobi on the rows of R): U_X = Q Z
// (O.qz, columns c < kept in sorted order; zero beyond), V_X = the rotated
// rows of R / S.  Tall: U = U_X, V^T = V_X^T; wide: U = V_X, V^T = U_X^T.
__global__ void bout(const double* __restrict__ Wv, const double* __restrict__ Zb,
                     const BJob* __restrict__ jobs, const OJob* __restrict__ outs,
                     const int64_t* __restrict__ sbase, const int* __restrict__ perm,
                     const double* __restrict__ Ss, const int* __restrict__ kept) {
  const int j = blockIdx.y;
  const BJob J = jobs[j];
  const OJob O = outs[j];
  const int64_t k = J.k;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t tot = (int64_t)O.m * k + k * O.n;
  const int* pm = perm + sbase[j];
  const double* S = Ss + sbase[j];
  if (O.qz) {
    const int64_t L = O.tall ? O.m : O.n, kp = kept[j];
    for (; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
      double v;
      if (idx < (int64_t)O.m * k) {
        const int64_t i = idx / k, c = idx - i * k;
        if (O.tall) v = c < kp ? O.qz[c * L + i] : 0.0;
        else v = S[c] > 0.0 ? Wv[J.wv + (int64_t)pm[c] * k + i] / S[c] : 0.0;
        if (O.f64) ((double*)O.U)[idx] = v;
        else ((float*)O.U)[idx] = (float)v;
      } else {
        const int64_t e = idx - (int64_t)O.m * k;
        const int64_t r = e / O.n, c = e - r * O.n;
        if (O.tall) v = S[r] > 0.0 ? Wv[J.wv + (int64_t)pm[r] * k + c] / S[r] : 0.0;
        else v = r < kp ? O.qz[r * L + c] : 0.0;
        if (O.f64) ((double*)O.Vt)[e] = v;
        else ((float*)O.Vt)[e] = (float)v;
      }
    }
    if (blockIdx.x == 0)
      for (int64_t c = threadIdx.x; c < k; c += blockDim.x) O.S[c] = S[c];
    return;
  }
  for (; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
    if (idx < (int64_t)O.m * k) {
      int64_t i = idx / k, c = idx - i * k;
      int64_t src = pm[c];
      double v = O.tall ? (S[c] > 0.0 ? Wv[J.wv + src * J.L + i] / S[c] : 0.0)
                        : Zb[J.z + src * k + i];
      if (O.f64) ((double*)O.U)[idx] = v;
      else ((float*)O.U)[idx] = (float)v;
    } else {
      int64_t e = idx - (int64_t)O.m * k;
      int64_t r = e / O.n, c = e - r * O.n;
      int64_t src = pm[r];
      double v = O.tall ? Zb[J.z + src * k + c]
                        : (S[r] > 0.0 ? Wv[J.wv + src * J.L + c] / S[r] : 0.0);
      if (O.f64) ((double*)O.Vt)[e] = v;
      else ((float*)O.Vt)[e] = (float)v;
    }
  }
  if (blockIdx.x == 0)
    for (int64_t c = threadIdx.x; c < k; c += blockDim.x) O.S[c] = S[c];
}

// ---------------------------------------------------------------- autotune
std::mutex g_tune_mu;
std::map<int, int> g_tune;  // size bucket -> FR (2: 64x64 tiles, 4: 128x128)

static int bucket_of(int M, int N) {
  int x = std::max(M, N);
  return x <= 128 ? 0 : x <= 256 ? 1 : x <= 512 ? 2 : 3;
}

static int launch_grouped(int FR, const std::vector<int>& ids, const std::vector<GJob>& hjobs,
                          GJob* djobs, int* dtile, int* dids, hipStream_t st) {
  if (ids.empty()) return TN_OK;
  const int BT = 32 * FR;
  std::vector<int> base(ids.size());
  int tot = 0;
  for (size_t i = 0; i < ids.size(); ++i) {
    base[i] = tot;
    const GJob& g = hjobs[ids[i]];
    tot += (int)(ceil_div(g.M, BT) * ceil_div(g.N, BT));
  }
  TN_HIP(hipMemcpyAsync(dtile, base.data(), sizeof(int) * base.size(), hipMemcpyHostToDevice, st));
  TN_HIP(hipMemcpyAsync(dids, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice, st));
  TN_HIP(hipStreamSynchronize(st));  // host tables are locals
  if (tot == 0) return TN_OK;
  if (FR == 2)
    hipLaunchKernelGGL((gemm_grouped<2, true>), dim3(tot), dim3(256), 0, st, djobs, dtile, dids,
                       (int)ids.size());
  else
    hipLaunchKernelGGL((gemm_grouped<4, true>), dim3(tot), dim3(256), 0, st, djobs, dtile, dids,
                       (int)ids.size());
  TN_CHECK_LAUNCH();
  return TN_OK;
}

static double time_grouped(int FR, const std::vector<int>& ids, const std::vector<GJob>& hjobs,
                           GJob* djobs, int* dtile, int* dids, hipStream_t st) {
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return 1e30;
  (void)hipEventRecord(a, st);
  int rc = launch_grouped(FR, ids, hjobs, djobs, dtile, dids, st);
  (void)hipEventRecord(b, st);
  (void)hipEventSynchronize(b);
  float ms = 1e30f;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc == TN_OK ? ms : 1e30;
}


// QR preconditioning (tn_bqr.h) for the jobs whose rows fit a workgroup's
// registers and that have more than one block pair
static bool qr_precond(int64_t L, int64_t k) { return k > BJ2 && L <= kQrRows; }

// Work items of a preconditioned job with k columns: one panel item per
// panel; trailing-update groups of 64 columns after each panel; Q-apply
// groups of 64 of the (at most k) kept columns per panel.
static void qr_item_counts(int64_t k, size_t* npan, size_t* ntrail, size_t* napply) {
  const int64_t np = ceil_div(k, QNB);
  *npan += (size_t)np;
  for (int64_t p = 0; p < np; ++p) *ntrail += (size_t)ceil_div(std::max<int64_t>(k - (p + 1) * QNB, 0), 64);
  *napply += (size_t)(np * ceil_div(k, 64));
}

static size_t batched_ws_bytes(const std::vector<int64_t>& ms, const std::vector<int64_t>& ns) {
  const int njobs = (int)ms.size();
  if (njobs <= 0) return 256;
  size_t wv = 0, z = 0, gi = 0, ei = 0, ai = 0, sv = 0, qz = 0, tf = 0, npre = 0;
  size_t npan = 0, ntrail = 0, napply = 0;
  for (int j = 0; j < njobs; ++j) {
    int64_t k = std::min(ms[j], ns[j]), L = std::max(ms[j], ns[j]);
    const bool pre = qr_precond(L, k);
    const int64_t Lj = pre ? k : L;  // Jacobi vector length
    int nblk = (int)ceil_div(std::max<int64_t>(k, 1), BJB);
    nblk += nblk % 2;
    int np = nblk / 2;
    wv += (size_t)(L * k) + (pre ? (size_t)(k * k) : 0);
    z += (size_t)(k * k);
    gi += (size_t)np * ceil_div(Lj, BCH);
    ei += (size_t)np;
    ai += (size_t)np * (ceil_div(Lj, BACH) + ceil_div(k, BACH));
    sv += (size_t)k;
    if (pre) {
      qz += (size_t)(L * k);
      tf += (size_t)ceil_div(k, QNB) * QNB * QNB;
      ++npre;
      qr_item_counts(k, &npan, &ntrail, &napply);
    }
  }
  Arena a{nullptr, 0, 0};
  a.take<double>(wv);
  a.take<double>(z);
  a.take<double>(qz);
  a.take<double>(tf);
  a.take<QJob>(npre);
  a.take<int>(npan);
  a.take<QItem>(ntrail);
  a.take<QItem>(napply);
  a.take<double>(gi * BJ2 * BJ2);
  a.take<double>(ei * BJ2 * BJ2);
  a.take<int>(ei);
  a.take<int>(gi * 3);
  a.take<int>(ei * 2);
  a.take<int>(ai * 3);
  a.take<BJob>((size_t)njobs);
  a.take<OJob>((size_t)njobs);
  a.take<GJob>((size_t)njobs);
  a.take<AJob>((size_t)njobs);
  a.take<int>((size_t)njobs * 2);  // tile base + ids
  a.take<int>((size_t)njobs * 4);  // done, sweeps, kept, active
  a.take<unsigned long long>((size_t)njobs);
  a.take<unsigned long long>(3);  // work counters
  a.take<double>(sv * 2);
  a.take<int>(sv);
  a.take<int64_t>((size_t)njobs);
  a.take<int>(1);  // device step base
  return a.used + 4096;
}

// One driver for both entry points: tj != nullptr forms each job's matrix as
// the FP32 product L R (grouped GEMM); sj != nullptr gathers FP64 matrices.
static int batched_driver(const tn_trunc_job* tj, const tn_svd_job* sj, int njobs,
                          double rel_threshold, int chi_max, int max_sweeps, int* kept_out,
                          int* sweeps_out, void* ws, size_t ws_bytes, void* stream) {
  const char* who = tj ? "tn_batched_truncate" : "tn_batched_svd";
  TN_REQUIRE(njobs >= 0 && (tj != nullptr || sj != nullptr), "%s: bad job list", who);
  if (njobs == 0) return TN_OK;
  std::vector<int64_t> ms(njobs), ns(njobs);
  for (int j = 0; j < njobs; ++j) {
    ms[j] = tj ? tj[j].m : sj[j].m;
    ns[j] = tj ? tj[j].n : sj[j].n;
  }
  TN_REQUIRE(ws && ws_bytes >= batched_ws_bytes(ms, ns), "%s: workspace too small", who);
  if (max_sweeps <= 0) max_sweeps = 40;
  hipStream_t st = as_stream(stream);
  // ---- host tables
  std::vector<BJob> bj(njobs);
  std::vector<OJob> oj(njobs);
  std::vector<GJob> gj(njobs);
  std::vector<int> g_job, g_pair, g_chunk, e_job, e_pair, a_job, a_pair, a_chunk;
  std::vector<int64_t> sbase(njobs), xoff(njobs);
  std::vector<QJob> qj;
  int64_t wv = 0, z = 0, sv = 0, qzs = 0, tfs = 0, rvs = 0;
  int maxsteps = 1, maxk = 1;
  std::vector<AJob> aj(njobs);
  for (int j = 0; j < njobs; ++j) {
    tn_trunc_job J{};
    if (tj) {
      J = tj[j];
      TN_REQUIRE(J.m > 0 && J.n > 0 && J.k > 0 && J.L && J.R && J.U && J.S && J.Vt,
                 "%s: job %d is invalid", who, j);
    } else {
      J.m = sj[j].m;
      J.n = sj[j].n;
      J.k = 1;
      J.U = sj[j].U;
      J.S = sj[j].S;
      J.Vt = sj[j].Vt;
      TN_REQUIRE(J.m > 0 && J.n > 0 && sj[j].A && J.U && J.S && J.Vt, "%s: job %d is invalid",
                 who, j);
      aj[j].A = (const double*)sj[j].A;
      aj[j].x = wv;
    }
    TN_REQUIRE(J.m <= (1 << 20) && J.n <= (1 << 20) && J.k <= (1 << 20), "%s: job %d too large",
               who, j);
    const int64_t k = std::min(J.m, J.n), L = std::max(J.m, J.n);
    const bool pre = qr_precond(L, k);
    const int64_t Lj = pre ? k : L;  // Jacobi vector length
    BJob b{};
    xoff[j] = wv;
    b.wv = wv;  // preconditioned: moved to the rows of R below
    b.z = z;
    b.L = (int)Lj;
    b.k = (int)k;
    b.nblk = (int)ceil_div(k, BJB);
    b.nblk += b.nblk % 2;
    b.nch = (int)ceil_div(Lj, BCH);
    b.ncha = (int)ceil_div(Lj, BACH);
    b.nchaz = (int)ceil_div(k, BACH);
    b.gram_base = (int)g_job.size();
    b.eig_base = (int)e_job.size();
    // Orthogonality to L * eps (LAPACK dgesvj's threshold) for the SVD jobs;
    // the truncation jobs, whose U and V^T are returned in FP32, stop at
    // max(1e-8, L eps): their kept singular values are then exact to ~1e-16
    // relative and their vectors to ~1e-8, below the FP32 outputs' rounding
    // (6e-8), while the clusters of tiny singular values that the FP32
    // rounding of the factors leaves (~1e-8 sigma_0) would need up to twice
    // the sweeps to reach L eps (measured: 1e-10 -> 1e-8, 2062 -> 2145 jobs/s
    // on the 1024-job batch, singular values still within 8e-14 sigma_0).
    b.tol = tj ? std::max(1e-8, (double)Lj * 2.220446049250313e-16)
               : (double)Lj * 2.220446049250313e-16;
    b.neg2 = 0.0;
    for (int p = 0; p < b.nblk / 2; ++p) {
      for (int c = 0; c < b.nch; ++c) { g_job.push_back(j); g_pair.push_back(p); g_chunk.push_back(c); }
      e_job.push_back(j);
      e_pair.push_back(p);
      for (int c = 0; c < b.ncha + b.nchaz; ++c) { a_job.push_back(j); a_pair.push_back(p); a_chunk.push_back(c); }
    }
    bj[j] = b;
    sbase[j] = sv;
    const bool tall = J.m >= J.n;
    gj[j] = GJob{(const float*)J.L, (const float*)J.R, nullptr, (int)J.m, (int)J.n, (int)J.k,
                 tall ? 1 : (int64_t)J.n, tall ? (int64_t)J.m : 1};
    oj[j] = OJob{J.U, J.Vt, J.S, nullptr, (int)J.m, (int)J.n, tall ? 1 : 0, tj ? 0 : 1};
    if (pre) {
      QJob q{};
      q.x = wv;
      q.r = rvs;  // relative to the R region, rebased below
      q.qz = qzs;
      q.tf = tfs;
      q.z = z;
      q.L = (int)L;
      q.k = (int)k;
      q.job = j;
      qj.push_back(q);
      rvs += k * k;
      qzs += L * k;
      tfs += ceil_div(k, QNB) * QNB * QNB;
    }
    wv += L * k;
    z += k * k;
    sv += k;
    maxsteps = std::max(maxsteps, b.nblk - 1);
    maxk = std::max<int>(maxk, (int)k);
  }
  const int ngi = (int)g_job.size(), nei = (int)e_job.size(), nai = (int)a_job.size();
  const int npre = (int)qj.size();
  // the rows of R follow the matrices in the vector buffer
  for (QJob& q : qj) {
    q.r += wv;
    bj[q.job].wv = q.r;
  }
  // ---- device workspace
  Arena a{(char*)ws, ws_bytes, 0};
  double* dWv = a.take<double>((size_t)(wv + rvs));
  double* dZ = a.take<double>((size_t)z);
  double* dQZ = a.take<double>((size_t)qzs);
  double* dTf = a.take<double>((size_t)tfs);
  QJob* dqj = a.take<QJob>((size_t)npre);
  // per-panel work lists of the QR (panels in order) and of the Q application
  int maxpan = 0;
  for (const QJob& q : qj) maxpan = std::max<int>(maxpan, (int)ceil_div(q.k, QNB));
  std::vector<int> pan_items;
  std::vector<QItem> trail_items, apply_items;
  std::vector<size_t> pan_off(maxpan + 1), trail_off(maxpan + 1), apply_off(maxpan + 1);
  for (int p = 0; p < maxpan; ++p) {
    pan_off[p] = pan_items.size();
    trail_off[p] = trail_items.size();
    apply_off[p] = apply_items.size();
    for (int i = 0; i < npre; ++i) {
      const int k = qj[i].k;
      if (p * QNB >= k) continue;
      pan_items.push_back(i);
      const int ng = (int)ceil_div(std::max(k - (p + 1) * QNB, 0), 64);
      for (int g = 0; g < ng; ++g) trail_items.push_back(QItem{i, g});
      for (int g = 0; g < (int)ceil_div(k, 64); ++g) apply_items.push_back(QItem{i, g});
    }
  }
  pan_off[maxpan] = pan_items.size();
  trail_off[maxpan] = trail_items.size();
  apply_off[maxpan] = apply_items.size();
  int* dpan = a.take<int>(pan_items.size());
  QItem* dtrail = a.take<QItem>(trail_items.size());
  QItem* dapply = a.take<QItem>(apply_items.size());
  double* dG = a.take<double>((size_t)ngi * BJ2 * BJ2);
  double* dQ = a.take<double>((size_t)nei * BJ2 * BJ2);
  int* dskip = a.take<int>((size_t)nei);
  int* dgt = a.take<int>((size_t)ngi * 3);
  int* det = a.take<int>((size_t)nei * 2);
  int* dat = a.take<int>((size_t)nai * 3);
  BJob* dbj = a.take<BJob>((size_t)njobs);
  OJob* doj = a.take<OJob>((size_t)njobs);
  GJob* dgj = a.take<GJob>((size_t)njobs);
  AJob* daj = a.take<AJob>((size_t)njobs);
  int* dtile = a.take<int>((size_t)njobs);
  int* dids = a.take<int>((size_t)njobs);
  int* ddone = a.take<int>((size_t)njobs);
  int* dsweeps = a.take<int>((size_t)njobs);
  int* dkept = a.take<int>((size_t)njobs);
  int* dactive = a.take<int>((size_t)njobs);
  unsigned long long* doff = a.take<unsigned long long>((size_t)njobs);
  unsigned long long* dwork = a.take<unsigned long long>(3);
  double* dsig = a.take<double>((size_t)sv);
  double* dS = a.take<double>((size_t)sv);
  int* dperm = a.take<int>((size_t)sv);
  int64_t* dsbase = a.take<int64_t>((size_t)njobs);
  int* dgbase = a.take<int>(1);
  TN_REQUIRE(dgbase != nullptr, "%s: workspace too small", who);
  for (int j = 0; j < njobs; ++j) gj[j].C = dWv + xoff[j];
  for (const QJob& q : qj) oj[q.job].qz = dQZ + q.qz;
  std::vector<int> gt(ngi * 3), et(nei * 2), at(nai * 3);
  for (int i = 0; i < ngi; ++i) { gt[i] = g_job[i]; gt[ngi + i] = g_pair[i]; gt[2 * ngi + i] = g_chunk[i]; }
  for (int i = 0; i < nei; ++i) { et[i] = e_job[i]; et[nei + i] = e_pair[i]; }
  for (int i = 0; i < nai; ++i) { at[i] = a_job[i]; at[nai + i] = a_pair[i]; at[2 * nai + i] = a_chunk[i]; }
  TN_HIP(hipMemcpyAsync(dgt, gt.data(), sizeof(int) * gt.size(), hipMemcpyHostToDevice, st));
  TN_HIP(hipMemcpyAsync(det, et.data(), sizeof(int) * et.size(), hipMemcpyHostToDevice, st));
  TN_HIP(hipMemcpyAsync(dat, at.data(), sizeof(int) * at.size(), hipMemcpyHostToDevice, st));
  TN_HIP(hipMemcpyAsync(dbj, bj.data(), sizeof(BJob) * njobs, hipMemcpyHostToDevice, st));
  TN_HIP(hipMemcpyAsync(doj, oj.data(), sizeof(OJob) * njobs, hipMemcpyHostToDevice, st));
  TN_HIP(hipMemcpyAsync(dgj, gj.data(), sizeof(GJob) * njobs, hipMemcpyHostToDevice, st));
  TN_HIP(hipMemcpyAsync(daj, aj.data(), sizeof(AJob) * njobs, hipMemcpyHostToDevice, st));
  TN_HIP(hipMemcpyAsync(dsbase, sbase.data(), sizeof(int64_t) * njobs, hipMemcpyHostToDevice, st));
  if (npre) {
    TN_HIP(hipMemcpyAsync(dqj, qj.data(), sizeof(QJob) * npre, hipMemcpyHostToDevice, st));
    TN_HIP(hipMemcpyAsync(dpan, pan_items.data(), sizeof(int) * pan_items.size(),
                          hipMemcpyHostToDevice, st));
    if (!trail_items.empty())
      TN_HIP(hipMemcpyAsync(dtrail, trail_items.data(), sizeof(QItem) * trail_items.size(),
                            hipMemcpyHostToDevice, st));
    TN_HIP(hipMemcpyAsync(dapply, apply_items.data(), sizeof(QItem) * apply_items.size(),
                          hipMemcpyHostToDevice, st));
  }
  TN_HIP(hipMemsetAsync(ddone, 0, sizeof(int) * njobs, st));
  TN_HIP(hipMemsetAsync(dsweeps, 0, sizeof(int) * njobs, st));
  TN_HIP(hipMemsetAsync(doff, 0, sizeof(unsigned long long) * njobs, st));
  TN_HIP(hipMemsetAsync(dgbase, 0, sizeof(int), st));
  // rows through the Gram and rotation kernels and eigensolves that rotated,
  // counted only while launch timing is on (they give the Jacobi rooflines)
  unsigned long long* cnt = prof_on() ? dwork : nullptr;
  if (cnt) TN_HIP(hipMemsetAsync(cnt, 0, 3 * sizeof(unsigned long long), st));
  int h_active = njobs;
  TN_HIP(hipMemcpyAsync(dactive, &h_active, sizeof(int), hipMemcpyHostToDevice, st));

  // ---- 1. the jobs' matrices into Wv: FP64 gather, or the grouped
  //         contraction with the tile shape per size bucket (autotuned once)
  if (sj) {
    int64_t maxe = 1;
    for (int j = 0; j < njobs; ++j) maxe = std::max<int64_t>(maxe, ms[j] * ns[j]);
    hipLaunchKernelGGL(bgather_f64, dim3((unsigned)std::min<int64_t>(ceil_div(maxe, 256), 4096), njobs),
                       dim3(256), 0, st, dWv, daj, doj);
    TN_CHECK_LAUNCH();
  } else {
    std::vector<int> byb[4];
    for (int j = 0; j < njobs; ++j) byb[bucket_of(gj[j].M, gj[j].N)].push_back(j);
    std::vector<int> use2, use4;
    for (int bkt = 0; bkt < 4; ++bkt) {
      if (byb[bkt].empty()) continue;
      int fr;
      {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(bkt);
        fr = it == g_tune.end() ? 0 : it->second;
      }
      if (fr == 0) {
        // time both shapes on this bucket (results are overwritten by the real launch)
        double t2 = time_grouped(2, byb[bkt], gj, dgj, dtile, dids, st);
        double t4 = time_grouped(4, byb[bkt], gj, dgj, dtile, dids, st);
        fr = t2 <= t4 ? 2 : 4;
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[bkt] = fr;
      }
      (fr == 2 ? use2 : use4).insert((fr == 2 ? use2 : use4).end(), byb[bkt].begin(), byb[bkt].end());
    }
    auto work = [&](const std::vector<int>& ids) {
      double f = 0.0;
      for (int j : ids) f += 2.0 * gj[j].M * (double)gj[j].N * gj[j].K;
      return f;
    };
    auto bytes = [&](const std::vector<int>& ids) {
      double b = 0.0;
      for (int j : ids)
        b += 4.0 * ((double)gj[j].M * gj[j].K + (double)gj[j].K * gj[j].N) +
             8.0 * (double)gj[j].M * gj[j].N;
      return b;
    };
    {
      ProfScope p(PROF_GEMM_F64, st, work(use2), bytes(use2));
      TN_TRY(launch_grouped(2, use2, gj, dgj, dtile, dids, st));
    }
    if (!use4.empty()) {
      // second grouped launch needs its own tables: the first launch's are in use
      TN_HIP(hipStreamSynchronize(st));
      ProfScope p(PROF_GEMM_F64, st, work(use4), bytes(use4));
      TN_TRY(launch_grouped(4, use4, gj, dgj, dtile, dids, st));
    }
  }
  // ---- 1b. QR preconditioning: X = Q R in place, the rows of R as the
  //          Jacobi vectors
  if (npre) {
    // Householder QR work 2 L k^2 - 2 k^3 / 3; each matrix read and written once
    double qf = 0.0, qb = 0.0;
    for (const QJob& q : qj) {
      qf += 2.0 * q.L * (double)q.k * q.k - 2.0 / 3.0 * (double)q.k * q.k * q.k;
      qb += 16.0 * q.L * (double)q.k;
    }
    ProfScope pr(PROF_QR_PANEL, st, qf, qb);
    for (int p = 0; p < maxpan; ++p) {
      const unsigned np_ = (unsigned)(pan_off[p + 1] - pan_off[p]);
      const unsigned nt_ = (unsigned)(trail_off[p + 1] - trail_off[p]);
      hipLaunchKernelGGL(bqr_panel, dim3(np_), dim3(256), 0, st, dWv, dTf, dqj, dpan + pan_off[p], p);
      if (nt_)
        hipLaunchKernelGGL(bqr_trail, dim3(nt_), dim3(256), 0, st, dWv, dTf, dqj,
                           dtrail + trail_off[p], p);
    }
    hipLaunchKernelGGL(bqr_rvec, dim3(64, (unsigned)npre), dim3(256), 0, st, dWv, dqj);
    TN_CHECK_LAUNCH();
  }
  // ---- 2. batched Jacobi
  hipLaunchKernelGGL(binit_identity, dim3((unsigned)ceil_div((int64_t)maxk * maxk, 256), njobs),
                     dim3(256), 0, st, dZ, dbj);
  // truncation jobs: vectors negligible against the keep threshold
  if (tj && rel_threshold > 0.0)
    hipLaunchKernelGGL(bjob_neg2, dim3((unsigned)njobs), dim3(256), 0, st, dWv, dbj, rel_threshold);
  TN_CHECK_LAUNCH();
  // every job within one row chunk: the one-launch step
  bool fused = true;
  for (int j = 0; j < njobs; ++j) fused = fused && bj[j].L <= BCH && bj[j].k <= BCH;
  // A step set of `period` steps is enqueued as one unit (captured once as a
  // HIP graph and replayed; launched kernel by kernel while launch timing is
  // on); each set ends by advancing the device step base.  The host reads
  // the active-job count of a set while the next set runs, so the GPU never
  // waits for the host; a set enqueued after the last job retired exits at
  // once in every block.
  const int period = std::max(8, maxsteps);
  auto enqueue_set = [&]() {
    for (int o = 0; o < period; ++o) {
      // work is counted on the device (cnt) and added once at the end
      if (fused) {
        ProfScope p(PROF_JAC_EIG, st, 0.0, 0.0);
        hipLaunchKernelGGL(bjac_fused, dim3(nei), dim3(256), 0, st, dWv, dZ, dbj, det, det + nei,
                           ddone, dgbase, o, doff, cnt);
      } else {
        {
          ProfScope p(PROF_JAC_GRAM, st, 0.0, 0.0);
          hipLaunchKernelGGL(bjac_gram, dim3(ngi), dim3(256), 0, st, dWv, dbj, dgt, dgt + ngi,
                             dgt + 2 * ngi, ddone, dgbase, o, dG, cnt);
        }
        {
          ProfScope p(PROF_JAC_EIG, st, 0.0, 0.0);
          hipLaunchKernelGGL(bjac_eig, dim3(nei), dim3(256), 0, st, dG, dbj, det, det + nei, ddone,
                             dQ, dskip, doff, cnt);
        }
        {
          ProfScope p(PROF_JAC_APPLY, st, 0.0, 0.0);
          hipLaunchKernelGGL(bjac_apply, dim3(nai), dim3(256), 0, st, dWv, dZ, dbj, dat, dat + nai,
                             dat + 2 * nai, ddone, dgbase, o, dQ, dskip, cnt);
        }
      }
      hipLaunchKernelGGL(bjac_sweep_end, dim3((unsigned)ceil_div(njobs, 256)), dim3(256), 0, st,
                         dbj, njobs, dgbase, o, max_sweeps, ddone, dsweeps, doff, dactive);
    }
    hipLaunchKernelGGL(badvance, dim3(1), dim3(64), 0, st, dgbase, period);
  };
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (!prof_on()) {
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
      enqueue_set();
      hipError_t e = hipStreamEndCapture(st, &graph);
      if (e != hipSuccess || hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        graph = nullptr;
        exec = nullptr;
        (void)hipGetLastError();
      }
    } else {
      (void)hipGetLastError();
    }
  }
  int* h_act = nullptr;  // pinned: the active count of the last two sets
  hipEvent_t ev[2] = {nullptr, nullptr};
  int rc_loop = TN_OK;
  if (hipHostMalloc((void**)&h_act, 2 * sizeof(int), hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) != hipSuccess) {
    set_error(TN_ERR_HIP, "%s: host buffer or event creation failed", who);
    rc_loop = TN_ERR_HIP;
  }
  const int nsets = (int)ceil_div((int64_t)maxsteps * max_sweeps, period);
  for (int p = 0; rc_loop == TN_OK && p < nsets; ++p) {
    if (exec) {
      if (hipGraphLaunch(exec, st) != hipSuccess) {
        set_error(TN_ERR_HIP, "%s: graph launch failed", who);
        rc_loop = TN_ERR_HIP;
        break;
      }
    } else {
      enqueue_set();
    }
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(h_act + (p & 1), dactive, sizeof(int), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipEventRecord(ev[p & 1], st) != hipSuccess) {
      set_error(TN_ERR_HIP, "%s: step set %d failed to enqueue", who, p);
      rc_loop = TN_ERR_HIP;
      break;
    }
    if (p >= 1) {
      if (hipEventSynchronize(ev[(p - 1) & 1]) != hipSuccess) {
        set_error(TN_ERR_HIP, "%s: step set %d failed", who, p - 1);
        rc_loop = TN_ERR_HIP;
        break;
      }
      if (h_act[(p - 1) & 1] <= 0) break;
    }
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (ev[0]) (void)hipEventDestroy(ev[0]);
  if (ev[1]) (void)hipEventDestroy(ev[1]);
  if (h_act) {
    (void)hipStreamSynchronize(st);  // the last copy into h_act has landed
    (void)hipHostFree(h_act);
  }
  if (rc_loop != TN_OK) return rc_loop;
  // ---- 3. sort, rank, outputs
  hipLaunchKernelGGL(bvec_norms, dim3((unsigned)maxk, njobs), dim3(256), 0, st, dWv, dbj, dsig,
                     dsbase);
  hipLaunchKernelGGL(bsort_rank, dim3(njobs), dim3(1024), 0, st, dsig, dbj, dsbase, dperm, dS,
                     rel_threshold, chi_max, dkept);
  if (npre) {
    // the left vectors of the preconditioned jobs: Q Z for the kept columns
    hipLaunchKernelGGL(bqr_qz_init, dim3(256, (unsigned)npre), dim3(256), 0, st, dQZ, dZ, dqj,
                       dperm, dsbase, dkept);
    for (int p = maxpan - 1; p >= 0; --p) {
      const unsigned na_ = (unsigned)(apply_off[p + 1] - apply_off[p]);
      hipLaunchKernelGGL(bqr_apply, dim3(na_), dim3(256), 0, st, dWv, dTf, dqj,
                         dapply + apply_off[p], dkept, p, dQZ);
    }
  }
  hipLaunchKernelGGL(bout, dim3(256, njobs), dim3(256), 0, st, dWv, dZ, dbj, doj, dsbase, dperm, dS,
                     dkept);
  TN_CHECK_LAUNCH();
  std::vector<int> done(njobs);
  if (kept_out)
    TN_HIP(hipMemcpyAsync(kept_out, dkept, sizeof(int) * njobs, hipMemcpyDeviceToHost, st));
  TN_HIP(hipMemcpyAsync(done.data(), ddone, sizeof(int) * njobs, hipMemcpyDeviceToHost, st));
  if (sweeps_out)
    TN_HIP(hipMemcpyAsync(sweeps_out, dsweeps, sizeof(int) * njobs, hipMemcpyDeviceToHost, st));
  unsigned long long hw[3] = {0, 0, 0};
  if (cnt) TN_HIP(hipMemcpyAsync(hw, cnt, sizeof(hw), hipMemcpyDeviceToHost, st));
  TN_HIP(hipStreamSynchronize(st));
  if (cnt) {
    // a Gram row of a 32-column block pair: 32 x 32 multiply-adds, 32 values
    // read; a rotated row: the same product, 32 values read and written; a
    // 32 x 32 eigensolve: 31 rotation rounds of about 9216 flops
    prof_add(PROF_JAC_GRAM, 2.0 * BJ2 * BJ2 * (double)hw[0], 8.0 * BJ2 * (double)hw[0]);
    prof_add(PROF_JAC_EIG, 31.0 * 9216.0 * (double)hw[1], 16.0 * BJ2 * BJ2 * (double)hw[1]);
    prof_add(PROF_JAC_APPLY, 2.0 * BJ2 * BJ2 * (double)hw[2], 16.0 * BJ2 * (double)hw[2]);
  }
  for (int j = 0; j < njobs; ++j)
    if (done[j] != 1) {
      set_error(TN_ERR_NOCONV, "%s: job %d did not converge", who, j);
      return TN_ERR_NOCONV;
    }
  return TN_OK;
}

}  // namespace tn

extern "C" {

size_t tn_batched_truncate_workspace(const tn_trunc_job* jobs, int njobs) {
  std::vector<int64_t> ms, ns;
  for (int j = 0; j < njobs && jobs; ++j) { ms.push_back(jobs[j].m); ns.push_back(jobs[j].n); }
  return tn::batched_ws_bytes(ms, ns);
}

int tn_batched_truncate(const tn_trunc_job* jobs, int njobs, double rel_threshold, int chi_max,
                        int max_sweeps, int* kept_out, int* sweeps_out, void* ws, size_t ws_bytes,
                        void* stream) {
  TN_REQUIRE(jobs != nullptr || njobs == 0, "tn_batched_truncate: bad job list");
  return tn::batched_driver(jobs, nullptr, njobs, rel_threshold, chi_max, max_sweeps, kept_out,
                            sweeps_out, ws, ws_bytes, stream);
}

size_t tn_batched_svd_workspace(const tn_svd_job* jobs, int njobs) {
  std::vector<int64_t> ms, ns;
  for (int j = 0; j < njobs && jobs; ++j) { ms.push_back(jobs[j].m); ns.push_back(jobs[j].n); }
  return tn::batched_ws_bytes(ms, ns);
}

int tn_batched_svd(const tn_svd_job* jobs, int njobs, int max_sweeps, int* sweeps_out, void* ws,
                   size_t ws_bytes, void* stream) {
  TN_REQUIRE(jobs != nullptr || njobs == 0, "tn_batched_svd: bad job list");
  return tn::batched_driver(nullptr, jobs, njobs, 0.0, 0, max_sweeps, nullptr, sweeps_out, ws,
                            ws_bytes, stream);
}

}  // extern "C"
