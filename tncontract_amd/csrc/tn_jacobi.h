// Pair-step building blocks of the one-sided block Jacobi (Hestenes) shared
// by the SVD and PSD eigensolver (tn_svd.hip) and the batched truncation
// (tn_batched.hip): vectors in blocks of 16, pairs of blocks (32 vectors)
// scheduled round-robin (circle method), per pair a 32 x 32 Gram matrix
// diagonalised by one cyclic sweep of two-sided Jacobi in LDS and applied
// to the pair's vectors by MFMA.
#pragma once
#include "tn_common.h"

namespace tn {

constexpr int JB = 16;       // vectors per block
constexpr int J2 = 2 * JB;   // pair width

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

// Workgroup barrier that orders LDS accesses only.  __syncthreads() also
// fences global memory, which on gfx9 means waiting for every outstanding
// global load (one counter for loads and stores); this barrier lets loads
// issued before it stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

struct EigLds {
  double G[2][J2][J2 + 4];   // row padding: fewer LDS bank conflicts in the rounds (measured)
  double Q[2][J2][J2 + 1];
  double red[4];
};

// Largest normalised off-diagonal over the pair, encoded for atomicMax.
__device__ __forceinline__ unsigned long long dbits(double x) {
  return (unsigned long long)__double_as_longlong(x);
}

// The pair's Gram entries G_pq, G_pp, G_qq scaled by the same power of two
// 2^-e (exactly), e the exponent of max(G_pp, G_qq): the rotation test and
// the convergence measure then square numbers of order one, never the tiny
// entries of nearly-null vectors, whose squares underflow FP64.
__device__ __forceinline__ void pair_scaled(double gpq, double gpp, double gqq, double& spq,
                                            double& spp, double& sqq) {
  const int e = __builtin_amdgcn_frexp_exp(fmax(gpp, gqq));
  spq = __builtin_amdgcn_ldexp(gpq, -e);
  spp = __builtin_amdgcn_ldexp(gpp, -e);
  sqq = __builtin_amdgcn_ldexp(gqq, -e);
}

// Rotation predicate of a pair (l, h) from G_lh, G_ll, G_hh: G_lh^2 > tol^2
// G_ll G_hh on the scaled entries, and one of the two vectors not negligible
// (squared norm > neg2).  rot_coef_v rotates exactly the pairs for which it
// holds, and the convergence measure eig_off counts exactly those.
__device__ __forceinline__ bool pair_rotates(double gpq, double gpp, double gqq, double tol,
                                             double neg2, double& spq, double& spp, double& sqq) {
  pair_scaled(gpq, gpp, gqq, spq, spp, sqq);
  return spq * spq > tol * tol * fmax(spp * sqq, 0.0) && fmax(gpp, gqq) > neg2;
}

// Convergence measure of the pair block: the largest |G_pq| / sqrt(G_pp G_qq)
// (an FP32 estimate from the scaled entries, at least just above tol) over
// the pairs that pair_rotates would rotate, 0 if there are none; so
// "off <= tol" holds exactly when a sweep would rotate nothing.  Pairs of two
// negligible vectors (squared norms <= neg2) are left out: the caller has
// shown that the space they span lies below its truncation threshold
// whether or not they are orthogonalised (neg2 = 0: every pair).
// Block-uniform.
__device__ inline double eig_off(EigLds& sm, double tol, double neg2 = 0.0) {
  const int tid = threadIdx.x;
  const float above = (float)tol * (1.0f + 0x1p-20f) + 0x1p-126f;
  float off = 0.0f;
  for (int e = tid; e < J2 * J2; e += 256) {
    const int p = e / J2, q = e % J2;
    if (p < q) {
      double spq, spp, sqq;
      const bool rot = pair_rotates(sm.G[0][p][q], sm.G[0][p][p], sm.G[0][q][q], tol, neg2, spq,
                                    spp, sqq);
      // (float)(spp sqq) may flush to 0: rsq then gives inf (or NaN for a
      // zero spq), which fminf clamps
      const float v = fminf(fabsf((float)spq) * __builtin_amdgcn_rsqf((float)(spp * sqq)), 1e30f);
      if (rot) off = fmaxf(off, fmaxf(v, above));
    }
  }
  off = wave_max_dpp(off);
  if ((tid & 63) == 0) sm.red[tid >> 6] = off;
  lds_barrier();
  off = fmaxf(fmaxf(sm.red[0], sm.red[1]), fmaxf(sm.red[2], sm.red[3]));
  lds_barrier();
  return (double)off;
}

// 1/sqrt(x) for FP64 from the hardware estimate (v_rsq_f64) refined by two
// Newton steps to full double precision: a far shorter dependent chain than
// an IEEE square root and division.
__device__ __forceinline__ double rsq_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  double e = fma(-hx * y, y, 0.5);
  y = fma(y, e, y);
  e = fma(-hx * y, y, 0.5);
  return fma(y, e, y);
}

// Rotation coefficients for index x against partner px: G' = J^T G J with
// J_ll = J_hh = c, J_lh = s, J_hl = -s (l < h); row x of J^T G is
// a*G[x] + b*G[px] with a = c and b = -s (x = l) or +s (x = h).
// t = tan(angle) = sgn(D P) |P| / (|D| + sqrt(D^2 + P^2)), D = gqq - gpp,
// P = 2 gpq (the sign from the operands' signs: their product may
// underflow).  The angle only has to shrink the off-diagonal, so
// 1 / (|D| + sqrt(D^2 + P^2)) is evaluated in FP32 on D, P scaled by a power
// of two (exact) into [0.5, 1), where it lies in (0.5, 2]: a 2^-24 error in
// t leaves a 2^-24 fraction of G_lh, which the next sweeps remove like any
// other (same sweep counts as an FP64 angle, measured).  t is that
// reciprocal times the scaled |P| in FP64, so an angle far below FP32's
// range (a nearly-null vector against a large one) is still |P| / (2 |D|)
// and rotates.  Orthogonality comes from c = 1/sqrt(1 + t^2) in FP64
// (v_rsq_f64 + two Newton steps) and s = t c, so c^2 + s^2 = 1 to FP64
// rounding.  A pair is rotated when pair_rotates holds; the coefficients are
// computed without a branch
// (a pair that is not rotated selects c = 1, s = 0, discarding whatever the
// formula gave).
// The coefficients from the three Gram entries of the pair (l, h): G_lh,
// G_ll, G_hh; x_is_l: whether the index asked for is l.
__device__ __forceinline__ void rot_coef_v(double gpq, double gpp, double gqq, bool x_is_l,
                                           double tol, double neg2, double& a, double& b) {
  double spq, spp, sqq;
  const bool rot = pair_rotates(gpq, gpp, gqq, tol, neg2, spq, spp, sqq);
  const double D = gqq - gpp, P = 2.0 * gpq;
  const int e = __builtin_amdgcn_frexp_exp(fmax(fabs(D), fabs(P)));
  const double dd = __builtin_amdgcn_ldexp(fabs(D), -e), pp = __builtin_amdgcn_ldexp(fabs(P), -e);
  const float d = (float)dd, p = (float)pp;
  const float hyp = __builtin_amdgcn_sqrtf(fmaf(d, d, p * p));
  const double tabs = pp * (double)__builtin_amdgcn_rcpf(d + hyp);
  const double t = ((D < 0.0) == (P < 0.0)) ? tabs : -tabs;
  const double c = rsq_nr(fma(t, t, 1.0));
  const double sn = t * c;
  a = rot ? c : 1.0;
  b = rot ? (x_is_l ? -sn : sn) : 0.0;
}

// v from the lane whose byte address (4 * lane) is addr, within the wave.
__device__ __forceinline__ double lane_fetch(int addr, double v) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)(unsigned)u);
  const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(u >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// v from lane + 1 (row_shl:1) / lane - 1 (row_shr:1) within the lane's row
// of 16 (DPP; the lane without a source gets 0)
__device__ __forceinline__ double dpp_from_next(double v) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, 0x101, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), 0x101, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double dpp_from_prev(double v) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, 0x111, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), 0x111, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// One cyclic sweep (31 rounds of 16 disjoint rotations, circle order) of the
// two-sided Jacobi on G, accumulating Q = J_1 ... J_31, one barrier per
// round.  Thread t (row pair rp = t/16, column pair cp = t%16) owns the 2x2
// block of G' = J^T G J at rows (x0, x1) and columns (y0, y1), the circle
// slots of rp and cp (G ping-pongs between two LDS buffers), and rows 2 rp,
// 2 rp + 1 of Q at the columns of its slots, in registers: Q' = Q J only
// mixes columns, and the circle moves a column between neighbouring column
// slots (first slots from lane cp + 1 to cp, second slots from cp - 1 to cp,
// slot 1-first to 0-second, 15-second to 15-first, vector 0 fixed), which
// stay within the lane's DPP row of 16 lanes.  Each lane computes the
// rotation of its column pair from the LDS values; the rotation of its row
// pair is the one computed by the lane of the same 16-lane row whose column
// pair it is (fetched across the wave, so row and column rotations of a pair
// are the same bits).  One inner sweep per outer step: the outer block
// iteration needs the same number of sweeps with one inner sweep as with
// full inner convergence (measured), at a fraction of the cost.  Q is left
// in sm.Q[0]; returns that buffer's index (0).
__device__ inline int eig_rotate(EigLds& sm, double tol, double neg2 = 0.0) {
  auto& G = sm.G;
  const int tid = threadIdx.x;
  const int rp = tid >> 4, cp = tid & 15;
  const int src = ((tid & 0x30) | rp) << 2;   // lane of this row with column pair rp
  int cur = 0;
  // circle positions, advanced by one (mod 31, in 1..31) each round; after
  // 31 rounds every slot holds its first vector again
  int x0 = rp == 0 ? 0 : rp, x1 = J2 - 1 - rp, y0 = cp == 0 ? 0 : cp, y1 = J2 - 1 - cp;
  const int R0 = 2 * rp, R1 = 2 * rp + 1;
  double qa0 = R0 == y0 ? 1.0 : 0.0, qa1 = R0 == y1 ? 1.0 : 0.0;   // Q[R0][y0], Q[R0][y1]
  double qb0 = R1 == y0 ? 1.0 : 0.0, qb1 = R1 == y1 ? 1.0 : 0.0;   // Q[R1][y0], Q[R1][y1]
  for (int r = 0; r < J2 - 1; ++r) {
    const int nxt = cur ^ 1;
    const int yl = min(y0, y1), yh = max(y0, y1);
    const double ypq = G[cur][yl][yh], ypp = G[cur][yl][yl], yqq = G[cur][yh][yh];
    const double g00 = G[cur][x0][y0], g01 = G[cur][x0][y1];
    const double g10 = G[cur][x1][y0], g11 = G[cur][x1][y1];
    // (c, b) for the first index of each pair; the partner's b is -b
    double cy, sy;
    rot_coef_v(ypq, ypp, yqq, y0 == yl, tol, neg2, cy, sy);
    const double cx = lane_fetch(src, cy), sx = lane_fetch(src, sy);
    // Q' = Q J on the column pair, then the circle shift of the columns
    const double na0 = cy * qa0 + sy * qa1, na1 = cy * qa1 - sy * qa0;
    const double nb0 = cy * qb0 + sy * qb1, nb1 = cy * qb1 - sy * qb0;
    const double fa = dpp_from_next(na0), fb = dpp_from_next(nb0);
    const double pa = dpp_from_prev(na1), pb = dpp_from_prev(nb1);
    qa0 = cp == 0 ? na0 : (cp == 15 ? na1 : fa);
    qb0 = cp == 0 ? nb0 : (cp == 15 ? nb1 : fb);
    qa1 = cp == 0 ? fa : pa;
    qb1 = cp == 0 ? fb : pb;
    const double bX0 = sx, bX1 = -sx, bY0 = sy, bY1 = -sy;
    const double r00 = cx * g00 + bX0 * g10, r01 = cx * g01 + bX0 * g11;
    const double r10 = cx * g10 + bX1 * g00, r11 = cx * g11 + bX1 * g01;
    G[nxt][x0][y0] = cy * r00 + bY0 * r01;
    G[nxt][x0][y1] = cy * r01 + bY1 * r00;
    G[nxt][x1][y0] = cy * r10 + bY0 * r11;
    G[nxt][x1][y1] = cy * r11 + bY1 * r10;
    cur = nxt;
    if (x0 != 0) x0 = x0 == J2 - 1 ? 1 : x0 + 1;
    x1 = x1 == J2 - 1 ? 1 : x1 + 1;
    if (y0 != 0) y0 = y0 == J2 - 1 ? 1 : y0 + 1;
    y1 = y1 == J2 - 1 ? 1 : y1 + 1;
    lds_barrier();
  }
  sm.Q[0][R0][y0] = qa0;
  sm.Q[0][R0][y1] = qa1;
  sm.Q[0][R1][y0] = qb0;
  sm.Q[0][R1][y1] = qb1;
  lds_barrier();
  return 0;
}

// Rows [rs, re) (re - rs <= 64 NT) of the pair's 32 vectors (vector v at
// base + v * ld, nv vectors) <- rows . Q, as the transposed product
// (Q^T X^T) on v_mfma_f64_16x16x4: the B operand of each 16 x 16 output tile
// is 16 consecutive rows of 4 vectors, read straight from memory (four
// 128-byte runs per load), the A operand is Q^T from LDS, and the output
// tile is 16 consecutive rows of 4 vectors again, stored straight back: no
// LDS staging and no barrier.  Wave w covers rows rs + 16 NT w + [0, 16 NT)
// as NT row tiles; lane (lr = lane % 16, lk = lane / 16).
//   apply_rows_fetch:  x[ks][nt] = row rs + 16 NT w + 16 nt + lr of vector
//                      4 ks + lk of the pair (issued early: in flight while
//                      the caller computes Q);
//   apply_rows_finish: 16 NT MFMAs per wave (2 vector tiles x NT row tiles x
//                      8 k-steps) and the stores.
// Whether the chunk is a whole 64 NT rows and all 32 vectors of the pair
// exist (block-uniform): then the loads and stores need no bounds test.
template <int NT>
__device__ __forceinline__ bool apply_rows_full(int64_t nv, int64_t rs, int64_t re, int I, int J) {
  return re - rs == 64 * NT && (int64_t)max(I, J) * JB + JB <= nv;
}

template <int NT>
__device__ __forceinline__ void apply_rows_fetch(const double* __restrict__ base, int64_t ld,
                                                 int64_t nv, int64_t rs, int64_t re, int I,
                                                 int J, double (&x)[8][NT]) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t r0 = rs + 16 * NT * wave + lr;
  if (apply_rows_full<NT>(nv, rs, re, I, J)) {
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const double* p = base + pair_vec(4 * ks + lk, I, J) * ld + r0;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) x[ks][nt] = p[16 * nt];
    }
    return;
  }
#pragma unroll
  for (int ks = 0; ks < 8; ++ks) {
    const int64_t v = pair_vec(4 * ks + lk, I, J);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int64_t r = r0 + 16 * nt;
      x[ks][nt] = (v < nv && r < re) ? base[v * ld + r] : 0.0;
    }
  }
}

template <int NT>
__device__ __forceinline__ void apply_rows_finish(double* __restrict__ base, int64_t ld,
                                                  int64_t nv, int64_t rs, int64_t re, int I,
                                                  int J, const double (&x)[8][NT],
                                                  const double (*Q)[J2 + 1]) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t r0 = rs + 16 * NT * wave + lr;
  f64x4 acc[2][NT];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f64x4{0, 0, 0, 0};
#pragma unroll
  for (int ks = 0; ks < 8; ++ks) {
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      // A[i][k] = Q^T[16 mt + i][4 ks + k] = Q[4 ks + k][16 mt + i]
      const double a = Q[4 * ks + lk][16 * mt + lr];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        acc[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, x[ks][nt], acc[mt][nt], 0, 0, 0);
    }
  }
  // output tile (mt, nt): lane holds vector 16 mt + lk + 4 e, row 16 nt + lr
  if (apply_rows_full<NT>(nv, rs, re, I, J)) {
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        double* p = base + pair_vec(16 * mt + lk + 4 * e, I, J) * ld + r0;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) p[16 * nt] = acc[mt][nt][e];
      }
    return;
  }
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t v = pair_vec(16 * mt + lk + 4 * e, I, J);
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const int64_t r = r0 + 16 * nt;
        if (v < nv && r < re) base[v * ld + r] = acc[mt][nt][e];
      }
    }
}

}  // namespace tn
