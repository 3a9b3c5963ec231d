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

struct EigLds {
  double G[2][J2][J2 + 1];
  double Q[2][J2][J2 + 1];
  double red[4];
};

// Largest normalised off-diagonal over the pair, encoded for atomicMax.
__device__ __forceinline__ unsigned long long dbits(double x) {
  return (unsigned long long)__double_as_longlong(x);
}

// max_{p<q} |G_pq| / sqrt(G_pp G_qq) over the 32 vectors (block-uniform).
__device__ inline double eig_off(EigLds& sm) {
  const int tid = threadIdx.x;
  double off = 0.0;
  for (int e = tid; e < J2 * J2; e += 256) {
    int p = e / J2, q = e % J2;
    if (p < q) {
      double d = sm.G[0][p][p] * sm.G[0][q][q];
      if (d > 0.0) off = fmax(off, fabs(sm.G[0][p][q]) / sqrt(d));
    }
  }
  off = wave_max(off);
  if ((tid & 63) == 0) sm.red[tid >> 6] = off;
  __syncthreads();
  off = fmax(fmax(sm.red[0], sm.red[1]), fmax(sm.red[2], sm.red[3]));
  __syncthreads();
  return off;
}

// 1/x and 1/sqrt(x) for FP64 from the hardware estimates (v_rcp_f64,
// v_rsq_f64) refined by two Newton steps to full double precision: far
// shorter dependent chains than IEEE division and square root.
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}
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
// P = 2 gpq, evaluated on D, P scaled by 1/max(|D|, |P|); c = 1/sqrt(1 + t^2),
// s = t c.  A pair is rotated when gpq^2 > tol^2 gpp gqq (then |gpq| is far
// above the underflow range, so the scaling is finite).
__device__ __forceinline__ void rot_coef(const double (*G)[J2 + 1], int x, int px, double tol,
                                         double& a, double& b) {
  const int l = min(x, px), h = max(x, px);
  const double gpq = G[l][h], gpp = G[l][l], gqq = G[h][h];
  double c = 1.0, sn = 0.0;
  if (gpq * gpq > tol * tol * fmax(gpp * gqq, 0.0)) {
    const double D = gqq - gpp, P = 2.0 * gpq;
    const double aD = fabs(D), aP = fabs(P);
    const double rm = rcp_nr(fmax(aD, aP));
    const double d = aD * rm, p = aP * rm;
    const double q = fma(d, d, p * p);
    const double hyp = q * rsq_nr(q);
    const double t = ((D * P >= 0.0) ? p : -p) * rcp_nr(d + hyp);
    c = rsq_nr(fma(t, t, 1.0));
    sn = t * c;
  }
  a = c;
  b = x == l ? -sn : sn;
}

// One cyclic sweep (31 rounds of 16 disjoint rotations, circle order) of the
// two-sided Jacobi on G, accumulated into Q, one barrier per round: thread t
// owns the 2x2 block (row pair t/16, column pair t%16) of G' = J^T G J and
// Q' = Q J and computes the two rotations it needs itself (the 16 threads
// that need a rotation compute it from the same LDS values, so they agree
// bit for bit).  One inner sweep per outer step: the outer block iteration
// needs the same number of sweeps with one inner sweep as with full inner
// convergence (measured), at a fraction of the cost.  Returns the buffer
// holding Q.
__device__ inline int eig_rotate(EigLds& sm, double tol) {
  auto& G = sm.G;
  auto& Q = sm.Q;
  const int tid = threadIdx.x;
  const int rp = tid >> 4, cp = tid & 15;
  int cur = 0;
  // circle positions, advanced by one (mod 31, in 1..31) each round
  int x0 = rp == 0 ? 0 : rp, x1 = J2 - 1 - rp, y0 = cp == 0 ? 0 : cp, y1 = J2 - 1 - cp;
  for (int r = 0; r < J2 - 1; ++r) {
    const int nxt = cur ^ 1;
    // (c, b) for the first index of each pair; the partner's b is -b
    double cx, sx, cy, sy;
    rot_coef(G[cur], x0, x1, tol, cx, sx);
    rot_coef(G[cur], y0, y1, tol, cy, sy);
    const double bX0 = sx, bX1 = -sx, bY0 = sy, bY1 = -sy;
    const double g00 = G[cur][x0][y0], g01 = G[cur][x0][y1];
    const double g10 = G[cur][x1][y0], g11 = G[cur][x1][y1];
    const double r00 = cx * g00 + bX0 * g10, r01 = cx * g01 + bX0 * g11;
    const double r10 = cx * g10 + bX1 * g00, r11 = cx * g11 + bX1 * g01;
    G[nxt][x0][y0] = cy * r00 + bY0 * r01;
    G[nxt][x0][y1] = cy * r01 + bY1 * r00;
    G[nxt][x1][y0] = cy * r10 + bY0 * r11;
    G[nxt][x1][y1] = cy * r11 + bY1 * r10;
    const double q00 = Q[cur][x0][y0], q01 = Q[cur][x0][y1];
    const double q10 = Q[cur][x1][y0], q11 = Q[cur][x1][y1];
    Q[nxt][x0][y0] = cy * q00 + bY0 * q01;
    Q[nxt][x0][y1] = cy * q01 + bY1 * q00;
    Q[nxt][x1][y0] = cy * q10 + bY0 * q11;
    Q[nxt][x1][y1] = cy * q11 + bY1 * q10;
    cur = nxt;
    if (x0 != 0) x0 = x0 == J2 - 1 ? 1 : x0 + 1;
    x1 = x1 == J2 - 1 ? 1 : x1 + 1;
    if (y0 != 0) y0 = y0 == J2 - 1 ? 1 : y0 + 1;
    y1 = y1 == J2 - 1 ? 1 : y1 + 1;
    __syncthreads();
  }
  return cur;
}

// Rows [rs, re) of the pair's 32 vectors (vector v at base + v * ld, nv
// vectors) <- rows . Q, Q 32 x 32 in LDS (row stride J2 + 1).  Each wave
// takes 64 rows per pass: it loads its 64 x 32 tile into registers (the A
// operands of v_mfma_f64_16x16x4), multiplies by Q (B operands from LDS) and
// stores the tile back; every store depends on all loads of its rows, so the
// update is in place.  256 threads.
__device__ inline void apply_mfma(double* __restrict__ base, int64_t ld, int64_t nv, int64_t rs,
                           int64_t re, int I, int J, const double (*Q)[J2 + 1]) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lr = lane & 15, lk = lane >> 4;
  for (int64_t r0 = rs + 64 * wave; r0 < re; r0 += 256) {
    double a[4][8];
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const int64_t v = pair_vec(4 * ks + lk, I, J);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int64_t r = r0 + 16 * t + lr;
        a[t][ks] = (v < nv && r < re) ? base[v * ld + r] : 0.0;
      }
    }
    f64x4 acc[4][2];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t][0] = acc[t][1] = f64x4{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) {
        const double b = Q[4 * ks + lk][16 * ct + lr];
#pragma unroll
        for (int t = 0; t < 4; ++t)
          acc[t][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t][ks], b, acc[t][ct], 0, 0, 0);
      }
    }
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
      const int64_t v = pair_vec(16 * ct + lr, I, J);
      if (v >= nv) continue;
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int64_t r = r0 + 16 * t + lk + 4 * e;
          if (r < re) base[v * ld + r] = acc[t][ct][e];
        }
    }
  }
}

}  // namespace tn
