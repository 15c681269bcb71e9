// In-workgroup Cholesky factorisation and inverse of one 64x64 FP64 tile.
//
// The tile is 4 x 4 blocks of 16; A (-> L) and E (-> L^-1) live in LDS.  For
// each block column j:
//   phase 1: wave 0 applies the update of block column j-1 to the diagonal
//            block, A_jj -= L_j,j-1 L_j,j-1^T (the one trailing update on the
//            critical path), then factors and inverts A_jj in registers
//            (chol16_inv16 below); meanwhile waves 1..3 apply the remaining
//            trailing updates of block column j-1, A_ik -= L_i,j-1 L_k,j-1^T
//            (j <= k <= i, (i, k) != (j, j)), and form
//            T_jk = sum_{m=k}^{j-1} L_jm E_mk for the inverse row j (k < j);
//   phase 2: the blocks below the diagonal, L_ij = A_ij E_jj^T (i > j; wave 0
//            takes i = j+1, whose update it applies at the start of the next
//            phase 1), and the inverse blocks E_jk = -E_jj T_jk (k < j):
//            always three 16x16 products.
// Two barriers per block column; every off-diagonal product is a 16x16x16
// FP64 MFMA product (4 v_mfma_f64_16x16x4f64) on LDS operands.  For a full
// tile the strict upper blocks of A and E are cleared by waves 1..3 during
// the first factorization (nothing else touches them); a partial tile (valid
// block mv < 64) is set to the identity outside the valid block first.
// Measured on MI355X, cycles of one tile (s_memtime): 8-wide panels 48 k;
// this blocking with v_readlane broadcasts in the 16x16 factorization and
// three barriers per block column 32 k; as described here 24 k.
#pragma once
#include "bta_device.h"

// 1/sqrt(x) from the hardware estimate plus two Newton steps (full FP64
// accuracy for normal positive x), and 1/x likewise.  A non-positive pivot
// propagates as a negative / NaN value and is reported by the caller.
__device__ __forceinline__ double rsq_f64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  return y * fma(-h * y, y, 1.5);
}
__device__ __forceinline__ double rcp2_f64(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(y, fma(-x, y, 1.0), y);
  return fma(y, fma(-x, y, 1.0), y);
}

struct TileCholResult {
  double logsum;  // sum of log(pivot) over the valid part (thread 0 only)
  int fail;       // first non-positive / NaN pivot in the valid part, or -1 (thread 0 only)
};

__device__ __forceinline__ f64x4 load16(const double* __restrict__ C, int r0, int c0) {
  const int lane = threadIdx.x & 63;
  f64x4 v;
#pragma unroll
  for (int r = 0; r < 4; ++r) v[r] = C[(r0 + 4 * r + (lane >> 4)) * BTA_LDS_LD + c0 + (lane & 15)];
  return v;
}

__device__ __forceinline__ void store16(double* __restrict__ C, int r0, int c0, f64x4 v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int r = 0; r < 4; ++r) C[(r0 + 4 * r + (lane >> 4)) * BTA_LDS_LD + c0 + (lane & 15)] = v[r];
}

// Store the lower triangle (i >= j) of a 16x16 tile (all of it off the diagonal).
__device__ __forceinline__ void store16_lower(double* __restrict__ C, int r0, int c0, f64x4 v) {
  const int lane = threadIdx.x & 63;
  const int j = c0 + (lane & 15);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = r0 + 4 * r + (lane >> 4);
    if (j <= i) C[i * BTA_LDS_LD + j] = v[r];
  }
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// acc +/-= X(rx.., kx..) * op(Y): NT form (ty == 0: Y rows ry.., K cols ky..),
// or NN form (ty == 1: Y rows ky.. (K), cols ry..), K = 16, LDS stride BTA_LDS_LD.
template <bool NEG, int TY>
__device__ __forceinline__ f64x4 mma16x16(const double* __restrict__ X, int rx, int kx, const double* __restrict__ Y,
                                          int ry, int ky, f64x4 acc) {
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  double a[4], b[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    a[k] = X[(rx + li) * BTA_LDS_LD + kx + 4 * k + lk];
    b[k] = TY == 0 ? Y[(ry + li) * BTA_LDS_LD + ky + 4 * k + lk] : Y[(ky + 4 * k + lk) * BTA_LDS_LD + ry + li];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) acc = mfma_f64(NEG ? -a[k] : a[k], b[k], acc);
  return acc;
}

// v moved from lane P of its 16-lane row to every lane of that row (DPP
// row_newbcast, a single 64-bit move on gfx950).
template <int P>
__device__ __forceinline__ double bcast_row16(double v) {
  return __longlong_as_double(__builtin_amdgcn_mov_dpp(__double_as_longlong(v), 0x150 + P, 0xf, 0xf, true));
}

// Pivot P of chol16_inv16 and, recursively, the pivots after it.  Lane c (of
// every 16-lane row) holds column c of the symmetric trailing matrix in a[]
// and column c of E in e[].  With q = 1 / A_PP, row i > P of both loses
// A_iP q times row P: A_ic -= A_iP A_Pc q (columns c > P only; the finished
// columns c <= P keep the unscaled column of L) and E_ic -= A_iP E_Pc q.
// A_iP is lane P's element of a[i], taken with a row broadcast, so the update
// of one row is one broadcast and two FMAs.  E ends as inv(U) of A = U D U^T
// (U unit lower); the caller scales it to inv(L) = D^-1/2 inv(U).
template <int P>
__device__ __forceinline__ void chol16_pivot(double (&a)[16], double (&e)[16], int c, double& myd) {
  const double d = readlane_f64(a[P], P);
  if (c == P) myd = d;
  const double nq = -rcp2_f64(d);
  const double m = (c > P) ? a[P] * nq : 0.0;
  const double me = e[P] * nq;
#pragma unroll
  for (int i = P + 1; i < 16; ++i) {
    const double b = bcast_row16<P>(a[i]);
    e[i] = fma(b, me, e[i]);
    a[i] = fma(b, m, a[i]);
  }
  // pin every register's update to its pivot (otherwise the compiler sinks the
  // E updates past later pivots and keeps every multiplier alive)
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    asm volatile("" : "+v"(a[i]));
    asm volatile("" : "+v"(e[i]));
  }
  if constexpr (P < 15) chol16_pivot<P + 1>(a, e, c, myd);
}

// 16x16 Cholesky and inverse of the block at (r0, r0) of s (lower triangle
// read, symmetric), by one wave; writes L (zero strict upper) to s and L^-1 to
// w at the same block.  Lane c records its own pivot d_c; at the end
// L_ic = A_ic / sqrt(d_c) (i > c), L_cc = sqrt(d_c), and row i of E is scaled
// by 1 / sqrt(d_i).
__device__ __forceinline__ void chol16_inv16(double* __restrict__ s, double* __restrict__ w, int r0) {
  const int lane = threadIdx.x & 63;
  const int c = lane & 15;
  double a[16], e[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    a[i] = (i >= c) ? s[(r0 + i) * BTA_LDS_LD + r0 + c] : s[(r0 + c) * BTA_LDS_LD + r0 + i];
    e[i] = (i == c) ? 1.0 : 0.0;
  }
  double myd = 1.0;
  chol16_pivot<0>(a, e, c, myd);
  const double myrs = rsq_f64(myd);
  if (lane < 16) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      s[(r0 + i) * BTA_LDS_LD + r0 + c] = (i > c) ? a[i] * myrs : (i == c ? myd * myrs : 0.0);
      w[(r0 + i) * BTA_LDS_LD + r0 + c] = e[i] * readlane_f64(myrs, i);
    }
  }
}

// s: 64x64 LDS tile (stride BTA_LDS_LD) holding the SPD matrix in its lower
// triangle (upper triangle ignored); rows / cols >= mv are treated as
// identity.  On exit s holds L (zero strict upper) and w holds L^-1 (zero
// strict upper), both with stride BTA_LDS_LD.
__device__ __forceinline__ TileCholResult tile_chol_inv(double* __restrict__ s, double* __restrict__ w, int mv) {
  const int tid = threadIdx.x;
  const int wv = tid >> 6, lane = tid & 63;
  if (mv < 64) {  // identity outside the valid block, zero strict upper triangle, E = 0
    for (int e = tid; e < 64 * 64; e += BTA_THREADS) {
      const int i = e >> 6, j = e & 63;
      if (i >= mv || j >= mv || j > i) s[i * BTA_LDS_LD + j] = (i == j) ? 1.0 : 0.0;
      w[i * BTA_LDS_LD + j] = 0.0;
    }
    __syncthreads();
  }
#pragma unroll 1
  for (int jb = 0; jb < 4; ++jb) {
    const int r0 = 16 * jb;
    // ---- phase 1
    if (wv == 0) {
      if (jb > 0) {
        f64x4 acc = load16(s, r0, r0);
        acc = mma16x16<true, 0>(s, r0, r0 - 16, s, r0, r0 - 16, acc);
        store16_lower(s, r0, r0, acc);
      }
      chol16_inv16(s, w, r0);
    } else if (jb == 0) {
      if (mv == 64) {  // strict upper blocks (0,1) (0,2) (0,3) (1,2) (1,3) (2,3) of s and w
        for (int e = tid - 64; e < 6 * 256; e += BTA_THREADS - 64) {
          const int blk = e >> 8, q = e & 255;
          const int bi = blk < 3 ? 0 : (blk < 5 ? 1 : 2);
          const int bk = blk < 3 ? blk + 1 : (blk < 5 ? blk - 1 : 3);
          const int idx = (16 * bi + (q >> 4)) * BTA_LDS_LD + 16 * bk + (q & 15);
          s[idx] = 0.0;
          w[idx] = 0.0;
        }
      }
    } else {
      // trailing tiles of block column jb-1 other than (jb, jb), then the T_jk
      const int nr = 4 - jb, nl = nr * (nr + 1) / 2 - 1;
      for (int t = wv - 1; t < nl + jb; t += 3) {
        if (t < nl) {
          int rr = 0, cc = t + 1;  // lower-triangle tile t+1 of the trailing block, row-major
          while (cc > rr) {
            cc -= rr + 1;
            ++rr;
          }
          const int ib = jb + rr, kb = jb + cc;
          f64x4 acc = load16(s, 16 * ib, 16 * kb);
          acc = mma16x16<true, 0>(s, 16 * ib, r0 - 16, s, 16 * kb, r0 - 16, acc);
          if (ib == kb) store16_lower(s, 16 * ib, 16 * kb, acc);
          else store16(s, 16 * ib, 16 * kb, acc);
        } else {
          const int k = t - nl;
          f64x4 acc = f64x4{0.0, 0.0, 0.0, 0.0};
          for (int m = k; m < jb; ++m) acc = mma16x16<false, 1>(s, r0, 16 * m, w, 16 * k, 16 * m, acc);
          store16(w, r0, 16 * k, acc);
        }
      }
    }
    __syncthreads();
    // ---- phase 2: product q < 3 - jb is L_(jb+1+q),jb, the others E_jb,(q - (3 - jb))
    {
      const int q = (jb < 3) ? (wv <= 2 ? wv : -1) : wv - 1;
      if (q >= 0) {
        f64x4 acc = f64x4{0.0, 0.0, 0.0, 0.0};
        if (q < 3 - jb) {
          const int ib = jb + 1 + q;
          acc = mma16x16<false, 0>(s, 16 * ib, r0, w, r0, r0, acc);  // A_ib,jb E_jj^T
          store16(s, 16 * ib, r0, acc);
        } else {
          const int k = q - (3 - jb);
          acc = mma16x16<true, 1>(w, r0, r0, w, 16 * k, r0, acc);    // -E_jj T_jk
          store16(w, r0, 16 * k, acc);
        }
      }
    }
    __syncthreads();
  }
  // sum of log pivots (log d_i = 2 log l_ii): fixed butterfly over wave 0
  // (deterministic order); first non-positive pivot from the diagonal of L
  TileCholResult res;
  res.fail = -1;
  res.logsum = 0.0;
  if (wv == 0) {
    const double lii = s[lane * BTA_LDS_LD + lane];
    const unsigned long long bad = __ballot(lane < mv && !(lii > 0.0));
    res.fail = bad ? __ffsll((long long)bad) - 1 : -1;
    double v = (lane < mv) ? 2.0 * log(lii) : 0.0;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    res.logsum = v;
  }
  __syncthreads();
  return res;
}
