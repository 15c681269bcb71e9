// In-workgroup Cholesky factorisation and inverse of one 64x64 FP64 tile.
//
// Blocked over four 16-column blocks.  For block c:
//   (a) wave 0 factors the 16x16 diagonal block and inverts it: lane i holds
//       row i of the block and row i of E = Lu^-1 (right-looking LDL^T
//       elimination applied to an identity), columns are broadcast with
//       v_readlane, so the 16 elimination steps need no barrier;
//       L = Lu D^1/2, L^-1 = D^-1/2 E;
//   (b) the rows below are solved, X = A inv(L_cc)^T, as 16x16x16 FP64 MFMA
//       products (one 16-row tile per wave);
//   (c) the trailing lower tiles are updated, A -= X X^T (MFMA).
// The inverse of the 64x64 factor is then completed block row by block row:
//   W_rc = -W_rr (sum_{c<=k<r} L_rk W_kc).
// About 20 workgroup barriers in total.
#pragma once
#include "bta_device.h"

#ifndef BTA_BASE_LDS_BROADCAST
#define BTA_BASE_LDS_BROADCAST 0
#endif

struct TileCholResult {
  double logsum;  // sum of log(pivot) over the valid part (thread 0 only)
  int fail;       // first non-positive / NaN pivot in the valid part, or -1 (thread 0 only)
};

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// 16x16 output tile at (r0, c0) of LDS matrix C (stride BTA_LDS_LD):
//   C += alpha * sum_{k in [0,K)} A[ra + i][ka + k] * B[kb + k][cb + j]   (bT == false)
//   C += alpha * sum_k A[ra + i][ka + k] * B[cb + j][kb + k]              (bT == true)
// computed by the calling wave with v_mfma_f64_16x16x4f64; K % 4 == 0.
template <bool BT>
__device__ __forceinline__ f64x4 mma16(const double* __restrict__ A, int ra, int ka, const double* __restrict__ B,
                                       int kb, int cb, int K, double alpha, f64x4 acc) {
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll 1
  for (int k = 0; k < K; k += 4) {
    const double a = alpha * A[(ra + li) * BTA_LDS_LD + ka + k + lk];
    const double b = BT ? B[(cb + li) * BTA_LDS_LD + kb + k + lk] : B[(kb + k + lk) * BTA_LDS_LD + cb + li];
    acc = mfma_f64(a, b, acc);
  }
  return acc;
}

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

// s: 64x64 LDS tile (stride BTA_LDS_LD) holding the SPD matrix in its lower
//    triangle (upper triangle ignored).  Rows/cols >= mv are treated as identity.
// On exit: s holds L (lower, zero strict upper, identity outside mv),
//          w holds L^-1 (lower), both with stride BTA_LDS_LD.
// scratch: LDS, >= 64 + 64 doubles (pivots and their logs; the LDS-broadcast
// base case also uses 32 doubles from offset 128 -- callers provide 320).
// Code size matters as much as instruction count here: the workgroup that runs
// this is on the critical path of every panel launch and starts with a cold
// instruction cache, so only the 16-step elimination is unrolled.
__device__ __forceinline__ TileCholResult tile_chol_inv(double* __restrict__ s, double* __restrict__ w,
                                                         double* __restrict__ scratch, int mv) {
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  double* piv = scratch;  // [64]
  int fail = -1;

  // sanitize: identity outside the valid part, zero strict upper triangle
  for (int e = threadIdx.x; e < 64 * 64; e += BTA_THREADS) {
    const int i = e >> 6, j = e & 63;
    if (i >= mv || j >= mv) s[i * BTA_LDS_LD + j] = (i == j) ? 1.0 : 0.0;
    else if (j > i) s[i * BTA_LDS_LD + j] = 0.0;
    w[i * BTA_LDS_LD + j] = 0.0;
  }
  __syncthreads();

#pragma unroll 1
  for (int blk = 0; blk < 4; ++blk) {
    const int c0 = 16 * blk;
    // (a) 16x16 diagonal block: factor + inverse on wave 0
    if (wv == 0) {
      const int i = lane & 15;
      double a[16], e[16];
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        a[c] = (c <= i) ? s[(c0 + i) * BTA_LDS_LD + c0 + c] : 0.0;
        e[c] = (c == i) ? 1.0 : 0.0;
      }
      double myd = 1.0;
#if BTA_BASE_LDS_BROADCAST
      // Column j of A and row j of E are broadcast through LDS (one store per
      // lane, conflict-free broadcast loads); wave-synchronous, no barrier.
      double* bc = scratch + 128;  // [16] column j of A, [16] row j of E
#pragma unroll 1
      for (int j = 0; j < 16; ++j) {
        const double ajj = a[j];
        if (lane < 16) bc[lane] = ajj;                     // A[lane][j]
        if (lane == j) {
#pragma unroll
          for (int c = 0; c < 16; ++c) bc[16 + c] = e[c];  // E[j][:]
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        const double d = bc[j];
        if (lane == 0) {
          piv[c0 + j] = d;
          if (c0 + j < mv && fail < 0 && !(d > 0.0)) fail = c0 + j;
        }
        if (i == j) myd = d;
        const double mi = (i > j) ? ajj / d : 0.0;
#pragma unroll 1
        for (int l = j + 1; l < 16; ++l) a[l] = fma(-mi, bc[l], a[l]);
#pragma unroll 1
        for (int c = 0; c <= j; ++c) e[c] = fma(-mi, bc[16 + c], e[c]);
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
      }
#else
      // Runtime loops with wave-uniform indices: the register arrays are
      // addressed indirectly (no scratch) and the code stays a few hundred
      // instructions long.
#pragma unroll 1
      for (int j = 0; j < 16; ++j) {
        const double ajj = a[j];
        const double d = readlane_f64(ajj, j);
        if (lane == 0) {
          piv[c0 + j] = d;
          if (c0 + j < mv && fail < 0 && !(d > 0.0)) fail = c0 + j;
        }
        if (i == j) myd = d;
        const double sinv = 1.0 / d;
        const double mi = (i > j) ? ajj * sinv : 0.0;  // multiplier of row i
#pragma unroll 1
        for (int l = j + 1; l < 16; ++l) a[l] = fma(-mi, readlane_f64(ajj, l), a[l]);
#pragma unroll 1
        for (int c = 0; c <= j; ++c) e[c] = fma(-mi, readlane_f64(e[c], j), e[c]);
      }
#endif
      const double rsi = 1.0 / sqrt(myd);
      if (lane < 16) {
#pragma unroll 4
        for (int c = 0; c < 16; ++c) {
          const double dc = piv[c0 + c];
          double lv = 0.0;
          if (c == i) lv = sqrt(dc);
          else if (c < i) lv = a[c] / sqrt(dc);
          s[(c0 + i) * BTA_LDS_LD + c0 + c] = lv;
          w[(c0 + i) * BTA_LDS_LD + c0 + c] = (c <= i) ? e[c] * rsi : 0.0;
        }
      }
    }
    __syncthreads();
    // (b) rows below: X = A inv(L_cc)^T  (one 16-row tile per wave)
    const int nbelow = 3 - blk;
    if (wv < nbelow) {
      const int r0 = c0 + 16 * (wv + 1);
      f64x4 x = {0.0, 0.0, 0.0, 0.0};
      // X[i][c] = sum_t A[i][c0+t] * Winv[c][t]  ->  B[k][j] = w[c0+j][c0+k]  (transposed read)
      x = mma16<true>(s, r0, c0, w, c0, c0, 16, 1.0, x);
      __builtin_amdgcn_s_waitcnt(0);
      store16(s, r0, c0, x);
    }
    __syncthreads();
    // (c) trailing lower tiles: A[ri][ci] -= X[ri] X[ci]^T, blk < ci <= ri <= 3
    if (nbelow > 0) {
      const int ntr = nbelow * (nbelow + 1) / 2;
#pragma unroll 1
      for (int q = wv; q < ntr; q += 4) {
        int rr = 0, cc = q;
        while (cc > rr) {
          cc -= rr + 1;
          ++rr;
        }
        const int ri = blk + 1 + rr, ci = blk + 1 + cc;
        f64x4 acc = load16(s, 16 * ri, 16 * ci);
        acc = mma16<true>(s, 16 * ri, c0, s, c0, 16 * ci, 16, -1.0, acc);
        store16(s, 16 * ri, 16 * ci, acc);
      }
    }
    __syncthreads();
  }
  // diagonal tiles above may hold trailing-update values in their strict upper
  // part: clear the strict upper triangle of s
  for (int e = threadIdx.x; e < 64 * 64; e += BTA_THREADS) {
    const int i = e >> 6, j = e & 63;
    if (j > i) s[i * BTA_LDS_LD + j] = 0.0;
  }
  // inverse, block row by block row: W_rc = -W_rr sum_{c<=k<r} L_rk W_kc
#pragma unroll 1
  for (int r = 1; r < 4; ++r) {
    __syncthreads();
    // T_rc = sum_{k=c}^{r-1} L_rk W_kc, parked in the (still empty) block (c, r) of w
    if (wv < r) {
      const int c = wv;
      f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
      for (int k = c; k < r; ++k) acc = mma16<false>(s, 16 * r, 16 * k, w, 16 * k, 16 * c, 16, 1.0, acc);
      store16(w, 16 * c, 16 * r, acc);
    }
    __syncthreads();
    if (wv < r) {
      const int c = wv;
      f64x4 acc = {0.0, 0.0, 0.0, 0.0};
      // W_rc = -W_rr T_rc, T_rc stored at block (c, r): B[k][j] = w[16c + k][16r + j]
      acc = mma16<false>(w, 16 * r, 16 * r, w, 16 * c, 16 * r, 16, -1.0, acc);
      __builtin_amdgcn_s_waitcnt(0);
      store16(w, 16 * r, 16 * c, acc);
    }
  }
  __syncthreads();
  // clear the parked T blocks (strict upper part of w)
  for (int e = threadIdx.x; e < 64 * 64; e += BTA_THREADS) {
    const int i = e >> 6, j = e & 63;
    if (j > i) w[i * BTA_LDS_LD + j] = 0.0;
  }
  // log of the pivots: one per lane of wave 0, then an ordered sum by thread 0
  if (threadIdx.x < 64) {
    const double p = piv[threadIdx.x];
    scratch[64 + threadIdx.x] = (threadIdx.x < mv) ? log(p) : 0.0;
  }
  __syncthreads();
  TileCholResult res;
  res.fail = fail;
  res.logsum = 0.0;
  if (threadIdx.x == 0) {
    double acc = 0.0;
    for (int j = 0; j < mv; ++j) acc += scratch[64 + j];
    res.logsum = acc;
  }
  __syncthreads();
  return res;
}
