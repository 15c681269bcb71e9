// In-workgroup Cholesky factorisation and inverse of one 64x64 FP64 tile.
//
// Right-looking over 16 panels of 4 columns; the matrix A (-> L) and
// E (-> L^-1, started at I) live in LDS.  One barrier per panel p:
//   - every thread factors the 4x4 diagonal block A_pp = l l^T itself (LDL^T,
//     one reciprocal per column on the dependent chain, square roots off it)
//     and forms row lk of inv(l) for its lane;
//   - every lane forms, for the four 16-row tiles, its MFMA operand values
//     X[16t + li][p + lk] = A[16t + li][p : p+4] inv(l)[lk]^T and
//     E_new[lk][16t + li] = inv(l)[lk] E[p : p+4][16t + li]
//     (redundantly across waves, so no wave waits for another);
//   - rank-4 updates with v_mfma_f64_16x16x4f64 (K = 4 is the panel width) of
//     every trailing lower 16x16 tile, A -= X X^T, and of the E rows below the
//     panel, E_r -= X E_new: at most 14 tile jobs, round-robin over 4 waves;
//   - the panel's own results (X in the panel columns, l, E_new in the panel
//     rows) are stored at the start of the next interval, because other waves
//     still read the old values of those locations in this one.
// Row r of E ends as row r of L^-1 because every panel applies the row
// operations of the forward substitution L W = I to it.  17 barriers in all.
#pragma once
#include "bta_device.h"

// 1/x and 1/sqrt(x) from the hardware estimates plus two Newton steps (full
// FP64 accuracy for normal positive x; a non-positive pivot still propagates
// as a negative / NaN value and is reported by the caller).
__device__ __forceinline__ double rcp_f64(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}
__device__ __forceinline__ double rsq_f64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  return y * fma(-h * y, y, 1.5);
}

struct TileCholResult {
  double logsum;  // sum of log(pivot) over the valid part (thread 0 only)
  int fail;       // first non-positive / NaN pivot in the valid part, or -1 (thread 0 only)
};

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

// Store only the elements with global row > rmin and global column > cmin
// (and, with lower, on or below the diagonal).
__device__ __forceinline__ void store16_below(double* __restrict__ C, int r0, int c0, f64x4 v, int rmin,
                                              int cmin, bool lower) {
  const int lane = threadIdx.x & 63;
  const int j = c0 + (lane & 15);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = r0 + 4 * r + (lane >> 4);
    if (i > rmin && j > cmin && (!lower || j <= i)) C[i * BTA_LDS_LD + j] = v[r];
  }
}

// s: 64x64 LDS tile (stride BTA_LDS_LD) holding the SPD matrix in its lower
//    triangle (upper triangle ignored).  Rows/cols >= mv are treated as identity.
// On exit: s holds L (lower, zero strict upper, identity outside mv),
//          w holds L^-1 (lower), both with stride BTA_LDS_LD.
// scratch: LDS, >= 64 doubles.
__device__ __forceinline__ TileCholResult tile_chol_inv(double* __restrict__ s, double* __restrict__ w,
                                                         double* __restrict__ scratch, int mv) {
  const int tid = threadIdx.x;
  const int wv = tid >> 6, lane = tid & 63;
  double* piv = scratch;  // [64] pivots (squared diagonal of L)
  int fail = -1;

  // identity outside the valid block, zero strict upper triangle (the
  // updates below never store above the diagonal), E = I
  for (int e = tid; e < 64 * 64; e += BTA_THREADS) {
    const int i = e >> 6, j = e & 63;
    if (i >= mv || j >= mv || j > i) s[i * BTA_LDS_LD + j] = (i == j) ? 1.0 : 0.0;
    w[i * BTA_LDS_LD + j] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();

  const int li = lane & 15, lk = lane >> 4;
  // Values of the previous panel whose stores are deferred by one interval
  // (their locations are still read, as old values, by other waves during the
  // panel that produced them): X[16 wv + li][pp + lk], E_new[lk][16 wv + li]
  // and, on threads 0..15, one entry of the 4x4 factor.
  double xw = 0.0, ew = 0.0, lw = 0.0;
  int pp = -1;
#pragma unroll 1
  for (int p = 0; p < 64; p += 4) {
    const int ts = p >> 4;
    if (pp >= 0) {
      if (16 * wv + li > pp + 3) s[(16 * wv + li) * BTA_LDS_LD + pp + lk] = xw;
      if (16 * wv + li <= pp + 3) w[(pp + lk) * BTA_LDS_LD + 16 * wv + li] = ew;
      if (tid < 16) s[(pp + (tid >> 2)) * BTA_LDS_LD + pp + (tid & 3)] = lw;
    }
    // operands: rows of the panel columns for the 4 tile rows, and the panel
    // rows of E at the 4 column groups (issued before the 4x4 factorization)
    double ar[4][4], er[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int r = 16 * t + li;
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        ar[t][m] = (r > p + 3) ? s[r * BTA_LDS_LD + p + m] : 0.0;
        er[t][m] = (16 * t <= p + 3) ? w[(p + m) * BTA_LDS_LD + 16 * t + li] : 0.0;
      }
    }
    const double* dg = s + p * BTA_LDS_LD + p;
    const double a00 = dg[0];
    const double a10 = dg[BTA_LDS_LD], a11 = dg[BTA_LDS_LD + 1];
    const double a20 = dg[2 * BTA_LDS_LD], a21 = dg[2 * BTA_LDS_LD + 1], a22 = dg[2 * BTA_LDS_LD + 2];
    const double a30 = dg[3 * BTA_LDS_LD], a31 = dg[3 * BTA_LDS_LD + 1], a32 = dg[3 * BTA_LDS_LD + 2];
    const double a33 = dg[3 * BTA_LDS_LD + 3];
    // LDL^T of the 4x4 block: the dependent chain holds one reciprocal per
    // column; square roots come after, off the chain.  l = u diag(sqrt d).
    const double d0 = a00, q0 = rcp_f64(d0);
    const double u10 = a10 * q0, u20 = a20 * q0, u30 = a30 * q0;
    const double d1 = fma(-u10, a10, a11), q1 = rcp_f64(d1);
    const double t21 = fma(-u20, a10, a21), t31 = fma(-u30, a10, a31);
    const double u21 = t21 * q1, u31 = t31 * q1;
    const double d2 = fma(-u21, t21, fma(-u20, a20, a22)), q2 = rcp_f64(d2);
    const double t32 = fma(-u31, t21, fma(-u30, a20, a32));
    const double u32 = t32 * q2;
    const double d3 = fma(-u32, t32, fma(-u31, t31, fma(-u30, a30, a33)));
    const double g0 = rsq_f64(d0), g1 = rsq_f64(d1), g2 = rsq_f64(d2), g3 = rsq_f64(d3);
    // inv(l) = diag(1/sqrt d) inv(u), inv(u) unit lower; this lane's row lk
    const double v10 = -u10, v21 = -u21, v32 = -u32;
    const double v20 = -fma(u21, v10, u20);
    const double v31 = -fma(u32, v21, u31);
    const double v30 = -fma(u32, v20, fma(u31, v10, u30));
    double iv0, iv1 = 0.0, iv2 = 0.0, iv3 = 0.0;
    if (lk == 0) iv0 = g0;
    else if (lk == 1) { iv0 = g1 * v10; iv1 = g1; }
    else if (lk == 2) { iv0 = g2 * v20; iv1 = g2 * v21; iv2 = g2; }
    else { iv0 = g3 * v30; iv1 = g3 * v31; iv2 = g3 * v32; iv3 = g3; }
    if (tid == 0) {
      piv[p] = d0;
      piv[p + 1] = d1;
      piv[p + 2] = d2;
      piv[p + 3] = d3;
      if (fail < 0) {
        if (p < mv && !(d0 > 0.0)) fail = p;
        else if (p + 1 < mv && !(d1 > 0.0)) fail = p + 1;
        else if (p + 2 < mv && !(d2 > 0.0)) fail = p + 2;
        else if (p + 3 < mv && !(d3 > 0.0)) fail = p + 3;
      }
    }
    if (tid < 16) {  // entry (i, j) of l, stored in the next interval
      const int i = tid >> 2, j = tid & 3;
      const double s0 = d0 * g0, s1 = d1 * g1, s2 = d2 * g2, s3 = d3 * g3;
      if (i == 0) lw = (j == 0) ? s0 : 0.0;
      else if (i == 1) lw = (j == 0) ? u10 * s0 : (j == 1 ? s1 : 0.0);
      else if (i == 2) lw = (j == 0) ? u20 * s0 : (j == 1 ? u21 * s1 : (j == 2 ? s2 : 0.0));
      else lw = (j == 0) ? u30 * s0 : (j == 1 ? u31 * s1 : (j == 2 ? u32 * s2 : s3));
    }
    // X[16 t + li][p + lk] = sum_m A[16 t + li][p + m] inv(l)[lk][m]   (rows > p + 3)
    // E_new[lk][16 t + li] = sum_m inv(l)[lk][m] E[p + m][16 t + li]
    double xv[4], ev[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      xv[t] = fma(ar[t][0], iv0, fma(ar[t][1], iv1, fma(ar[t][2], iv2, ar[t][3] * iv3)));
      ev[t] = fma(iv0, er[t][0], fma(iv1, er[t][1], fma(iv2, er[t][2], iv3 * er[t][3])));
    }
    xw = xv[0];
    ew = ev[0];
#pragma unroll
    for (int t = 1; t < 4; ++t)
      if (wv == t) {
        xw = xv[t];
        ew = ev[t];
      }
    // rank-4 jobs, round-robin over the waves (<= 4 each):
    //   A: lower tiles (ri, ci), ci >= ts: C -= X_ri X_ci^T (rows, cols > p + 3)
    //   E: tiles (ri, ci), ri >= ts, ci <= ts: E_r -= X_r E_new (rows > p + 3)
    const int nA = (4 - ts) * (5 - ts) / 2, nE = (4 - ts) * (ts + 1);
    f64x4 acc[4];
    int jr[4], jc[4], jk[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j = wv + 4 * q;
      jk[q] = (j < nA) ? 0 : (j < nA + nE ? 1 : 2);
      int ri = 0, ci = 0;
      if (jk[q] == 0) {
        int t = j;
        ci = ts;
        while (t >= 4 - ci) {
          t -= 4 - ci;
          ++ci;
        }
        ri = ci + t;
      } else if (jk[q] == 1) {
        const int t = j - nA;
        ri = ts + t / (ts + 1);
        ci = t % (ts + 1);
      }
      jr[q] = ri;
      jc[q] = ci;
      if (jk[q] < 2) {
        double xa = xv[0], xb = (jk[q] == 0) ? xv[0] : ev[0];
#pragma unroll
        for (int t = 1; t < 4; ++t) {
          if (ri == t) xa = xv[t];
          if (ci == t) xb = (jk[q] == 0) ? xv[t] : ev[t];
        }
        acc[q] = mfma_f64(-xa, xb, load16(jk[q] == 0 ? s : w, 16 * ri, 16 * ci));
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (jk[q] == 0) store16_below(s, 16 * jr[q], 16 * jc[q], acc[q], p + 3, p + 3, jr[q] == jc[q]);
      else if (jk[q] == 1) store16_below(w, 16 * jr[q], 16 * jc[q], acc[q], p + 3, -1, false);
    }
    pp = p;
    __syncthreads();
  }
  // deferred stores of the last panel
  if (16 * wv + li > pp + 3) s[(16 * wv + li) * BTA_LDS_LD + pp + lk] = xw;
  if (16 * wv + li <= pp + 3) w[(pp + lk) * BTA_LDS_LD + 16 * wv + li] = ew;
  if (tid < 16) s[(pp + (tid >> 2)) * BTA_LDS_LD + pp + (tid & 3)] = lw;
  __syncthreads();
  // sum of log pivots: fixed butterfly over wave 0 (deterministic order)
  TileCholResult res;
  res.fail = fail;
  res.logsum = 0.0;
  if (wv == 0) {
    double v = (lane < mv) ? log(piv[lane]) : 0.0;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    res.logsum = v;
  }
  __syncthreads();
  return res;
}
