// In-workgroup Cholesky factorisation and inverse of one 64x64 FP64 tile.
//
// Right-looking over 16 panels of 4 columns, grouped in four 16-column tile
// columns; the matrix A (-> L) and E (-> L^-1, started at I) live in LDS.
// Panel p = [p, p+4):
//   phase A (all 256 threads, no cross-wave dependence):
//     every thread factors the 4x4 diagonal block A_pp = l l^T itself and
//     forms inv(l) (IEEE sqrt / divide, so the pivots are correctly rounded);
//     the 4 (60 - p) threads of the rows below solve X = A_rp inv(l)^T, one
//     element each; the other 4 (p + 4) threads replace the panel rows of E
//     by inv(l) E_p (columns <= p + 3, the only non-zero ones);
//   phase B (after one barrier):
//     rank-4 updates with v_mfma_f64_16x16x4f64 (K = 4 is the panel width):
//     the rest of the current 16-column tile column, A -= X X^T, and the
//     rows of E below the panel, E_r -= X E_p;
// and after the four panels of a tile column, the tiles to its right receive
// the rank-16 update in one pass.  Row r of E ends as row r of L^-1 because
// every panel step applies the row operations of the forward substitution
// L W = I to it.  Two barriers per panel, 37 in all.  (Measured alternatives
// -- a balanced job list over the waves, or one barrier per panel with
// redundant per-lane operands -- compiled to branchier code and ran 1.5-2x
// slower on MI355X.)
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

#pragma unroll 1
  for (int ts = 0; ts < 4; ++ts) {
#pragma unroll 1
    for (int s4 = 0; s4 < 4; ++s4) {
      const int p = 16 * ts + 4 * s4;
      // ---- phase A.  Each thread first issues the LDS reads of its own
      // operands (a row of the panel below, or a column of the panel rows of
      // E), so that their latency overlaps the 4x4 factorization.
      const int nrow = 60 - p;  // rows below the panel
      const int q = tid & 3;
      const bool solve_row = tid < 4 * nrow;
      double* opnd = solve_row ? s + (p + 4 + (tid >> 2)) * BTA_LDS_LD + p
                               : w + p * BTA_LDS_LD + ((tid - 4 * nrow) >> 2);
      const int ostep = solve_row ? 1 : BTA_LDS_LD;
      const double x0 = opnd[0], x1 = opnd[ostep], x2 = opnd[2 * ostep], x3 = opnd[3 * ostep];
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
      const double s0 = d0 * g0, s1 = d1 * g1, s2 = d2 * g2, s3 = d3 * g3;  // sqrt(d)
      // inv(l) = diag(1/sqrt d) inv(u), inv(u) unit lower
      const double v10 = -u10, v21 = -u21, v32 = -u32;
      const double v20 = -fma(u21, v10, u20);
      const double v31 = -fma(u32, v21, u31);
      const double v30 = -fma(u32, v20, fma(u31, v10, u30));
      const double r0 = g0, r1 = g1, r2 = g2, r3 = g3;
      const double i10 = g1 * v10, i20 = g2 * v20, i21 = g2 * v21;
      const double i30 = g3 * v30, i31 = g3 * v31, i32 = g3 * v32;
      const double l00 = s0, l10 = u10 * s0, l20 = u20 * s0, l30 = u30 * s0;
      const double l11 = s1, l21 = u21 * s1, l31 = u31 * s1;
      const double l22 = s2, l32 = u32 * s2, l33 = s3;
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
      // rows below: X[r][q] = sum_{m<=q} A[r][p+m] inv(l)[q][m];
      // panel rows of E: E[p+q][c] = sum_{m<=q} inv(l)[q][m] E[p+m][c] (c <= p+3).
      // The 4 threads sharing a row / column are adjacent lanes of one wave,
      // so all of their reads precede the write.
      {
        double v;
        if (q == 0) v = x0 * r0;
        else if (q == 1) v = fma(x0, i10, x1 * r1);
        else if (q == 2) v = fma(x0, i20, fma(x1, i21, x2 * r2));
        else v = fma(x0, i30, fma(x1, i31, fma(x2, i32, x3 * r3)));
        opnd[q * ostep] = v;
      }
      __syncthreads();
      // ---- phase B: write l; rank-4 updates of the tile column and of E
      if (tid < 16) {
        const int i = tid >> 2, j = tid & 3;
        double v;
        if (i == 0) v = (j == 0) ? l00 : 0.0;
        else if (i == 1) v = (j == 0) ? l10 : (j == 1 ? l11 : 0.0);
        else if (i == 2) v = (j == 0) ? l20 : (j == 1 ? l21 : (j == 2 ? l22 : 0.0));
        else v = (j == 0) ? l30 : (j == 1 ? l31 : (j == 2 ? l32 : l33));
        s[(p + i) * BTA_LDS_LD + p + j] = v;
      }
      if (wv >= ts) {
        const int li = lane & 15, lk = lane >> 4;
        const int ra = 16 * wv + li;  // row of the A operand (this wave's tile row)
        const int rb = 16 * ts + li;  // row of the B operand (tile column ts)
        const double xa = (ra > p + 3) ? s[ra * BTA_LDS_LD + p + lk] : 0.0;
        const double xb = (rb > p + 3) ? s[rb * BTA_LDS_LD + p + lk] : 0.0;
        f64x4 acc = mfma_f64(-xa, xb, load16(s, 16 * wv, 16 * ts));
        // rows of E in this tile row below the panel: E_r -= X_r E_p, tiles (wv, 0..ts)
        f64x4 eacc[4];
#pragma unroll
        for (int ci = 0; ci < 4; ++ci) {
          if (ci <= ts) {
            const double eb = w[(p + lk) * BTA_LDS_LD + 16 * ci + li];
            eacc[ci] = mfma_f64(-xa, eb, load16(w, 16 * wv, 16 * ci));
          }
        }
        store16_below(s, 16 * wv, 16 * ts, acc, p + 3, p + 3, wv == ts);
#pragma unroll
        for (int ci = 0; ci < 4; ++ci)
          if (ci <= ts) store16_below(w, 16 * wv, 16 * ci, eacc[ci], p + 3, -1, false);
      }
      __syncthreads();
    }
    // tiles right of tile column ts: A[ri][ci] -= X[ri] X[ci]^T (K = 16)
    const int nr = 3 - ts;
    if (nr > 0) {
      const int ntr = nr * (nr + 1) / 2;
#pragma unroll 1
      for (int t = wv; t < ntr; t += 4) {
        int rr = 0, cc = t;
        while (cc > rr) {
          cc -= rr + 1;
          ++rr;
        }
        const int ri = ts + 1 + rr, ci = ts + 1 + cc;
        f64x4 acc = load16(s, 16 * ri, 16 * ci);
        acc = mma16<true>(s, 16 * ri, 16 * ts, s, 16 * ts, 16 * ci, 16, -1.0, acc);
        if (ri == ci) store16_lower(s, 16 * ri, 16 * ci, acc);
        else store16(s, 16 * ri, 16 * ci, acc);
      }
      __syncthreads();
    }
  }
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
