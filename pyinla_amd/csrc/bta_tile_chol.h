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
// L W = I to it.  Two barriers per panel, 38 in all.
#pragma once
#include "bta_device.h"

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

// Store only the elements with global row > rmin and global column > cmin.
__device__ __forceinline__ void store16_below(double* __restrict__ C, int r0, int c0, f64x4 v, int rmin,
                                              int cmin) {
  const int lane = threadIdx.x & 63;
  const int j = c0 + (lane & 15);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = r0 + 4 * r + (lane >> 4);
    if (i > rmin && j > cmin) C[i * BTA_LDS_LD + j] = v[r];
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

  for (int e = tid; e < 64 * 64; e += BTA_THREADS) {
    const int i = e >> 6, j = e & 63;
    if (i >= mv || j >= mv) s[i * BTA_LDS_LD + j] = (i == j) ? 1.0 : 0.0;
    w[i * BTA_LDS_LD + j] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();

#pragma unroll 1
  for (int ts = 0; ts < 4; ++ts) {
#pragma unroll 1
    for (int s4 = 0; s4 < 4; ++s4) {
      const int p = 16 * ts + 4 * s4;
      // ---- phase A: 4x4 diagonal factor (every thread) and its inverse
      const double* dg = s + p * BTA_LDS_LD + p;
      const double a00 = dg[0];
      const double a10 = dg[BTA_LDS_LD], a11 = dg[BTA_LDS_LD + 1];
      const double a20 = dg[2 * BTA_LDS_LD], a21 = dg[2 * BTA_LDS_LD + 1], a22 = dg[2 * BTA_LDS_LD + 2];
      const double a30 = dg[3 * BTA_LDS_LD], a31 = dg[3 * BTA_LDS_LD + 1], a32 = dg[3 * BTA_LDS_LD + 2];
      const double a33 = dg[3 * BTA_LDS_LD + 3];
      const double l00 = sqrt(a00), r0 = 1.0 / l00;
      const double l10 = a10 * r0, l20 = a20 * r0, l30 = a30 * r0;
      const double d1 = a11 - l10 * l10;
      const double l11 = sqrt(d1), r1 = 1.0 / l11;
      const double l21 = (a21 - l20 * l10) * r1, l31 = (a31 - l30 * l10) * r1;
      const double d2 = a22 - l20 * l20 - l21 * l21;
      const double l22 = sqrt(d2), r2 = 1.0 / l22;
      const double l32 = (a32 - l30 * l20 - l31 * l21) * r2;
      const double d3 = a33 - l30 * l30 - l31 * l31 - l32 * l32;
      const double l33 = sqrt(d3), r3 = 1.0 / l33;
      // inv(l), lower: i_qq = r_q, i_qm = -r_q sum_{m<=k<q} l_qk i_km
      const double i10 = -r1 * (l10 * r0);
      const double i21 = -r2 * (l21 * r1);
      const double i20 = -r2 * (l20 * r0 + l21 * i10);
      const double i32 = -r3 * (l32 * r2);
      const double i31 = -r3 * (l31 * r1 + l32 * i21);
      const double i30 = -r3 * (l30 * r0 + l31 * i10 + l32 * i20);
      if (tid == 0) {
        piv[p] = a00;
        piv[p + 1] = d1;
        piv[p + 2] = d2;
        piv[p + 3] = d3;
        if (fail < 0) {
          if (p < mv && !(a00 > 0.0)) fail = p;
          else if (p + 1 < mv && !(d1 > 0.0)) fail = p + 1;
          else if (p + 2 < mv && !(d2 > 0.0)) fail = p + 2;
          else if (p + 3 < mv && !(d3 > 0.0)) fail = p + 3;
        }
      }
      const int nrow = 60 - p;  // rows below the panel
      const int q = tid & 3;
      if (tid < 4 * nrow) {
        // X[r][q] = sum_{m<=q} A[r][p+m] inv(l)[q][m]; the 4 threads of a row
        // are adjacent lanes of one wave, so all reads precede the write
        double* row = s + (p + 4 + (tid >> 2)) * BTA_LDS_LD + p;
        const double x0 = row[0], x1 = row[1], x2 = row[2], x3 = row[3];
        double v;
        if (q == 0) v = x0 * r0;
        else if (q == 1) v = x0 * i10 + x1 * r1;
        else if (q == 2) v = x0 * i20 + x1 * i21 + x2 * r2;
        else v = x0 * i30 + x1 * i31 + x2 * i32 + x3 * r3;
        row[q] = v;
      } else {
        // E[p+q][c] = sum_{m<=q} inv(l)[q][m] E[p+m][c], c <= p + 3
        const int c = (tid - 4 * nrow) >> 2;
        double* col = w + p * BTA_LDS_LD + c;
        const double e0 = col[0], e1 = col[BTA_LDS_LD], e2 = col[2 * BTA_LDS_LD], e3 = col[3 * BTA_LDS_LD];
        double v;
        if (q == 0) v = r0 * e0;
        else if (q == 1) v = i10 * e0 + r1 * e1;
        else if (q == 2) v = i20 * e0 + i21 * e1 + r2 * e2;
        else v = i30 * e0 + i31 * e1 + i32 * e2 + r3 * e3;
        col[q * BTA_LDS_LD] = v;
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
        store16_below(s, 16 * wv, 16 * ts, acc, p + 3, p + 3);
#pragma unroll
        for (int ci = 0; ci < 4; ++ci)
          if (ci <= ts) store16_below(w, 16 * wv, 16 * ci, eacc[ci], p + 3, -1);
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
        store16(s, 16 * ri, 16 * ci, acc);
      }
      __syncthreads();
    }
  }
  // the rank updates also wrote the strict upper part of diagonal tiles
  for (int e = tid; e < 64 * 64; e += BTA_THREADS) {
    const int i = e >> 6, j = e & 63;
    if (j > i) s[i * BTA_LDS_LD + j] = 0.0;
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
