// In-workgroup Cholesky factorisation and inverse of one 64x64 FP64 tile.
//
// Right-looking over 8 panels of 8 columns (two per 16-column tile column);
// the matrix A (-> L) and E (-> L^-1, started at I) live in LDS.  For panel
// p = [p, p+8) the 8x8 diagonal block has already been factored,
// A_pp = u diag(d) u^T, by one wave (its unit lower factor u and 1/sqrt(d)
// are in LDS scratch).  Then
//   phase A (all 256 threads, 64 units of 4 threads): a unit owns a row r
//     below the panel and solves X_r = A_rp inv(L_pp)^T, or a column c < p+8
//     of the panel rows of E and replaces it by inv(L_pp) E_pc (forward
//     substitution with u, then scaling by 1/sqrt(d)); the units of
//     c in [p, p+8) also write row c - p of L_pp = u diag(sqrt d);
//   phase B (after one barrier):
//     the first panel of a tile column applies its rank-8 update to the
//     rest of the tile column (A -= X X^T, v_mfma_f64_16x16x4f64, K = 8);
//     the second applies the rank-16 update of the whole tile column to
//     the tiles right of it.  The wave that has just updated the next
//     panel's diagonal block factors it (LDL^T with one reciprocal per
//     column on the dependent chain); the three other waves meanwhile update
//     the rows of E below the panel, E_r -= X_r E_p.
// Row r of E ends as row r of L^-1 because every panel applies the row
// operations of the forward substitution L W = I to it.  Two barriers per
// panel.  The log pivots and the first failing pivot are read off the
// diagonal of L at the end (a non-positive pivot gives NaN there).
// (Measured alternatives on MI355X, cycles of one tile: 4-wide panels with
// the 4x4 factorization repeated by every thread 56 k; 8-wide panels with
// the 8x8 factorization repeated by every thread 48 k; this scheme 45 k.)
#pragma once
#include "bta_device.h"

// 1/sqrt(x) from the hardware estimate plus two Newton steps (full FP64
// accuracy for normal positive x), and 1/x with one Newton step (a few ulp).
// A non-positive pivot propagates as a negative / NaN value and is reported
// by the caller.
__device__ __forceinline__ double rsq_f64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  return y * fma(-h * y, y, 1.5);
}
__device__ __forceinline__ double rcp1_f64(double x) {
  const double y = __builtin_amdgcn_rcp(x);
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

// 16x16 tile C(ra, rb) -= sum_{k<16} A[ra+i][ka+k] A[rb+j][ka+k] (LDS stride
// BTA_LDS_LD), K = 16 unrolled so that all operand reads are issued first.
__device__ __forceinline__ f64x4 mma16_k16_nt(const double* __restrict__ A, int ra, int rb, int ka, f64x4 acc) {
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  double a[4], b[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    a[k] = A[(ra + li) * BTA_LDS_LD + ka + 4 * k + lk];
    b[k] = A[(rb + li) * BTA_LDS_LD + ka + 4 * k + lk];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) acc = mfma_f64(-a[k], b[k], acc);
  return acc;
}

// Rows of E in the 16x16 tile (rt, ct) below panel p: E_r -= X_r E_p, with X
// the panel columns [p, p+8) of s below row p+7 and E_p the panel rows of E.
__device__ __forceinline__ void e_rows_update(const double* __restrict__ s, double* __restrict__ w, int p, int rt,
                                              int ct) {
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int ra = 16 * rt + li;
  const double xa0 = (ra > p + 7) ? s[ra * BTA_LDS_LD + p + lk] : 0.0;
  const double xa1 = (ra > p + 7) ? s[ra * BTA_LDS_LD + p + 4 + lk] : 0.0;
  const double eb0 = w[(p + lk) * BTA_LDS_LD + 16 * ct + li];
  const double eb1 = w[(p + 4 + lk) * BTA_LDS_LD + 16 * ct + li];
  f64x4 e = load16(w, 16 * rt, 16 * ct);
  e = mfma_f64(-xa0, eb0, e);
  e = mfma_f64(-xa1, eb1, e);
  store16_below(w, 16 * rt, 16 * ct, e, p + 7, -1, false);
}

// LDL^T of the 8x8 block at (p, p) of s (lower part), computed by the calling
// wave; lane 0 stores the unit lower factor to ub[8 i + j] (i > j), 1/sqrt(d)
// to gb[0..8) and sqrt(d) to gb[8..16).
__device__ __forceinline__ void panel_ldlt8(const double* __restrict__ s, int p, double* __restrict__ ub,
                                            double* __restrict__ gb) {
  const double* dg = s + p * BTA_LDS_LD + p;
  double a[8][8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) a[i][j] = dg[i * BTA_LDS_LD + j];
  double d[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    d[j] = a[j][j];
    const double qj = rcp1_f64(d[j]);
    double tc[8];
#pragma unroll
    for (int i = j + 1; i < 8; ++i) {
      tc[i] = a[i][j];
      a[i][j] = tc[i] * qj;
    }
#pragma unroll
    for (int i = j + 1; i < 8; ++i)
#pragma unroll
      for (int c = j + 1; c <= i; ++c) a[i][c] = fma(-a[i][j], tc[c], a[i][c]);
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int i = 1; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < i; ++j) ub[8 * i + j] = a[i][j];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const double g = rsq_f64(d[j]);
      gb[j] = g;
      gb[8 + j] = d[j] * g;
    }
  }
}

// s: 64x64 LDS tile (stride BTA_LDS_LD) holding the SPD matrix in its lower
//    triangle (upper triangle ignored).  Rows/cols >= mv are treated as identity.
// On exit: s holds L (lower, zero strict upper, identity outside mv),
//          w holds L^-1 (lower), both with stride BTA_LDS_LD.
// scratch: LDS, >= 80 doubles.
__device__ __forceinline__ TileCholResult tile_chol_inv_p8(double* __restrict__ s, double* __restrict__ w,
                                                            double* __restrict__ scratch, int mv) {
  const int tid = threadIdx.x;
  const int wv = tid >> 6, lane = tid & 63;
  double* ub = scratch;       // [8][8] unit lower factor of the panel's diagonal block
  double* gb = scratch + 64;  // [8] 1 / sqrt(d), then [8] sqrt(d)

  // identity outside the valid block, zero strict upper triangle (the
  // updates below never store above the diagonal), E = I
  for (int e = tid; e < 64 * 64; e += BTA_THREADS) {
    const int i = e >> 6, j = e & 63;
    if (i >= mv || j >= mv || j > i) s[i * BTA_LDS_LD + j] = (i == j) ? 1.0 : 0.0;
    w[i * BTA_LDS_LD + j] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  if (wv == 0) panel_ldlt8(s, 0, ub, gb);
  __syncthreads();

#pragma unroll 1
  for (int k = 0; k < 8; ++k) {
    const int ts = k >> 1, h = k & 1;
    const int p = 8 * k;
    // ---- phase A.  Units 0 .. 55-p: rows p+8+u; the rest: columns
    // c = u - (56-p) of the panel rows of E.  The operands are read first so
    // that their latency overlaps the reads of the factor.
    {
      const int nrow = 56 - p;
      const int u = tid >> 2, t = tid & 3;
      const bool solve_row = u < nrow;
      double* opnd = solve_row ? s + (p + 8 + u) * BTA_LDS_LD + p : w + p * BTA_LDS_LD + (u - nrow);
      const int ostep = solve_row ? 1 : BTA_LDS_LD;
      double x[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) x[m] = opnd[m * ostep];
      double uu[8][8], g[8];
#pragma unroll
      for (int i = 1; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j) uu[i][j] = ub[8 * i + j];
#pragma unroll
      for (int j = 0; j < 8; ++j) g[j] = gb[j];
#pragma unroll
      for (int c = 0; c < 8; ++c)
#pragma unroll
        for (int i = c + 1; i < 8; ++i) x[i] = fma(-uu[i][c], x[c], x[i]);
      // thread t stores outputs 2t and 2t+1 of its unit; the 4 threads of a
      // unit are adjacent lanes of one wave, so all of their reads precede
      // the stores
      const double v0 = t == 0 ? x[0] * g[0] : (t == 1 ? x[2] * g[2] : (t == 2 ? x[4] * g[4] : x[6] * g[6]));
      const double v1 = t == 0 ? x[1] * g[1] : (t == 1 ? x[3] * g[3] : (t == 2 ? x[5] * g[5] : x[7] * g[7]));
      opnd[(2 * t) * ostep] = v0;
      opnd[(2 * t + 1) * ostep] = v1;
      if (u >= 56) {  // column c = p + di of E: also row di of L_pp
        const int di = u - 56;
        double* lrow = s + (p + di) * BTA_LDS_LD + p;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int j = 2 * t + e;
          lrow[j] = j < di ? ub[8 * di + j] * gb[8 + j] : (j == di ? gb[8 + j] : 0.0);
        }
      }
    }
    __syncthreads();
    // ---- phase B
    const int li = lane & 15, lk = lane >> 4;
    const int dw = h == 0 ? ts : 0;  // wave that updates and factors the next diagonal block
    if (h == 0) {
      // rank-8 update of the rest of tile column ts
      if (wv >= ts) {
        const int ra = 16 * wv + li, rb = 16 * ts + li;
        const double xa0 = (ra > p + 7) ? s[ra * BTA_LDS_LD + p + lk] : 0.0;
        const double xa1 = (ra > p + 7) ? s[ra * BTA_LDS_LD + p + 4 + lk] : 0.0;
        const double xb0 = (rb > p + 7) ? s[rb * BTA_LDS_LD + p + lk] : 0.0;
        const double xb1 = (rb > p + 7) ? s[rb * BTA_LDS_LD + p + 4 + lk] : 0.0;
        f64x4 acc = load16(s, 16 * wv, 16 * ts);
        acc = mfma_f64(-xa0, xb0, acc);
        acc = mfma_f64(-xa1, xb1, acc);
        store16_below(s, 16 * wv, 16 * ts, acc, p + 7, p + 7, wv == ts);
      }
    } else if (ts < 3) {
      // tiles right of tile column ts: A[ri][ci] -= X[ri] X[ci]^T (K = 16);
      // wave 0 takes the diagonal tile (ts+1, ts+1) alone, waves 1-3 the rest
      const int nr = 3 - ts;
      const int ntr = nr * (nr + 1) / 2;
      const int t0 = wv, tstep = wv == 0 ? ntr : 3;
#pragma unroll 1
      for (int tt = t0; tt < ntr; tt += tstep) {
        int rr = 0, cc = tt;
        while (cc > rr) {
          cc -= rr + 1;
          ++rr;
        }
        const int ri = ts + 1 + rr, ci = ts + 1 + cc;
        f64x4 acc = load16(s, 16 * ri, 16 * ci);
        acc = mma16_k16_nt(s, 16 * ri, 16 * ci, 16 * ts, acc);
        if (ri == ci) store16_lower(s, 16 * ri, 16 * ci, acc);
        else store16(s, 16 * ri, 16 * ci, acc);
      }
    }
    if (k < 7 && wv == dw) {
      panel_ldlt8(s, p + 8, ub, gb);  // reads back this wave's own stores
    } else {
      // rows of E below the panel, over the three other waves
      const int rt0 = h == 0 ? ts : ts + 1;
      const int ncol = ts + 1;
      const int njobs = (4 - rt0) * ncol;
      const int slot = wv < dw ? wv : wv - 1;
#pragma unroll 1
      for (int j = slot; j < njobs; j += 3) e_rows_update(s, w, p, rt0 + j / ncol, j % ncol);
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

// ---------------------------------------------------------------------------
// Blocked version: the 64x64 tile as 4 x 4 blocks of 16.  For each panel j,
//   phase 1: wave 0 factors and inverts the 16x16 diagonal block A_jj in
//            registers (lane c holds column c; the pivot column is broadcast
//            with v_readlane, so no LDS round trip or barrier sits inside the
//            16 pivots); meanwhile waves 1..3 form T_jk = sum_{m=k}^{j-1}
//            L_jm E_mk for the inverse row j (k < j);
//   phase 2: waves 1..3 form the panel's blocks below the diagonal,
//            L_ij = A_ij E_jj^T (i > j), and the inverse blocks
//            E_jk = -E_jj T_jk (k < j) -- always three 16x16 products;
//   phase 3: all waves apply the trailing updates A_ik -= L_ij L_kj^T
//            (j < k <= i).
// Every off-diagonal product is a 16x16x16 FP64 MFMA product (4
// v_mfma_f64_16x16x4f64) on LDS operands.  Three barriers per panel.
// Inside the 16x16 factorization every column c of E = L^-1 receives the
// same row operations as the matrix (E starts at I): row p is scaled by
// 1 / L_pp and row i > p loses L_ip times row p.

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

// 16x16 Cholesky and inverse of the block at (r0, r0) of s (lower triangle
// read, symmetric), by one wave; writes L (zero strict upper) to s and L^-1 to
// w at the same block.  Lane c (of every 16-lane row) holds column c of the
// symmetric matrix and of E.  At pivot p the finished columns c <= p take a
// zero multiplier (their registers keep the unscaled column of L), and each
// lane scales its own column once, at the end, with the 1 / sqrt(pivot) it
// recorded at its own pivot.  The empty asm statements pin every register's
// update to its pivot (otherwise the compiler sinks the E updates past later
// pivots and keeps every multiplier alive).
__device__ __forceinline__ void chol16_inv16(double* __restrict__ s, double* __restrict__ w, int r0) {
  const int lane = threadIdx.x & 63;
  const int c = lane & 15;
  double a[16], e[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    a[i] = (i >= c) ? s[(r0 + i) * BTA_LDS_LD + r0 + c] : s[(r0 + c) * BTA_LDS_LD + r0 + i];
    e[i] = (i == c) ? 1.0 : 0.0;
  }
  double myrs = 1.0, mysq = 1.0;
#pragma unroll
  for (int p = 0; p < 16; ++p) {
    const double d = readlane_f64(a[p], p);  // pivot A_pp (lane p holds column p)
    const double rs = rsq_f64(d);
    if (c == p) {
      myrs = rs;
      mysq = d * rs;
    }
    const double mylp = (c > p) ? a[p] * rs : 0.0;  // L_cp = A_pc / sqrt(d) (symmetry)
    e[p] *= rs;                                     // row p of E is final
#pragma unroll
    for (int i = p + 1; i < 16; ++i) {
      const double bl = readlane_f64(a[i], p) * rs;  // L_ip
      a[i] = fma(-bl, mylp, a[i]);
      e[i] = fma(-bl, e[p], e[i]);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      asm volatile("" : "+v"(a[i]));
      asm volatile("" : "+v"(e[i]));
    }
  }
  if (lane < 16) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      s[(r0 + i) * BTA_LDS_LD + r0 + c] = (i > c) ? a[i] * myrs : (i == c ? mysq : 0.0);
      w[(r0 + i) * BTA_LDS_LD + r0 + c] = e[i];
    }
  }
}

__device__ __forceinline__ TileCholResult tile_chol_inv16(double* __restrict__ s, double* __restrict__ w, int mv) {
  const int tid = threadIdx.x;
  const int wv = tid >> 6, lane = tid & 63;
  // identity outside the valid block, zero strict upper triangle, E = 0
  for (int e = tid; e < 64 * 64; e += BTA_THREADS) {
    const int i = e >> 6, j = e & 63;
    if (i >= mv || j >= mv || j > i) s[i * BTA_LDS_LD + j] = (i == j) ? 1.0 : 0.0;
    w[i * BTA_LDS_LD + j] = 0.0;
  }
  __syncthreads();
#pragma unroll 1
  for (int jb = 0; jb < 4; ++jb) {
    const int r0 = 16 * jb;
    // ---- phase 1
    if (wv == 0) {
      chol16_inv16(s, w, r0);
    } else if (jb > 0) {
      // T_jk = sum_{m=k}^{j-1} L_jm E_mk into w(j, k), k < j; jobs spread
      // over waves 1..3 (at most 3 for jb == 3: k = 0 (3 terms), 1 (2), 2 (1))
      for (int k = wv - 1; k < jb; k += 3) {
        f64x4 acc = f64x4{0.0, 0.0, 0.0, 0.0};
        for (int m = k; m < jb; ++m) acc = mma16x16<false, 1>(s, r0, 16 * m, w, 16 * k, 16 * m, acc);
        store16(w, r0, 16 * k, acc);
      }
    }
    __syncthreads();
    // ---- phase 2: waves 1..3, one 16x16 product each
    if (wv > 0) {
      const int q = wv - 1;  // q < 3 - jb: L_(jb+1+q), jb;  else E_jb, (q - (3 - jb))
      f64x4 acc = f64x4{0.0, 0.0, 0.0, 0.0};
      if (q < 3 - jb) {
        const int ib = jb + 1 + q;
        acc = mma16x16<false, 0>(s, 16 * ib, r0, w, r0, r0, acc);  // A_ib,jb E_jb,jb^T
        store16(s, 16 * ib, r0, acc);
      } else {
        const int k = q - (3 - jb);
        acc = mma16x16<true, 1>(w, r0, r0, w, 16 * k, r0, acc);    // -E_jj T_jk
        store16(w, r0, 16 * k, acc);
      }
    }
    __syncthreads();
    // ---- phase 3: trailing updates A_ik -= L_ij L_kj^T, jb < k <= i
    if (jb < 3) {
      const int nr = 3 - jb, nt = nr * (nr + 1) / 2;
      for (int t = wv; t < nt; t += 4) {
        int rr = 0, cc = t;
        while (cc > rr) {
          cc -= rr + 1;
          ++rr;
        }
        const int ib = jb + 1 + rr, kb = jb + 1 + cc;
        f64x4 acc = load16(s, 16 * ib, 16 * kb);
        acc = mma16x16<true, 0>(s, 16 * ib, r0, s, 16 * kb, r0, acc);
        if (ib == kb) store16_lower(s, 16 * ib, 16 * kb, acc);
        else store16(s, 16 * ib, 16 * kb, acc);
      }
      __syncthreads();
    }
  }
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

#ifndef BTA_TILE_CHOL16
#define BTA_TILE_CHOL16 1
#endif

// s: 64x64 LDS tile (stride BTA_LDS_LD) holding the SPD matrix in its lower
// triangle; rows / cols >= mv are treated as identity.  On exit s holds L
// (zero strict upper) and w holds L^-1 (lower).  scratch: >= 80 doubles of
// LDS (used by the 8-wide panel variant only).
__device__ __forceinline__ TileCholResult tile_chol_inv(double* __restrict__ s, double* __restrict__ w,
                                                         double* __restrict__ scratch, int mv) {
#if BTA_TILE_CHOL16
  (void)scratch;
  return tile_chol_inv16(s, w, mv);
#else
  return tile_chol_inv_p8(s, w, scratch, mv);
#endif
}
