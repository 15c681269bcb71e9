// In-workgroup Cholesky factorisation and inverse of one 64x64 FP64 tile.
//
// Method: right-looking LDL^T elimination with the row operations applied to
// an identity at the same time (A = Lu D Lu^T, E = Lu^-1), then
// L = Lu D^1/2 and L^-1 = D^-1/2 E.  Register layout: thread t (lane = t%64,
// q = t/64) holds A[lane][q+4m] and E[q+4m][lane] for m = 0..15, so column j of
// A and row j of E are both owned by wave j%4 one element per lane and can be
// broadcast through LDS with a single store per lane and one barrier per
// column (the broadcast buffers alternate between two LDS slots).
#pragma once
#include "bta_device.h"

struct TileCholResult {
  double logsum;  // sum of log(pivot) over the valid part (thread 0 only)
  int fail;       // first non-positive / NaN pivot in the valid part, or -1 (thread 0 only)
};

// s: 64x64 LDS tile (stride BTA_LDS_LD) holding the SPD matrix in its lower
//    triangle (upper triangle ignored).  Rows/cols >= mv are treated as identity.
// On exit: s holds L (lower, zero strict upper, identity outside mv),
//          w holds L^-1 (lower), both with stride BTA_LDS_LD.
// scratch: LDS, >= 4*64 + 128 doubles.
__device__ __forceinline__ TileCholResult tile_chol_inv(double* __restrict__ s, double* __restrict__ w,
                                                         double* __restrict__ scratch, int mv) {
  const int lane = threadIdx.x & 63;
  const int q = threadIdx.x >> 6;
  double* col = scratch;          // [2][64] broadcast of A[:, j]
  double* erow = scratch + 128;   // [2][64] broadcast of E[j, :]
  double* piv = scratch + 256;    // [64] pivots
  double A[16], E[16], dl[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const int l = q + 4 * m;
    double v;
    if (lane < mv && l < mv) v = (lane >= l) ? s[lane * BTA_LDS_LD + l] : 0.0;
    else v = (lane == l) ? 1.0 : 0.0;
    A[m] = v;
    E[m] = (l == lane) ? 1.0 : 0.0;
    dl[m] = 1.0;
  }
  int fail = -1;
#pragma unroll
  for (int j = 0; j < 64; ++j) {
    const int buf = (j & 1) * 64;
    if (q == (j & 3)) {
      col[buf + lane] = A[j >> 2];
      erow[buf + lane] = E[j >> 2];
    }
    __syncthreads();
    const double d = col[buf + j];
    if (threadIdx.x == 0) {
      piv[j] = d;
      if (j < mv && fail < 0 && !(d > 0.0)) fail = j;
    }
    const double sinv = 1.0 / d;
    const double ai = col[buf + lane] * sinv;
    const double er = erow[buf + lane];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int l = q + 4 * m;
      if (l > j) {
        const double al = col[buf + l];
        A[m] = fma(-ai, al, A[m]);
        E[m] = fma(-(al * sinv), er, E[m]);
      }
    }
    if (q == (j & 3)) dl[j >> 2] = d;
  }
  // Final scaling: L[i][l] = A[i][l] / sqrt(d_l),  Linv[l][r] = E[l][r] / sqrt(d_l).
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const int l = q + 4 * m;
    const double sq = sqrt(dl[m]);
    const double rs = 1.0 / sq;
    double lv = (lane > l) ? A[m] * rs : ((lane == l) ? sq : 0.0);
    double ev = (l >= lane) ? E[m] * rs : 0.0;
    s[lane * BTA_LDS_LD + l] = lv;
    w[l * BTA_LDS_LD + lane] = ev;
  }
  __syncthreads();
  TileCholResult res;
  res.fail = fail;
  res.logsum = 0.0;
  // log of the pivots: one per lane of wave 0, then an ordered sum by thread 0.
  if (threadIdx.x < 64) {
    const double p = piv[threadIdx.x];
    piv[threadIdx.x] = (threadIdx.x < mv) ? log(p) : 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double acc = 0.0;
    for (int j = 0; j < mv; ++j) acc += piv[j];
    res.logsum = acc;
  }
  __syncthreads();
  return res;
}
