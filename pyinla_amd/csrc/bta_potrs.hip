// Triangular sweeps with the BTA factor (pobtas / pobts) on gfx950.
//
// With the block inverses W_i = D_i^-1 (built once per factor), both sweeps
// are chains of memory-bound matrix-vector products, fully parallel inside a
// block:
//   forward  ("N"):  y_i = W_i (r_i - L_{i-1} y_{i-1});  r_a -= sum_i B_i y_i;
//                    y_a = T^-1 r_a
//   backward ("C"):  x_a = T^-T y_a;
//                    x_i = W_i^T (y_i - L_i^T x_{i+1} - B_i^T x_a)
// The right-hand side is overwritten in place (N x nrhs, row-major, leading
// dimension ldrhs), as the reference's in-place pobtas does.
#include <algorithm>

#include "bta_internal.h"

namespace bta {

__global__ void __launch_bounds__(BTA_THREADS) k_copy_rows(double* dst, int ldd, const double* src, int lds,
                                                          int rows, int cols) {
  const int64_t tot = (int64_t)rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * BTA_THREADS + threadIdx.x; e < tot; e += (int64_t)gridDim.x * BTA_THREADS) {
    const int r = (int)(e / cols), c = (int)(e % cols);
    dst[(int64_t)r * ldd + c] = src[(int64_t)r * lds + c];
  }
}

static void copy_rows(double* dst, int ldd, const double* src, int lds, int rows, int cols, hipStream_t st) {
  const int64_t tot = (int64_t)rows * cols;
  const int grid = (int)std::min<int64_t>((tot + BTA_THREADS - 1) / BTA_THREADS, 1024);
  if (grid > 0) hipLaunchKernelGGL(k_copy_rows, dim3(grid), dim3(BTA_THREADS), 0, st, dst, ldd, src, lds, rows, cols);
}

static GemvSeg gseg(const double* A, int lda, int ta, const double* x, int ldx, int k, double alpha, int tri = 0) {
  GemvSeg s;
  s.A = A;
  s.x = x;
  s.alpha = alpha;
  s.k = k;
  s.lda = lda;
  s.ldx = ldx;
  s.ta = ta;
  s.tri = tri;
  return s;
}

static GemvJob gjob(double* y, int m, int ldy, int nrhs, double beta) {
  GemvJob g;
  g.y = y;
  g.beta = beta;
  g.m = m;
  g.ldy = ldy;
  g.nrhs = nrhs;
  g.nseg = 0;
  return g;
}

// op(A) = A^T products through the workspace scratch: right-hand sides are
// taken in groups small enough for the partial sums of the two-pass kernel.
static void gemv_t(Workspace& ws, GemvJob g, hipStream_t st) {
  GemvJob one = g;
  one.nrhs = 1;
  const size_t per = std::max<size_t>(1, gemv_scratch(one, 1));
  const int group = (int)std::max<size_t>(1, std::min<size_t>((size_t)g.nrhs, ws.work_doubles / per));
  for (int c0 = 0; c0 < g.nrhs; c0 += group) {
    GemvJob h = g;
    h.nrhs = std::min(group, g.nrhs - c0);
    h.y = g.y + c0;
    for (int s = 0; s < h.nseg; ++s) h.seg[s].x = g.seg[s].x + c0;
    TimedScope ts_(ws, KC_GEMV, st);
    gemv_launch(h, 1, ws.work, st);
  }
}

int potrs_enqueue(Workspace& ws, const double* D, const double* L, const double* B, const double* T,
                  int n, int b, int a, double* rhs, int nrhs, int ldrhs, int trans, hipStream_t st) {
  (void)D;
  (void)T;
  double* tmp = ws.work;  // b x nrhs, ld nrhs
  double* xa = rhs + (int64_t)n * b * ldrhs;
  if (trans == 0) {
    for (int i = 0; i < n; ++i) {
      double* xi = rhs + (int64_t)i * b * ldrhs;
      if (i > 0) {
        const double* xp = rhs + (int64_t)(i - 1) * b * ldrhs;
        GemvJob g = gjob(xi, b, ldrhs, nrhs, 1.0);
        g.seg[g.nseg++] = gseg(L + (int64_t)(i - 1) * b * b, b, 0, xp, ldrhs, b, -1.0);
        { TimedScope ts_(ws, KC_GEMV, st); gemv_launch(g, 0, nullptr, st); }
      }
      // y_i = W_i x_i through the scratch, in column groups that fit it
      const int group = (int)std::max<size_t>(1, std::min<size_t>((size_t)nrhs, ws.work_doubles / b));
      for (int c0 = 0; c0 < nrhs; c0 += group) {
        const int nc = std::min(group, nrhs - c0);
        GemvJob w = gjob(tmp, b, nc, nc, 0.0);
        w.seg[w.nseg++] = gseg(ws.wfull + (int64_t)i * b * b, b, 0, xi + c0, ldrhs, b, 1.0, 1);
        { TimedScope ts_(ws, KC_GEMV, st); gemv_launch(w, 0, nullptr, st); }
        copy_rows(xi + c0, ldrhs, tmp, nc, b, nc, st);
      }
      if (a > 0) {
        GemvJob ga = gjob(xa, a, ldrhs, nrhs, 1.0);
        ga.seg[ga.nseg++] = gseg(B + (int64_t)i * a * b, b, 0, xi, ldrhs, b, -1.0);
        { TimedScope ts_(ws, KC_GEMV, st); gemv_launch(ga, 0, nullptr, st); }
      }
    }
    if (a > 0) {
      const int group = (int)std::max<size_t>(1, std::min<size_t>((size_t)nrhs, ws.work_doubles / a));
      for (int c0 = 0; c0 < nrhs; c0 += group) {
        const int nc = std::min(group, nrhs - c0);
        GemvJob t = gjob(tmp, a, nc, nc, 0.0);
        t.seg[t.nseg++] = gseg(ws.tinv, 64, 0, xa + c0, ldrhs, a, 1.0, 1);
        { TimedScope ts_(ws, KC_GEMV, st); gemv_launch(t, 0, nullptr, st); }
        copy_rows(xa + c0, ldrhs, tmp, nc, a, nc, st);
      }
    }
  } else {
    // the two-pass T-form products may write y over their own x
    if (a > 0) {
      GemvJob t = gjob(xa, a, ldrhs, nrhs, 0.0);
      t.seg[t.nseg++] = gseg(ws.tinv, 64, 1, xa, ldrhs, a, 1.0, 2);
      gemv_t(ws, t, st);
    }
    for (int i = n - 1; i >= 0; --i) {
      double* xi = rhs + (int64_t)i * b * ldrhs;
      GemvJob g = gjob(xi, b, ldrhs, nrhs, 1.0);
      if (i < n - 1)
        g.seg[g.nseg++] = gseg(L + (int64_t)i * b * b, b, 1, rhs + (int64_t)(i + 1) * b * ldrhs, ldrhs, b, -1.0);
      if (a > 0) g.seg[g.nseg++] = gseg(B + (int64_t)i * a * b, b, 1, xa, ldrhs, a, -1.0);
      if (g.nseg > 0) gemv_t(ws, g, st);
      GemvJob w = gjob(xi, b, ldrhs, nrhs, 0.0);
      w.seg[w.nseg++] = gseg(ws.wfull + (int64_t)i * b * b, b, 1, xi, ldrhs, b, 1.0, 2);
      gemv_t(ws, w, st);
    }
  }
  return (int)hipGetLastError();
}

}  // namespace bta
