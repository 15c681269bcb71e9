// Grouped FP64 GEMM on 64x64 output tiles (FP64 MFMA), driven by a job table
// in device memory.  One launch covers any number of independent jobs of any
// shapes; a workgroup finds its job by binary search over the tile prefix
// sums.  Each job computes
//     C = beta * C + sum_s alpha_s * op(A_s) * op(B_s)
// over up to three segments (K-concatenation), optionally only on the tiles on
// or below the diagonal (symmetric output), optionally mirroring them so the
// full symmetric matrix is written.  Triangular operands restrict each tile's
// K range so structurally-zero products are skipped.
#pragma once
#include <vector>

#include "bta_device.h"

namespace bta {

enum : int {
  TRI_A_LOWER = 1,  // op(A)[i][k] == 0 for k > i   -> k < (tr+1)*64
  TRI_A_UPPER = 2,  // op(A)[i][k] == 0 for k < i   -> k >= tr*64
  TRI_B_LOWER = 4,  // op(B)[k][j] == 0 for k < j   -> k >= tc*64
  TRI_B_UPPER = 8,  // op(B)[k][j] == 0 for k > j   -> k < (tc+1)*64
};

struct GemmSeg {
  const double* A;
  const double* B;
  double alpha;
  int k;       // inner dimension
  int lda, ldb;
  int ta, tb;  // 0: op(X) = X, 1: op(X) = X^T
  int tri;     // TRI_* flags
};

struct GemmJob {
  double* C;
  double beta;   // 0: C is not read
  int m, n, ldc;
  int lower;     // 1: only tiles with tr >= tc
  int mirror;    // with lower: also write the transpose into the upper triangle
  int nseg;
  // split = 2 (lower jobs of the 64-tile kernel only): two workgroups per
  // output tile, one per half of the tile's K chunks; each writes its partial
  // sum to part (two 64 x 64 slots per tile) and counts itself in cnt (one
  // counter per tile, zero between launches), and the second to arrive adds
  // the other's partial and writes the tile.  Set by DevPlan::build.
  int split;
  double* part;
  int* cnt;
  GemmSeg seg[3];
};

// Host-side helper: builds the job table and tile prefix for one launch.
struct GemmLaunch {
  std::vector<GemmJob> jobs;
  std::vector<int> prefix;  // njobs + 1
  int ts = 64;              // output tile edge: 64 or 128
  explicit GemmLaunch(int tile = 64) : ts(tile) {}
  int tiles() const { return prefix.empty() ? 0 : prefix.back(); }
  void add(const GemmJob& j) {
    if (j.m <= 0 || j.n <= 0) return;
    if (prefix.empty()) prefix.push_back(0);
    const int tm = (j.m + ts - 1) / ts, tn = (j.n + ts - 1) / ts;
    jobs.push_back(j);
    prefix.push_back(prefix.back() + (j.lower ? tm * (tm + 1) / 2 : tm * tn) * j.split);
  }
};

// Default output tile edge.  64x64 tiles run two workgroups per CU (one hides
// the other's load latency) and keep 256 CUs busy down to 2000x2000 outputs;
// measured on MI355X at b = 2000 and 4500 they beat the 128x128 kernel for
// every product of the factorization and the selected inversion (1.44x on
// the selected inversion at b = 2000).  128 stays available to the
// autotuner.
inline int pick_tile(int m, int n) {
  (void)m;
  (void)n;
  return 64;
}

inline GemmSeg seg(const double* A, int lda, int ta, const double* B, int ldb, int tb, int k,
                   double alpha, int tri = 0) {
  GemmSeg s;
  s.A = A;
  s.B = B;
  s.alpha = alpha;
  s.k = k;
  s.lda = lda;
  s.ldb = ldb;
  s.ta = ta;
  s.tb = tb;
  s.tri = tri;
  return s;
}

inline GemmJob job(double* C, int m, int n, int ldc, double beta, int lower = 0, int mirror = 0) {
  GemmJob j;
  j.C = C;
  j.beta = beta;
  j.m = m;
  j.n = n;
  j.ldc = ldc;
  j.lower = lower;
  j.mirror = mirror;
  j.nseg = 0;
  j.split = 1;
  j.part = nullptr;
  j.cnt = nullptr;
  return j;
}

inline GemmJob& add_seg(GemmJob& j, const GemmSeg& s) {
  if (s.k > 0) j.seg[j.nseg++] = s;
  return j;
}

// y = beta*y + sum_s alpha_s * op(A_s) x_s, with x and y row-major k x nrhs /
// m x nrhs blocks (leading dimensions ldx / ldy: one column per right-hand side).
struct GemvSeg {
  const double* A;
  const double* x;
  double alpha;
  int k;      // length of x (columns of op(A))
  int lda, ldx;
  int ta;     // 0: A (m x k), 1: A^T (A is k x m)
  int tri;    // 1: op(A) lower (k <= i), 2: op(A) upper (k >= i)
};

struct GemvJob {
  double* y;
  double beta;
  int m, ldy, nrhs;
  int nseg;
  GemvSeg seg[3];
};


// A set of grouped-GEMM launches whose job tables live in device memory.  Built
// once (hipMalloc + hipMemcpy, never during stream capture), then launched any
// number of times, including from inside a captured hipGraph.
struct DevPlan {
  // vec: every operand has an even leading dimension and a 16-byte aligned
  // base (the pipelined kernel's paired loads need both)
  // split: the launch holds split jobs (run by the 64-tile kernel whatever
  // the variant asked for)
  struct Entry { int job_off, pre_off, njobs, tiles, ts, vec, split; };
  GemmJob* d_jobs = nullptr;
  int* d_prefix = nullptr;
  double* d_part = nullptr;  // partial sums of split jobs, shared by the launches (one stream)
  int* d_cnt = nullptr;      // their per-tile counters
  size_t bytes = 0;  // device bytes of the job table, prefix sums and split scratch
  std::vector<Entry> launches;
  DevPlan() = default;
  DevPlan(const DevPlan&) = delete;
  DevPlan& operator=(const DevPlan&) = delete;
  ~DevPlan();
  void release();
  int build(const std::vector<GemmLaunch>& ls);
  void launch(int idx, hipStream_t st, int variant = 0) const;
  bool empty() const { return launches.empty(); }
};

// form_t = 1 (op(A) = A^T) needs gemv_scratch(g, 1) doubles of scratch; y may
// then alias an x of the job (the partial sums are complete before y is written).
size_t gemv_scratch(const GemvJob& g, int form_t);
void gemv_launch(const GemvJob& g, int form_t, double* scratch, hipStream_t st);

}  // namespace bta
