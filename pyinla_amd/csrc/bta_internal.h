// Host-side internals shared by the BTA kernel files and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <deque>
#include <map>
#include <memory>
#include <utility>
#include <vector>

#include "bta_device.h"
#include "bta_flow.h"
#include "bta_gemm.h"

namespace bta {

static constexpr unsigned long long kInfoNone = ~0ull;

// Kernel classes for the optional event timing of bta_ws_set_timing.
enum KernelClass : int {
  KC_PANEL = 0,     // k_panel_step (tile-column factorization)
  KC_UPDATE = 1,    // k_gemm_nt_sub (block updates D_{i+1}, B_{i+1}, T)
  KC_GROUPED = 2,   // k_gemm_grouped (block inverses, selected inversion)
  KC_GEMV = 3,      // GEMV kernels (triangular sweeps)
  KC_BLAS = 4,      // rocBLAS dgemm of the selected inversion (Sigma_{i+1,i} products)
  KC_COUNT = 5
};

// Event pairs recorded around launches while timing is enabled.
struct TimingLog {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[KC_COUNT];
  size_t used[KC_COUNT] = {0, 0, 0, 0, 0};
  hipEvent_t begin(int cls, hipStream_t st);
  void end(int cls, hipStream_t st);
};

// Device copy of the helper task lists of the persistent panel launch for one
// (nT, nTa, helpers, delta); variant 0 is a block with L, variant 1 the last
// block.
struct FlowSched {
  uint32_t* d_tasks = nullptr;
  int* d_off = nullptr;
  int off_base[2] = {0, 0};
  int helpers = 0, delta = 0, nTa = 0, near_div = 16, fold = 1;
  size_t bytes = 0;
};

// Exchange buffers of the persistent panel launch for one panel shape, plus the
// schedules built for it (one per helper count / group size in use).
// The exchange buffers are sized for nTa_max arrow tile rows, so one plan
// serves the BT (nTa = 0) and BTA (nTa = nTa_max) factorizations of a
// workspace; each shape has its own task lists.
struct FlowPlan {
  double* xbuf = nullptr;   // (2 nT + nTa_max) * nT tiles, then 2 nT hand-off tiles
  double* hacc = nullptr;
  unsigned long long* flags = nullptr;
  int nT = -1, nTa = 0;     // nTa: arrow tile rows the buffers hold (nTa_max)
  unsigned long long spin_ticks = 0;
  size_t buf_bytes = 0;
  std::deque<FlowSched> scheds;  // deque: pointers to entries stay valid
  FlowPlan() = default;
  FlowPlan(const FlowPlan&) = delete;
  FlowPlan& operator=(const FlowPlan&) = delete;
  ~FlowPlan();
  void release();
  static size_t bytes_for(int nT, int nTa);
  int build(int nT, int nTa);
  // schedule for (nTa, helpers, delta) with nTa <= the buffers' nTa, built on
  // first request (host; never during capture)
  int sched(int nTa, int helpers, int delta, int near_div, int fold, const FlowSched** out);
};

int flow_helpers_default(int device);
int flow_capacity(int device);  // co-resident panel workgroups on the device
FlowArgs flow_args(const FlowPlan& fp, const FlowSched& sc, double* D, double* L, double* B, double* T, double* winv,
                   double* logpart, const unsigned long long* epoch, unsigned long long* info, int* err, int i,
                   int n, int b, int a);
void flow_launch_batched(const FlowBatch& fb, int total, hipStream_t st);
unsigned long long flow_spin_ticks(int device, double seconds);

// Launch plans are cached per key (block pointers, arrow size, tile, variant)
// for as long as the workspace lives, so that a hipGraph recorded with a plan's
// job tables never outlives them.  A cache that exceeds kPlanCap entries is
// emptied (after a device synchronisation) and the workspace's graphs are
// dropped with it (graphs_stale).
struct PlanKey {
  const void* p[4];
  int a, tile, variant;
  bool operator<(const PlanKey& o) const {
    for (int q = 0; q < 4; ++q)
      if (p[q] != o.p[q]) return p[q] < o.p[q];
    if (a != o.a) return a < o.a;
    if (tile != o.tile) return tile < o.tile;
    return variant < o.variant;
  }
};
static constexpr size_t kPlanCap = 16;

// Selected-inversion plan: the grouped launches, and (variant with library
// products) the order of grouped launches (>= 0) and dgemm calls (-1 - index).
struct PotriPlan {
  struct BlasGemm {
    const double* A;  // Sigma_{i+1,i+1}
    const double* B;  // G_i
    double* C;        // Sigma_{i+1,i}
    double beta;      // 1 when the arrow term is already in C
  };
  DevPlan grouped;    // every product in the grouped kernel
  DevPlan with_blas;  // grouped launches of the variant with dgemm calls
  std::vector<BlasGemm> gemms;
  std::vector<int> order;
  bool has_blas = false;
};

// Reservation flags of a workspace (bta_ws_create / bta_ws_reserve).
enum : int { WS_GRAPH = 1, WS_SOLVE = 2, WS_SELINV = 4 };

// Per-solver device workspace.  Every buffer a hot call touches is allocated
// by bta_ws_create / bta_ws_reserve; the enqueue functions only read and
// write it.  Launch plans (small job tables) are built on the first call with
// a given set of block pointers, or ahead of time by bta_ws_prepare.
struct Workspace {
  int device = 0;
  int n = 0, b = 0, a = 0, nT = 0;
  double* winv = nullptr;             // n * nT * 64 * 64: inverse of every diagonal tile
  double* logpart = nullptr;          // n * nT + 1 partial log sums
  unsigned long long* flags = nullptr;  // n * nT readiness flags
  unsigned long long* epoch = nullptr;  // call counter (device)
  unsigned long long* info = nullptr;   // first failing pivot (global row), kInfoNone if none
  int* err = nullptr;                   // nonzero: a bounded wait timed out
  double* logdet = nullptr;             // last log-determinant
  double* work = nullptr;               // scratch for solve / selected inversion
  size_t work_doubles = 0;
  double* tinv = nullptr;               // 64 x 64: inverse of the tip factor
  unsigned long long spin_limit = 1ull << 24;
  size_t base_bytes = 0;                // bytes of the buffers above
  // Full inverses W_i = D_i^-1 of the diagonal factor blocks (n * b * b),
  // reserved with WS_SOLVE or WS_SELINV.
  double* wfull = nullptr;
  // G_i = L_i W_i (n b^2) and Ga_i = B_i W_i (n max(a,1) b) of the selected
  // inversion, reserved with WS_SELINV.
  double* gfull = nullptr;
  int reserved = 0;                // WS_SOLVE | WS_SELINV bits held
  bool w_valid = false;            // W matches the factor currently held
  const double* w_of = nullptr;    // D pointer W was computed from
  // cached launch plans
  std::map<PlanKey, std::unique_ptr<DevPlan>> trtri_plans, update_plans;
  std::map<PlanKey, std::unique_ptr<PotriPlan>> potri_plans;
  DevPlan* trtri_cur = nullptr;     // plans selected by the last prepare call
  DevPlan* update_cur = nullptr;
  PotriPlan* potri_cur = nullptr;
  bool graphs_stale = false;        // a plan cache was emptied: drop recorded graphs
  void* blas = nullptr;  // rocblas_handle, created with the first plan that uses it
  int use_blas = 1;      // 0: grouped kernel for every product
  TimingLog* tlog = nullptr;  // non-null: bracket launches with events
  // Output-tile edge per GEMM class: 0 = heuristic (pick_tile), 64 or 128.
  int tile_update = 0;   // block updates of the factorization
  int tile_grouped = 0;  // block inverses and selected inversion
  // Panel factorization: 0 = automatic, 1 = one launch per tile column,
  // 2 = persistent dataflow launch per block (bta_flow.hip).
  int panel_mode = 0;
  int flow_helpers = 0;  // 0 = automatic
  int flow_delta = 0;    // 0: list-scheduled task order; >= 1: level order, far-tile update groups
  int flow_near = 16;    // near pool = helpers / flow_near (see flow_near_pool)
  int flow_fold = 0;     // block update: 0 = automatic, 1 = in the persistent panel, 2 = a grouped GEMM after it
  int gemm_variant = 0;  // 64x64 grouped GEMM: 0 = single LDS buffer, 1 = two LDS buffers
  int batch_lanes = 0;   // problems per wave of a batched factorization (0 = 6)
  // persistent-panel exchange buffers, sized for the workspace's arrowhead
  // (allocated by bta_ws_create when the flow panel can run)
  std::unique_ptr<FlowPlan> flow;
  const FlowSched* flow_sched = nullptr;
  // bytes currently held by plans (job tables, task lists)
  size_t plan_bytes() const;
  void clear_plans();
};

// Looks up or builds (through `build`) the plan for `key` in `cache`; empties
// the cache first when it is full (synchronising the device, marking graphs
// stale).  Returns the plan or nullptr with *rc set.
template <class P, class F>
P* plan_lookup(Workspace& ws, std::map<PlanKey, std::unique_ptr<P>>& cache, const PlanKey& key, F&& build,
               int* rc) {
  auto it = cache.find(key);
  if (it != cache.end()) {
    *rc = 0;
    return it->second.get();
  }
  if (cache.size() >= kPlanCap) {
    hipDeviceSynchronize();
    cache.clear();
    ws.graphs_stale = true;
  }
  std::unique_ptr<P> pl(new P());
  *rc = build(*pl);
  if (*rc) return nullptr;
  P* raw = pl.get();
  cache.emplace(key, std::move(pl));
  return raw;
}

inline int choose_tile(int pref, int m, int n) { return pref ? pref : pick_tile(m, n); }

// RAII bracket: records a start event now and the stop event at scope exit.
struct TimedScope {
  Workspace& ws;
  int cls;
  hipStream_t st;
  TimedScope(Workspace& w, int c, hipStream_t s) : ws(w), cls(c), st(s) {
    if (ws.tlog) ws.tlog->begin(cls, st);
  }
  ~TimedScope() {
    if (ws.tlog) ws.tlog->end(cls, st);
  }
};

// One problem of a batched factorization (pointers into its own buffers and
// workspace).
struct ProbDesc {
  double *D, *L, *B, *T;
  double* winv;
  double* logpart;
  unsigned long long* flags;
  unsigned long long* epoch;
  unsigned long long* info;
  int* err;
  double* logdet;
  double* tinv;
  unsigned long long spin_limit;
  int n, b, a, nT, nTa;
  FlowPlan* fplan;  // the workspace's exchange buffers (nullptr: the batch builds its own)
};

// Launch sequence of a batched factorization: for every block step i and
// tile column k, one panel launch over all problems that have it, and after
// the last column of step i one grouped update launch.
struct BatchPlan {
  struct Launch { int i, k, nent, total, off_prob, off_begin, update; };
  struct FLaunch { FlowBatch fb; int total; };
  std::vector<ProbDesc> probs;
  std::vector<Launch> launches;
  ProbDesc* d_probs = nullptr;
  int* d_ent = nullptr;
  DevPlan updates;
  // persistent dataflow form: one launch per block step over all problems
  bool flow = false;
  std::vector<FlowPlan*> fplans;                   // per problem (borrowed or owned)
  std::vector<std::unique_ptr<FlowPlan>> owned;    // built for problems without one
  std::vector<FLaunch> flaunches;
  BatchPlan() = default;
  BatchPlan(const BatchPlan&) = delete;
  BatchPlan& operator=(const BatchPlan&) = delete;
  ~BatchPlan();
  // lanes: problems factored concurrently in the persistent form (0 = 6,
  // at most FLOW_BATCH_MAX)
  int build(const std::vector<ProbDesc>& ps, int panel_mode, int lanes, int delta, int near_div, int device);
  int enqueue(hipStream_t st) const;
};

// Exchange buffers of the persistent panel for a workspace of block size b and
// arrowhead a (allocated by bta_ws_create; nullptr when the flow panel cannot
// run for this shape).
bool flow_wanted(const Workspace& ws, int b);
int flow_alloc(Workspace& ws);
// Builds the grouped block-update plan when the blocks are large enough for
// 128x128 tiles (else the by-value update kernel is used) and selects the
// flow schedule; host only.
int potrf_prepare(Workspace& ws, double* D, double* L, double* B, double* T, int n, int b, int a);
// One persistent launch: panel of block i plus its updates of D_{i+1},
// B_{i+1} and T.
void flow_launch_block(const Workspace& ws, const FlowPlan& fp, const FlowSched& sc, double* D, double* L,
                       double* B, double* T, int i, int n, int b, int a, hipStream_t st);
int potrf_enqueue(Workspace& ws, double* D, double* L, double* B, double* T, int n, int b, int a,
                  hipStream_t st);
// Host-side preparation (allocations, job tables); never called during capture.
int trtri_prepare(Workspace& ws, const double* D, int n, int b);
int trtri_enqueue(Workspace& ws, const double* D, int n, int b, hipStream_t st);
int potri_prepare(Workspace& ws, double* D, double* L, double* B, double* T, int n, int b, int a);
int potri_enqueue(Workspace& ws, double* D, double* L, double* B, double* T, int n, int b, int a,
                  hipStream_t st);
int potrs_enqueue(Workspace& ws, const double* D, const double* L, const double* B, const double* T,
                  int n, int b, int a, double* rhs, int nrhs, int ldrhs, int trans, hipStream_t st);
// 2 sum log diag(D_i) (+ 2 sum log diag(T) if a > 0) of the blocks as they are
// held, in a fixed summation order, written to out (device).
int logdet_enqueue(const double* D, const double* T, int n, int b, int a, double* part, double* out,
                   hipStream_t st);

// Writes the 64 x 4 accumulator outputs of one v_mfma_f64_16x16x4f64 with
// A[i][k] = i + 100 k and B[k][j] = j + 1000 k (lane-major) to out[256].
int selftest_mfma_enqueue(double* out, hipStream_t st);

}  // namespace bta
