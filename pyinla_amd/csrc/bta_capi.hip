// extern "C" entry points of libpyinla_bta.so (declared in include/pyinla_bta.h).
#include <rocblas/rocblas.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "bta_internal.h"
#include "../../include/pyinla_bta.h"

namespace {

struct GraphKey {
  int op;
  std::vector<uintptr_t> ptrs;
  std::vector<long long> ints;
  bool operator<(const GraphKey& o) const { return std::tie(op, ptrs, ints) < std::tie(o.op, o.ptrs, o.ints); }
};

}  // namespace

struct bta_ws {
  bta::Workspace w;
  int use_graph = 0;
  int saved_use_graph = 0;
  hipStream_t gstream = nullptr;  // capture stream
  std::map<GraphKey, hipGraphExec_t> graphs;
  std::mutex mu;
};

namespace bta {

hipEvent_t TimingLog::begin(int cls, hipStream_t st) {
  auto& v = ev[cls];
  if (used[cls] == v.size()) {
    hipEvent_t a, b;
    hipEventCreate(&a);
    hipEventCreate(&b);
    v.emplace_back(a, b);
  }
  hipEvent_t e = v[used[cls]].first;
  hipEventRecord(e, st);
  return e;
}

void TimingLog::end(int cls, hipStream_t st) {
  hipEventRecord(ev[cls][used[cls]].second, st);
  used[cls] += 1;
}

}  // namespace bta

namespace {

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

size_t ws_doubles(int n, int b, int a, size_t* work_out) {
  const int nT = (b + 63) / 64;
  // scratch for solve / selected inversion: 3 b x b blocks + arrow temporaries
  const size_t work = 3 * (size_t)b * b + 4 * (size_t)a * b + 2 * (size_t)a * a + 64;
  if (work_out) *work_out = work;
  return (size_t)n * nT * 4096 + ((size_t)n * nT + 1) + 1 + 4096 + work;
}

// Run `enqueue` either directly on `st` or through a cached hipGraph.
void drop_graphs(bta_ws* ws) {
  if (!ws->graphs.empty()) {
    hipDeviceSynchronize();  // a recorded graph may still be running
    for (auto& kv : ws->graphs) hipGraphExecDestroy(kv.second);
    ws->graphs.clear();
  }
  ws->w.graphs_stale = false;
}

template <class F>
int run_maybe_graph(bta_ws* ws, const GraphKey& key, hipStream_t st, F&& enqueue) {
  if (ws->w.graphs_stale) drop_graphs(ws);  // a plan some graph recorded was freed
  if (!ws->use_graph) return enqueue(st);
  std::lock_guard<std::mutex> g(ws->mu);
  hipError_t e;
  auto it = ws->graphs.find(key);
  if (it == ws->graphs.end()) {
    hipGraph_t graph;
    e = hipStreamBeginCapture(ws->gstream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return (int)e;
    const int rc = enqueue(ws->gstream);
    e = hipStreamEndCapture(ws->gstream, &graph);
    if (rc != 0) return rc;
    if (e != hipSuccess) return (int)e;
    hipGraphExec_t exec;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return (int)e;
    it = ws->graphs.emplace(key, exec).first;
  }
  // Replayed on the caller's stream: solvers driven from different streams
  // then run concurrently (a private replay stream per workspace can share a
  // hardware queue with another workspace's, which serializes them).
  if ((e = hipGraphLaunch(it->second, st)) != hipSuccess) return (int)e;
  return 0;
}

int check_shape(bta_ws* ws, int n, int b, int a) {
  if (!ws) return BTA_E_ARG;
  // a == 0 (BT) is accepted on a BTA workspace: the arrow is then ignored.
  if (n != ws->w.n || b != ws->w.b || (a != ws->w.a && a != 0)) return BTA_E_SHAPE;
  return 0;
}

}  // namespace

extern "C" {

int bta_version(void) { return 200; }

static size_t w_bytes(int n, int b) { return (size_t)n * b * b * sizeof(double); }
static size_t g_bytes(int n, int b, int a) {
  return ((size_t)n * b * b + (size_t)n * std::max(a, 1) * b) * sizeof(double);
}

size_t bta_ws_bytes(int n, int b, int a, int flags) {
  if (n < 1 || b < 1 || a < 0) return 0;
  size_t bytes = 8 * ws_doubles(n, b, a, nullptr) + ((size_t)n * ((b + 63) / 64) + 2) * 8 + 64;
  if (b > 64) bytes += bta::FlowPlan::bytes_for((b + 63) / 64, (a + 63) / 64);
  if (flags & (BTA_WS_SOLVE | BTA_WS_SELINV)) bytes += w_bytes(n, b);
  if (flags & BTA_WS_SELINV) bytes += g_bytes(n, b, a);
  return bytes;
}

int bta_ws_reserve(bta_ws_t ws, int flags) {
  if (!ws || (flags & ~(BTA_WS_SOLVE | BTA_WS_SELINV))) return BTA_E_ARG;
  bta::Workspace& w = ws->w;
  hipError_t e = hipSetDevice(w.device);
  if (e != hipSuccess) return (int)e;
  if ((flags & (BTA_WS_SOLVE | BTA_WS_SELINV)) && !w.wfull) {
    if ((e = hipMalloc(&w.wfull, w_bytes(w.n, w.b))) != hipSuccess) return (int)e;
    w.w_valid = false;
  }
  if ((flags & BTA_WS_SELINV) && !w.gfull) {
    if ((e = hipMalloc(&w.gfull, g_bytes(w.n, w.b, w.a))) != hipSuccess) return (int)e;
  }
  w.reserved |= flags;
  return 0;
}

size_t bta_ws_allocated(bta_ws_t ws) {
  if (!ws) return 0;
  const bta::Workspace& w = ws->w;
  size_t bytes = w.base_bytes;
  if (w.wfull) bytes += w_bytes(w.n, w.b);
  if (w.gfull) bytes += g_bytes(w.n, w.b, w.a);
  return bytes + w.plan_bytes();
}

int bta_ws_create(int device, int n, int b, int a, int flags, bta_ws_t* out) {
  if (!out || n < 1 || b < 1 || a < 0 || a > 64) return BTA_E_ARG;
  if (flags & ~(BTA_WS_GRAPH | BTA_WS_SOLVE | BTA_WS_SELINV)) return BTA_E_ARG;
  *out = nullptr;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return (int)e;
  bta_ws* ws = new bta_ws();
  bta::Workspace& w = ws->w;
  w.device = device;
  w.n = n;
  w.b = b;
  w.a = a;
  w.nT = (b + 63) / 64;
  size_t work = 0;
  const size_t nd = ws_doubles(n, b, a, &work);
  double* base = nullptr;
  if ((e = hipMalloc(&base, nd * sizeof(double))) != hipSuccess) { delete ws; return (int)e; }
  unsigned long long* u = nullptr;
  const size_t nflags = (size_t)n * w.nT;
  if ((e = hipMalloc(&u, (nflags + 2) * sizeof(unsigned long long) + 64)) != hipSuccess) {
    hipFree(base);
    delete ws;
    return (int)e;
  }
  hipMemset(u, 0, (nflags + 2) * sizeof(unsigned long long) + 64);
  hipMemset(base, 0, nd * sizeof(double));
  w.winv = base;
  w.logpart = w.winv + (size_t)n * w.nT * 4096;
  w.logdet = w.logpart + (size_t)n * w.nT + 1;
  w.tinv = w.logdet + 1;
  w.work = w.tinv + 4096;
  w.work_doubles = work;
  w.flags = u;
  w.epoch = u + nflags;
  w.info = u + nflags + 1;
  w.err = reinterpret_cast<int*>(u + nflags + 2);
  w.base_bytes = nd * sizeof(double) + (nflags + 2) * sizeof(unsigned long long) + 64;
  int rc = 0;
  if (bta::flow_wanted(w, b)) rc = bta::flow_alloc(w);
  if (!rc && (flags & (BTA_WS_SOLVE | BTA_WS_SELINV))) rc = bta_ws_reserve(ws, flags & (BTA_WS_SOLVE | BTA_WS_SELINV));
  if (rc) {
    bta_ws_destroy(ws);
    return rc;
  }
  ws->use_graph = (flags & BTA_WS_GRAPH) ? 1 : 0;
  if (ws->use_graph) {
    hipStreamCreateWithFlags(&ws->gstream, hipStreamNonBlocking);
  }
  if ((e = hipDeviceSynchronize()) != hipSuccess) {
    bta_ws_destroy(ws);
    return (int)e;
  }
  *out = ws;
  return 0;
}

int bta_ws_destroy(bta_ws_t ws) {
  if (!ws) return 0;
  hipSetDevice(ws->w.device);
  hipDeviceSynchronize();
  if (ws->w.tlog) bta_ws_set_timing(ws, 0);
  for (auto& kv : ws->graphs) hipGraphExecDestroy(kv.second);
  if (ws->gstream) hipStreamDestroy(ws->gstream);
  ws->w.clear_plans();
  ws->w.flow.reset();
  if (ws->w.blas) rocblas_destroy_handle(static_cast<rocblas_handle>(ws->w.blas));
  if (ws->w.winv) hipFree(ws->w.winv);
  if (ws->w.gfull) hipFree(ws->w.gfull);
  if (ws->w.flags) hipFree(ws->w.flags);
  if (ws->w.wfull) hipFree(ws->w.wfull);
  delete ws;
  return 0;
}

int bta_potrf(bta_ws_t ws, double* D, double* L, double* B, double* T, int n, int b, int a, void* stream) {
  int rc = check_shape(ws, n, b, a);
  if (rc) return rc;
  if (!D || (n > 1 && !L) || (a > 0 && (!B || !T))) return BTA_E_ARG;
  GraphKey key{1, {(uintptr_t)D, (uintptr_t)L, (uintptr_t)B, (uintptr_t)T}, {n, b, a}};
  ws->w.w_valid = false;
  rc = bta::potrf_prepare(ws->w, D, L, B, T, n, b, a);
  if (rc) return rc;
  return run_maybe_graph(ws, key, as_stream(stream), [&](hipStream_t st) {
    return bta::potrf_enqueue(ws->w, D, L, B, T, n, b, a, st);
  });
}

// Make the block inverses W_i of the current factor available (enqueued on st).
static int ensure_inverses(bta_ws* ws, const double* D, int n, int b, hipStream_t st) {
  if (ws->w.w_valid && ws->w.w_of == D) return 0;
  int rc = bta::trtri_prepare(ws->w, D, n, b);
  if (rc) return rc;
  GraphKey key{3, {(uintptr_t)D}, {n, b}};
  rc = run_maybe_graph(ws, key, st, [&](hipStream_t s2) { return bta::trtri_enqueue(ws->w, D, n, b, s2); });
  if (rc) return rc;
  ws->w.w_valid = true;
  ws->w.w_of = D;
  return 0;
}

int bta_potrs(bta_ws_t ws, double* D, double* L, double* B, double* T, int n, int b, int a, double* rhs,
              int nrhs, int ldrhs, int trans, void* stream) {
  int rc = check_shape(ws, n, b, a);
  if (rc) return rc;
  if (!D || (n > 1 && !L) || (a > 0 && (!B || !T)) || !rhs || nrhs < 1 || ldrhs < nrhs) return BTA_E_ARG;
  if (trans != 0 && trans != 1) return BTA_E_ARG;
  hipStream_t st = as_stream(stream);
  rc = ensure_inverses(ws, D, n, b, st);
  if (rc) return rc;
  // the sweeps are a few GEMV launches per block: launched directly (a graph
  // per right-hand-side buffer would grow without bound)
  std::lock_guard<std::mutex> g(ws->mu);
  return bta::potrs_enqueue(ws->w, D, L, B, T, n, b, a, rhs, nrhs, ldrhs, trans, st);
}

int bta_logdet(bta_ws_t ws, const double* D, const double* T, int n, int b, int a, double* out, void* stream) {
  int rc = check_shape(ws, n, b, a);
  if (rc) return rc;
  if (!D || !out || (a > 0 && !T)) return BTA_E_ARG;
  std::lock_guard<std::mutex> g(ws->mu);
  return bta::logdet_enqueue(D, T, n, b, a, ws->w.logpart, out, as_stream(stream));
}

int bta_ws_prepare(bta_ws_t ws, double* D, double* L, double* B, double* T, int n, int b, int a, int ops) {
  int rc = check_shape(ws, n, b, a);
  if (rc) return rc;
  if (!D || (n > 1 && !L) || (a > 0 && (!B || !T)) || (ops & ~7)) return BTA_E_ARG;
  std::lock_guard<std::mutex> g(ws->mu);
  if (ops & 1) {
    rc = bta::potrf_prepare(ws->w, D, L, B, T, n, b, a);
    if (rc) return rc;
  }
  if (ops & 2) {
    rc = bta::trtri_prepare(ws->w, D, n, b);
    if (rc) return rc;
  }
  if (ops & 4) {
    rc = bta::potri_prepare(ws->w, D, L, B, T, n, b, a);
    if (rc) return rc;
  }
  return 0;
}

int bta_block_inverses(bta_ws_t ws, double* D, int n, int b, void* stream) {
  int rc = check_shape(ws, n, b, 0);
  if (rc) return rc;
  if (!D) return BTA_E_ARG;
  return ensure_inverses(ws, D, n, b, as_stream(stream));
}

int bta_potri(bta_ws_t ws, double* D, double* L, double* B, double* T, int n, int b, int a, void* stream) {
  int rc = check_shape(ws, n, b, a);
  if (rc) return rc;
  if (!D || (n > 1 && !L) || (a > 0 && (!B || !T))) return BTA_E_ARG;
  hipStream_t st = as_stream(stream);
  rc = ensure_inverses(ws, D, n, b, st);
  if (rc) return rc;
  rc = bta::potri_prepare(ws->w, D, L, B, T, n, b, a);
  if (rc) return rc;
  GraphKey key{4, {(uintptr_t)D, (uintptr_t)L, (uintptr_t)B, (uintptr_t)T}, {n, b, a}};
  rc = run_maybe_graph(ws, key, st, [&](hipStream_t s2) {
    return bta::potri_enqueue(ws->w, D, L, B, T, n, b, a, s2);
  });
  // The factor has been overwritten by the selected inverse.
  ws->w.w_valid = false;
  return rc;
}

}  // extern "C"

struct bta_batch {
  bta::BatchPlan plan;
  std::vector<bta_ws*> ws;
  int device = 0;
  int use_graph = 0;
  hipStream_t gstream = nullptr;  // capture stream
  hipGraphExec_t exec = nullptr;
};

extern "C" {

int bta_batch_create(int nprob, bta_ws_t* ws, double** D, double** L, double** B, double** T, const int* a,
                     int use_graph, bta_batch_t* out) {
  if (!out || nprob < 1 || !ws || !D || !L || !a) return BTA_E_ARG;
  *out = nullptr;
  std::vector<bta::ProbDesc> ps;
  for (int q = 0; q < nprob; ++q) {
    bta_ws* w = ws[q];
    if (!w || !D[q]) return BTA_E_ARG;
    if (a[q] != 0 && a[q] != w->w.a) return BTA_E_SHAPE;
    if (w->w.n > 1 && !L[q]) return BTA_E_ARG;
    if (a[q] > 0 && (!B || !T || !B[q] || !T[q])) return BTA_E_ARG;
    if (w->w.device != ws[0]->w.device) return BTA_E_ARG;
    for (int r = 0; r < q; ++r)
      if (ws[r] == w) return BTA_E_ARG;  // one workspace per problem
    bta::ProbDesc d;
    d.D = D[q];
    d.L = L[q];
    d.B = a[q] > 0 ? B[q] : nullptr;
    d.T = a[q] > 0 ? T[q] : nullptr;
    d.winv = w->w.winv;
    d.logpart = w->w.logpart;
    d.flags = w->w.flags;
    d.epoch = w->w.epoch;
    d.info = w->w.info;
    d.err = w->w.err;
    d.logdet = w->w.logdet;
    d.tinv = w->w.tinv;
    d.spin_limit = w->w.spin_limit;
    d.n = w->w.n;
    d.b = w->w.b;
    d.a = a[q];
    d.nT = w->w.nT;
    d.nTa = (a[q] + 63) / 64;
    d.fplan = w->w.flow.get();
    ps.push_back(d);
  }
  hipError_t e = hipSetDevice(ws[0]->w.device);
  if (e != hipSuccess) return (int)e;
  bta_batch* bt = new bta_batch();
  bt->device = ws[0]->w.device;
  bt->ws.assign(ws, ws + nprob);
  const int rc = bt->plan.build(ps, ws[0]->w.panel_mode, ws[0]->w.batch_lanes, ws[0]->w.flow_delta,
                                ws[0]->w.flow_near, ws[0]->w.device);
  if (rc) {
    delete bt;
    return rc;
  }
  bt->use_graph = use_graph;
  if (use_graph) {
    hipStreamCreateWithFlags(&bt->gstream, hipStreamNonBlocking);
  }
  *out = bt;
  return 0;
}

int bta_batch_potrf(bta_batch_t bt, void* stream) {
  if (!bt) return BTA_E_ARG;
  for (bta_ws* w : bt->ws) w->w.w_valid = false;
  hipStream_t st = as_stream(stream);
  if (!bt->use_graph) return bt->plan.enqueue(st);
  hipError_t e;
  if (!bt->exec) {
    hipGraph_t graph;
    if ((e = hipStreamBeginCapture(bt->gstream, hipStreamCaptureModeThreadLocal)) != hipSuccess) return (int)e;
    const int rc = bt->plan.enqueue(bt->gstream);
    e = hipStreamEndCapture(bt->gstream, &graph);
    if (rc) return rc;
    if (e != hipSuccess) return (int)e;
    e = hipGraphInstantiate(&bt->exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return (int)e;
  }
  if ((e = hipGraphLaunch(bt->exec, st)) != hipSuccess) return (int)e;
  return 0;
}

int bta_batch_destroy(bta_batch_t bt) {
  if (!bt) return 0;
  hipSetDevice(bt->device);
  hipDeviceSynchronize();
  if (bt->exec) hipGraphExecDestroy(bt->exec);
  if (bt->gstream) hipStreamDestroy(bt->gstream);
  delete bt;
  return 0;
}

int bta_ws_status(bta_ws_t ws, int64_t* info, double* logdet, int* err, void* stream) {
  if (!ws) return BTA_E_ARG;
  hipStream_t st = as_stream(stream);
  hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) return (int)e;
  unsigned long long inf = 0;
  int er = 0;
  double ld = 0.0;
  if ((e = hipMemcpy(&inf, ws->w.info, sizeof(inf), hipMemcpyDeviceToHost)) != hipSuccess) return (int)e;
  if ((e = hipMemcpy(&er, ws->w.err, sizeof(er), hipMemcpyDeviceToHost)) != hipSuccess) return (int)e;
  if ((e = hipMemcpy(&ld, ws->w.logdet, sizeof(ld), hipMemcpyDeviceToHost)) != hipSuccess) return (int)e;
  if (info) *info = (inf == bta::kInfoNone) ? 0 : (int64_t)inf + 1;
  if (logdet) *logdet = ld;
  if (err) *err = er ? BTA_E_TIMEOUT : 0;
  return 0;
}

double* bta_ws_logdet_ptr(bta_ws_t ws) { return ws ? ws->w.logdet : nullptr; }

int bta_flow_schedule(int nT, int nTL, int nTa, int helpers, int delta, int near_div, uint32_t* tasks, int64_t cap,
                      int* off, int64_t* ntasks) {
  if (nT < 1 || nT > 1023 || nTL < 0 || nTa < 0 || nTa > 1 || helpers < 1 || helpers > 1023 || delta < 0 ||
      2 * nT + nTa > 1023 || near_div < 0 || near_div > 64)
    return BTA_E_ARG;
  bta::FlowSchedule sc;
  bta::FlowShape shape{nT, nTL, nTa};
  if (near_div > 0) shape.near_div = near_div;
  if (!bta::flow_build_schedule(shape, helpers, delta, sc)) return BTA_E_ARG;
  if (ntasks) *ntasks = (int64_t)sc.tasks.size();
  if (!tasks && !off) return 0;
  if (!tasks || !off || cap < (int64_t)sc.tasks.size()) return BTA_E_ARG;
  for (size_t q = 0; q < sc.tasks.size(); ++q) tasks[q] = sc.tasks[q];
  for (int h = 0; h <= helpers; ++h) off[h] = sc.off[h];
  return 0;
}

int bta_ws_set_option(bta_ws_t ws, int option, int value) {
  if (!ws) return BTA_E_ARG;
  const bool tile_opt = (option == BTA_OPT_TILE_UPDATE || option == BTA_OPT_TILE_GROUPED);
  if (tile_opt && value != 0 && value != 64 && value != 128) return BTA_E_ARG;
  if (option == BTA_OPT_PANEL && (value < 0 || value > 2)) return BTA_E_ARG;
  if (option == BTA_OPT_FLOW_HELPERS && (value < 0 || value > 1023)) return BTA_E_ARG;
  if (option == BTA_OPT_FLOW_DELTA && (value < 0 || value > 64)) return BTA_E_ARG;
  if (option == BTA_OPT_BATCH_LANES && (value < 0 || value > 512)) return BTA_E_ARG;
  if (option == BTA_OPT_BLAS && value != 0 && value != 1) return BTA_E_ARG;
  if (option == BTA_OPT_FLOW_NEAR && (value < 1 || value > 64)) return BTA_E_ARG;
  if (option == BTA_OPT_GEMM_KERNEL && value != 0 && value != 1) return BTA_E_ARG;
  if (option == BTA_OPT_FLOW_FOLD && (value < 0 || value > 2)) return BTA_E_ARG;
  std::lock_guard<std::mutex> g(ws->mu);
  if (option == BTA_OPT_TILE_UPDATE) ws->w.tile_update = value;
  else if (option == BTA_OPT_TILE_GROUPED) ws->w.tile_grouped = value;
  else if (option == BTA_OPT_PANEL) {
    ws->w.panel_mode = value;
    if (bta::flow_wanted(ws->w, ws->w.b)) {
      const int rc = bta::flow_alloc(ws->w);
      if (rc) return rc;
    }
  }
  else if (option == BTA_OPT_FLOW_HELPERS) ws->w.flow_helpers = value;
  else if (option == BTA_OPT_FLOW_DELTA) ws->w.flow_delta = value;
  else if (option == BTA_OPT_BATCH_LANES) ws->w.batch_lanes = value;
  else if (option == BTA_OPT_BLAS) ws->w.use_blas = value;  // part of the plan key
  else if (option == BTA_OPT_FLOW_NEAR) ws->w.flow_near = value;
  else if (option == BTA_OPT_GEMM_KERNEL) ws->w.gemm_variant = value;
  else if (option == BTA_OPT_FLOW_FOLD) ws->w.flow_fold = value;
  else return BTA_E_ARG;
  // recorded graphs embed the old launch shapes
  drop_graphs(ws);
  return 0;
}

int bta_ws_set_timing(bta_ws_t ws, int on) {
  if (!ws) return BTA_E_ARG;
  if (on) {
    if (!ws->w.tlog) ws->w.tlog = new bta::TimingLog();
    for (int c = 0; c < bta::KC_COUNT; ++c) ws->w.tlog->used[c] = 0;
    ws->saved_use_graph = ws->use_graph;
    ws->use_graph = 0;  // events bracket individual launches, so no capture
  } else if (ws->w.tlog) {
    for (int c = 0; c < bta::KC_COUNT; ++c)
      for (auto& pr : ws->w.tlog->ev[c]) {
        hipEventDestroy(pr.first);
        hipEventDestroy(pr.second);
      }
    delete ws->w.tlog;
    ws->w.tlog = nullptr;
    ws->use_graph = ws->saved_use_graph;
  }
  return 0;
}

int bta_ws_get_timing(bta_ws_t ws, int kclass, double* total_ms, int64_t* count) {
  if (!ws || kclass < 0 || kclass >= bta::KC_COUNT || !ws->w.tlog) return BTA_E_ARG;
  bta::TimingLog* t = ws->w.tlog;
  double tot = 0.0;
  for (size_t q = 0; q < t->used[kclass]; ++q) {
    hipError_t e = hipEventSynchronize(t->ev[kclass][q].second);
    if (e != hipSuccess) return (int)e;
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, t->ev[kclass][q].first, t->ev[kclass][q].second);
    if (e != hipSuccess) return (int)e;
    tot += ms;
  }
  if (total_ms) *total_ms = tot;
  if (count) *count = (int64_t)t->used[kclass];
  return 0;
}

int bta_ws_reset_timing(bta_ws_t ws) {
  if (!ws || !ws->w.tlog) return BTA_E_ARG;
  hipDeviceSynchronize();
  for (int c = 0; c < bta::KC_COUNT; ++c) ws->w.tlog->used[c] = 0;
  return 0;
}

int bta_gemm(const double* A, int lda, int ta, const double* B, int ldb, int tb, double* C, int ldc, int m, int n,
             int k, double alpha, double beta, int variant, void* stream) {
  if (!A || !B || !C || m < 0 || n < 0 || k < 0 || (ta != 0 && ta != 1) || (tb != 0 && tb != 1)) return BTA_E_ARG;
  if (variant < 0 || variant > 2) return BTA_E_ARG;
  bta::GemmLaunch l(variant == 2 ? 128 : 64);
  bta::GemmJob j = bta::job(C, m, n, ldc, beta);
  bta::add_seg(j, bta::seg(A, lda, ta, B, ldb, tb, k, alpha));
  l.add(j);
  bta::DevPlan plan;
  std::vector<bta::GemmLaunch> ls;
  ls.push_back(std::move(l));
  int rc = plan.build(ls);
  if (rc) return rc;
  hipStream_t st = as_stream(stream);
  plan.launch(0, st, variant == 1 ? 1 : 0);
  hipError_t e = hipStreamSynchronize(st);  // the plan is freed on return
  return e == hipSuccess ? (int)hipGetLastError() : (int)e;
}

int bta_selftest_mfma(double* out, void* stream) {
  if (!out) return BTA_E_ARG;
  return bta::selftest_mfma_enqueue(out, as_stream(stream));
}

}  // extern "C"
