// Persistent dataflow panel factorization (see bta_flow.h for the scheme).
//
// Replaces the nT launches per block of the column-by-column panel with one
// launch of 1 + P resident workgroups: the chain (workgroup 0) runs the
// diagonal Cholesky + inverse, TRSM of the sub-diagonal tile and the rank-64
// update of the next diagonal tile back to back, while P helpers work through
// their task lists.  Same algorithm and outputs as the per-column path
// (the published pobtaf block loop: POTRF of D_i, TRSM of L_i and B_i, then the
// updates of D_{i+1}, B_{i+1} and T; called by the reference at
// structured_solver.py:83-88): L_ii in D_i,
// L_{i+1,i} in L_i, L_{a,i} in B_i, inverse of every diagonal tile in winv and
// per-tile log-pivot sums in logpart.
#include <algorithm>
#include <queue>

#include "bta_internal.h"
#include "bta_flow.h"
#include "bta_tile_chol.h"

namespace bta {

// ---------------------------------------------------------------- schedule

// Ownership.  The tiles on the chain's path -- (c, c), (c+1, c) and (c+2, c)
// of D, whose updates and solves the chain waits for one or two columns later
// -- are dealt round-robin over a small pool of "near" helpers that own
// nothing else, so none of their tasks ever queues behind bulk work.  All
// other tiles are dealt round-robin in column-major order (column c: D rows
// c+3..nT-1, then the L and arrow rows) over the remaining helpers, so each of
// them owns tiles spread evenly over the columns and its share of every
// column's updates stays balanced; the next-block updates go to the same
// pool.  Cleared upper tiles of D go by row-major index.
// Pool size: helpers / near_div, but at least min(10, helpers / 8).  Measured
// on MI355X (factorization pair on two streams, 255 helpers each): 31 near
// helpers 1.42 ms per block pair at b = 2000 and 15.8 ms at b = 4500; 15-21
// helpers 1.34 / 15.3; 10 helpers 1.33 / 15.0; 7 helpers 1.65 (too few: the
// chain waits on them).  The floor keeps about 10 for the smaller helper
// counts of batched launches.
static int flow_near_pool(const FlowShape& s, int helpers) {
  const int div = s.near_div > 0 ? s.near_div : 16;
  if (helpers < 8) return 0;
  return std::min(3 * s.nT, std::max(helpers / div, std::min(10, helpers / 8)));
}

static bool flow_is_near(int nT, int rho, int c) { return rho < nT && rho >= c && rho - c <= 2; }

int flow_owner(const FlowShape& s, int rho, int c, int helpers) {
  const int nT = s.nT, below = s.nTL + s.nTa;
  if (rho < nT && rho < c) return (int)(((long long)rho * nT + c) % helpers);
  const int hn = flow_near_pool(s, helpers);
  if (hn > 0 && flow_is_near(nT, rho, c)) return (3 * c + (rho - c)) % hn;
  long long t = 0;  // column-major index among the other tiles
  for (int cc = 0; cc < c; ++cc) {
    const int nd = nT - cc;
    t += (hn > 0 ? nd - std::min(3, nd) : nd) + below;
  }
  if (rho < nT) t += hn > 0 ? rho - c - 3 : rho - c;
  else t += (hn > 0 ? std::max(0, nT - c - 3) : nT - c) + (rho - nT);
  return hn + (int)(t % (helpers - hn));
}

// Owner of a next-block update target (never a near helper).
static int flow_next_owner(const FlowShape& s, int rho, int c, int helpers) {
  const int hn = flow_near_pool(s, helpers);
  const long long t = (long long)rho * (s.nT + 1) + c;
  return hn + (int)(t % (helpers - hn));
}

// Level order (delta >= 1): each helper runs its tasks by dependency level;
// far tiles take their updates in groups of `delta` columns.
static void flow_build_levels(const FlowShape& s, int helpers, int delta, FlowSchedule& out) {
  struct Item {
    unsigned long long key;
    uint32_t word;
    int h;
  };
  std::vector<Item> v;
  const int nT = s.nT, R = s.rows();
  const int hn = flow_near_pool(s, helpers);  // near tiles are never grouped
  auto push = [&](int level, int prio, int op, int rho, int c, int k) {
    const unsigned long long key = ((unsigned long long)level << 42) | ((unsigned long long)prio << 31) |
                                   ((unsigned long long)c << 21) | ((unsigned long long)rho << 10) |
                                   (unsigned long long)k;
    v.push_back({key, flow_encode(op, rho, c, k), flow_owner(s, rho, c, helpers)});
  };
  // next-block updates: one task per target tile and span of S columns of
  // L_i / B_i, at the level of the span's last column, lowest priority
  auto push_next = [&](int rho, int c) {
    for (int k = 0; k < nT; k = flow_next_end(k, nT)) {
      const int klast = flow_next_end(k, nT) - 1;
      const unsigned long long key = ((unsigned long long)(3 * klast + 2) << 42) | (1023ull << 31) |
                                     ((unsigned long long)c << 21) | ((unsigned long long)rho << 10) |
                                     (unsigned long long)k;
      v.push_back({key, flow_encode(FLOW_NEXT, rho, c, k), flow_next_owner(s, rho, c, helpers)});
    }
  };
  if (s.fold && s.nTL > 0) {
    for (int r = 0; r < nT; ++r)
      for (int c = 0; c <= r; ++c) push_next(r, c);
    if (s.nTa > 0)
      for (int c = 0; c < nT; ++c) push_next(nT, c);
  }
  if (s.fold && s.nTa > 0) push_next(nT + 1, 0);
  for (int rho = 0; rho < R; ++rho) {
    for (int c = 0; c < nT; ++c) {
      if (rho < nT && rho < c) {  // strict upper tile of D: cleared
        push(0, 0, FLOW_ZERO, rho, c, 0);
        continue;
      }
      const bool diag = (rho == c);
      const bool hand = flow_is_handoff(nT, rho, c);
      const int cap = diag ? c - 2 : c - 1;
      for (int j = 0; j < c; ++j) {
        if (diag && j == c - 1) continue;  // applied by the chain
        // far tiles take their updates in groups of `delta` columns at the
        // group's last column: consecutive in the owner's list, so the
        // accumulator stays in registers across the group
        const int grp = (hn > 0 && flow_is_near(nT, rho, c)) ? 1 : std::max(delta, 1);
        const int round = hand ? j : std::min((j / grp + 1) * grp - 1, cap);
        const int prio = hand ? 0 : 1 + (c - round);
        push(3 * round + 2, prio, FLOW_UPD, rho, c, j);
      }
      if (!diag && !(rho < nT && rho == c + 1)) push(3 * c + 1, rho - c, FLOW_TRSM, rho, c, c);
    }
  }
  std::stable_sort(v.begin(), v.end(), [](const Item& x, const Item& y) {
    return x.h != y.h ? x.h < y.h : x.key < y.key;
  });
  out.helpers = helpers;
  out.tasks.resize(v.size());
  out.off.assign(helpers + 1, 0);
  for (size_t q = 0; q < v.size(); ++q) {
    out.tasks[q] = v[q].word;
    out.off[v[q].h + 1] += 1;
  }
  for (int h = 0; h < helpers; ++h) out.off[h + 1] += out.off[h];
}

// List order (delta == 0).  The launch is simulated on the host with the task
// costs below (core clock cycles, from s_memtime stamps on MI355X): whenever a
// helper is idle it starts the ready task of highest upward rank (its cost
// plus the longest path of costs and flag latencies from it to the end of the
// launch), so the solves and updates the chain is about to wait for come
// before bulk updates of tiles needed much later and the next-block updates
// fill the gaps.  Each helper's list is the order of its tasks in that
// simulation.  The simulated start times order all tasks, the chain's
// included, consistently with every dependency and with every list, so with
// all workgroups resident the earliest unfinished task in that order can
// always run, whatever the real durations: no interleaving can deadlock.
#ifndef FLOW_COST_CHOL
#define FLOW_COST_CHOL 30000  // chain: factor + inverse of a diagonal tile, stores, publish
#endif
#ifndef FLOW_COST_CTRSM
#define FLOW_COST_CTRSM 10000  // chain: solve of the sub-diagonal tile, stores
#endif
#ifndef FLOW_COST_CUPD
#define FLOW_COST_CUPD 4600  // chain: update of the next diagonal tile
#endif
#ifndef FLOW_COST_UPD
#define FLOW_COST_UPD 12000  // rank-64 update, accumulator loaded and stored
#endif
#ifndef FLOW_COST_CARRY
#define FLOW_COST_CARRY 3000  // saved when the previous task left the accumulator in registers
#endif
#ifndef FLOW_COST_TRSM
#define FLOW_COST_TRSM 14500  // tile solve, exchange copy and flag
#endif
#ifndef FLOW_COST_NEXTCOL
#define FLOW_COST_NEXTCOL 10800  // one column of a next-block task
#endif
#ifndef FLOW_COST_ZERO
#define FLOW_COST_ZERO 500
#endif
#ifndef FLOW_UPD_GROUP
#define FLOW_UPD_GROUP 2  // columns per held-back update group
#endif
#ifndef FLOW_UPD_MARGIN
#define FLOW_UPD_MARGIN 2  // the updates by the last columns before a tile's own are not held back
#endif
#ifndef FLOW_COST_LAT
#define FLOW_COST_LAT 1500  // flag store to flag seen by another workgroup
#endif

static bool flow_build_list(const FlowShape& s, int helpers, FlowSchedule& out) {
  const int nT = s.nT, R = s.rows();
  const int chain = helpers;  // agent id of the chain
  std::vector<uint32_t> word;
  std::vector<int> agent, cost, tile, op;
  auto add = [&](int a, int c, int o, uint32_t w, int t) {
    agent.push_back(a);
    cost.push_back(c);
    op.push_back(o);
    word.push_back(w);
    tile.push_back(t);
    return (int)agent.size() - 1;
  };
  std::vector<std::vector<int>> succ;
  std::vector<int> npred;
  auto edge = [&](int from, int to) {
    if (from < 0 || to < 0) return;
    if ((int)succ.size() < (int)agent.size()) succ.resize(agent.size());
    succ[from].push_back(to);
  };
  // chain: CF(k) factor, CT(k) solve of (k+1, k), CU(k) update of (k+1, k+1)
  std::vector<int> cf(nT, -1), ct(nT, -1), cu(nT, -1);
  for (int k = 0; k < nT; ++k) {
    cf[k] = add(chain, FLOW_COST_CHOL, -1, 0, -1);
    if (k + 1 < nT) {
      ct[k] = add(chain, FLOW_COST_CTRSM, -1, 0, -1);
      cu[k] = add(chain, FLOW_COST_CUPD, -1, 0, -1);
    }
  }
  // producer of X(rho, c); last helper update of each tile
  std::vector<int> xprod((size_t)R * nT, -1), last_upd((size_t)R * nT, -1);
  std::vector<std::pair<int, int>> upd_in;  // (task, j) for the X edges
  for (int k = 0; k + 1 < nT; ++k) xprod[(size_t)(k + 1) * nT + k] = ct[k];
  for (int rho = 0; rho < R; ++rho)
    for (int c = 0; c < nT; ++c) {
      const int owner = flow_owner(s, rho, c, helpers);
      const int tid = rho * nT + c;
      if (rho < nT && rho < c) {
        add(owner, FLOW_COST_ZERO, FLOW_ZERO, flow_encode(FLOW_ZERO, rho, c, 0), -1);
        continue;
      }
      const bool diag = (rho == c);
      const bool hand = flow_is_handoff(nT, rho, c);
      int prev = -1;
      for (int j = 0; j < c; ++j) {
        if (diag && j == c - 1) continue;  // applied by the chain
        const int t = add(owner, FLOW_COST_UPD, FLOW_UPD, flow_encode(FLOW_UPD, rho, c, j), hand ? -1 : tid);
        upd_in.push_back({t, j});
        edge(prev, t);  // one accumulator: updates of a tile in column order
        prev = t;
      }
      last_upd[tid] = prev;
      if (!diag && !(rho < nT && rho == c + 1)) {
        const int t = add(owner, FLOW_COST_TRSM, FLOW_TRSM, flow_encode(FLOW_TRSM, rho, c, c), tid);
        edge(prev, t);
        // inv(L_cc) is published right after the chain's factorization
        edge(cf[c], t);
        xprod[tid] = t;
      }
    }
  // next-block updates: one task per target and span of columns.  Each
  // target goes to the helper outside the near pool with the least work so
  // far.  (Given to the near pool as gap filler, they ran when the chain's
  // solves were already due -- the device takes longer than the simulated
  // costs -- and the chain waited on them: measured on MI355X at b = 2000,
  // keeping them off the near pool cut the chain's wait per column from
  // 19 k to 8 k cycles.)
  std::vector<double> load(helpers + 1, 0.0);
  for (size_t t = 0; t < agent.size(); ++t) load[agent[t]] += cost[t];
  const int hn_next = flow_near_pool(s, helpers);
  auto push_next = [&](int rho, int c) {
    const int ra = rho < nT ? nT + rho : nT + s.nTL;
    const int rb = rho < nT ? nT + c : (rho == nT ? nT + c : nT + s.nTL);
    int owner = hn_next;
    for (int h = owner + 1; h < helpers; ++h)
      if (load[h] < load[owner]) owner = h;
    load[owner] += (double)FLOW_COST_NEXTCOL * nT;
    int prev = -1;
    for (int k = 0; k < nT; k = flow_next_end(k, nT)) {
      const int kend = flow_next_end(k, nT);
      const int t = add(owner, FLOW_COST_NEXTCOL * (kend - k), FLOW_NEXT, flow_encode(FLOW_NEXT, rho, c, k), -1);
      edge(prev, t);
      prev = t;
      for (int kk = k; kk < kend; ++kk) {
        edge(xprod[(size_t)ra * nT + kk], t);
        if (rb != ra) edge(xprod[(size_t)rb * nT + kk], t);
      }
    }
  };
  succ.resize(agent.size());
  for (auto& u : upd_in) {
    int o, rho, c, j;
    flow_decode(word[u.first], o, rho, c, j);
    edge(xprod[(size_t)rho * nT + j], u.first);
    if (rho != c) edge(xprod[(size_t)c * nT + j], u.first);
    // Updates of a tile by columns far from its own are held back to the
    // last column of their group of FLOW_UPD_GROUP: the helper then applies
    // the group back to back with the accumulator kept in registers (one
    // load and one store of the tile per group instead of per column).  Only
    // the simulated order sees these edges; the device waits for the real
    // flags.
    const int g1 = std::min((j / FLOW_UPD_GROUP + 1) * FLOW_UPD_GROUP - 1, c - 1 - FLOW_UPD_MARGIN);
    if (g1 > j && !flow_is_handoff(nT, rho, c)) {
      edge(xprod[(size_t)rho * nT + g1], u.first);
      if (rho != c) edge(xprod[(size_t)c * nT + g1], u.first);
    }
  }
  if (s.fold && s.nTL > 0) {
    for (int r = 0; r < nT; ++r)
      for (int c = 0; c <= r; ++c) push_next(r, c);
    if (s.nTa > 0)
      for (int c = 0; c < nT; ++c) push_next(nT, c);
  }
  if (s.fold && s.nTa > 0) push_next(nT + 1, 0);
  succ.resize(agent.size());
  // chain order and its waits on the hand-off tiles
  for (int k = 0; k < nT; ++k) {
    if (k + 1 < nT) {
      edge(cf[k], ct[k]);
      edge(ct[k], cu[k]);
      edge(cu[k], cf[k + 1]);
      if (k >= 1) {  // the chain fetches both hand-off tiles before the solve
        edge(last_upd[(size_t)(k + 1) * nT + k], ct[k]);
        edge(last_upd[(size_t)(k + 1) * nT + k + 1], ct[k]);
      }
    }
  }
  const int N = (int)agent.size();
  npred.assign(N, 0);
  for (int t = 0; t < N; ++t)
    for (int v : succ[t]) ++npred[v];
  // topological order, then upward ranks
  std::vector<int> topo;
  topo.reserve(N);
  {
    std::vector<int> np = npred;
    for (int t = 0; t < N; ++t)
      if (np[t] == 0) topo.push_back(t);
    for (size_t q = 0; q < topo.size(); ++q)
      for (int v : succ[topo[q]])
        if (--np[v] == 0) topo.push_back(v);
  }
  if ((int)topo.size() != N) return false;  // dependency cycle
  std::vector<double> rank(N, 0.0);
  for (int q = N - 1; q >= 0; --q) {
    const int t = topo[q];
    double m = 0.0;
    for (int v : succ[t]) m = std::max(m, rank[v] + (agent[v] != agent[t] ? FLOW_COST_LAT : 0));
    rank[t] = cost[t] + m;
  }
  // list-scheduling simulation
  const int A = helpers + 1;
  using RQ = std::priority_queue<std::pair<double, int>>;                       // (rank, -task)
  using FQ = std::priority_queue<std::pair<double, int>, std::vector<std::pair<double, int>>,
                                 std::greater<std::pair<double, int>>>;         // (ready time, task)
  std::vector<RQ> ready(A);
  std::vector<FQ> future(A);
  std::vector<double> rtime(N, 0.0), busy(A, -1.0);
  std::vector<int> carried(A, -1), left = npred;
  std::vector<std::vector<int>> order(A);
  struct Ev {
    double t;
    int kind, x;  // kind 0: task x finishes, 1: agent x looks for work
    bool operator>(const Ev& o) const { return t != o.t ? t > o.t : kind > o.kind; }
  };
  std::priority_queue<Ev, std::vector<Ev>, std::greater<Ev>> ev;
  for (int t = 0; t < N; ++t)
    if (left[t] == 0) future[agent[t]].push({0.0, t});
  for (int a = 0; a < A; ++a) ev.push({0.0, 1, a});
  int done = 0;
  while (!ev.empty()) {
    const Ev e = ev.top();
    ev.pop();
    if (e.kind == 0) {
      const int t = e.x, a = agent[t];
      ++done;
      for (int v : succ[t]) {
        rtime[v] = std::max(rtime[v], e.t + (agent[v] != a ? FLOW_COST_LAT : 0));
        if (--left[v] == 0) {
          future[agent[v]].push({rtime[v], v});
          ev.push({rtime[v], 1, agent[v]});
        }
      }
      ev.push({e.t, 1, a});
      continue;
    }
    const int a = e.x;
    if (busy[a] > e.t) continue;  // still running a task; its finish re-posts
    while (!future[a].empty() && future[a].top().first <= e.t) {
      const int t = future[a].top().second;
      future[a].pop();
      ready[a].push({rank[t], -t});
    }
    if (ready[a].empty()) continue;  // woken again when a task becomes ready
    const int t = -ready[a].top().second;
    ready[a].pop();
    int c = cost[t];
    if (tile[t] >= 0 && tile[t] == carried[a]) c -= FLOW_COST_CARRY;
    carried[a] = (op[t] == FLOW_UPD && tile[t] >= 0) ? tile[t] : -1;
    busy[a] = e.t + c;
    order[a].push_back(t);
    ev.push({busy[a], 0, t});
  }
  if (done != N) return false;
  out.helpers = helpers;
  out.tasks.clear();
  out.off.assign(helpers + 1, 0);
  for (int h = 0; h < helpers; ++h) {
    for (int t : order[h]) out.tasks.push_back(word[t]);
    out.off[h + 1] = (int)out.tasks.size();
  }
  return true;
}

bool flow_build_schedule(const FlowShape& s, int helpers, int delta, FlowSchedule& out) {
  if (delta >= 1) {
    flow_build_levels(s, helpers, delta, out);
    return true;
  }
  return flow_build_list(s, helpers, out);
}

// ------------------------------------------------------------------ device


__device__ __forceinline__ double* flow_tile(const FlowArgs& p, int rho, int c, int& rows, int& cols) {
  cols = min(64, p.b - 64 * c);
  if (rho < p.nT) {
    rows = min(64, p.b - 64 * rho);
    return p.D + (int64_t)(64 * rho) * p.b + 64 * c;
  }
  rho -= p.nT;
  if (rho < p.nTL) {
    rows = min(64, p.b - 64 * rho);
    return p.L + (int64_t)(64 * rho) * p.b + 64 * c;
  }
  rho -= p.nTL;
  rows = min(64, p.a - 64 * rho);
  return p.B + (int64_t)(64 * rho) * p.b + 64 * c;
}

__device__ __forceinline__ int flow_rows(const FlowArgs& p) { return p.nT + p.nTL + p.nTa; }
__device__ __forceinline__ double* xtile(const FlowArgs& p, int rho, int c) {
  return p.xbuf + ((int64_t)rho * p.nT + c) * 4096;
}
__device__ __forceinline__ unsigned long long* xflag(const FlowArgs& p, int rho, int c) {
  return p.flags + (int64_t)rho * p.nT + c;
}
__device__ __forceinline__ unsigned long long* wflag(const FlowArgs& p, int k) {
  return p.flags + (int64_t)flow_rows(p) * p.nT + k;
}
__device__ __forceinline__ unsigned long long* hflag(const FlowArgs& p, int k, int s) {
  return p.flags + (int64_t)flow_rows(p) * p.nT + p.nT + 2 * k + s;
}
__device__ __forceinline__ double* hsub(const FlowArgs& p, int c) { return p.hacc + (int64_t)c * 4096; }
__device__ __forceinline__ double* hdiag(const FlowArgs& p, int c) {
  return p.hacc + ((int64_t)p.nT + c) * 4096;
}

// Bounded wait by thread 0 for *flag == tag (nullptr: nothing to wait for).
// Gives up (returns 0 and raises err) after spin_ticks, or at once when another
// workgroup has raised err, so that every workgroup still reaches the end of
// its list.
__device__ __forceinline__ int flow_poll(const FlowArgs& p, unsigned long long* flag, unsigned long long tag) {
  if (flag == nullptr) return 1;
  if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) == tag) return 1;
  const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
  while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != tag) {
    if (__hip_atomic_load(p.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0) return 0;
    if (__builtin_amdgcn_s_memrealtime() - t0 > p.spin_ticks) {
      atomicOr(p.err, 1);
      return 0;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return 1;
}

// The workgroup waits until both flags are set (f2 may be nullptr).  The
// barriers order LDS only: the guarded data are read after the second one in
// program order, and a full barrier would also wait for every global load the
// workgroup has in flight (the prefetched accumulator of the step).
__device__ __forceinline__ bool flow_wait2(const FlowArgs& p, unsigned long long* f1, unsigned long long* f2,
                                           unsigned long long tag, int* ok_sh) {
  if (threadIdx.x == 0) {
    int ok = flow_poll(p, f1, tag);
    if (ok) ok = flow_poll(p, f2, tag);
    *ok_sh = ok;
  }
  lds_barrier();
  const int ok = *ok_sh;
  lds_barrier();
  return ok != 0;
}

// 64x64 contiguous tile <-> registers / LDS with system-scope accesses.
__device__ __forceinline__ void lds_load_sys(double* __restrict__ s, const double* __restrict__ g) {
  const int c = threadIdx.x & 63;
#pragma unroll 4
  for (int r = threadIdx.x >> 6; r < 64; r += 4) s[r * BTA_LDS_LD + c] = ld_sys(&g[r * 64 + c]);
}

// Stores, after the caller's st_sys data stores, the flag once they completed.
__device__ __forceinline__ void flow_publish(unsigned long long* flag, unsigned long long tag) {
  flag_publish(flag, tag);
}

// X = C inv(L_kk)^T from C in sC and inv(L_kk) (lower) in sW.
__device__ __forceinline__ void flow_trsm(TileAcc& x, const double* sC, const double* sW) {
  acc_zero(x);
  const int kend = ((threadIdx.x >> 6) & 1) * 32 + 32;
  mma_nt<false>(x, sC, sW, 0, kend);
}

// Chain: store the factored diagonal tile k held in sA (L_kk) / sB (inverse).
// The flag is published separately (chain_flag), so that the write-through
// stores drain while the chain already computes.
__device__ __forceinline__ void chain_store_inv(const FlowArgs& p, int k, const double* sB) {
  double* w = p.winv + (int64_t)k * 4096;
  const int c = threadIdx.x & 63, r0 = threadIdx.x >> 6;
#pragma unroll
  for (int h = 0; h < 16; h += 8) {  // two halves of 8 rows: reads first, then stores
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = sB[(r0 + 4 * (h + j)) * BTA_LDS_LD + c];
#pragma unroll
    for (int j = 0; j < 8; ++j) st_sys(&w[(r0 + 4 * (h + j)) * 64 + c], v[j]);
  }
}

// ... then, after the inverse's flag, L_kk to the factor (read by no helper)
// and the tile's log pivots.
__device__ __forceinline__ void chain_store_l(const FlowArgs& p, int k, const TileCholResult& res,
                                              const double* sA) {
  int rows, cols;
  double* g = flow_tile(p, k, k, rows, cols);
  const int c = threadIdx.x & 63, r0 = threadIdx.x >> 6;
  for (int r = r0; r < 64; r += 4)
    if (r < rows && c < cols) g[(int64_t)r * p.b + c] = sA[r * BTA_LDS_LD + c];
  if (threadIdx.x == 0) {
    p.logpart[k] = res.logsum;
    if (res.fail >= 0) atomicMin(p.info, (unsigned long long)(p.row0 + (long long)k * 64 + res.fail));
  }
}

// Per column k the chain
//   1. waits for the hand-off tiles (k+1, k) and (k+1, k+1) (updated by their
//      helpers through column k-1) and loads both;
//   2. solves X = C(k+1, k) inv(L_kk)^T and issues the stores of X;
//   3. applies C(k+1, k+1) -= X X^T, then publishes X (its stores drained
//      during the update);
//   4. factors C(k+1, k+1), stores its inverse and publishes it at once (the
//      helpers' solves of column k+1 wait for it; the stores drain while
//      they are issued, a few hundred cycles of wait measured on MI355X),
//      then stores L into the factor (no helper reads it; those stores drain
//      during the next step's wait).
__device__ void flow_chain(const FlowArgs& p, unsigned long long tag, double* sA, double* sB,
                           int* ok_sh) {
  const int nT = p.nT;
  int rows, cols;
  TileAcc acc, accd;
  {
    const double* g = flow_tile(p, 0, 0, rows, cols);
    acc_load(acc, g, p.b, rows, cols);
    acc_to_lds(acc, sA);
    __syncthreads();
    const TileCholResult res = tile_chol_inv(sA, sB, rows);
    chain_store_inv(p, 0, sB);
    flow_publish(wflag(p, 0), tag);
    chain_store_l(p, 0, res, sA);
  }
  for (int k = 0; k + 1 < nT; ++k) {
    if (k == 0) {
      acc_load(acc, flow_tile(p, 1, 0, rows, cols), p.b, rows, cols);
      acc_load(accd, flow_tile(p, 1, 1, rows, cols), p.b, rows, cols);
    } else {
      if (!flow_wait2(p, hflag(p, k, 0), hflag(p, k, 1), tag, ok_sh) && threadIdx.x == 0) atomicOr(p.err, 2);
      acc_load_sys(acc, hsub(p, k));
      acc_load_sys(accd, hdiag(p, k + 1));
    }
    acc_to_lds(acc, sA);
    __syncthreads();
    flow_trsm(acc, sA, sB);  // X(k+1, k) in acc
    lds_barrier();           // the reads of sA are done
    {
      double* g = flow_tile(p, k + 1, k, rows, cols);
      acc_store(acc, g, p.b, rows, cols);
      acc_store_sys(acc, xtile(p, k + 1, k));
      acc_to_lds(acc, sA);
    }
    lds_barrier();
    mma_nt64<true>(accd, sA, sA);
    flow_publish(xflag(p, k + 1, k), tag);
    acc_to_lds(accd, sA);
    __syncthreads();
    const int mv = min(64, p.b - 64 * (k + 1));
    const TileCholResult res = tile_chol_inv(sA, sB, mv);
    chain_store_inv(p, k + 1, sB);
    flow_publish(wflag(p, k + 1), tag);
    chain_store_l(p, k + 1, res, sA);
  }
}

// One step of a helper: a task, or one column kk of a next-block task.
struct HStep {
  int t, kk;            // task index, column (FLOW_NEXT: current column of the span)
  int op, rho, c, k;    // decoded task
  const double* srcA;   // operand tiles (64 x 64, contiguous); TRSM: A is the accumulator
  const double* srcB;
  unsigned long long* f1;
  unsigned long long* f2;
};

__device__ __forceinline__ void next_rows(const FlowArgs& p, int rho, int c, int& ra, int& rb) {
  const int nT = p.nT;
  if (rho < nT) { ra = nT + rho; rb = nT + c; }
  else if (rho == nT) { ra = nT + p.nTL; rb = nT + c; }
  else { ra = rb = nT + p.nTL; }
}

// Decodes task t at column kk (kk < 0: the task's first column).
__device__ __forceinline__ HStep make_step(const FlowArgs& p, int t, int kk) {
  HStep st;
  st.t = t;
  flow_decode(p.tasks[t], st.op, st.rho, st.c, st.k);
  st.kk = kk < 0 ? st.k : kk;
  st.srcA = st.srcB = nullptr;
  st.f1 = st.f2 = nullptr;
  if (st.op == FLOW_UPD) {
    st.srcA = xtile(p, st.rho, st.k);
    st.srcB = xtile(p, st.c, st.k);
    st.f1 = xflag(p, st.rho, st.k);
    if (st.rho != st.c) st.f2 = xflag(p, st.c, st.k);
  } else if (st.op == FLOW_TRSM) {
    st.srcB = p.winv + (int64_t)st.k * 4096;
    st.f1 = wflag(p, st.k);
  } else if (st.op == FLOW_NEXT) {
    int ra, rb;
    next_rows(p, st.rho, st.c, ra, rb);
    st.srcA = xtile(p, ra, st.kk);
    st.srcB = xtile(p, rb, st.kk);
    st.f1 = xflag(p, ra, st.kk);
    if (rb != ra) st.f2 = xflag(p, rb, st.kk);
  }
  return st;
}

// Current value of a flag (`tag` for an absent flag).
__device__ __forceinline__ unsigned long long flag_value(const unsigned long long* f, unsigned long long tag) {
  return f == nullptr ? tag
                      : __hip_atomic_load(const_cast<unsigned long long*>(f), __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_SYSTEM);
}

// Issue the operand loads of a step into registers.  The exchange tiles and
// tile inverses are written once per launch (before their flag), so ordinary
// cached loads after the flag see them; the L2 lines cannot be older than the
// launch.
__device__ __forceinline__ void step_fetch(const HStep& st, TilePrefetch& pa, TilePrefetch& pb) {
  if (st.srcA) pa.load_full(st.srcA);
  if (st.srcB && st.srcB != st.srcA) pb.load_full(st.srcB);
}

__device__ void flow_helper(const FlowArgs& p, int h, unsigned long long tag, double* sA, double* sB,
                            int* ok_sh) {
  const int nT = p.nT;
  const int t0 = p.off[h], t1 = p.off[h + 1];
  TileAcc acc;
  TilePrefetch pa, pb;
  int carry_rho = -1, carry_c = -1;  // in-panel tile currently held in acc
  bool fetched = false;              // operands of `cur` already in pa / pb
  if (t0 >= t1) return;
  HStep cur = make_step(p, t0, -1);
  for (;;) {
    int rows = 0, cols = 0;
    double* g = nullptr;
    if (cur.op != FLOW_NEXT) g = flow_tile(p, cur.rho, cur.c, rows, cols);
    if (cur.op == FLOW_ZERO) {
      const int cc = threadIdx.x & 63;
      for (int r = threadIdx.x >> 6; r < rows; r += 4)
        if (cc < cols) g[(int64_t)r * p.b + cc] = 0.0;
      if (cur.t + 1 >= t1) break;
      cur = make_step(p, cur.t + 1, -1);
      fetched = false;
      continue;
    }
    // accumulator of the step's target: its data come from this helper's own
    // earlier steps, so its loads are issued before the wait for the operands
    // and their latency overlaps it
    double* tg = g;
    int64_t ld = p.b;
    int trows = rows, tcols = cols;
    bool hand = false;
    if (cur.op == FLOW_NEXT) {
      if (cur.rho < nT) {
        tg = p.Dn + (int64_t)(64 * cur.rho) * p.b + 64 * cur.c;
        trows = min(64, p.b - 64 * cur.rho);
        tcols = min(64, p.b - 64 * cur.c);
      } else if (cur.rho == nT) {
        tg = p.Bn + 64 * cur.c;
        trows = p.a;
        tcols = min(64, p.b - 64 * cur.c);
      } else {
        tg = p.T;
        ld = p.a;
        trows = tcols = p.a;
      }
      if (cur.kk == cur.k) acc_load(acc, tg, ld, trows, tcols);
    } else {
      hand = (cur.op == FLOW_UPD) && flow_is_handoff(nT, cur.rho, cur.c);
      if (!(carry_rho == cur.rho && carry_c == cur.c)) {
        if (hand && cur.k > 0) acc_load_sys(acc, cur.rho == cur.c ? hdiag(p, cur.c) : hsub(p, cur.c));
        else acc_load(acc, g, p.b, rows, cols);
      }
    }
    if (!fetched) {
      if (!flow_wait2(p, cur.f1, cur.f2, tag, ok_sh) && threadIdx.x == 0) atomicOr(p.err, 4);
      step_fetch(cur, pa, pb);
    }
    // the following step.  Thread 0 issues the loads of its flags now, both
    // at once, so that their latency overlaps the staging of this step.
    HStep nxt;
    bool has_next = true;
    if (cur.op == FLOW_NEXT && cur.kk + 1 < flow_next_end(cur.k, nT)) nxt = make_step(p, cur.t, cur.kk + 1);
    else if (cur.t + 1 < t1) nxt = make_step(p, cur.t + 1, -1);
    else has_next = false;
    const bool probe = has_next && nxt.op != FLOW_ZERO;
    unsigned long long fv1 = 0, fv2 = 0;
    if (threadIdx.x == 0 && probe) {
      fv1 = flag_value(nxt.f1, tag);
      fv2 = flag_value(nxt.f2, tag);
    }
    // stage operands in LDS
    if (cur.op == FLOW_TRSM) acc_to_lds(acc, sA);
    else pa.store_full(sA);
    if (cur.srcB != cur.srcA) pb.store_full(sB);
    // the following step's operands are fetched now if they are ready
    if (threadIdx.x == 0) *ok_sh = (probe && fv1 == tag && fv2 == tag) ? 1 : 0;
    lds_barrier();
    fetched = (*ok_sh != 0);
    // not ready yet: thread 0 loads the flags again now and looks at them
    // after the MFMAs (a second chance to fetch ahead of the next step)
    const bool second = probe && !fetched;
    unsigned long long gv1 = 0, gv2 = 0;
    if (second && threadIdx.x == 0) {
      gv1 = flag_value(nxt.f1, tag);
      gv2 = flag_value(nxt.f2, tag);
    }
    // an update followed by an update: its operand loads ride in the MFMA loop
    const bool pf_mma = fetched && cur.op != FLOW_TRSM && nxt.srcA != nullptr;
    if (fetched && !pf_mma) step_fetch(nxt, pa, pb);
    const double* sBop = (cur.srcB == cur.srcA) ? sA : sB;
    if (cur.op == FLOW_TRSM) {
      flow_trsm(acc, sA, sB);  // the accumulator was staged in sA: reuse its registers
      lds_barrier();
      acc_store(acc, g, p.b, rows, cols);
      acc_store_sys(acc, xtile(p, cur.rho, cur.c));
      carry_rho = carry_c = -1;
      flow_publish(xflag(p, cur.rho, cur.c), tag);
    } else {
      if (pf_mma) mma_nt64_neg_pf(acc, sA, sBop, pa.r, nxt.srcA, pb.r, nxt.srcB);
      else mma_nt64<true>(acc, sA, sBop);
      lds_barrier();
      if (cur.op == FLOW_NEXT) {
        if (!(has_next && nxt.op == FLOW_NEXT && nxt.t == cur.t)) acc_store(acc, tg, ld, trows, tcols);
        carry_rho = carry_c = -1;
      } else if (hand) {
        // hand-off tiles are never carried: every update goes to the buffer
        acc_store_sys(acc, cur.rho == cur.c ? hdiag(p, cur.c) : hsub(p, cur.c));
        carry_rho = carry_c = -1;
        if (cur.rho == cur.c + 1 && cur.k == cur.c - 1) flow_publish(hflag(p, cur.c, 0), tag);
        else if (cur.rho == cur.c && cur.k == cur.c - 2) flow_publish(hflag(p, cur.c - 1, 1), tag);
      } else if (has_next && nxt.rho == cur.rho && nxt.c == cur.c &&
                 (nxt.op == FLOW_UPD || nxt.op == FLOW_TRSM)) {
        carry_rho = cur.rho;  // the next task on this tile keeps the accumulator
        carry_c = cur.c;
      } else {
        acc_store(acc, g, p.b, rows, cols);
        carry_rho = carry_c = -1;
      }
    }
    if (second) {  // ok_sh[1] is written once per step, after a barrier that follows its last read
      if (threadIdx.x == 0) ok_sh[1] = (gv1 == tag && gv2 == tag) ? 1 : 0;
      lds_barrier();
      fetched = (ok_sh[1] != 0);
      if (fetched) step_fetch(nxt, pa, pb);
    }
    if (!has_next) break;
    cur = nxt;
  }
}

__global__ void __launch_bounds__(BTA_THREADS, 2) k_panel_flow(FlowArgs p) {
  __shared__ double sA[64 * BTA_LDS_LD];
  __shared__ double sB[64 * BTA_LDS_LD];
  __shared__ int ok_sh[2];
  const unsigned long long tag = (p.epoch[0] << 20) | (unsigned long long)p.blk;
  if (blockIdx.x == 0) flow_chain(p, tag, sA, sB, ok_sh);
  else flow_helper(p, blockIdx.x - 1, tag, sA, sB, ok_sh);
}

// -------------------------------------------------------------------- host

int flow_capacity(int device);

FlowPlan::~FlowPlan() { release(); }

void FlowPlan::release() {
  for (auto& sc : scheds) {
    if (sc.d_tasks) hipFree(sc.d_tasks);
    if (sc.d_off) hipFree(sc.d_off);
  }
  scheds.clear();
  if (xbuf) hipFree(xbuf);
  if (flags) hipFree(flags);
  xbuf = nullptr;
  flags = nullptr;
  hacc = nullptr;
  nT = -1;
}

size_t FlowPlan::bytes_for(int nT, int nTa) {
  const size_t R = 2 * (size_t)nT + nTa;
  return ((R * nT + 2 * (size_t)nT) * 4096) * sizeof(double) + (R * nT + 3 * (size_t)nT) * 8;
}

int FlowPlan::build(int nT_, int nTa_) {
  release();
  nT = nT_;
  nTa = nTa_;
  const int R = 2 * nT + nTa;
  hipError_t e;
  const size_t xd = ((size_t)R * nT + 2 * (size_t)nT) * 4096;
  if ((e = hipMalloc(&xbuf, xd * sizeof(double))) != hipSuccess) return (int)e;
  hacc = xbuf + (size_t)R * nT * 4096;
  const size_t nf = (size_t)R * nT + 3 * (size_t)nT;
  if ((e = hipMalloc(&flags, nf * sizeof(unsigned long long))) != hipSuccess) return (int)e;
  if ((e = hipMemset(flags, 0, nf * sizeof(unsigned long long))) != hipSuccess) return (int)e;
  buf_bytes = bytes_for(nT, nTa);
  return 0;
}

int FlowPlan::sched(int nTa_s, int helpers, int delta, int near_div, int fold, const FlowSched** out) {
  if (nTa_s > nTa) return (int)hipErrorInvalidValue;
  for (const auto& sc : scheds)
    if (sc.helpers == helpers && sc.delta == delta && sc.nTa == nTa_s && sc.near_div == near_div &&
        sc.fold == fold) {
      *out = &sc;
      return 0;
    }
  FlowSched sc;
  sc.helpers = helpers;
  sc.delta = delta;
  sc.nTa = nTa_s;
  sc.near_div = near_div;
  sc.fold = fold;
  std::vector<uint32_t> all;
  std::vector<int> off;
  for (int v = 0; v < 2; ++v) {  // 0: block with L, 1: last block (no L)
    FlowShape s{nT, v == 0 ? nT : 0, nTa_s, near_div, fold};
    FlowSchedule h;
    if (!flow_build_schedule(s, helpers, delta, h)) return -1;
    const int base = (int)all.size();
    sc.off_base[v] = (int)off.size();
    for (int q = 0; q <= helpers; ++q) off.push_back(h.off[q] + base);
    all.insert(all.end(), h.tasks.begin(), h.tasks.end());
  }
  hipError_t e;
  if ((e = hipMalloc(&sc.d_tasks, std::max<size_t>(1, all.size()) * sizeof(uint32_t))) != hipSuccess) return (int)e;
  if ((e = hipMalloc(&sc.d_off, off.size() * sizeof(int))) != hipSuccess) return (int)e;
  if (!all.empty() && (e = hipMemcpy(sc.d_tasks, all.data(), all.size() * 4, hipMemcpyHostToDevice)) != hipSuccess)
    return (int)e;
  if ((e = hipMemcpy(sc.d_off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
    return (int)e;
  sc.bytes = std::max<size_t>(1, all.size()) * sizeof(uint32_t) + off.size() * sizeof(int);
  scheds.push_back(sc);
  *out = &scheds.back();
  return 0;
}

int flow_helpers_default(int device) {
  // room for two factorizations in flight (prior and conditional chains)
  const int cap = flow_capacity(device);
  return cap > 0 ? std::max(1, cap / 2 - 1) : 0;
}

unsigned long long flow_spin_ticks(int device, double seconds) {
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || khz <= 0) khz = 100000;
  return (unsigned long long)(seconds * 1e3 * khz);
}

FlowArgs flow_args(const FlowPlan& fp, const FlowSched& sc, double* D, double* L, double* B, double* T, double* winv,
                   double* logpart, const unsigned long long* epoch, unsigned long long* info, int* err, int i,
                   int n, int b, int a) {
  FlowArgs p;
  const int nT = fp.nT;
  p.D = D + (int64_t)i * b * b;
  p.L = (i < n - 1) ? L + (int64_t)i * b * b : nullptr;
  p.B = (a > 0) ? B + (int64_t)i * a * b : nullptr;
  p.Dn = (i < n - 1) ? D + (int64_t)(i + 1) * b * b : nullptr;
  p.Bn = (a > 0 && i < n - 1) ? B + (int64_t)(i + 1) * a * b : nullptr;
  p.T = (a > 0) ? T : nullptr;
  p.winv = winv + (int64_t)i * nT * 4096;
  p.xbuf = fp.xbuf;
  p.hacc = fp.hacc;
  p.logpart = logpart + (int64_t)i * nT;
  p.flags = fp.flags;
  const int v = (i < n - 1) ? 0 : 1;
  p.tasks = sc.d_tasks;
  p.off = sc.d_off + sc.off_base[v];
  p.epoch = epoch;
  p.info = info;
  p.err = err;
  p.row0 = (long long)i * b;
  p.spin_ticks = fp.spin_ticks;
  p.b = b;
  p.a = a;
  p.nT = nT;
  p.nTL = (i < n - 1) ? nT : 0;
  p.nTa = sc.nTa;
  p.blk = i + 1;
  p.helpers = sc.helpers;
  return p;
}

void flow_launch_block(const Workspace& ws, const FlowPlan& fp, const FlowSched& sc, double* D, double* L,
                       double* B, double* T, int i, int n, int b, int a, hipStream_t st) {
  const FlowArgs p = flow_args(fp, sc, D, L, B, T, ws.winv, ws.logpart, ws.epoch, ws.info, ws.err, i, n, b, a);
  hipLaunchKernelGGL(k_panel_flow, dim3(1 + sc.helpers), dim3(BTA_THREADS), 0, st, p);
}

// Batched form: every entry's chain and helpers in one launch (layout in
// FlowBatch).  All entries' workgroups are resident together (their sum is
// within the device's capacity).
__global__ void __launch_bounds__(BTA_THREADS, 2) k_panel_flow_batched(FlowBatch fb) {
  __shared__ double sA[64 * BTA_LDS_LD];
  __shared__ double sB[64 * BTA_LDS_LD];
  __shared__ int ok_sh[2];
  const int wg = (int)blockIdx.x;
  int pos = wg;  // position in the entries' ranges of `begin`
  if (fb.iso) {
    if (wg >= fb.cus && wg < fb.cus + fb.nent) return;  // the CU partner of a chain
    if (wg < fb.nent) pos = -1;                         // chain of entry wg
    else pos = wg < fb.cus ? wg - fb.nent : wg - 2 * fb.nent;
  }
  int lo = 0, local = 0;
  if (pos < 0) {
    lo = wg;
  } else {
    int hi = fb.nent - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (fb.begin[mid] <= pos) lo = mid;
      else hi = mid - 1;
    }
    local = pos - fb.begin[lo] + (fb.iso ? 1 : 0);
  }
  const FlowArgs p = fb.e[lo];
  const unsigned long long tag = (p.epoch[0] << 20) | (unsigned long long)p.blk;
  if (local == 0) flow_chain(p, tag, sA, sB, ok_sh);
  else flow_helper(p, local - 1, tag, sA, sB, ok_sh);
}

void flow_launch_batched(const FlowBatch& fb, int total, hipStream_t st) {
  hipLaunchKernelGGL(k_panel_flow_batched, dim3(total), dim3(BTA_THREADS), 0, st, fb);
}

int flow_capacity(int device) {
  int cus = 0, per = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
    return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_panel_flow, BTA_THREADS, 0) != hipSuccess || per <= 0)
    return 0;
  int per_b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_b, k_panel_flow_batched, BTA_THREADS, 0) == hipSuccess &&
      per_b > 0)
    per = std::min(per, per_b);
  return cus * per;
}

}  // namespace bta
