// Persistent dataflow factorization of one BTA block column (host + device).
//
// The stacked panel P_i = [D_i; L_i; B_i] of block i is split into 64x64
// tiles (ρ, c): stacked tile row ρ (D rows 0..nT-1, L rows nT..nT+nTL-1, the
// arrow row after them), tile column c.  Its right-looking tile Cholesky is
// run by ONE launch of 1 + P workgroups that stay resident:
//
//   * workgroup 0, the chain, walks the diagonal: with L_kk and inv(L_kk) of
//     step k in LDS it solves the sub-diagonal tile (k+1, k), applies that
//     tile's rank-64 product to the diagonal tile (k+1, k+1) and factors it;
//   * workgroups 1..P, the helpers, run every other tile task -- the solve
//     X(ρ, k) = C(ρ, k) inv(L_kk)^T and the rank-64 updates
//     C(ρ, c) -= X(ρ, k) X(c, k)^T -- from a task list built on the host.
//
// Every task has a level: Winv_k (chain) 3k, solves of column k 3k+1, updates
// with column j at 3 min(g(j), cap) + 2, where g(j) is the last column of j's
// group of δ columns and the cap keeps a tile's updates before its own solve
// (cap = c - 1) or before the chain takes it over (cap = c - 2 on the
// diagonal).  Each helper runs its tasks in increasing level and every task
// depends only on lower levels, so, with all workgroups resident, the lowest
// unfinished task can always run: no ordering can deadlock.  Grouping (δ > 1)
// applies a far tile's updates from δ columns back to back with the
// accumulator in registers (one load and one store per group); the hand-off
// tiles on the chain's path are never grouped.
//
// The block update that couples block i to the next one is folded into the
// same launch: as soon as a span of columns of L_i (and B_i) is final,
// helpers apply D_{i+1} -= L_i L_i^T, B_{i+1} -= B_i L_i^T and T -= B_i B_i^T
// for those columns to the tiles they own (op FLOW_NEXT, flow_next_end).
// Nothing in the launch waits on these tasks, so they take the lowest
// priority of their level and fill the time the helpers would otherwise
// spend waiting for the chain; the next block's launch reads the results.
//
// Memory protocol (the L2 caches of the XCDs are not coherent): every address
// is stored to by one workgroup only.  Helpers accumulate their tiles in place
// with ordinary loads and stores; a finished tile is also written once, with
// system-scope (write-through) stores, to an exchange buffer that readers load
// with system-scope loads, and a flag carrying the call tag is stored after
// the data has completed.  The diagonal and sub-diagonal tiles the chain takes
// over are accumulated by their helpers in a separate hand-off buffer so that
// only the chain ever stores to their place in the factor.
#pragma once
#include <stdint.h>

#include <vector>

namespace bta {

enum : int { FLOW_UPD = 0, FLOW_TRSM = 1, FLOW_ZERO = 2, FLOW_NEXT = 3 };

// Columns of L_i per next-block update task.
#define FLOW_SPAN 4

// A next-block task covers columns [k, flow_next_end(k, nT)) of L_i: spans of
// FLOW_SPAN columns, except that each of the last FLOW_SPAN columns is a task
// of its own.  Those columns are solved only at the end of the panel, when
// every target still waits for them; single-column tasks keep the work left
// after the chain's last step small.
__host__ __device__ inline int flow_next_end(int k, int nT) {
  const int tail = nT - FLOW_SPAN;
  return k < tail ? (k + FLOW_SPAN < tail ? k + FLOW_SPAN : tail) : k + 1;
}

// Task word: op (2 bits) | ρ (10) | c (10) | k (10).
__host__ __device__ inline uint32_t flow_encode(int op, int rho, int c, int k) {
  return ((uint32_t)op << 30) | ((uint32_t)rho << 20) | ((uint32_t)c << 10) | (uint32_t)k;
}
__host__ __device__ inline void flow_decode(uint32_t w, int& op, int& rho, int& c, int& k) {
  op = (int)(w >> 30);
  rho = (int)((w >> 20) & 1023u);
  c = (int)((w >> 10) & 1023u);
  k = (int)(w & 1023u);
}

// Shape of a panel: nT tile columns (and D tile rows), nTL L tile rows (0 for
// the last block), nTa arrow tile rows (0 or 1).  With fold set, nTL > 0 the
// launch also updates D_{i+1} (and B_{i+1}); with nTa > 0 it updates the tip T.
// FLOW_NEXT targets: row rho < nT is D_{i+1} tile (rho, c); rho == nT is the
// arrow tile (0, c) of B_{i+1}; rho == nT + 1 is the tip T (c = 0).
struct FlowShape {
  int nT, nTL, nTa;
  int near_div = 16;  // near pool: helpers / near_div, at least min(10, helpers / 8), at most 3 nT
  int fold = 1;       // 0: no next-block tasks (the caller applies the block update as a GEMM)
  int rows() const { return nT + nTL + nTa; }
};

// Tiles the chain takes over: the sub-diagonal (k+1, k) and diagonal
// (k+1, k+1) tiles for k + 1 < nT.  Their helper updates go through the
// hand-off buffer.
__host__ __device__ inline bool flow_is_handoff(int nT, int rho, int c) {
  return rho < nT && ((rho == c && c >= 1) || (rho == c + 1));
}

// Helper task lists for one panel shape.  tasks[off[h] .. off[h+1]) belong to
// helper h (0-based), in execution order.
struct FlowSchedule {
  std::vector<uint32_t> tasks;
  std::vector<int> off;
  int helpers = 0;
};

// Arguments of one panel launch (one problem's block i).
struct FlowArgs {
  double* D;                 // D_i (b x b)
  double* L;                 // L_i or nullptr
  double* B;                 // B_i (a x b) or nullptr
  double* Dn;                // D_{i+1} or nullptr (last block)
  double* Bn;                // B_{i+1} or nullptr
  double* T;                 // tip (a x a) or nullptr
  double* winv;              // nT tiles: inverse of each diagonal tile of D_i
  double* xbuf;              // R * nT tiles: exchange copy of every finished tile
  double* hacc;              // 2 * nT tiles: hand-off accumulators (sub-diagonal, diagonal)
  double* logpart;           // nT partial log sums
  unsigned long long* flags; // R*nT finished tiles, nT tile inverses, 2*nT hand-offs
  const uint32_t* tasks;
  const int* off;            // helpers + 1
  const unsigned long long* epoch;
  unsigned long long* info;
  int* err;
  long long row0;
  unsigned long long spin_ticks;  // bound on one wait, in s_memrealtime ticks
  int b, a, nT, nTL, nTa;
  int blk;                   // block index + 1 (part of the flag tag)
  int helpers;               // helper workgroups of this launch entry
};

// Entries of one batched launch, passed by value: pointers in kernel
// arguments are known to the compiler to be global (read from a table in
// memory they would be generic and every access a FLAT one).
#define FLOW_BATCH_MAX 8
// Workgroup layout of a batched launch.  iso == 0: entry q owns workgroups
// [begin[q], begin[q+1]), its chain first.  iso == 1 (two workgroups per CU,
// workgroup p sharing its CU with workgroup p + cus): the chains are
// workgroups 0 .. nent-1, their CU partners cus .. cus+nent-1 stay idle, so
// no helper slows a chain; the other workgroups, in order, are the helpers,
// entry q owning helper numbers [begin[q], begin[q+1]).
struct FlowBatch {
  FlowArgs e[FLOW_BATCH_MAX];
  int begin[FLOW_BATCH_MAX + 1];
  int nent;
  int iso, cus;
};

int flow_owner(const FlowShape& s, int rho, int c, int helpers);
// delta >= 1: level order, far tiles' updates grouped by delta columns;
// delta == 0: list order from a simulated launch.  False on an internal
// inconsistency (a dependency cycle).
bool flow_build_schedule(const FlowShape& s, int helpers, int delta, FlowSchedule& out);

}  // namespace bta
