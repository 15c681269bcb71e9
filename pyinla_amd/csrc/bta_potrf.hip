// BTA / BT block Cholesky (pobtaf / pobtf) and its log-determinant on gfx950.
//
// The factorisation of block column i is a right-looking tile Cholesky of the
// stacked panel [D_i; L_i; B_i] (64x64 tiles), one launch per tile column k:
//   launch k:  every trailing tile (column >= k) of the panel receives the
//              rank-64 update from the already factored column k-1, then the
//              tiles of column k are finished:
//                - workgroup 0 owns the diagonal tile (k,k): update, Cholesky +
//                  inverse in registers/LDS, store, publish a flag;
//                - the workgroups owning the tiles below it wait on that flag
//                  and apply X <- X * inv(L_kk)^T with FP64 MFMA.
// Workgroups are dispatched in index order, so the flag producer (index 0) is
// always resident before any consumer; consumers additionally spin with a
// bounded limit and report a timeout instead of hanging.
// After the panel, one grouped launch applies the block updates
//   D_{i+1} -= L_i L_i^T (lower),  B_{i+1} -= B_i L_i^T,  T -= B_i B_i^T (lower)
// and the arrowhead tip is factored at the end.  Per-tile log-pivot sums are
// reduced in a fixed order, so the log-determinant is deterministic.
#include <algorithm>
#include <mutex>

#include "bta_internal.h"
#include "../../include/pyinla_bta.h"
#include "bta_tile_chol.h"

namespace bta {

// ---------------------------------------------------------------- bump kernel

__global__ void k_begin_call(unsigned long long* epoch, unsigned long long* info, int* err) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    epoch[0] = epoch[0] + 1ull;
    info[0] = kInfoNone;
    err[0] = 0;
  }
}

// ------------------------------------------------------------ panel column k

struct PanelArgs {
  double* D;                 // D_i, b x b, ld b
  double* L;                 // L_i (b x b) or nullptr (last block)
  double* B;                 // B_i (a x b) or nullptr (a == 0)
  double* Winv;              // nT tiles of 64x64: inverse of each diagonal tile
  double* logpart;           // nT partial log sums for this block
  unsigned long long* flags; // nT flags for this block
  const unsigned long long* epoch;
  unsigned long long* info;
  int* err;
  long long row0;            // global row of the block's first row
  unsigned long long spin_limit;
  unsigned long long blk;    // mixed into the flag value: flags shared by several blocks stay distinct
  int b, a, nT, nTL, nTa, k;
};

struct TileRef {
  int kind;  // 0 = D, 1 = L, 2 = B, 3 = zero-fill (strict upper tile of D)
  int r, j;  // tile row within its part, tile column
};

// Decode workgroup index -> tile for launch k (see file header for the order).
__device__ __forceinline__ TileRef decode_panel_tile(const PanelArgs& p, int idx) {
  TileRef t;
  const int k = p.k;
  if (idx == 0) { t.kind = 0; t.r = k; t.j = k; return t; }
  idx -= 1;
  // column-k tiles below the diagonal
  const int nDb = p.nT - 1 - k;
  if (idx < nDb) { t.kind = 0; t.r = k + 1 + idx; t.j = k; return t; }
  idx -= nDb;
  if (idx < p.nTL) { t.kind = 1; t.r = idx; t.j = k; return t; }
  idx -= p.nTL;
  if (idx < p.nTa) { t.kind = 2; t.r = idx; t.j = k; return t; }
  idx -= p.nTa;
  // trailing D tiles: columns j in (k, nT), rows r in [j, nT), column-major
  for (int j = k + 1; j < p.nT; ++j) {
    const int cnt = p.nT - j;
    if (idx < cnt) { t.kind = 0; t.r = j + idx; t.j = j; return t; }
    idx -= cnt;
  }
  const int ntr = p.nT - 1 - k;
  if (idx < p.nTL * ntr) { t.kind = 1; t.r = idx / ntr; t.j = k + 1 + idx % ntr; return t; }
  idx -= p.nTL * ntr;
  if (idx < p.nTa * ntr) { t.kind = 2; t.r = idx / ntr; t.j = k + 1 + idx % ntr; return t; }
  idx -= p.nTa * ntr;
  // launch 0 only: zero the strict upper tiles (r < j) of D, row-major
  t.kind = 3;
  for (int r = 0; r < p.nT; ++r) {
    const int cnt = p.nT - 1 - r;
    if (idx < cnt) { t.r = r; t.j = r + 1 + idx; return t; }
    idx -= cnt;
  }
  t.r = 0; t.j = 0; t.kind = 4;  // out of range (never launched)
  return t;
}

__host__ __device__ inline int panel_grid(int nT, int nTL, int nTa, int k) {
  const int ntr = nT - 1 - k;
  int n = 1 + ntr + nTL + nTa + ntr * (ntr + 1) / 2 + (nTL + nTa) * ntr;
  if (k == 0) n += nT * (nT - 1) / 2;
  return n;
}

__device__ __forceinline__ void panel_body(const PanelArgs& p, int idx, double* __restrict__ sA,
                                           double* __restrict__ sB) {
  const TileRef t = decode_panel_tile(p, idx);
  const unsigned long long ep = (*p.epoch << 20) | p.blk;
  const int b = p.b, k = p.k;

  if (t.kind == 3) {  // zero a strict-upper tile of D
    const int rows = min(64, b - t.r * 64), cols = min(64, b - t.j * 64);
    double* g = p.D + (int64_t)(t.r * 64) * b + t.j * 64;
    const int c = threadIdx.x & 63;
    for (int r = threadIdx.x >> 6; r < rows; r += 4)
      if (c < cols) g[(int64_t)r * b + c] = 0.0;
    return;
  }
  if (t.kind > 3) return;

  // Pointer and extent of the target tile.
  double* base = (t.kind == 0) ? p.D : (t.kind == 1 ? p.L : p.B);
  const int prow = (t.kind == 2) ? p.a : b;  // rows of the part
  const int rows = min(64, prow - t.r * 64);
  const int cols = min(64, b - t.j * 64);
  double* ct = base + (int64_t)(t.r * 64) * b + t.j * 64;

  TileAcc acc;
  acc_load(acc, ct, b, rows, cols);
  if (k > 0) {
    // rank-64 update from the factored column k-1: C -= P(r, k-1) * D(j, k-1)^T
    const double* pa = base + (int64_t)(t.r * 64) * b + (k - 1) * 64;
    const double* pb = p.D + (int64_t)(t.j * 64) * b + (k - 1) * 64;
    lds_load_tile(sA, pa, b, rows, 64);
    lds_load_tile(sB, pb, b, cols, 64);
    __syncthreads();
    mma_nt64<true>(acc, sA, sB);
    __syncthreads();
  }

  if (t.j != k) {  // trailing tile: only the update
    acc_store(acc, ct, b, rows, cols);
    return;
  }

  if (t.kind == 0 && t.r == k) {
    // Diagonal tile: Cholesky + inverse, then publish.
    acc_to_lds(acc, sA);
    __syncthreads();
    const int mv = rows;  // == cols
    const TileCholResult res = tile_chol_inv(sA, sB, mv);
    // store L_kk (full tile: zero strict upper) and inv(L_kk)
    {
      const int c = threadIdx.x & 63;
      double* wk = p.Winv + (int64_t)k * 4096;
      for (int r = threadIdx.x >> 6; r < 64; r += 4) {
        if (r < mv && c < mv) ct[(int64_t)r * b + c] = sA[r * BTA_LDS_LD + c];
        st_sys(&wk[r * 64 + c], sB[r * BTA_LDS_LD + c]);  // read by this launch's TRSM tiles
      }
    }
    if (threadIdx.x == 0) {
      p.logpart[k] = res.logsum;
      if (res.fail >= 0)
        atomicMin(p.info, (unsigned long long)(p.row0 + (long long)k * 64 + res.fail));
    }
    flag_publish(&p.flags[k], ep);
    return;
  }

  // Tile below the diagonal in column k: X = C * inv(L_kk)^T.
  acc_to_lds(acc, sA);
  if (!flag_wait(&p.flags[k], ep, p.spin_limit)) {
    if (threadIdx.x == 0) atomicOr(p.err, 1);
  }
  {
    const double* wk = p.Winv + (int64_t)k * 4096;
    const int c = threadIdx.x & 63;
    for (int r = threadIdx.x >> 6; r < 64; r += 4) sB[r * BTA_LDS_LD + c] = ld_sys(&wk[r * 64 + c]);
  }
  __syncthreads();
  TileAcc x;
  acc_zero(x);
  // inv(L_kk) is lower triangular: output column c only needs k-index <= c.
  const int kend = ((threadIdx.x >> 6) & 1) * 32 + 32;
  mma_nt<false>(x, sA, sB, 0, kend);
  acc_store(x, ct, b, rows, cols);
}

__global__ void __launch_bounds__(BTA_THREADS, 2) k_panel_step(PanelArgs p) {
  __shared__ double sA[64 * BTA_LDS_LD];
  __shared__ double sB[64 * BTA_LDS_LD];
  panel_body(p, blockIdx.x, sA, sB);
}

// Batched form: one launch covers tile column k of block step i of every
// problem in the batch that has such a step.  ent_begin[e] is the first flat
// workgroup of entry e (ent_begin[nent] = total), ent_prob[e] its problem.
// Each problem's diagonal-tile workgroup precedes its consumers in the flat
// order, so the in-order dispatch argument of k_panel_step still holds.
__device__ __forceinline__ PanelArgs panel_args_of(const ProbDesc& d, int i, int k) {
  PanelArgs p;
  const int b = d.b, a = d.a, n = d.n;
  p.D = d.D + (int64_t)i * b * b;
  p.L = (i < n - 1) ? d.L + (int64_t)i * b * b : nullptr;
  p.B = (a > 0) ? d.B + (int64_t)i * a * b : nullptr;
  p.Winv = d.winv + (int64_t)i * d.nT * 4096;
  p.logpart = d.logpart + (int64_t)i * d.nT;
  p.flags = d.flags + (int64_t)i * d.nT;
  p.epoch = d.epoch;
  p.info = d.info;
  p.err = d.err;
  p.row0 = (long long)i * b;
  p.spin_limit = d.spin_limit;
  p.blk = 0;
  p.b = b;
  p.a = a;
  p.nT = d.nT;
  p.nTL = p.L ? d.nT : 0;
  p.nTa = p.B ? d.nTa : 0;
  p.k = k;
  return p;
}

__global__ void __launch_bounds__(BTA_THREADS, 2) k_panel_step_batched(const ProbDesc* __restrict__ probs,
                                                                   const int* __restrict__ ent_prob,
                                                                   const int* __restrict__ ent_begin, int nent,
                                                                   int i, int k) {
  __shared__ double sA[64 * BTA_LDS_LD];
  __shared__ double sB[64 * BTA_LDS_LD];
  int lo = 0, hi = nent - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ent_begin[mid] <= (int)blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const PanelArgs p = panel_args_of(probs[ent_prob[lo]], i, k);
  panel_body(p, blockIdx.x - ent_begin[lo], sA, sB);
}

__global__ void k_begin_batch(const ProbDesc* probs, int nprob) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nprob) {
    probs[q].epoch[0] = probs[q].epoch[0] + 1ull;
    probs[q].info[0] = kInfoNone;
    probs[q].err[0] = 0;
  }
}

// ------------------------------------------------------- block update GEMMs

// One grouped launch: a list of "C -= A * B^T" jobs over 64x64 output tiles,
// each with its own pointers, extents and K.  lower != 0 restricts the job to
// tiles on or below the diagonal (symmetric update).
struct UpdJob {
  double* C;
  const double* A;
  const double* B;
  int m, n, kdim;     // C is m x n, A is m x kdim, B is n x kdim (all row-major)
  int ldc, lda, ldb;
  int lower;
  int tile_begin;     // first flat tile index of this job
  int tiles_n;        // tiles along n
};

#define BTA_MAX_JOBS 8
struct UpdJobs {
  UpdJob job[BTA_MAX_JOBS];
  int njobs;
};

__device__ __forceinline__ void decode_job_tile(const UpdJob& g, int idx, int& tr, int& tc) {
  if (g.lower) {
    // row-major over the lower triangle: row r has r+1 tiles
    int r = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= idx) ++r;
    while (r * (r + 1) / 2 > idx) --r;
    tr = r;
    tc = idx - r * (r + 1) / 2;
  } else {
    tr = idx / g.tiles_n;
    tc = idx % g.tiles_n;
  }
}

__global__ void __launch_bounds__(BTA_THREADS) k_gemm_nt_sub(UpdJobs jobs) {
  __shared__ double sA[64 * BTA_LDS_LD];
  __shared__ double sB[64 * BTA_LDS_LD];
  int jid = 0;
  for (int q = 1; q < jobs.njobs; ++q)
    if ((int)blockIdx.x >= jobs.job[q].tile_begin) jid = q;
  const UpdJob g = jobs.job[jid];
  int tr, tc;
  decode_job_tile(g, blockIdx.x - g.tile_begin, tr, tc);
  const int rows = min(64, g.m - tr * 64), cols = min(64, g.n - tc * 64);
  if (rows <= 0 || cols <= 0) return;
  double* ct = g.C + (int64_t)(tr * 64) * g.ldc + tc * 64;
  const double* at = g.A + (int64_t)(tr * 64) * g.lda;
  const double* bt = g.B + (int64_t)(tc * 64) * g.ldb;
  TileAcc acc;
  acc_load(acc, ct, g.ldc, rows, cols);
  TilePrefetch pa, pb;
  if (g.kdim > 0) {
    pa.load(at, g.lda, 0, rows, min(64, g.kdim));
    pb.load(bt, g.ldb, 0, cols, min(64, g.kdim));
  }
  for (int k0 = 0; k0 < g.kdim; k0 += 64) {
    const int kw = min(64, g.kdim - k0);
    pa.store(sA, 0);
    pb.store(sB, 0);
    __syncthreads();
    if (k0 + 64 < g.kdim) {  // fetch the next chunk while this one is multiplied
      const int kn = min(64, g.kdim - k0 - 64);
      pa.load(at + k0 + 64, g.lda, 0, rows, kn);
      pb.load(bt + k0 + 64, g.ldb, 0, cols, kn);
    }
    mma_nt<true>(acc, sA, sB, 0, (kw + 3) & ~3);
    __syncthreads();
  }
  if (g.lower && tr == tc) acc_store_lower(acc, ct, g.ldc, rows, cols);
  else acc_store(acc, ct, g.ldc, rows, cols);
}

// ------------------------------------------------------------- tip / logdet

__global__ void __launch_bounds__(BTA_THREADS) k_tip_potrf(double* T, int a, double* logpart_tip,
                                                          unsigned long long* info, long long row0,
                                                          double* tinv) {
  __shared__ double sA[64 * BTA_LDS_LD];
  __shared__ double sB[64 * BTA_LDS_LD];
  lds_load_tile(sA, T, a, a, a);
  __syncthreads();
  const TileCholResult res = tile_chol_inv(sA, sB, a);
  const int c = threadIdx.x & 63;
  for (int r = threadIdx.x >> 6; r < 64; r += 4) {
    if (r < a && c < a) T[(int64_t)r * a + c] = sA[r * BTA_LDS_LD + c];
    tinv[r * 64 + c] = sB[r * BTA_LDS_LD + c];
  }
  if (threadIdx.x == 0) {
    logpart_tip[0] = res.logsum;
    if (res.fail >= 0) atomicMin(info, (unsigned long long)(row0 + res.fail));
  }
}

// Ordered reduction of npart partial sums: thread t sums t, t+256, ... in
// order, then a fixed pairwise tree.  Each partial is a sum of log pivots,
// and a pivot is the square of the factor's diagonal entry, so the plain sum
// is the log-determinant 2 sum log diag L.
__global__ void __launch_bounds__(BTA_THREADS) k_logdet_reduce(const double* part, int npart, double* out) {
  __shared__ double s[BTA_THREADS];
  double acc = 0.0;
  for (int i = threadIdx.x; i < npart; i += BTA_THREADS) acc += part[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = BTA_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

__global__ void __launch_bounds__(BTA_THREADS) k_tip_potrf_batched(const ProbDesc* probs) {
  __shared__ double sA[64 * BTA_LDS_LD];
  __shared__ double sB[64 * BTA_LDS_LD];
  const ProbDesc& d = probs[blockIdx.x];
  const int a = d.a;
  if (a <= 0) return;
  lds_load_tile(sA, d.T, a, a, a);
  __syncthreads();
  const TileCholResult res = tile_chol_inv(sA, sB, a);
  const int c = threadIdx.x & 63;
  for (int r = threadIdx.x >> 6; r < 64; r += 4) {
    if (r < a && c < a) d.T[(int64_t)r * a + c] = sA[r * BTA_LDS_LD + c];
    d.tinv[r * 64 + c] = sB[r * BTA_LDS_LD + c];
  }
  if (threadIdx.x == 0) {
    d.logpart[(int64_t)d.n * d.nT] = res.logsum;
    if (res.fail >= 0) atomicMin(d.info, (unsigned long long)((long long)d.n * d.b + res.fail));
  }
}

__global__ void __launch_bounds__(BTA_THREADS) k_logdet_reduce_batched(const ProbDesc* probs) {
  __shared__ double s[BTA_THREADS];
  const ProbDesc& d = probs[blockIdx.x];
  const int npart = d.n * d.nT + (d.a > 0 ? 1 : 0);
  double acc = 0.0;
  for (int i = threadIdx.x; i < npart; i += BTA_THREADS) acc += d.logpart[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = BTA_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) d.logdet[0] = s[0];
}

// 2 sum_j log X[j][j] of block blockIdx.x (the tip T for blockIdx.x == n),
// strided per thread then a fixed tree: part[blockIdx.x].
__global__ void __launch_bounds__(BTA_THREADS) k_block_logdiag(const double* D, const double* T, int n, int b, int a,
                                                              double* part) {
  __shared__ double s[BTA_THREADS];
  const int i = blockIdx.x;
  const double* X = i < n ? D + (int64_t)i * b * b : T;
  const int m = i < n ? b : a;
  double acc = 0.0;
  for (int j = threadIdx.x; j < m; j += BTA_THREADS) acc += log(X[(int64_t)j * m + j]);
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = BTA_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[i] = 2.0 * s[0];
}

int logdet_enqueue(const double* D, const double* T, int n, int b, int a, double* part, double* out,
                   hipStream_t st) {
  const int np = n + (a > 0 ? 1 : 0);
  hipLaunchKernelGGL(k_block_logdiag, dim3(np), dim3(BTA_THREADS), 0, st, D, T, n, b, a, part);
  hipLaunchKernelGGL(k_logdet_reduce, dim3(1), dim3(BTA_THREADS), 0, st, part, np, out);
  return (int)hipGetLastError();
}

// ------------------------------------------------------------ layout check

__global__ void k_selftest_mfma(double* out) {
  const int l = threadIdx.x;
  const double a = (double)(l % 16) + 100.0 * (l / 16);   // A[i][k], i = l%16, k = l/16
  const double b = (double)(l % 16) + 1000.0 * (l / 16);  // B[k][j], j = l%16, k = l/16
  f64x4 c = {0.0, 0.0, 0.0, 0.0};
  c = mfma_f64(a, b, c);
  for (int r = 0; r < 4; ++r) out[l * 4 + r] = c[r];
}

int selftest_mfma_enqueue(double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_selftest_mfma, dim3(1), dim3(64), 0, st, out);
  return (int)hipGetLastError();
}

// -------------------------------------------------------------- host driver

static void launch_gemm_jobs(UpdJobs& jobs, hipStream_t st) {
  int total = 0;
  for (int q = 0; q < jobs.njobs; ++q) {
    UpdJob& g = jobs.job[q];
    const int tm = (g.m + 63) / 64, tn = (g.n + 63) / 64;
    g.tiles_n = tn;
    g.tile_begin = total;
    total += g.lower ? tm * (tm + 1) / 2 : tm * tn;
  }
  if (total > 0) hipLaunchKernelGGL(k_gemm_nt_sub, dim3(total), dim3(BTA_THREADS), 0, st, jobs);
}

static bool use_flow(const Workspace& ws, int b) {
  if (ws.panel_mode == 1) return false;
  if (ws.panel_mode == 2) return true;
  return b > 64;  // a single tile column has no panel dataflow to overlap
}

bool flow_wanted(const Workspace& ws, int b) { return use_flow(ws, b); }

int flow_alloc(Workspace& ws) {
  if (ws.flow) return 0;
  const int nT = (ws.b + 63) / 64, nTa = (ws.a + 63) / 64;
  std::unique_ptr<FlowPlan> fp(new FlowPlan());
  const int rc = fp->build(nT, nTa);
  if (rc) return rc;
  fp->spin_ticks = flow_spin_ticks(ws.device, 20.0);
  ws.flow = std::move(fp);
  return 0;
}

// Whether the persistent panel applies the block update itself.  Measured on
// MI355X (factorizations of Q_prior and Q_cond, ms per block, lone / pair on
// two streams): b = 2000 (32 tile columns) in the panel 0.918 / 1.318, as a
// GEMM after it 0.941 / 1.397; b = 4500 (71 tile columns) in the panel
// 9.10 / 15.47, as a GEMM 7.46 / 14.16.  The helpers are the bound in both
// cases: for wide panels the separate update (one large grouped GEMM) is the
// cheaper way to do its share of the work.
static int flow_fold_of(const Workspace& ws, int b) {
  if (ws.flow_fold == 1 || ws.flow_fold == 2) return ws.flow_fold == 1;
  return (b + 63) / 64 <= 48;
}

static int flow_prepare(Workspace& ws, int b, int a) {
  const int nT = (b + 63) / 64, nTa = (a + 63) / 64;
  int helpers = ws.flow_helpers > 0 ? ws.flow_helpers : flow_helpers_default(ws.device);
  if (helpers <= 0) return (int)hipErrorInvalidValue;
  helpers = std::min(helpers, 1023);
  // the exchange buffers come with the workspace (bta_ws_create, or the
  // option call that selected the flow panel)
  if (!ws.flow || ws.flow->nT != nT || ws.flow->nTa < nTa) return (int)hipErrorNotReady;
  return ws.flow->sched(nTa, helpers, ws.flow_delta, ws.flow_near, flow_fold_of(ws, b), &ws.flow_sched);
}

int potrf_prepare(Workspace& ws, double* D, double* L, double* B, double* T, int n, int b, int a) {
  if (use_flow(ws, b)) {
    const int rc = flow_prepare(ws, b, a);
    if (rc) return rc;
  }
  ws.update_cur = nullptr;
  // grouped-GEMM block updates: with 128-wide tiles, or after a persistent
  // panel that does not apply them itself
  const bool sep = use_flow(ws, b) && !flow_fold_of(ws, b);
  const int tsu = choose_tile(ws.tile_update, b, b);
  if (tsu != 128 && !sep) return 0;
  const PlanKey key{{D, L, B, T}, a, tsu, 0};
  int rc = 0;
  ws.update_cur = plan_lookup(ws, ws.update_plans, key, [&](DevPlan& plan) {
    std::vector<GemmLaunch> ls;
    for (int i = 0; i < n; ++i) {
      GemmLaunch g(tsu);
      double* Li = (i < n - 1) ? L + (int64_t)i * b * b : nullptr;
      double* Bi = (a > 0) ? B + (int64_t)i * a * b : nullptr;
      if (i < n - 1) {
        GemmJob j = job(D + (int64_t)(i + 1) * b * b, b, b, b, 1.0, 1, 0);
        add_seg(j, seg(Li, b, 0, Li, b, 1, b, -1.0));
        g.add(j);
        if (a > 0) {
          GemmJob jb = job(B + (int64_t)(i + 1) * a * b, a, b, b, 1.0);
          add_seg(jb, seg(Bi, b, 0, Li, b, 1, b, -1.0));
          g.add(jb);
        }
      }
      if (a > 0) {
        GemmJob jt = job(T, a, a, a, 1.0, 1, 0);
        add_seg(jt, seg(Bi, b, 0, Bi, b, 1, b, -1.0));
        g.add(jt);
      }
      ls.push_back(std::move(g));
    }
    return plan.build(ls);
  }, &rc);
  return rc;
}

int potrf_enqueue(Workspace& ws, double* D, double* L, double* B, double* T, int n, int b, int a,
                  hipStream_t st) {
  const bool use_plan = ws.update_cur != nullptr;
  const int nT = (b + 63) / 64;
  const int nTa = (a + 63) / 64;
  const bool flow = use_flow(ws, b) && ws.flow && ws.flow->nT == nT && ws.flow_sched && ws.flow_sched->nTa == nTa;
  if (use_flow(ws, b) && !flow) return (int)hipErrorNotReady;  // potrf_prepare not run for this shape
  hipLaunchKernelGGL(k_begin_call, dim3(1), dim3(64), 0, st, ws.epoch, ws.info, ws.err);
  for (int i = 0; i < n; ++i) {
    PanelArgs p;
    p.D = D + (int64_t)i * b * b;
    p.L = (i < n - 1) ? L + (int64_t)i * b * b : nullptr;
    p.B = (a > 0) ? B + (int64_t)i * a * b : nullptr;
    p.Winv = ws.winv + (int64_t)i * nT * 4096;
    p.logpart = ws.logpart + (int64_t)i * nT;
    p.flags = ws.flags + (int64_t)i * nT;
    p.epoch = ws.epoch;
    p.info = ws.info;
    p.err = ws.err;
    p.row0 = (long long)i * b;
    p.spin_limit = ws.spin_limit;
    p.blk = 0;
    p.b = b;
    p.a = a;
    p.nT = nT;
    p.nTL = p.L ? nT : 0;
    p.nTa = p.B ? nTa : 0;
    if (flow) {
      {
        TimedScope ts(ws, KC_PANEL, st);
        flow_launch_block(ws, *ws.flow, *ws.flow_sched, D, L, B, T, i, n, b, a, st);
      }
      if (ws.flow_sched->fold) continue;  // the launch also applied the block update
    } else {
      for (int k = 0; k < nT; ++k) {
        p.k = k;
        TimedScope ts(ws, KC_PANEL, st);
        hipLaunchKernelGGL(k_panel_step, dim3(panel_grid(nT, p.nTL, p.nTa, k)), dim3(BTA_THREADS), 0, st, p);
      }
    }
    UpdJobs jobs;
    jobs.njobs = 0;
    if (i < n - 1) {
      double* Dn = D + (int64_t)(i + 1) * b * b;
      jobs.job[jobs.njobs++] = UpdJob{Dn, p.L, p.L, b, b, b, b, b, b, 1, 0, 0};
      if (a > 0) {
        double* Bn = B + (int64_t)(i + 1) * a * b;
        jobs.job[jobs.njobs++] = UpdJob{Bn, p.B, p.L, a, b, b, b, b, b, 0, 0, 0};
      }
    }
    if (a > 0) jobs.job[jobs.njobs++] = UpdJob{T, p.B, p.B, a, a, b, a, b, b, 1, 0, 0};
    TimedScope ts(ws, KC_UPDATE, st);
    if (use_plan) ws.update_cur->launch(i, st);
    else launch_gemm_jobs(jobs, st);
  }
  int npart = n * nT;
  if (a > 0) {
    hipLaunchKernelGGL(k_tip_potrf, dim3(1), dim3(BTA_THREADS), 0, st, T, a, ws.logpart + npart, ws.info,
                       (long long)n * b, ws.tinv);
    npart += 1;
  }
  hipLaunchKernelGGL(k_logdet_reduce, dim3(1), dim3(BTA_THREADS), 0, st, ws.logpart, npart, ws.logdet);
  return (int)hipGetLastError();
}

BatchPlan::~BatchPlan() {
  if (d_probs) hipFree(d_probs);
  if (d_ent) hipFree(d_ent);
}

int BatchPlan::build(const std::vector<ProbDesc>& ps, int panel_mode, int lanes, int delta, int near_div,
                     int device) {
  probs = ps;
  // Persistent dataflow form when every problem has more than one tile column
  // (or when asked for): each problem gets a chain and an equal share of the
  // device's resident workgroups as helpers, all in one launch per block step.
  bool all_wide = true;
  for (const auto& d : ps) all_wide = all_wide && d.b > 64;
  flow = (panel_mode == 2) || (panel_mode == 0 && all_wide);
  if (flow) {
    // helpers per problem: the device's resident capacity split evenly over
    // the problems still active at each block step
    const int cap = flow_capacity(device);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) cus = 0;
    const int np = (int)ps.size();
    const int nl = std::min(lanes > 0 ? std::min(lanes, np) : std::min(np, 6), FLOW_BATCH_MAX);
    if (cap / nl - 1 < 1) {
      flow = false;  // too many problems to give each a resident chain + helper
    } else {
      int nmax = 0;
      for (const auto& d : ps) nmax = std::max(nmax, d.n);
      // each problem uses its workspace's exchange buffers when it has them
      // (problems run one after another within a lane, and every launch tags
      // its flags with the call epoch, so no two launches share a live flag)
      fplans.clear();
      owned.clear();
      for (const auto& d : ps) {
        const int nTa = d.a > 0 ? d.nTa : 0;
        if (d.fplan && d.fplan->nT == d.nT && d.fplan->nTa >= nTa) {
          fplans.push_back(d.fplan);
          continue;
        }
        owned.emplace_back(new FlowPlan());
        const int rc = owned.back()->build(d.nT, nTa);
        if (rc) return rc;
        owned.back()->spin_ticks = flow_spin_ticks(device, 20.0);
        fplans.push_back(owned.back().get());
      }
      // Problems run in waves of at most `lanes` (longest first), so that each
      // chain gets enough helpers; within a wave every launch covers the
      // wave's problems that still have block step i.
      std::vector<int> order(np);
      for (int q = 0; q < np; ++q) order[q] = q;
      std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return ps[x].n > ps[y].n; });
      for (int w0 = 0; w0 < np; w0 += nl) {
        const int w1 = std::min(np, w0 + nl);
        int wmax = 0;
        for (int t = w0; t < w1; ++t) wmax = std::max(wmax, ps[order[t]].n);
        for (int i = 0; i < wmax; ++i) {
          int nact = 0;
          for (int t = w0; t < w1; ++t) nact += (i < ps[order[t]].n) ? 1 : 0;
          // two workgroups per CU: each chain gets a CU of its own (its
          // partner slot idle), measured on MI355X to shorten the chain's
          // steps as much as a co-resident helper had lengthened them
          const int h_iso = (cap - 2 * nact) / nact;
          const bool iso = cus > 0 && cap == 2 * cus && h_iso >= 1 && nact * (2 + h_iso) >= cus + nact;
          const int helpers = std::min(1023, iso ? h_iso : cap / nact - 1);
          FLaunch L;
          L.fb.nent = 0;
          L.fb.iso = iso ? 1 : 0;
          L.fb.cus = cus;
          int total = 0;
          for (int t = w0; t < w1; ++t) {
            const int q = order[t];
            const ProbDesc& d = ps[q];
            if (i >= d.n) continue;
            const FlowSched* sc = nullptr;
            const int rc = fplans[q]->sched(d.a > 0 ? d.nTa : 0, helpers, delta, near_div, 1, &sc);
            if (rc) return rc;
            L.fb.e[L.fb.nent] = flow_args(*fplans[q], *sc, d.D, d.L, d.B, d.T, d.winv, d.logpart, d.epoch,
                                          d.info, d.err, i, d.n, d.b, d.a);
            const int j = L.fb.nent;
            L.fb.begin[j] = iso ? j * helpers : total;  // iso: helper numbers
            L.fb.nent = j + 1;
            total += 1 + helpers;
          }
          L.fb.begin[L.fb.nent] = iso ? L.fb.nent * helpers : total;
          L.total = iso ? total + L.fb.nent : total;  // + the chains' idle partners
          flaunches.push_back(L);
        }
      }
      (void)nmax;
      hipError_t e;
      if ((e = hipMalloc(&d_probs, ps.size() * sizeof(ProbDesc))) != hipSuccess) return (int)e;
      if ((e = hipMemcpy(d_probs, ps.data(), ps.size() * sizeof(ProbDesc), hipMemcpyHostToDevice)) != hipSuccess)
        return (int)e;
      return 0;
    }
  }
  int nmax = 0;
  for (const auto& d : ps) nmax = std::max(nmax, d.n);
  std::vector<int> ent;  // per launch: [prob..., begin..., total]
  launches.clear();
  std::vector<GemmLaunch> upd;
  int bmin = 1 << 30;
  for (const auto& d : ps) bmin = std::min(bmin, d.b);
  const int ts = pick_tile(bmin, bmin);
  for (int i = 0; i < nmax; ++i) {
    int ntmax = 0;
    for (const auto& d : ps)
      if (i < d.n) ntmax = std::max(ntmax, d.nT);
    for (int k = 0; k < ntmax; ++k) {
      std::vector<int> pr, be;
      int total = 0;
      for (int q = 0; q < (int)ps.size(); ++q) {
        const ProbDesc& d = ps[q];
        if (i >= d.n || k >= d.nT) continue;
        const int nTL = (i < d.n - 1) ? d.nT : 0;
        const int nTa = d.a > 0 ? d.nTa : 0;
        pr.push_back(q);
        be.push_back(total);
        total += panel_grid(d.nT, nTL, nTa, k);
      }
      Launch L;
      L.i = i;
      L.k = k;
      L.nent = (int)pr.size();
      L.total = total;
      L.off_prob = (int)ent.size();
      ent.insert(ent.end(), pr.begin(), pr.end());
      L.off_begin = (int)ent.size();
      ent.insert(ent.end(), be.begin(), be.end());
      ent.push_back(total);
      L.update = -1;
      launches.push_back(L);
    }
    GemmLaunch g(ts);
    for (const auto& d : ps) {
      if (i >= d.n) continue;
      const int b = d.b, a = d.a, n = d.n;
      double* Li = d.L + (int64_t)i * b * b;
      double* Bi = a > 0 ? d.B + (int64_t)i * a * b : nullptr;
      if (i < n - 1) {
        GemmJob j = job(d.D + (int64_t)(i + 1) * b * b, b, b, b, 1.0, 1, 0);
        add_seg(j, seg(Li, b, 0, Li, b, 1, b, -1.0));
        g.add(j);
        if (a > 0) {
          GemmJob jb = job(d.B + (int64_t)(i + 1) * a * b, a, b, b, 1.0);
          add_seg(jb, seg(Bi, b, 0, Li, b, 1, b, -1.0));
          g.add(jb);
        }
      }
      if (a > 0) {
        GemmJob jt = job(d.T, a, a, a, 1.0, 1, 0);
        add_seg(jt, seg(Bi, b, 0, Bi, b, 1, b, -1.0));
        g.add(jt);
      }
    }
    launches.back().update = (int)upd.size();
    upd.push_back(std::move(g));
  }
  hipError_t e;
  if ((e = hipMalloc(&d_probs, ps.size() * sizeof(ProbDesc))) != hipSuccess) return (int)e;
  if ((e = hipMemcpy(d_probs, ps.data(), ps.size() * sizeof(ProbDesc), hipMemcpyHostToDevice)) != hipSuccess)
    return (int)e;
  if ((e = hipMalloc(&d_ent, std::max<size_t>(1, ent.size()) * sizeof(int))) != hipSuccess) return (int)e;
  if (!ent.empty() &&
      (e = hipMemcpy(d_ent, ent.data(), ent.size() * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
    return (int)e;
  return updates.build(upd);
}

int BatchPlan::enqueue(hipStream_t st) const {
  const int np = (int)probs.size();
  hipLaunchKernelGGL(k_begin_batch, dim3((np + 63) / 64), dim3(64), 0, st, d_probs, np);
  if (flow) {  // block updates are folded into the panel launches
    for (const FLaunch& L : flaunches)
      flow_launch_batched(L.fb, L.total, st);
    hipLaunchKernelGGL(k_tip_potrf_batched, dim3(np), dim3(BTA_THREADS), 0, st, d_probs);
    hipLaunchKernelGGL(k_logdet_reduce_batched, dim3(np), dim3(BTA_THREADS), 0, st, d_probs);
    return (int)hipGetLastError();
  }
  for (const Launch& L : launches) {
    if (L.total > 0)
      hipLaunchKernelGGL(k_panel_step_batched, dim3(L.total), dim3(BTA_THREADS), 0, st, d_probs,
                         d_ent + L.off_prob, d_ent + L.off_begin, L.nent, L.i, L.k);
    if (L.update >= 0) updates.launch(L.update, st);
  }
  hipLaunchKernelGGL(k_tip_potrf_batched, dim3(np), dim3(BTA_THREADS), 0, st, d_probs);
  hipLaunchKernelGGL(k_logdet_reduce_batched, dim3(np), dim3(BTA_THREADS), 0, st, d_probs);
  return (int)hipGetLastError();
}

}  // namespace bta

// ------------------------------------------------ factorization without allocation
//
// bta_potrf_noalloc runs the per-column panel path above (k_panel_step and the
// 64x64-tile block updates, whose job lists travel as kernel arguments) with
// every buffer beyond the caller's blocks and outputs in one file-scope device
// array, `scratch`, of fixed size.  Layout, in doubles:
//   [0] call counter (epoch)  [1] first failing pivot  [2] error bits  [3] logdet
//   [8, 8 + NA_MAX_NT)                    readiness flags of the tile columns,
//                                         shared by all blocks of a call (the
//                                         flag value carries the block index)
//   NA_TINV  (4096)                        inverse of the 64x64 tip tile
//   NA_WINV  (NA_MAX_NT * 4096)            inverses of the current block's
//                                         diagonal tiles (rewritten per block)
//   NA_PART  (NA_MAX_PARTS)                per-tile log-pivot sums of the call
// The counter starts at zero (static storage) and only grows, so flags never
// need clearing; the pivot and error words are reset by k_begin_call.
namespace bta {

static constexpr int NA_MAX_NT = 64;          // b <= 4096
static constexpr int NA_MAX_PARTS = 1 << 17;  // n * ceil(b / 64) + 1 tile log-sums
static constexpr size_t NA_TINV = 8 + NA_MAX_NT;
static constexpr size_t NA_WINV = NA_TINV + 4096;
static constexpr size_t NA_PART = NA_WINV + (size_t)NA_MAX_NT * 4096;
static constexpr size_t NA_TOTAL = NA_PART + NA_MAX_PARTS;  // 3.1 MiB

static __device__ __align__(256) double scratch[NA_TOTAL];

// Device address of `scratch` per device, taken once with that device current.
static std::mutex na_mu;
static double* na_ptr[256];

static int noalloc_scratch(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= 256) return BTA_E_ARG;
  std::lock_guard<std::mutex> lk(na_mu);
  if (!na_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    if (!ptr) return (int)hipErrorInvalidSymbol;
    na_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = na_ptr[dev];
  return 0;
}

// Caller-visible results of the call: logdet, LAPACK-style info, error code.
__global__ void k_noalloc_results(const double* sc, double* logdet, int64_t* info, int* err) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const unsigned long long inf = reinterpret_cast<const unsigned long long*>(sc)[1];
  const int e = reinterpret_cast<const int*>(sc + 2)[0];
  if (logdet) logdet[0] = sc[3];
  if (info) info[0] = (inf == kInfoNone) ? 0 : (int64_t)inf + 1;
  if (err) err[0] = e ? BTA_E_TIMEOUT : 0;
}

}  // namespace bta

extern "C" int bta_potrf_noalloc(double* D, double* L, double* B, double* T, int n, int b, int a, double* logdet,
                                 int64_t* info, int* err, void* stream) {
  using namespace bta;
  if (!D || n < 1 || b < 1 || a < 0 || a > 64) return BTA_E_ARG;
  if (n > 1 && !L) return BTA_E_ARG;
  if (a > 0 && (!B || !T)) return BTA_E_ARG;
  const int nT = (b + 63) / 64, nTa = (a + 63) / 64;
  if (nT > NA_MAX_NT || (long long)n * nT + 1 > NA_MAX_PARTS || n >= (1 << 20)) return BTA_E_SCRATCH;
  double* sc = nullptr;
  int rc = noalloc_scratch(&sc);
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  auto* u = reinterpret_cast<unsigned long long*>(sc);
  unsigned long long* epoch = u;
  unsigned long long* inf = u + 1;
  int* er = reinterpret_cast<int*>(u + 2);
  hipLaunchKernelGGL(k_begin_call, dim3(1), dim3(64), 0, st, epoch, inf, er);
  for (int i = 0; i < n; ++i) {
    PanelArgs p;
    p.D = D + (int64_t)i * b * b;
    p.L = (i < n - 1) ? L + (int64_t)i * b * b : nullptr;
    p.B = (a > 0) ? B + (int64_t)i * a * b : nullptr;
    p.Winv = sc + NA_WINV;
    p.logpart = sc + NA_PART + (int64_t)i * nT;
    p.flags = u + 8;
    p.epoch = epoch;
    p.info = inf;
    p.err = er;
    p.row0 = (long long)i * b;
    p.spin_limit = 1ull << 24;
    p.blk = (unsigned long long)i;
    p.b = b;
    p.a = a;
    p.nT = nT;
    p.nTL = p.L ? nT : 0;
    p.nTa = p.B ? nTa : 0;
    for (int k = 0; k < nT; ++k) {
      p.k = k;
      hipLaunchKernelGGL(k_panel_step, dim3(panel_grid(nT, p.nTL, p.nTa, k)), dim3(BTA_THREADS), 0, st, p);
    }
    UpdJobs jobs;
    jobs.njobs = 0;
    if (i < n - 1) {
      double* Dn = D + (int64_t)(i + 1) * b * b;
      jobs.job[jobs.njobs++] = UpdJob{Dn, p.L, p.L, b, b, b, b, b, b, 1, 0, 0};
      if (a > 0) {
        double* Bn = B + (int64_t)(i + 1) * a * b;
        jobs.job[jobs.njobs++] = UpdJob{Bn, p.B, p.L, a, b, b, b, b, b, 0, 0, 0};
      }
    }
    if (a > 0) jobs.job[jobs.njobs++] = UpdJob{T, p.B, p.B, a, a, b, a, b, b, 1, 0, 0};
    launch_gemm_jobs(jobs, st);
  }
  int npart = n * nT;
  if (a > 0) {
    hipLaunchKernelGGL(k_tip_potrf, dim3(1), dim3(BTA_THREADS), 0, st, T, a, sc + NA_PART + npart, inf,
                       (long long)n * b, sc + NA_TINV);
    npart += 1;
  }
  hipLaunchKernelGGL(k_logdet_reduce, dim3(1), dim3(BTA_THREADS), 0, st, sc + NA_PART, npart, sc + 3);
  hipLaunchKernelGGL(k_noalloc_results, dim3(1), dim3(64), 0, st, sc, logdet, info, err);
  return (int)hipGetLastError();
}
