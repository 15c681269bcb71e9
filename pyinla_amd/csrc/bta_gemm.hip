// Grouped FP64 GEMM and GEMV kernels + device job-table plans.
#include <algorithm>
#include <stdexcept>

#include "bta_gemm.h"
#include "bta_internal.h"

namespace bta {

// K chunk of the 64x64-tile kernel: 32 keeps the LDS footprint at 34 KB, so
// four workgroups share a CU (latency hiding, and 1024 resident tiles)
#ifndef BTA_GEMM_KC
#define BTA_GEMM_KC 32
#endif
#define BTA_GEMM_LD (BTA_GEMM_KC + 1)  // odd stride: conflict-free 16-row fragment reads
#ifndef BTA_GEMM_PAIRS
#define BTA_GEMM_PAIRS 1  // 16-byte loads of interior chunks where aligned
#endif
#ifndef BTA_GEMM_MINB
#define BTA_GEMM_MINB 4  // resident workgroups per CU the register budget is cut for
#endif

// acc += sA(64 x K) * sB(64 x K)^T over k in [0, kend) (LDS stride LD; any
// alpha was applied to sA when it was stored).
template <int LD>
__device__ __forceinline__ void mma_nt_ld(TileAcc& t, const double* __restrict__ sA,
                                          const double* __restrict__ sB, int kend) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  const double* a0p = sA + (wm + li) * LD + lk;
  const double* a1p = sA + (wm + 16 + li) * LD + lk;
  const double* b0p = sB + (wn + li) * LD + lk;
  const double* b1p = sB + (wn + 16 + li) * LD + lk;
#pragma unroll 4
  for (int k = 0; k < kend; k += 4) {
    const double a0 = a0p[k], a1 = a1p[k], b0 = b0p[k], b1 = b1p[k];
    t.c[0][0] = mfma_f64(a0, b0, t.c[0][0]);
    t.c[0][1] = mfma_f64(a0, b1, t.c[0][1]);
    t.c[1][0] = mfma_f64(a1, b0, t.c[1][0]);
    t.c[1][1] = mfma_f64(a1, b1, t.c[1][1]);
  }
}

__device__ __forceinline__ int find_job(const int* __restrict__ prefix, int njobs, int bid) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= bid) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Output tile of workgroup `local` of a tm x tn job.  Workgroups are
// dispatched to the 8 XCDs round-robin (blockIdx % 8); when the job's tiles
// start on an XCD boundary and split evenly, XCD x takes a contiguous run of
// the tiles ordered in bands of 8 tile rows (column by column inside a band),
// i.e. a compact 8 x (T/64) rectangle whose A and B panels its L2 shares.
#ifndef BTA_GEMM_XCD
#define BTA_GEMM_XCD 1
#endif
__device__ __forceinline__ void xcd_tile(int local, int first, int tm, int tn, int& tr, int& tc) {
  const int T = tm * tn;
  int t = local;
  if (BTA_GEMM_XCD && (first & 7) == 0 && (T & 7) == 0 && tm % 8 == 0) t = (local & 7) * (T >> 3) + (local >> 3);
  if (BTA_GEMM_XCD && tm % 8 == 0) {
    const int band = t / (8 * tn), idx = t % (8 * tn);
    tr = band * 8 + (idx & 7);
    tc = idx >> 3;
  } else {
    tr = t / tn;
    tc = t % tn;
  }
}

__device__ __forceinline__ void decode_lower(int idx, int& tr, int& tc) {
  int r = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
  while ((r + 1) * (r + 2) / 2 <= idx) ++r;
  while (r * (r + 1) / 2 > idx) --r;
  tr = r;
  tc = idx - r * (r + 1) / 2;
}

// A 64 x BTA_GEMM_KC operand chunk staged global -> registers -> LDS (same
// conventions as TilePrefetch: LDS layout s[i][k] of op(X)).  Thread t holds
// pairs of contiguous elements: with H = KC / 2 threads per chunk row,
// r[2q + u] is (i, k) = (t / H + (256 / H) q, 2 (t % H) + u) of a row-major
// chunk (trans 0) or (k, i) = (t / 32 + 8q, 2 (t % 32) + u) of a transposed
// one (trans 1), so an interior chunk with an even leading dimension and a
// 16-byte aligned origin is fetched with 16-byte loads.
struct ChunkPrefetch {
  static_assert(BTA_THREADS == 256 && BTA_GEMM_KC % 8 == 0 && BTA_GEMM_KC <= 64, "pair mapping");
  static constexpr int N = 64 * BTA_GEMM_KC / BTA_THREADS;
  static constexpr int H = BTA_GEMM_KC / 2;     // threads per row of a row-major chunk
  static constexpr int RQ = BTA_THREADS / H;    // its rows per pair step
  double r[N];
  __device__ __forceinline__ void load(const double* __restrict__ g0, int64_t ld, int trans, int rows, int kw,
                                       int mask, int i0, int k0) {
    const gdouble* g = gview(g0);
    mask = chunk_mask(mask, i0, rows, k0, kw);
    const int64_t step = trans == 0 ? RQ * ld : 8 * ld;
    const gdouble* t = trans == 0 ? g + (int64_t)(threadIdx.x / H) * ld + 2 * (threadIdx.x % H)
                                  : g + (int64_t)(threadIdx.x >> 5) * ld + 2 * (threadIdx.x & 31);
    if (mask == MASK_NONE && rows >= 64 && kw >= BTA_GEMM_KC) {  // interior chunk: no tests
      if (BTA_GEMM_PAIRS && (ld & 1) == 0 && (reinterpret_cast<uintptr_t>(g0) & 15) == 0) {
#pragma unroll
        for (int q = 0; q < N / 2; ++q) {
          const f64x2 x = *(const gf64x2*)(t + q * step);
          r[2 * q] = x.x;
          r[2 * q + 1] = x.y;
        }
      } else {
#pragma unroll
        for (int q = 0; q < N / 2; ++q) {
          r[2 * q] = t[q * step];
          r[2 * q + 1] = t[q * step + 1];
        }
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < N / 2; ++q)
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        int i, k;
        if (trans == 0) {
          i = threadIdx.x / H + RQ * q;
          k = 2 * (threadIdx.x % H) + u;
        } else {
          k = (threadIdx.x >> 5) + 8 * q;
          i = 2 * (threadIdx.x & 31) + u;
        }
        r[2 * q + u] = (i < rows && k < kw && !masked_out(mask, i0 + i, k0 + k)) ? t[q * step + u] : 0.0;
      }
  }
  __device__ __forceinline__ void store(double* __restrict__ s, int trans, double alpha = 1.0) const {
#pragma unroll
    for (int q = 0; q < N / 2; ++q) {
      if (trans == 0) {
        const int i = threadIdx.x / H + RQ * q, k = 2 * (threadIdx.x % H);
        s[i * BTA_GEMM_LD + k] = alpha * r[2 * q];
        s[i * BTA_GEMM_LD + k + 1] = alpha * r[2 * q + 1];
      } else {
        const int k = (threadIdx.x >> 5) + 8 * q, i = 2 * (threadIdx.x & 31);
        s[i * BTA_GEMM_LD + k] = alpha * r[2 * q];
        s[(i + 1) * BTA_GEMM_LD + k] = alpha * r[2 * q + 1];
      }
    }
  }
};

__global__ void __launch_bounds__(BTA_THREADS, BTA_GEMM_MINB) k_gemm_grouped(const GemmJob* __restrict__ jobs,
                                                             const int* __restrict__ prefix, int njobs) {
  __shared__ double sA[64 * BTA_GEMM_LD];
  __shared__ double sB[64 * BTA_GEMM_LD];
  __shared__ int last_sh;
  const int jid = find_job(prefix, njobs, blockIdx.x);
  const GemmJob& g = jobs[jid];
  const bool split = g.split == 2;
  const int local = split ? (blockIdx.x - prefix[jid]) >> 1 : blockIdx.x - prefix[jid];
  const int half = split ? (blockIdx.x - prefix[jid]) & 1 : 0;
  int tr, tc;
  if (g.lower) decode_lower(local, tr, tc);
  else {
    const int tm = (g.m + 63) / 64, tn = (g.n + 63) / 64;
    xcd_tile(local, prefix[jid], tm, tn, tr, tc);
  }
  const int rows = min(64, g.m - tr * 64), cols = min(64, g.n - tc * 64);
  // A tile at a ragged edge of a job at least 64 wide works on the 64-wide
  // window that ends at the edge (origin r0 / c0, overlapping the previous
  // tile) and writes only its own rows / columns: every chunk of it then
  // takes the same unguarded loads as an interior tile (measured: guarded
  // edge tiles held a 2000 x 2000 product at 397 us against 339 us for
  // 1984 x 1984).  Overlap rows and columns may miss K terms outside the
  // tile's own triangular range; they are never stored.
  const int r0 = (rows < 64 && g.m >= 64) ? g.m - 64 : tr * 64;
  const int c0 = (cols < 64 && g.n >= 64) ? g.n - 64 : tc * 64;
  const int wrows = min(64, g.m - r0), wcols = min(64, g.n - c0);
  double* ct = g.C + (int64_t)r0 * g.ldc + c0;
  TileAcc acc;
  if (g.beta != 0.0 && half == 0) {  // the second half of a split tile starts from zero
    acc_load(acc, ct, g.ldc, wrows, wcols);
    if (g.beta != 1.0) {
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc.c[x][y] *= g.beta;
    }
  } else {
    acc_zero(acc);
  }
  // K chunks of all segments form one pipeline: chunk q+1 is fetched into
  // registers while chunk q is multiplied out of LDS.
  int seg_id = 0, k0 = 0, kmax = 0;
  auto seg_range = [&](int sid, int& kb, int& ke) {
    const GemmSeg& sg = g.seg[sid];
    kb = 0;
    ke = sg.k;
    if (sg.tri & TRI_A_LOWER) ke = min(ke, (tr + 1) * 64);
    if (sg.tri & TRI_A_UPPER) kb = max(kb, tr * 64);
    if (sg.tri & TRI_B_LOWER) kb = max(kb, tc * 64);
    if (sg.tri & TRI_B_UPPER) ke = min(ke, (tc + 1) * 64);
  };
  // advance (seg_id, k0) to the next non-empty chunk; returns false at the end
  auto next_chunk = [&](bool first) -> bool {
    if (!first) k0 += BTA_GEMM_KC;
    while (seg_id < g.nseg) {
      if (first || k0 >= kmax) {
        if (!first) ++seg_id;
        if (seg_id >= g.nseg) return false;
        int kb;
        seg_range(seg_id, kb, kmax);
        k0 = kb;
        first = false;
        if (k0 < kmax) return true;
        continue;
      }
      return true;
    }
    return false;
  };
  ChunkPrefetch pa, pb;
  auto fetch = [&]() {
    const GemmSeg& sg = g.seg[seg_id];
    const int kw = min(BTA_GEMM_KC, kmax - k0);
    const int ma = (sg.tri & TRI_A_LOWER) ? MASK_K_GT_I : ((sg.tri & TRI_A_UPPER) ? MASK_K_LT_I : MASK_NONE);
    const int mb = (sg.tri & TRI_B_LOWER) ? MASK_K_LT_I : ((sg.tri & TRI_B_UPPER) ? MASK_K_GT_I : MASK_NONE);
    if (sg.ta == 0) pa.load(sg.A + (int64_t)r0 * sg.lda + k0, sg.lda, 0, wrows, kw, ma, r0, k0);
    else pa.load(sg.A + (int64_t)k0 * sg.lda + r0, sg.lda, 1, wrows, kw, ma, r0, k0);
    if (sg.tb == 1) pb.load(sg.B + (int64_t)c0 * sg.ldb + k0, sg.ldb, 0, wcols, kw, mb, c0, k0);
    else pb.load(sg.B + (int64_t)k0 * sg.ldb + c0, sg.ldb, 1, wcols, kw, mb, c0, k0);
  };
  // a split tile's halves take the first ceil(nch / 2) and the remaining of
  // its nch chunks
  int qend = 0x7fffffff;
  int q = 0;
  if (split) {
    int nch = 0;
    for (int sid = 0; sid < g.nseg; ++sid) {
      int kb, ke;
      seg_range(sid, kb, ke);
      if (ke > kb) nch += (ke - kb + BTA_GEMM_KC - 1) / BTA_GEMM_KC;
    }
    qend = half ? nch : (nch + 1) / 2;
  }
  bool have = (g.nseg > 0) && next_chunk(true);
  if (half)
    for (const int q0 = qend - qend / 2; have && q < q0; ++q) have = next_chunk(false);
  if (q >= qend) have = false;
  if (have) fetch();
  while (have) {
    const GemmSeg& sg = g.seg[seg_id];
    const int kw = min(BTA_GEMM_KC, kmax - k0);
    pa.store(sA, sg.ta, sg.alpha);
    pb.store(sB, sg.tb == 1 ? 0 : 1);
    __syncthreads();
    have = next_chunk(false) && ++q < qend;
    if (have) fetch();
    mma_nt_ld<BTA_GEMM_LD>(acc, sA, sB, (kw + 3) & ~3);
    lds_barrier();  // not a wait for the next chunk's loads: its stores wait for them
  }
  if (split) {
    // both halves publish their partial sums (write-through stores that have
    // completed before the counter moves); the second to count adds the other
    // half's partial (a + b == b + a: the tile is the same whichever comes
    // second), resets the counter for the next launch and writes the tile
    double* mine = g.part + ((int64_t)local * 2 + half) * 4096;
    acc_store_sys(acc, mine);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (threadIdx.x == 0) {
      const int before = __hip_atomic_fetch_add(g.cnt + local, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      if (before == 1) __hip_atomic_store(g.cnt + local, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      last_sh = before;
    }
    __syncthreads();
    if (last_sh != 1) return;
    const double* other = g.part + ((int64_t)local * 2 + (half ^ 1)) * 4096;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc.c[x][y][r] += ld_sys(&other[acc_row(x, r) * 64 + acc_col(y)]);
  }
  // epilogue
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = acc_row(x, r), j = acc_col(y);
        const int gi = r0 + i, gj = c0 + j;  // global row / column
        if (i >= wrows || j >= wcols || gi < tr * 64 || gj < tc * 64) continue;
        const double v = acc.c[x][y][r];
        gdouble* cg = gview(g.C);
        if (g.lower && tr == tc) {
          if (gi >= gj) {
            cg[(int64_t)gi * g.ldc + gj] = v;
            if (g.mirror && gi > gj) cg[(int64_t)gj * g.ldc + gi] = v;
          }
        } else {
          cg[(int64_t)gi * g.ldc + gj] = v;
          if (g.lower && g.mirror) cg[(int64_t)gj * g.ldc + gi] = v;
        }
      }
}

// ------------------------------------------- 64x64 tiles, pipelined K loop
//
// Variant of the grouped GEMM (BTA_OPT_GEMM_KERNEL 1; measured slower on
// MI355X than the single-buffered k_gemm_grouped above, 470 us against 395 us
// for NN 2000^3, and kept for comparison).  K advances in chunks of 16; the
// A chunk (64 rows of op(A) x 16) is kept in LDS as sA[i][k] (odd stride 17)
// and the B chunk (16 x 64 columns of op(B)) as sB[k][j] (stride 68), so the
// MFMA operand reads of a wave (16 rows x 4 k) hit distinct banks.  Two LDS
// buffers and two register sets: while chunk t is multiplied out of one
// buffer, chunk t+1 (loaded two chunks earlier) is stored to the other and
// the loads of chunk t+3 are issued -- each global load has two chunks of
// MFMA work to arrive.
//
// Every chunk is loaded the same way, as pairs of doubles from addresses
// clamped into the operand (rows past the edge of the matrix, K past the end
// of the segment's range), and everything else -- zero fill at ragged edges
// and partial chunks, triangular masks, alpha -- is applied when the chunk is
// stored to LDS.  The load sequence is then the same straight-line code for
// every chunk, so the compiler can count the loads in flight and wait for one
// register set only.  Past the last chunk the loop re-loads the first one
// into the idle set (never stored).  Needs even leading dimensions and
// 16-byte aligned operands (DevPlan::launch falls back to the single-buffered
// kernel otherwise): a clamped pair then never leaves its matrix row.

#define G2_KC 16
#define G2_LDA (G2_KC + 1)
#define G2_LDB (64 + 4)


// Four elements of one 64 x 16 operand chunk per thread.  form 0 ("K
// contiguous"): element (r, k) at g[r * ld + k], the thread owns rows t/8 and
// t/8 + 32, columns 2 (t%8) and 2 (t%8) + 1; form 1 ("rows contiguous"):
// element (r, k) at g[k * ld + r], the thread owns k = t/32 and t/32 + 8, rows
// 2 (t%32) and 2 (t%32) + 1.  rows / kw: valid rows / K columns of the chunk.
struct G2Op {
  double r[4];
  // (row, k) inside the chunk of element u of pair q, and the memory offset
  // (in elements, from the chunk origin) of pair q before clamping
  __device__ __forceinline__ static void coord(int form, int q, int u, int& i, int& k) {
    const int t = threadIdx.x;
    const int i0 = (t >> 3) + 32 * q, k0 = 2 * (t & 7) + u;  // form 0
    const int i1 = 2 * (t & 31) + u, k1 = (t >> 5) + 8 * q;  // form 1
    i = form ? i1 : i0;
    k = form ? k1 : k0;
  }
  // The same two paired loads for either form: only the addresses differ.
  __device__ __forceinline__ void load(const double* __restrict__ g0, int64_t ld, int form, int rows, int kw) {
    const int t = threadIdx.x;
    const gdouble* g = gview(g0);
    const int kc = min(2 * (t & 7), (kw - 1) & ~1);
    const int rc = min(2 * (t & 31), (rows - 1) & ~1);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int rr = min((t >> 3) + 32 * q, rows - 1);
      const int kq = min((t >> 5) + 8 * q, kw - 1);
      const int64_t off = form ? (int64_t)kq * ld + rc : (int64_t)rr * ld + kc;
      const f64x2 v = *(const gf64x2*)(g + off);
      r[2 * q] = v.x;
      r[2 * q + 1] = v.y;
    }
  }
  // value of element (q, u) after the zero fill and the triangular mask,
  // evaluated without branches (bitwise, not short-circuit)
  __device__ __forceinline__ double val(int form, int q, int u, int rows, int kw, int mask, int i0, int k0) const {
    int i, k;
    coord(form, q, u, i, k);
    const int ig = i0 + i, kg = k0 + k;
    const bool out = (mask == MASK_K_GT_I) & (kg > ig) | (mask == MASK_K_LT_I) & (kg < ig);
    const bool keep = (i < rows) & (k < kw) & !out;
    return keep ? r[2 * q + u] : 0.0;
  }
  // A operand: sA[i][k] scaled by alpha
  __device__ __forceinline__ void store_a(double* __restrict__ sA, int form, double alpha, int rows, int kw, int mask,
                                          int i0, int k0) const {
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        int i, k;
        coord(form, q, u, i, k);
        sA[i * G2_LDA + k] = alpha * val(form, q, u, rows, kw, mask, i0, k0);
      }
  }
  // B operand: sB[k][j]
  __device__ __forceinline__ void store_b(double* __restrict__ sB, int form, int cols, int kw, int mask, int j0,
                                          int k0) const {
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        int i, k;
        coord(form, q, u, i, k);
        sB[k * G2_LDB + i] = val(form, q, u, cols, kw, mask, j0, k0);
      }
  }
};

// One staged chunk: both operands and the chunk they hold.
struct G2Stage {
  G2Op a, b;
  int seg, k0, kw;
};

__global__ void __launch_bounds__(BTA_THREADS, 4) k_gemm_grouped_db(const GemmJob* __restrict__ jobs,
                                                                   const int* __restrict__ prefix, int njobs) {
  __shared__ double sA[2][64 * G2_LDA];
  __shared__ double sB[2][G2_KC * G2_LDB];
  const int jid = find_job(prefix, njobs, blockIdx.x);
  const GemmJob& g = jobs[jid];
  const int local = blockIdx.x - prefix[jid];
  int tr, tc;
  if (g.lower) decode_lower(local, tr, tc);
  else {
    const int tm = (g.m + 63) / 64, tn = (g.n + 63) / 64;
    xcd_tile(local, prefix[jid], tm, tn, tr, tc);
  }
  const int rows = min(64, g.m - tr * 64), cols = min(64, g.n - tc * 64);
  double* ct = g.C + (int64_t)(tr * 64) * g.ldc + tc * 64;
  TileAcc acc;
  if (g.beta != 0.0) {
    acc_load(acc, ct, g.ldc, rows, cols);
    if (g.beta != 1.0) {
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc.c[x][y] *= g.beta;
    }
  } else {
    acc_zero(acc);
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  auto mma = [&](int buf) {
    const double* a0p = &sA[buf][(wm + li) * G2_LDA + lk];
    const double* a1p = a0p + 16 * G2_LDA;
    const double* b0p = &sB[buf][lk * G2_LDB + wn + li];
    const double* b1p = b0p + 16;
#pragma unroll
    for (int k = 0; k < G2_KC; k += 4) {
      const double a0 = a0p[k], a1 = a1p[k], b0 = b0p[k * G2_LDB], b1 = b1p[k * G2_LDB];
      acc.c[0][0] = mfma_f64(a0, b0, acc.c[0][0]);
      acc.c[0][1] = mfma_f64(a0, b1, acc.c[0][1]);
      acc.c[1][0] = mfma_f64(a1, b0, acc.c[1][0]);
      acc.c[1][1] = mfma_f64(a1, b1, acc.c[1][1]);
    }
  };
  // One pipeline per segment, its fields read once: the K loop then touches
  // no memory but the operands (segments are few, so refilling the pipeline
  // at a segment boundary costs little).
  for (int sid = 0; sid < g.nseg; ++sid) {
    const GemmSeg sg = g.seg[sid];
    int kb = 0, ke = sg.k;  // K range of this tile (triangular operands skip zero chunks)
    if (sg.tri & TRI_A_LOWER) ke = min(ke, (tr + 1) * 64);
    if (sg.tri & TRI_A_UPPER) kb = max(kb, tr * 64);
    if (sg.tri & TRI_B_LOWER) kb = max(kb, tc * 64);
    if (sg.tri & TRI_B_UPPER) ke = min(ke, (tc + 1) * 64);
    if (ke <= kb) continue;
    const int nch = (ke - kb + G2_KC - 1) / G2_KC;
    const int ma = (sg.tri & TRI_A_LOWER) ? MASK_K_GT_I : ((sg.tri & TRI_A_UPPER) ? MASK_K_LT_I : MASK_NONE);
    const int mb = (sg.tri & TRI_B_LOWER) ? MASK_K_LT_I : ((sg.tri & TRI_B_UPPER) ? MASK_K_GT_I : MASK_NONE);
    // chunk q (past the last one: the last one again, loaded and never used)
    auto fetch = [&](int q, G2Stage& st) {
      const int c0 = kb + G2_KC * min(q, nch - 1);
      const int kw = min(G2_KC, ke - c0);
      const int64_t oa = sg.ta ? (int64_t)c0 * sg.lda + tr * 64 : (int64_t)(tr * 64) * sg.lda + c0;
      const int64_t ob = sg.tb ? (int64_t)(tc * 64) * sg.ldb + c0 : (int64_t)c0 * sg.ldb + tc * 64;
      st.a.load(sg.A + oa, sg.lda, sg.ta, rows, kw);
      st.b.load(sg.B + ob, sg.ldb, 1 - sg.tb, cols, kw);
      st.k0 = c0;
      st.kw = kw;
    };
    auto stage = [&](const G2Stage& st, int buf) {
      st.a.store_a(sA[buf], sg.ta, sg.alpha, rows, st.kw, ma, tr * 64, st.k0);
      st.b.store_b(sB[buf], 1 - sg.tb, cols, st.kw, mb, tc * 64, st.k0);
    };
    // Chunk t sits in LDS buffer t % 2 when it is multiplied; every pass
    // issues the same loads and stores, only the MFMAs of a missing odd
    // chunk are skipped.
    G2Stage s0, s1;
    fetch(0, s0);
    fetch(1, s1);
    stage(s0, 0);
    fetch(2, s0);
    __syncthreads();
    for (int t = 0; t < nch; t += 2) {
      mma(0);           // chunk t
      stage(s1, 1);     // chunk t + 1
      fetch(t + 3, s1);
      lds_barrier();
      if (t + 1 < nch) mma(1);
      stage(s0, 0);     // chunk t + 2
      fetch(t + 4, s0);
      lds_barrier();
    }
  }
  // epilogue
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = acc_row(x, r), j = acc_col(y);
        if (i >= rows || j >= cols) continue;
        const double v = acc.c[x][y][r];
        gdouble* cg = gview(ct);
        if (g.lower && tr == tc) {
          if (i >= j) {
            cg[(int64_t)i * g.ldc + j] = v;
            if (g.mirror && i > j) cg[(int64_t)j * g.ldc + i] = v;
          }
        } else {
          cg[(int64_t)i * g.ldc + j] = v;
          if (g.lower && g.mirror) gview(g.C)[(int64_t)(tc * 64 + j) * g.ldc + tr * 64 + i] = v;
        }
      }
}

// ------------------------------------------------- 128x128 output tiles
//
// Same job semantics as k_gemm_grouped with 128x128 output tiles: wave w owns
// the 64x64 quadrant (64*(w/2), 64*(w%2)) as 4x4 MFMA 16x16 accumulators, K
// advances in chunks of 16 staged through LDS, the next chunk is fetched into
// registers while the current one is multiplied.

#define BIG_TS 128
#ifndef BIG_KC
#define BIG_KC 16
#endif
// odd row stride in doubles: the 16 rows of an MFMA operand fetch start on
// 16 distinct even dword banks, so the 64-bit reads are conflict-free
#define BIG_LD (BIG_KC + 1)

struct BigOp {
  static constexpr int N = BIG_TS * BIG_KC / BTA_THREADS;
  double r[N];
  // trans == 0: X row-major rows x K -> element (i, k) = X[i][k]
  // trans == 1: X row-major K x rows -> element (i, k) = X[k][i]
  __device__ __forceinline__ void load(const double* __restrict__ g0, int64_t ld, int trans, int rows, int kw,
                                       int mask, int i0, int k0) {
    const gdouble* g = gview(g0);
    mask = chunk_mask(mask, i0, rows, k0, kw);
    if (mask == MASK_NONE && rows >= BIG_TS && kw >= BIG_KC) {  // interior chunk: no tests
      if (trans == 0) {
        const gdouble* t = g + (int64_t)(threadIdx.x / BIG_KC) * ld + threadIdx.x % BIG_KC;
#pragma unroll
        for (int p = 0; p < N; ++p) r[p] = t[(int64_t)p * (BTA_THREADS / BIG_KC) * ld];
      } else {
        const gdouble* t = g + (int64_t)(threadIdx.x >> 7) * ld + (threadIdx.x & 127);
#pragma unroll
        for (int p = 0; p < N; ++p) r[p] = t[(int64_t)p * (BTA_THREADS / BIG_TS) * ld];
      }
      return;
    }
#pragma unroll
    for (int p = 0; p < N; ++p) {
      const int e = threadIdx.x + BTA_THREADS * p;
      if (trans == 0) {
        const int i = e / BIG_KC, k = e % BIG_KC;
        r[p] = (i < rows && k < kw && !masked_out(mask, i0 + i, k0 + k)) ? g[(int64_t)i * ld + k] : 0.0;
      } else {
        const int k = e >> 7, i = e & 127;
        r[p] = (i < rows && k < kw && !masked_out(mask, i0 + i, k0 + k)) ? g[(int64_t)k * ld + i] : 0.0;
      }
    }
  }
  __device__ __forceinline__ void store(double* __restrict__ s, int trans) const {
#pragma unroll
    for (int p = 0; p < N; ++p) {
      const int e = threadIdx.x + BTA_THREADS * p;
      if (trans == 0) s[(e / BIG_KC) * BIG_LD + (e % BIG_KC)] = r[p];
      else s[(e & 127) * BIG_LD + (e >> 7)] = r[p];
    }
  }
};

__global__ void __launch_bounds__(BTA_THREADS, 1) k_gemm_grouped_big(const GemmJob* __restrict__ jobs,
                                                                    const int* __restrict__ prefix, int njobs) {
  // two LDS buffers per operand: chunk q is multiplied out of one while chunk
  // q+1, fetched into registers before the multiply, is stored to the other
  __shared__ double sA[2][BIG_TS * BIG_LD];
  __shared__ double sB[2][BIG_TS * BIG_LD];
  const int jid = find_job(prefix, njobs, blockIdx.x);
  const GemmJob& g = jobs[jid];
  const int local = blockIdx.x - prefix[jid];
  int tr, tc;
  if (g.lower) decode_lower(local, tr, tc);
  else {
    const int tn = (g.n + BIG_TS - 1) / BIG_TS;
    tr = local / tn;
    tc = local % tn;
  }
  const int rows = min(BIG_TS, g.m - tr * BIG_TS), cols = min(BIG_TS, g.n - tc * BIG_TS);
  double* ct = g.C + (int64_t)(tr * BIG_TS) * g.ldc + tc * BIG_TS;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wm = (w >> 1) * 64, wn = (w & 1) * 64;
  const int li = lane & 15, lk = lane >> 4;
  f64x4 acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      if (g.beta != 0.0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = wm + 16 * x + 4 * r + lk, j = wn + 16 * y + li;
          acc[x][y][r] = (i < rows && j < cols) ? g.beta * gview(ct)[(int64_t)i * g.ldc + j] : 0.0;
        }
      } else {
        acc[x][y] = f64x4{0.0, 0.0, 0.0, 0.0};
      }
    }
  int seg_id = 0, k0 = 0, kmax = 0;
  auto seg_range = [&](int sid, int& kb, int& ke) {
    const GemmSeg& sg = g.seg[sid];
    kb = 0;
    ke = sg.k;
    if (sg.tri & TRI_A_LOWER) ke = min(ke, (tr + 1) * BIG_TS);
    if (sg.tri & TRI_A_UPPER) kb = max(kb, tr * BIG_TS);
    if (sg.tri & TRI_B_LOWER) kb = max(kb, tc * BIG_TS);
    if (sg.tri & TRI_B_UPPER) ke = min(ke, (tc + 1) * BIG_TS);
  };
  auto next_chunk = [&](bool first) -> bool {
    if (!first) k0 += BIG_KC;
    while (seg_id < g.nseg) {
      if (first || k0 >= kmax) {
        if (!first) ++seg_id;
        if (seg_id >= g.nseg) return false;
        int kb;
        seg_range(seg_id, kb, kmax);
        k0 = kb;
        first = false;
        if (k0 < kmax) return true;
        continue;
      }
      return true;
    }
    return false;
  };
  BigOp pa, pb;
  auto fetch = [&]() {
    const GemmSeg& sg = g.seg[seg_id];
    const int kw = min(BIG_KC, kmax - k0);
    const int ma = (sg.tri & TRI_A_LOWER) ? MASK_K_GT_I : ((sg.tri & TRI_A_UPPER) ? MASK_K_LT_I : MASK_NONE);
    const int mb = (sg.tri & TRI_B_LOWER) ? MASK_K_LT_I : ((sg.tri & TRI_B_UPPER) ? MASK_K_GT_I : MASK_NONE);
    if (sg.ta == 0) pa.load(sg.A + (int64_t)(tr * BIG_TS) * sg.lda + k0, sg.lda, 0, rows, kw, ma, tr * BIG_TS, k0);
    else pa.load(sg.A + (int64_t)k0 * sg.lda + tr * BIG_TS, sg.lda, 1, rows, kw, ma, tr * BIG_TS, k0);
    if (sg.tb == 1) pb.load(sg.B + (int64_t)(tc * BIG_TS) * sg.ldb + k0, sg.ldb, 0, cols, kw, mb, tc * BIG_TS, k0);
    else pb.load(sg.B + (int64_t)k0 * sg.ldb + tc * BIG_TS, sg.ldb, 1, cols, kw, mb, tc * BIG_TS, k0);
  };
  bool have = (g.nseg > 0) && next_chunk(true);
  int buf = 0;
  if (have) {
    fetch();
    const GemmSeg& sg = g.seg[seg_id];
    pa.store(sA[0], sg.ta);
    pb.store(sB[0], (sg.tb == 1) ? 0 : 1);
    __syncthreads();
  }
  while (have) {
    const int kw = min(BIG_KC, kmax - k0);
    const double alpha = g.seg[seg_id].alpha;
    const bool more = next_chunk(false);
    if (more) fetch();
    const double* cA = sA[buf];
    const double* cB = sB[buf];
#pragma unroll
    for (int k = 0; k < BIG_KC; k += 4) {
      if (k < kw) {
        double av[4], bv[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) av[x] = alpha * cA[(wm + 16 * x + li) * BIG_LD + k + lk];
#pragma unroll
        for (int y = 0; y < 4; ++y) bv[y] = cB[(wn + 16 * y + li) * BIG_LD + k + lk];
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
          for (int y = 0; y < 4; ++y) acc[x][y] = mfma_f64(av[x], bv[y], acc[x][y]);
      }
    }
    if (more) {
      const GemmSeg& sg = g.seg[seg_id];
      pa.store(sA[buf ^ 1], sg.ta);
      pb.store(sB[buf ^ 1], (sg.tb == 1) ? 0 : 1);
    }
    // one barrier per chunk: the buffer stored above was last read before
    // the previous barrier, the one read next was stored before this one
    __syncthreads();
    buf ^= 1;
    have = more;
  }
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = wm + 16 * x + 4 * r + lk, j = wn + 16 * y + li;
        if (i >= rows || j >= cols) continue;
        const double v = acc[x][y][r];
        gdouble* cg = gview(ct);
        if (g.lower && tr == tc) {
          if (i >= j) {
            cg[(int64_t)i * g.ldc + j] = v;
            if (g.mirror && i > j) cg[(int64_t)j * g.ldc + i] = v;
          }
        } else {
          cg[(int64_t)i * g.ldc + j] = v;
          if (g.lower && g.mirror) gview(g.C)[(int64_t)(tc * BIG_TS + j) * g.ldc + tr * BIG_TS + i] = v;
        }
      }
}

// ------------------------------------------------------------------- GEMV
//
// Memory-bound sweeps: every matrix element is read once, so the kernels are
// shaped for bandwidth.  N form (op(A) = A): one wave per output row; with a
// single unit-stride right-hand side each lane reads two consecutive doubles
// per access, four accesses in flight, and a butterfly combines the lanes.
// T form (op(A) = A^T): two passes.  Workgroup (64-column block, K chunk of
// 128 rows) sums its rows for 64 columns (one per lane, the four waves take
// every fourth row, four rows in flight) into a partial vector; a second
// kernel adds the partials of each column in chunk order into y.  Both
// orders of summation are fixed, so results are reproducible run to run.

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

#define GEMV_TK 128  // rows of A per T-form chunk

__global__ void __launch_bounds__(BTA_THREADS) k_gemv_n(GemvJob g) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = blockIdx.x * 4 + w;
  if (i >= g.m) return;
  for (int c = 0; c < g.nrhs; ++c) {
    double acc = 0.0;
    for (int s = 0; s < g.nseg; ++s) {
      const GemvSeg& sg = g.seg[s];
      int kmin = 0, kmax = sg.k;
      if (sg.tri == 1) kmax = min(kmax, i + 1);
      if (sg.tri == 2) kmin = i;
      const double* a = sg.A + (int64_t)i * sg.lda;
      double p[4] = {0.0, 0.0, 0.0, 0.0};
      const bool vec = sg.ldx == 1 && g.nrhs == 1 && (sg.lda & 1) == 0 &&
                       ((reinterpret_cast<uintptr_t>(sg.A) | reinterpret_cast<uintptr_t>(sg.x)) & 15) == 0;
      if (vec) {
        // pairs [k0, ke2) as double2; all 16 loads of a 2048-wide stretch are
        // issued before the first FMA (addresses clamped into the row, the
        // FMAs masked), so a row of up to 2048 costs one memory round trip
        const int k0 = kmin & ~1, ke2 = kmax & ~1;
        for (int base = k0; base < ke2; base += 2048) {
          double2 av[16], xv[16];
#pragma unroll
          for (int u = 0; u < 16; ++u) {
            const int kk = min(base + 128 * u + 2 * lane, ke2 - 2);
            av[u] = *reinterpret_cast<const double2*>(a + kk);
            xv[u] = *reinterpret_cast<const double2*>(sg.x + kk);
          }
#pragma unroll
          for (int u = 0; u < 16; ++u) {
            const int kk = base + 128 * u + 2 * lane;
            if (kk < ke2) {
              const double t = fma(av[u].y, xv[u].y, p[u & 3]);
              p[u & 3] = (kk >= kmin) ? fma(av[u].x, xv[u].x, t) : t;
            }
          }
        }
        if ((kmax & 1) && kmax - 1 >= kmin && lane == 0) p[0] = fma(a[kmax - 1], sg.x[kmax - 1], p[0]);
      } else {
        for (int base = kmin; base < kmax; base += 256) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int kk = base + 64 * u + lane;
            if (kk < kmax) p[u] = fma(a[kk], sg.x[(int64_t)kk * sg.ldx + c], p[u]);
          }
        }
      }
      acc = fma(sg.alpha, wave_sum((p[0] + p[1]) + (p[2] + p[3])), acc);
    }
    if (lane == 0) {
      double* yp = g.y + (int64_t)i * g.ldy + c;
      *yp = (g.beta != 0.0) ? fma(g.beta, *yp, acc) : acc;
    }
  }
}

__host__ __device__ inline int gemv_t_chunks(const GemvJob& g) {
  int n = 0;
  for (int s = 0; s < g.nseg; ++s) n += (g.seg[s].k + GEMV_TK - 1) / GEMV_TK;
  return n;
}

// part[(q * m + j) * nrhs + rc]: alpha_s * sum over chunk q's rows of A[k][j] x[k].
__global__ void __launch_bounds__(BTA_THREADS) k_gemv_t_part(GemvJob g, double* __restrict__ part) {
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + lane;
  int q = blockIdx.y, s = 0;
  while (s < g.nseg - 1) {
    const int nc = (g.seg[s].k + GEMV_TK - 1) / GEMV_TK;
    if (q < nc) break;
    q -= nc;
    ++s;
  }
  const GemvSeg& sg = g.seg[s];
  const int kb = q * GEMV_TK, ke = min(sg.k, kb + GEMV_TK);
  int klo = kb, khi = ke;  // this lane's rows (triangular operands)
  if (sg.tri == 1) khi = min(khi, j + 1);
  if (sg.tri == 2) klo = max(klo, j);
  const bool col = j < g.m;
  const int jc = min(j, g.m - 1);
  const int wu = __builtin_amdgcn_readfirstlane(w);
  // a chunk outside the triangle for every column of this workgroup loads nothing
  const int j0 = blockIdx.x * 64, j1 = min(j0 + 64, g.m) - 1;
  const int nh = ((sg.tri == 1 && kb > j1) || (sg.tri == 2 && ke - 1 < j0)) ? 0 : GEMV_TK / 64;
  for (int rc = 0; rc < g.nrhs; ++rc) {
    double p[4] = {0.0, 0.0, 0.0, 0.0};
    // wave w takes rows kb + w + 4r (r < 32) in two batches of 16 whose loads
    // are all issued before their FMAs (addresses clamped into the chunk,
    // the FMAs masked)
#pragma unroll 1
    for (int h = 0; h < nh; ++h) {
      double av[16], xv[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int kc = min(kb + wu + 4 * (16 * h + r), ke - 1);  // uniform: x from scalar loads
        av[r] = sg.A[(int64_t)kc * sg.lda + jc];
        xv[r] = sg.x[(int64_t)kc * sg.ldx + rc];
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int k = kb + w + 4 * (16 * h + r);
        if (col && k < ke && k >= klo && k < khi) p[r & 3] = fma(av[r], xv[r], p[r & 3]);
      }
    }
    red[w][lane] = (p[0] + p[1]) + (p[2] + p[3]);
    __syncthreads();
    if (w == 0 && col)
      part[((int64_t)blockIdx.y * g.m + j) * g.nrhs + rc] =
          sg.alpha * ((red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]));
    __syncthreads();
  }
}

__global__ void __launch_bounds__(BTA_THREADS) k_gemv_t_reduce(GemvJob g, const double* __restrict__ part,
                                                              int nchunk) {
  const int64_t e = (int64_t)blockIdx.x * BTA_THREADS + threadIdx.x;
  if (e >= (int64_t)g.m * g.nrhs) return;
  const int j = (int)(e / g.nrhs), rc = (int)(e % g.nrhs);
  double acc = 0.0;
  for (int q = 0; q < nchunk; ++q) acc += part[((int64_t)q * g.m + j) * g.nrhs + rc];
  double* yp = g.y + (int64_t)j * g.ldy + rc;
  *yp = (g.beta != 0.0) ? fma(g.beta, *yp, acc) : acc;
}

// ------------------------------------------------------------------- plans

DevPlan::~DevPlan() { release(); }

void DevPlan::release() {
  if (d_jobs) hipFree(d_jobs);
  if (d_prefix) hipFree(d_prefix);
  if (d_part) hipFree(d_part);
  if (d_cnt) hipFree(d_cnt);
  d_jobs = nullptr;
  d_prefix = nullptr;
  d_part = nullptr;
  d_cnt = nullptr;
  bytes = 0;
  launches.clear();
}

int DevPlan::build(const std::vector<GemmLaunch>& ls) {
  release();
  std::vector<GemmJob> all;
  std::vector<int> pre;
  // split jobs: tile offsets into one scratch shared by all launches (they run
  // one after another on the plan's stream), sized for the largest launch
  int64_t split_max = 0;
  std::vector<int64_t> split_off;
  for (const auto& l : ls) {
    Entry e;
    e.job_off = (int)all.size();
    e.pre_off = (int)pre.size();
    e.njobs = (int)l.jobs.size();
    e.tiles = l.tiles();
    e.ts = l.ts;
    e.vec = 1;
    e.split = 0;
    int64_t split_tiles = 0;
    for (const GemmJob& j : l.jobs) {
      for (int q = 0; q < j.nseg; ++q) {
        const GemmSeg& sg = j.seg[q];
        if ((sg.lda & 1) || (sg.ldb & 1) || (reinterpret_cast<uintptr_t>(sg.A) & 15) ||
            (reinterpret_cast<uintptr_t>(sg.B) & 15))
          e.vec = 0;
      }
      if (j.split != 1 && (j.split != 2 || !j.lower || l.ts != 64)) return (int)hipErrorInvalidValue;
      split_off.push_back(j.split == 2 ? split_tiles : -1);
      if (j.split == 2) {
        const int tm = (j.m + 63) / 64;
        split_tiles += tm * (tm + 1) / 2;
        e.split = 1;
      }
    }
    split_max = std::max(split_max, split_tiles);
    launches.push_back(e);
    all.insert(all.end(), l.jobs.begin(), l.jobs.end());
    pre.insert(pre.end(), l.prefix.begin(), l.prefix.end());
  }
  if (all.empty()) return 0;
  hipError_t e;
  if (split_max > 0) {
    if ((e = hipMalloc(&d_part, (size_t)split_max * 2 * 4096 * sizeof(double))) != hipSuccess) return (int)e;
    if ((e = hipMalloc(&d_cnt, (size_t)split_max * sizeof(int))) != hipSuccess) return (int)e;
    if ((e = hipMemset(d_cnt, 0, (size_t)split_max * sizeof(int))) != hipSuccess) return (int)e;
    for (size_t q = 0; q < all.size(); ++q)
      if (split_off[q] >= 0) {
        all[q].part = d_part + split_off[q] * 2 * 4096;
        all[q].cnt = d_cnt + split_off[q];
      }
  }
  if ((e = hipMalloc(&d_jobs, all.size() * sizeof(GemmJob))) != hipSuccess) return (int)e;
  if ((e = hipMalloc(&d_prefix, pre.size() * sizeof(int))) != hipSuccess) return (int)e;
  if ((e = hipMemcpy(d_jobs, all.data(), all.size() * sizeof(GemmJob), hipMemcpyHostToDevice)) != hipSuccess)
    return (int)e;
  if ((e = hipMemcpy(d_prefix, pre.data(), pre.size() * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
    return (int)e;
  bytes = all.size() * sizeof(GemmJob) + pre.size() * sizeof(int) +
          (size_t)split_max * (2 * 4096 * sizeof(double) + sizeof(int));
  return 0;
}

void DevPlan::launch(int idx, hipStream_t st, int variant) const {
  const Entry& e = launches[idx];
  if (e.tiles <= 0) return;
  if (e.split)
    hipLaunchKernelGGL(k_gemm_grouped, dim3(e.tiles), dim3(BTA_THREADS), 0, st, d_jobs + e.job_off,
                       d_prefix + e.pre_off, e.njobs);
  else if (e.ts == BIG_TS)
    hipLaunchKernelGGL(k_gemm_grouped_big, dim3(e.tiles), dim3(BTA_THREADS), 0, st, d_jobs + e.job_off,
                       d_prefix + e.pre_off, e.njobs);
  else if (variant == 1 && e.vec)
    hipLaunchKernelGGL(k_gemm_grouped_db, dim3(e.tiles), dim3(BTA_THREADS), 0, st, d_jobs + e.job_off,
                       d_prefix + e.pre_off, e.njobs);
  else
    hipLaunchKernelGGL(k_gemm_grouped, dim3(e.tiles), dim3(BTA_THREADS), 0, st, d_jobs + e.job_off,
                       d_prefix + e.pre_off, e.njobs);
}

size_t gemv_scratch(const GemvJob& g, int form_t) {
  return form_t ? (size_t)gemv_t_chunks(g) * g.m * g.nrhs : 0;
}

void gemv_launch(const GemvJob& g, int form_t, double* scratch, hipStream_t st) {
  if (g.m <= 0) return;
  if (!form_t) {
    hipLaunchKernelGGL(k_gemv_n, dim3((g.m + 3) / 4), dim3(BTA_THREADS), 0, st, g);
    return;
  }
  const int nchunk = gemv_t_chunks(g);
  hipLaunchKernelGGL(k_gemv_t_part, dim3((g.m + 63) / 64, nchunk), dim3(BTA_THREADS), 0, st, g, scratch);
  const int64_t tot = (int64_t)g.m * g.nrhs;
  hipLaunchKernelGGL(k_gemv_t_reduce, dim3((unsigned)((tot + BTA_THREADS - 1) / BTA_THREADS)), dim3(BTA_THREADS), 0,
                     st, g, scratch, nchunk);
}

}  // namespace bta
