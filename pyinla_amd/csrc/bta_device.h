// Device-side building blocks for the BTA kernels (gfx950 / CDNA4, FP64).
//
// Everything works on 64x64 FP64 tiles held row-major in global memory with an
// explicit leading dimension, staged through LDS, and multiplied with the
// FP64 matrix instruction v_mfma_f64_16x16x4f64.  A workgroup is 256 threads =
// 4 wave64s; wave w owns the 32x32 quadrant (rows 32*(w/2), cols 32*(w%2)) of a
// 64x64 output tile as 2x2 MFMA 16x16 accumulators.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BTA_TILE 64
#define BTA_THREADS 256
#define BTA_LDS_LD 66  // padded LDS row stride in doubles (bank-conflict-free fragment reads)

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// Global-address-space views of pointers.  Pointers read from memory (job
// tables) are generic to the compiler, which then emits FLAT accesses; a FLAT
// load also counts against lgkmcnt, so every wait for an LDS read would wait
// for the outstanding global prefetch as well.  Through these views the
// compiler emits global_load / global_store.
typedef __attribute__((address_space(1))) double gdouble;
__device__ __forceinline__ const gdouble* gview(const double* p) { return (const gdouble*)p; }
__device__ __forceinline__ gdouble* gview(double* p) { return (gdouble*)p; }
typedef __attribute__((address_space(1))) f64x2 gf64x2;

// Operand layout of v_mfma_f64_16x16x4f64 (D = A*B + C, A 16x4, B 4x16):
//   A[i][k] is supplied by lane i + 16*k, B[k][j] by lane j + 16*k,
//   D[i][j] is returned by lane j + 16*(i%4) in element i/4
// (measured on gfx950 by bta_selftest_mfma).
__device__ __forceinline__ f64x4 mfma_f64(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
// D = -A*B + C through the instruction's negation modifier (neg:[1,0,0]; the
// blgp field of the F64 MFMA), so no VALU negation sits between an LDS read
// and its MFMA.
__device__ __forceinline__ f64x4 mfma_f64_nega(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 1);
}

struct TileAcc {
  f64x4 c[2][2];
};

__device__ __forceinline__ void acc_zero(TileAcc& t) {
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) t.c[x][y] = f64x4{0.0, 0.0, 0.0, 0.0};
}

// Row/col of accumulator element (x, y, r) for the calling thread.
__device__ __forceinline__ int acc_row(int x, int r) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  return (w >> 1) * 32 + 16 * x + 4 * r + (lane >> 4);
}
__device__ __forceinline__ int acc_col(int y) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  return (w & 1) * 32 + 16 * y + (lane & 15);
}

// Load a rows x cols (<= 64 x 64) block of a row-major global matrix into the
// 64x64 LDS tile s (stride BTA_LDS_LD), zero-filling outside the valid block.
__device__ __forceinline__ void lds_load_tile(double* __restrict__ s, const double* __restrict__ g,
                                              int64_t ld, int rows, int cols) {
  const int c = threadIdx.x & 63;
  const int r0 = threadIdx.x >> 6;
#pragma unroll 4
  for (int r = r0; r < 64; r += 4) {
    double v = 0.0;
    if (r < rows && c < cols) v = g[(int64_t)r * ld + c];
    s[r * BTA_LDS_LD + c] = v;
  }
}

// Same, but store the transpose: s[c][r] = g[r][c].
__device__ __forceinline__ void lds_load_tile_t(double* __restrict__ s, const double* __restrict__ g,
                                                int64_t ld, int rows, int cols) {
  const int c = threadIdx.x & 63;
  const int r0 = threadIdx.x >> 6;
#pragma unroll 4
  for (int r = r0; r < 64; r += 4) {
    double v = 0.0;
    if (r < rows && c < cols) v = g[(int64_t)r * ld + c];
    s[c * BTA_LDS_LD + r] = v;
  }
}

// A 64x64 operand tile staged global -> registers -> LDS, so that the global
// loads of the next K chunk are in flight while the current chunk is
// multiplied.  The LDS layout is always s[i][k] (row i of op(X), column k).
//   trans == 0: op(X) = X, X row-major (rows x K): s[i][k] = X[i][k]
//   trans == 1: op(X) = X^T, X row-major (K x rows): s[i][k] = X[k][i]
// Element-exact triangular masks for operand tiles (row index i of op(X),
// K index k, both global): MASK_K_GT_I zeroes k > i, MASK_K_LT_I zeroes k < i.
enum : int { MASK_NONE = 0, MASK_K_GT_I = 1, MASK_K_LT_I = 2 };

__device__ __forceinline__ bool masked_out(int mask, int ig, int kg) {
  return (mask == MASK_K_GT_I && kg > ig) || (mask == MASK_K_LT_I && kg < ig);
}

// The mask a rows x kw chunk at global (i0, k0) needs: MASK_NONE when none of
// its elements is masked, so that its loads need no per-element test.
__device__ __forceinline__ int chunk_mask(int mask, int i0, int rows, int k0, int kw) {
  if (mask == MASK_K_GT_I) return (k0 + kw - 1 <= i0) ? MASK_NONE : mask;
  if (mask == MASK_K_LT_I) return (k0 >= i0 + rows - 1) ? MASK_NONE : mask;
  return mask;
}

struct TilePrefetch {
  double r[16];
  // i0 / k0: global row / K index of the tile origin (used by the mask only)
  __device__ __forceinline__ void load(const double* __restrict__ g, int64_t ld, int trans, int rows, int kw,
                                       int mask = MASK_NONE, int i0 = 0, int k0 = 0) {
    const int c = threadIdx.x & 63;
    const int r0 = threadIdx.x >> 6;
    const gdouble* gg = gview(g);
    if (trans == 0) {
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int i = r0 + 4 * j;
        r[j] = (i < rows && c < kw && !masked_out(mask, i0 + i, k0 + c)) ? gg[(int64_t)i * ld + c] : 0.0;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int k = r0 + 4 * j;
        r[j] = (k < kw && c < rows && !masked_out(mask, i0 + c, k0 + k)) ? gg[(int64_t)k * ld + c] : 0.0;
      }
    }
  }
  __device__ __forceinline__ void store(double* __restrict__ s, int trans) const {
    const int c = threadIdx.x & 63;
    const int r0 = threadIdx.x >> 6;
    if (trans == 0) {
#pragma unroll
      for (int j = 0; j < 16; ++j) s[(r0 + 4 * j) * BTA_LDS_LD + c] = r[j];
    } else {
#pragma unroll
      for (int j = 0; j < 16; ++j) s[c * BTA_LDS_LD + r0 + 4 * j] = r[j];
    }
  }
  // A full, contiguous 64 x 64 tile (ld 64) as 128-bit pairs: pair q of the
  // thread is columns 2 (t % 32), +1 of row t / 32 + 8 q (one wave: two whole
  // rows per load), half the load and LDS-store instructions of load/store.
  __device__ __forceinline__ static const gf64x2* full_pair(const double* __restrict__ g, int q) {
    const int t = threadIdx.x;
    return (const gf64x2*)(gview(g) + ((t >> 5) + 8 * q) * 64 + 2 * (t & 31));
  }
  __device__ __forceinline__ void set_pair(int q, f64x2 v) {
    r[2 * q] = v.x;
    r[2 * q + 1] = v.y;
  }
  __device__ __forceinline__ void load_full(const double* __restrict__ g) {
#pragma unroll
    for (int q = 0; q < 8; ++q) set_pair(q, *full_pair(g, q));
  }
  __device__ __forceinline__ void store_full(double* __restrict__ s) const {
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      *(f64x2*)&s[((t >> 5) + 8 * q) * BTA_LDS_LD + 2 * (t & 31)] = f64x2{r[2 * q], r[2 * q + 1]};
  }
};

// acc += sign * sA(64 x K) * sB(64 x K)^T over k in [0, kmax), kmax % 4 == 0.
// sA[i][k] and sB[j][k] are LDS tiles with stride BTA_LDS_LD.
template <bool NEG>
__device__ __forceinline__ void mma_nt(TileAcc& t, const double* __restrict__ sA,
                                       const double* __restrict__ sB, int kbeg, int kend) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  const double* a0p = sA + (wm + li) * BTA_LDS_LD + lk;
  const double* a1p = sA + (wm + 16 + li) * BTA_LDS_LD + lk;
  const double* b0p = sB + (wn + li) * BTA_LDS_LD + lk;
  const double* b1p = sB + (wn + 16 + li) * BTA_LDS_LD + lk;
#pragma unroll 4
  for (int k = kbeg; k < kend; k += 4) {
    const double a0 = a0p[k], a1 = a1p[k], b0 = b0p[k], b1 = b1p[k];
    if (NEG) {
      t.c[0][0] = mfma_f64_nega(a0, b0, t.c[0][0]);
      t.c[0][1] = mfma_f64_nega(a0, b1, t.c[0][1]);
      t.c[1][0] = mfma_f64_nega(a1, b0, t.c[1][0]);
      t.c[1][1] = mfma_f64_nega(a1, b1, t.c[1][1]);
    } else {
      t.c[0][0] = mfma_f64(a0, b0, t.c[0][0]);
      t.c[0][1] = mfma_f64(a0, b1, t.c[0][1]);
      t.c[1][0] = mfma_f64(a1, b0, t.c[1][0]);
      t.c[1][1] = mfma_f64(a1, b1, t.c[1][1]);
    }
  }
}

// Workgroup barrier ordering LDS accesses only.  __syncthreads() also waits
// for every outstanding global load of the wave (its release fence covers all
// address spaces), which would stall a wave at the barrier until an operand
// prefetch meant for the next step has arrived.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// mma_nt over the full K = 64 of two LDS tiles, unrolled, with the operand
// fragments of k step t+1 read while the MFMAs of step t run (two fragment
// register sets; the scheduling barriers keep each step's reads ahead of the
// previous step's MFMAs, which LLVM otherwise sinks behind them and then
// waits for).  A helper runs one wave per SIMD, so nothing else hides the
// LDS latency.
template <bool NEG>
__device__ __forceinline__ void mma_nt64(TileAcc& t, const double* __restrict__ sA,
                                         const double* __restrict__ sB) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  const double* a0p = sA + (wm + li) * BTA_LDS_LD + lk;
  const double* a1p = sA + (wm + 16 + li) * BTA_LDS_LD + lk;
  const double* b0p = sB + (wn + li) * BTA_LDS_LD + lk;
  const double* b1p = sB + (wn + 16 + li) * BTA_LDS_LD + lk;
  double a0 = a0p[0], a1 = a1p[0], b0 = b0p[0], b1 = b1p[0];
#pragma unroll
  for (int k = 0; k < 64; k += 4) {
    double n0 = 0.0, n1 = 0.0, m0 = 0.0, m1 = 0.0;
    if (k + 4 < 64) {
      n0 = a0p[k + 4];
      n1 = a1p[k + 4];
      m0 = b0p[k + 4];
      m1 = b1p[k + 4];
    }
    if (NEG) {
      t.c[0][0] = mfma_f64_nega(a0, b0, t.c[0][0]);
      t.c[0][1] = mfma_f64_nega(a0, b1, t.c[0][1]);
      t.c[1][0] = mfma_f64_nega(a1, b0, t.c[1][0]);
      t.c[1][1] = mfma_f64_nega(a1, b1, t.c[1][1]);
    } else {
      t.c[0][0] = mfma_f64(a0, b0, t.c[0][0]);
      t.c[0][1] = mfma_f64(a0, b1, t.c[0][1]);
      t.c[1][0] = mfma_f64(a1, b0, t.c[1][0]);
      t.c[1][1] = mfma_f64(a1, b1, t.c[1][1]);
    }
    if (k + 4 < 64) {
      __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);  // the 4 LDS reads of step k+4
      __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);  // then the 4 MFMAs of step k
    }
    a0 = n0;
    a1 = n1;
    b0 = m0;
    b1 = m1;
  }
}

// mma_nt64<true> that also loads the next step's two operand tiles (64 x 64,
// contiguous) into the prefetch registers, one 128-bit load per k step, so
// that the load issue overlaps the MFMAs instead of preceding them.  Same
// registers and layout as TilePrefetch::load_full.
__device__ __forceinline__ void mma_nt64_neg_pf(TileAcc& t, const double* __restrict__ sA,
                                                const double* __restrict__ sB, double (&ra)[16],
                                                const double* __restrict__ srcA, double (&rb)[16],
                                                const double* __restrict__ srcB) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  const double* a0p = sA + (wm + li) * BTA_LDS_LD + lk;
  const double* a1p = sA + (wm + 16 + li) * BTA_LDS_LD + lk;
  const double* b0p = sB + (wn + li) * BTA_LDS_LD + lk;
  const double* b1p = sB + (wn + 16 + li) * BTA_LDS_LD + lk;
  double a0 = a0p[0], a1 = a1p[0], b0 = b0p[0], b1 = b1p[0];
#pragma unroll
  for (int k = 0; k < 64; k += 4) {
    double n0 = 0.0, n1 = 0.0, m0 = 0.0, m1 = 0.0;
    if (k + 4 < 64) {
      n0 = a0p[k + 4];
      n1 = a1p[k + 4];
      m0 = b0p[k + 4];
      m1 = b1p[k + 4];
    }
    const int j = k >> 2;  // k step: pair j / 2 of operand A (even) or B (odd)
    const f64x2 v = *TilePrefetch::full_pair((j & 1) ? srcB : srcA, j >> 1);
    if (j & 1) {
      rb[j - 1] = v.x;
      rb[j] = v.y;
    } else {
      ra[j] = v.x;
      ra[j + 1] = v.y;
    }
    t.c[0][0] = mfma_f64_nega(a0, b0, t.c[0][0]);
    t.c[0][1] = mfma_f64_nega(a0, b1, t.c[0][1]);
    t.c[1][0] = mfma_f64_nega(a1, b0, t.c[1][0]);
    t.c[1][1] = mfma_f64_nega(a1, b1, t.c[1][1]);
    if (k + 4 < 64) {
      __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);  // LDS reads of step k+4
      __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // one prefetch load
      __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);  // MFMAs of step k
    }
    a0 = n0;
    a1 = n1;
    b0 = m0;
    b1 = m1;
  }
}

// Load / store the accumulator from / to a row-major global block (masked).
// Addresses are formed as one per-lane base pointer plus wave-uniform offsets
// (16x + 4r) * ld + 16y, so that only the base occupies vector registers.
__device__ __forceinline__ void acc_load(TileAcc& t, const double* __restrict__ g, int64_t ld,
                                         int rows, int cols) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rb = (w >> 1) * 32 + (lane >> 4), cb = (w & 1) * 32 + (lane & 15);
  const gdouble* base = gview(g) + (int64_t)rb * ld + cb;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int di = 16 * x + 4 * r;
        t.c[x][y][r] = (rb + di < rows && cb + 16 * y < cols) ? base[(int64_t)di * ld + 16 * y] : 0.0;
      }
}

__device__ __forceinline__ void acc_store(const TileAcc& t, double* __restrict__ g, int64_t ld,
                                          int rows, int cols) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rb = (w >> 1) * 32 + (lane >> 4), cb = (w & 1) * 32 + (lane & 15);
  gdouble* base = gview(g) + (int64_t)rb * ld + cb;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int di = 16 * x + 4 * r;
        if (rb + di < rows && cb + 16 * y < cols) base[(int64_t)di * ld + 16 * y] = t.c[x][y][r];
      }
}

// Store only the lower triangle (i >= j) of a diagonal tile.
__device__ __forceinline__ void acc_store_lower(const TileAcc& t, double* __restrict__ g, int64_t ld,
                                                int rows, int cols) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rb = (w >> 1) * 32 + (lane >> 4), cb = (w & 1) * 32 + (lane & 15);
  gdouble* base = gview(g) + (int64_t)rb * ld + cb;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int di = 16 * x + 4 * r;
        const int i = rb + di, j = cb + 16 * y;
        if (i < rows && j < cols && i >= j) base[(int64_t)di * ld + 16 * y] = t.c[x][y][r];
      }
}

__device__ __forceinline__ void acc_to_lds(const TileAcc& t, double* __restrict__ s) {
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) s[acc_row(x, r) * BTA_LDS_LD + acc_col(y)] = t.c[x][y][r];
}

// ------------------------------------------------------------ cross-WG flags
//
// Flags are 64-bit words holding the call epoch when the guarded data is ready.
// The XCDs' L2 caches are not coherent with each other, and agent-scope
// release / acquire fences write back / invalidate a whole L2.  The data a
// flag guards is therefore written with system-scope relaxed atomic stores
// (write-through to the coherence point) and read with system-scope relaxed
// atomic loads (which bypass the non-coherent caches); the producer waits for
// its stores to complete before it stores the flag the same way, and the
// consumer polls the flag with the same cache-bypassing loads.

__device__ __forceinline__ void st_sys(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ double ld_sys(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// A 64 x 64 accumulator tile to / from a contiguous 64 x 64 buffer with
// system-scope stores / loads.
__device__ __forceinline__ void acc_load_sys(TileAcc& t, const double* __restrict__ g) {
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) t.c[x][y][r] = ld_sys(&g[acc_row(x, r) * 64 + acc_col(y)]);
}

__device__ __forceinline__ void acc_store_sys(const TileAcc& t, double* __restrict__ g) {
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) st_sys(&g[acc_row(x, r) * 64 + acc_col(y)], t.c[x][y][r]);
}

// All threads call after their st_sys stores of the guarded data.
__device__ __forceinline__ void flag_publish(unsigned long long* flag, unsigned long long epoch) {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);  // this thread's stores have completed
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __syncthreads();                // ... and every other thread's
  if (threadIdx.x == 0) __hip_atomic_store(flag, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Returns false if the spin limit was hit (the caller records an error and
// carries on so that every wave still exits).  Guarded data must then be read
// with ld_sys.
__device__ __forceinline__ bool flag_wait(unsigned long long* flag, unsigned long long epoch,
                                          unsigned long long spin_limit) {
  __shared__ int ok_sh;
  if (threadIdx.x == 0) {
    int ok = 1;
    unsigned long long it = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != epoch) {
      if (++it > spin_limit) { ok = 0; break; }
      __builtin_amdgcn_s_sleep(1);
    }
    ok_sh = ok;
  }
  __syncthreads();
  const int ok = ok_sh;
  __syncthreads();
  return ok != 0;
}
