// Q(theta) assembly straight into dense BTA blocks, sparse scatter / gather
// between COO matrices and the blocks, and the Kronecker trace / quadratic
// contractions the adjoint gradient needs (gfx950, FP64).
//
// Spatio-temporal precision: Q_st = sum_{m,k} c[m][k] * (M_m (x) S_k) with
// tridiagonal temporal M_m and sparse spatial S_k sharing one union CSR
// pattern (values vals[k][e]).  Block t of the diagonal and sub-diagonal is
//     D_t = sum_k wD[t][k] S_k,   wD[t][k] = sum_m c[m][k] M_m[t][t]
//     L_t = sum_k wL[t][k] S_k,   wL[t][k] = sum_m c[m][k] M_m[t+1][t]
// Every block row is written in full (zeros included) with coalesced stores,
// so one pass produces the dense blocks the factorization reads.
#include "bta_internal.h"

namespace bta {

#define BTA_MAX_KM 128  // nv * nv * nM * nK coefficient table entries

struct KronArgs {
  double* D;
  double* L;            // nullptr if n == 1
  double* B;            // arrow blocks to clear (nullptr: untouched)
  double* T;            // tip to set to tip_diag * I (nullptr: untouched)
  const int* rowptr;    // b + 1
  const int* colidx;    // nnz
  const double* vals;   // nK x nnz
  const double* Mdiag;  // nM x n
  const double* Moff;   // nM x (n-1)
  double coef[BTA_MAX_KM];  // [v][w][m][k], nv x nv x nM x nK
  double tip_diag;
  int n, b, a, nnz, nM, nK, nv, ns;  // b = nv * ns
};

// grid: n * rows_chunks workgroups; each handles 4 rows of block t (one per wave).
// Row r of block t belongs to field v = r / ns, mesh node s = r % ns; its
// entries in field w sit at columns w*ns + colidx[e] with value
//   sum_k (sum_m coef[v][w][m][k] M_m[t][t]) vals[k][e]   (diagonal block)
//   sum_k (sum_m coef[v][w][m][k] M_m[t+1][t]) vals[k][e] (sub-diagonal block).
__global__ void __launch_bounds__(BTA_THREADS) k_assemble_kron(KronArgs p) {
  const int chunks = (p.b + 3) / 4;
  const int t = blockIdx.x / chunks;
  const int r = (blockIdx.x % chunks) * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r < p.b) {
    const int v = r / p.ns, srow = r % p.ns;
    double* drow = p.D + ((int64_t)t * p.b + r) * p.b;
    double* lrow = (t < p.n - 1) ? p.L + ((int64_t)t * p.b + r) * p.b : nullptr;
    for (int c = lane; c < p.b; c += 64) {
      drow[c] = 0.0;
      if (lrow) lrow[c] = 0.0;
    }
    // the zero stores above and the value stores below come from the same
    // wave; the fence keeps the values last
    __threadfence_block();
    const int e0 = p.rowptr[srow], e1 = p.rowptr[srow + 1];
    for (int w = 0; w < p.nv; ++w) {
      double wD[4], wL[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        wD[k] = 0.0;
        wL[k] = 0.0;
        if (k < p.nK)
          for (int m = 0; m < p.nM; ++m) {
            const double cf = p.coef[((v * p.nv + w) * p.nM + m) * p.nK + k];
            wD[k] = fma(cf, p.Mdiag[m * p.n + t], wD[k]);
            if (t < p.n - 1) wL[k] = fma(cf, p.Moff[m * (p.n - 1) + t], wL[k]);
          }
      }
      for (int e = e0 + lane; e < e1; e += 64) {
        const int c = w * p.ns + p.colidx[e];
        double vd = 0.0, vl = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          if (k < p.nK) {
            const double val = p.vals[(int64_t)k * p.nnz + e];
            vd = fma(wD[k], val, vd);
            vl = fma(wL[k], val, vl);
          }
        }
        drow[c] = vd;
        if (lrow) lrow[c] = vl;
      }
    }
  }
  // arrow rows of block t and the tip (block 0 handles the tip)
  if (p.B && (blockIdx.x % chunks) == 0) {
    double* bt = p.B + (int64_t)t * p.a * p.b;
    for (int e = threadIdx.x; e < p.a * p.b; e += BTA_THREADS) bt[e] = 0.0;
  }
  if (p.T && blockIdx.x == 0) {
    for (int e = threadIdx.x; e < p.a * p.a; e += BTA_THREADS) p.T[e] = (e / p.a == e % p.a) ? p.tip_diag : 0.0;
  }
}

// dst[buf(e)][off(e)] += scale * val[e]  (each destination appears once).
// code = (buffer id << 60) | offset, buffer id 0..3 = D, L, B, T.
__global__ void __launch_bounds__(BTA_THREADS) k_scatter(double* D, double* L, double* B, double* T,
                                                         const unsigned long long* code, const double* val,
                                                         const int* src_idx, int64_t count, double scale,
                                                         int accumulate) {
  for (int64_t e = (int64_t)blockIdx.x * BTA_THREADS + threadIdx.x; e < count;
       e += (int64_t)gridDim.x * BTA_THREADS) {
    const unsigned long long c = code[e];
    const int buf = (int)(c >> 60);
    const int64_t off = (int64_t)(c & ((1ull << 60) - 1));
    double* base = buf == 0 ? D : (buf == 1 ? L : (buf == 2 ? B : T));
    const double v = scale * val[src_idx ? src_idx[e] : e];
    if (accumulate) base[off] += v;
    else base[off] = v;
  }
}

// out[e] = block value at code[e] (gather of a BTA pattern into COO data).
__global__ void __launch_bounds__(BTA_THREADS) k_gather(const double* D, const double* L, const double* B,
                                                        const double* T, const unsigned long long* code,
                                                        double* out, int64_t count) {
  for (int64_t e = (int64_t)blockIdx.x * BTA_THREADS + threadIdx.x; e < count;
       e += (int64_t)gridDim.x * BTA_THREADS) {
    const unsigned long long c = code[e];
    const int buf = (int)(c >> 60);
    const int64_t off = (int64_t)(c & ((1ull << 60) - 1));
    const double* base = buf == 0 ? D : (buf == 1 ? L : (buf == 2 ? B : T));
    out[e] = base[off];
  }
}

// Per-workgroup partial sums of  sum_e w[e] * val[e] * G(e)  where G is either a
// BTA block entry (mode 0) or x[row]*x[col] (mode 1, rows/cols are global).
// partial[blockIdx] is written; a fixed-order reduction follows.
__global__ void __launch_bounds__(BTA_THREADS) k_bilinear_partial(const double* D, const double* L,
                                                                  const double* B, const double* T,
                                                                  const unsigned long long* code,
                                                                  const int* grow, const int* gcol,
                                                                  const double* x, const double* wval,
                                                                  int64_t count, int mode, double* partial) {
  __shared__ double red[BTA_THREADS];
  double acc = 0.0;
  const int64_t per = (count + gridDim.x - 1) / gridDim.x;
  const int64_t e0 = (int64_t)blockIdx.x * per, e1 = min(count, e0 + per);
  for (int64_t e = e0 + threadIdx.x; e < e1; e += BTA_THREADS) {
    double g;
    if (mode == 0) {
      const unsigned long long c = code[e];
      const int buf = (int)(c >> 60);
      const int64_t off = (int64_t)(c & ((1ull << 60) - 1));
      const double* base = buf == 0 ? D : (buf == 1 ? L : (buf == 2 ? B : T));
      g = base[off];
    } else {
      g = x[grow[e]] * x[gcol[e]];
    }
    acc = fma(wval[e], g, acc);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = BTA_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Entry (k, l) of a symmetric matrix held on the BTA pattern (D_i full
// symmetric blocks, L_i the block below D_i, B_i the arrow row, T the tip);
// NaN for a pair outside the pattern.
__device__ __forceinline__ double bta_sym_at(const double* __restrict__ D, const double* __restrict__ L,
                                             const double* __restrict__ B, const double* __restrict__ T,
                                             int b, int a, int64_t N, int64_t k, int64_t l) {
  if (k < l) {
    const int64_t t = k;
    k = l;
    l = t;
  }
  // now k >= l
  if (k >= N) {
    const int64_t r = k - N;
    if (l >= N) return T[r * a + (l - N)];
    const int64_t bl = l / b;
    return B[(bl * a + r) * b + (l - bl * b)];
  }
  const int64_t bk = k / b, bl = l / b;
  const int64_t ik = k - bk * b, il = l - bl * b;
  if (bk == bl) return D[(bk * b + ik) * b + il];
  if (bk == bl + 1) return L[(bl * b + ik) * b + il];
  return __builtin_nan("");
}

// out[j] = a_j^T S a_j for the rows a_j of a CSR matrix (one thread per row),
// S on the BTA pattern: the marginal variances diag(A Sigma A^T) of the
// observations once S holds the selected inverse.
__global__ void __launch_bounds__(BTA_THREADS) k_quadform_rows(const double* D, const double* L,
                                                               const double* B, const double* T, int b, int a,
                                                               int64_t N, const int* rowptr, const int* colidx,
                                                               const double* val, int64_t m, double* out) {
  for (int64_t j = (int64_t)blockIdx.x * BTA_THREADS + threadIdx.x; j < m; j += (int64_t)gridDim.x * BTA_THREADS) {
    const int p0 = rowptr[j], p1 = rowptr[j + 1];
    double acc = 0.0;
    for (int p = p0; p < p1; ++p) {
      const int64_t kp = colidx[p];
      const double vp = val[p];
      acc = fma(vp * vp, bta_sym_at(D, L, B, T, b, a, N, kp, kp), acc);
      for (int q = p + 1; q < p1; ++q) acc = fma(2.0 * vp * val[q], bta_sym_at(D, L, B, T, b, a, N, kp, colidx[q]), acc);
    }
    out[j] = acc;
  }
}

// Kronecker contractions, per block t, field pair (v, w) and spatial matrix k:
//   mode 0 (traces):     PD = <D_t[v,w], S_k>_F,  PL = <L_t[v,w], S_k>_F
//   mode 1 (quadratic):  PD = x_{t,v}^T S_k x_{t,w},  PL = x_{t+1,v}^T S_k x_{t,w}
// where X[v,w] is the (v, w) ns x ns sub-block of a b x b block (b = nv ns).
// grid: n * nv * chunks workgroups (64 mesh rows of one field each); partials
// per (t, v, chunk) are reduced afterwards in chunk order.
// part layout: [((t*nv + v)*chunks + ch) * width + (w*nK + k)*2 + {0: D, 1: L}],
// width = nv * nK * 2.
__global__ void __launch_bounds__(BTA_THREADS) k_kron_contract(const double* D, const double* L,
                                                               const double* x, const int* rowptr,
                                                               const int* colidx, const double* vals, int n,
                                                               int nv, int ns, int nnz, int nK, int mode,
                                                               double* part) {
  __shared__ double red[4][32];
  const int b = nv * ns;
  const int chunks = (ns + 63) / 64;
  const int t = blockIdx.x / (nv * chunks);
  const int v = (blockIdx.x / chunks) % nv;
  const int ch = blockIdx.x % chunks;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double accD[4][4], accL[4][4];
#pragma unroll
  for (int w = 0; w < 4; ++w)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      accD[w][k] = 0.0;
      accL[w][k] = 0.0;
    }
  for (int rr = wv; rr < 64; rr += 4) {
    const int s = ch * 64 + rr;
    if (s >= ns) break;
    const int r = v * ns + s;
    const int e0 = rowptr[s], e1 = rowptr[s + 1];
    for (int e = e0 + lane; e < e1; e += 64) {
      const int cs = colidx[e];
      double vk[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) vk[k] = (k < nK) ? vals[(int64_t)k * nnz + e] : 0.0;
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        if (w < nv) {
          const int c = w * ns + cs;
          double gD, gL = 0.0;
          if (mode == 0) {
            gD = D[((int64_t)t * b + r) * b + c];
            if (t < n - 1) gL = L[((int64_t)t * b + r) * b + c];
          } else {
            gD = x[(int64_t)t * b + r] * x[(int64_t)t * b + c];
            if (t < n - 1) gL = x[(int64_t)(t + 1) * b + r] * x[(int64_t)t * b + c];
          }
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            accD[w][k] = fma(vk[k], gD, accD[w][k]);
            accL[w][k] = fma(vk[k], gL, accL[w][k]);
          }
        }
      }
    }
  }
  const int width = nv * nK * 2;
#pragma unroll
  for (int w = 0; w < 4; ++w)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double d = accD[w][k], l = accL[w][k];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        d += __shfl_xor(d, off, 64);
        l += __shfl_xor(l, off, 64);
      }
      if (lane == 0 && w < nv && k < nK) {
        red[wv][(w * nK + k) * 2] = d;
        red[wv][(w * nK + k) * 2 + 1] = l;
      }
    }
  __syncthreads();
  if ((int)threadIdx.x < width) {
    const int q = threadIdx.x;
    part[((int64_t)(t * nv + v) * chunks + ch) * width + q] = (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
  }
}

// Sum partials over chunks in order: out[t][q] = sum_ch part[t][ch][q].
__global__ void k_reduce_chunks(const double* part, int n, int chunks, int width, double* out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * width) return;
  const int t = idx / width, q = idx % width;
  double s = 0.0;
  for (int ch = 0; ch < chunks; ++ch) s += part[((int64_t)t * chunks + ch) * width + q];
  out[idx] = s;
}

__global__ void k_reduce_all(const double* part, int npart, double* out) {
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

}  // namespace bta

// ------------------------------------------------------------ extern "C"

#include "../../include/pyinla_bta.h"

extern "C" {

int bta_assemble_kron(double* D, double* L, double* B, double* T, int n, int b, int a, int nv,
                      const int* rowptr, const int* colidx, const double* vals, int nnz, int nK,
                      const double* Mdiag, const double* Moff, int nM, const double* coef, double tip_diag,
                      void* stream) {
  if (!D || (n > 1 && !L) || !rowptr || !colidx || !vals || !Mdiag || (n > 1 && !Moff) || !coef) return BTA_E_ARG;
  if (nv < 1 || nv > 4 || b % nv != 0) return BTA_E_ARG;
  if (nK < 1 || nK > 4 || nM < 1 || nv * nv * nM * nK > BTA_MAX_KM || n < 1 || b < 1 || a < 0) return BTA_E_ARG;
  bta::KronArgs p;
  p.D = D;
  p.L = n > 1 ? L : nullptr;
  p.B = (a > 0) ? B : nullptr;
  p.T = (a > 0) ? T : nullptr;
  p.rowptr = rowptr;
  p.colidx = colidx;
  p.vals = vals;
  p.Mdiag = Mdiag;
  p.Moff = Moff;
  for (int q = 0; q < BTA_MAX_KM; ++q) p.coef[q] = q < nv * nv * nM * nK ? coef[q] : 0.0;
  p.tip_diag = tip_diag;
  p.n = n;
  p.b = b;
  p.a = a;
  p.nnz = nnz;
  p.nM = nM;
  p.nK = nK;
  p.nv = nv;
  p.ns = b / nv;
  const int chunks = (b + 3) / 4;
  hipLaunchKernelGGL(bta::k_assemble_kron, dim3(n * chunks), dim3(BTA_THREADS), 0,
                     reinterpret_cast<hipStream_t>(stream), p);
  return (int)hipGetLastError();
}

int bta_scatter_coo(double* D, double* L, double* B, double* T, const uint64_t* code, const double* val,
                    const int* src_idx, int64_t count, double scale, int accumulate, void* stream) {
  if (count < 0 || (count > 0 && (!code || !val))) return BTA_E_ARG;
  if (count == 0) return 0;
  const int grid = (int)std::min<int64_t>((count + BTA_THREADS - 1) / BTA_THREADS, 4096);
  hipLaunchKernelGGL(bta::k_scatter, dim3(grid), dim3(BTA_THREADS), 0, reinterpret_cast<hipStream_t>(stream), D, L,
                     B, T, reinterpret_cast<const unsigned long long*>(code), val, src_idx, count, scale, accumulate);
  return (int)hipGetLastError();
}

int bta_gather_coo(const double* D, const double* L, const double* B, const double* T, const uint64_t* code,
                   double* out, int64_t count, void* stream) {
  if (count < 0 || (count > 0 && (!code || !out))) return BTA_E_ARG;
  if (count == 0) return 0;
  const int grid = (int)std::min<int64_t>((count + BTA_THREADS - 1) / BTA_THREADS, 4096);
  hipLaunchKernelGGL(bta::k_gather, dim3(grid), dim3(BTA_THREADS), 0, reinterpret_cast<hipStream_t>(stream), D, L,
                     B, T, reinterpret_cast<const unsigned long long*>(code), out, count);
  return (int)hipGetLastError();
}

int bta_quadform_rows(const double* D, const double* L, const double* B, const double* T, int n, int b, int a,
                      const int* rowptr, const int* colidx, const double* val, int64_t m, double* out,
                      void* stream) {
  if (n < 1 || b < 1 || a < 0 || m < 0 || !D || (n > 1 && !L) || (a > 0 && (!B || !T))) return BTA_E_ARG;
  if (m == 0) return 0;
  if (!rowptr || !colidx || !val || !out) return BTA_E_ARG;
  const int grid = (int)std::min<int64_t>((m + BTA_THREADS - 1) / BTA_THREADS, 8192);
  hipLaunchKernelGGL(bta::k_quadform_rows, dim3(grid), dim3(BTA_THREADS), 0, reinterpret_cast<hipStream_t>(stream),
                     D, L, B, T, b, a, (int64_t)n * b, rowptr, colidx, val, m, out);
  return (int)hipGetLastError();
}

int bta_bilinear_coo(const double* D, const double* L, const double* B, const double* T, const uint64_t* code,
                     const int* grow, const int* gcol, const double* x, const double* wval, int64_t count,
                     int mode, double* work, int nwork, double* out, void* stream) {
  if (!wval || !out || !work || nwork < 1 || (mode == 0 && !code) || (mode == 1 && (!grow || !gcol || !x)))
    return BTA_E_ARG;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(nwork, (count + BTA_THREADS - 1) / BTA_THREADS));
  hipLaunchKernelGGL(bta::k_bilinear_partial, dim3(grid), dim3(BTA_THREADS), 0, st, D, L, B, T,
                     reinterpret_cast<const unsigned long long*>(code), grow, gcol, x, wval, count, mode, work);
  hipLaunchKernelGGL(bta::k_reduce_all, dim3(1), dim3(BTA_THREADS), 0, st, work, grid, out);
  return (int)hipGetLastError();
}

int bta_kron_contract(const double* D, const double* L, const double* x, int n, int b, int nv, const int* rowptr,
                      const int* colidx, const double* vals, int nnz, int nK, int mode, double* work,
                      double* out, void* stream) {
  if (nK < 1 || nK > 4 || nv < 1 || nv > 4 || b % nv != 0 || !rowptr || !colidx || !vals || !work || !out)
    return BTA_E_ARG;
  if ((mode == 0 && !D) || (mode == 1 && !x)) return BTA_E_ARG;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const int ns = b / nv;
  const int chunks = (ns + 63) / 64;
  hipLaunchKernelGGL(bta::k_kron_contract, dim3(n * nv * chunks), dim3(BTA_THREADS), 0, st, D, L, x, rowptr,
                     colidx, vals, n, nv, ns, nnz, nK, mode, work);
  const int width = nv * nK * 2;
  hipLaunchKernelGGL(bta::k_reduce_chunks, dim3((n * nv * width + 255) / 256), dim3(256), 0, st, work, n * nv,
                     chunks, width, out);
  return (int)hipGetLastError();
}

}  // extern "C"
