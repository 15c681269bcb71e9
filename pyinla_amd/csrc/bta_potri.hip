// Block inverses of the diagonal factor blocks and the BTA selected inversion
// (pobtasi / pobtsi) on gfx950.
//
// W_i = D_i^-1 for every diagonal factor block is built bottom-up from the
// 64x64 tile inverses the factorization already produced: ranges of 2^s tiles
// are merged pairwise with
//     [[W11, 0], [W21, W22]],  W21 = -W22 (L21 W11),
// where T1^T = W11^T L21^T is parked in the (still unused) upper triangle of
// W_i.  Every level is two grouped-GEMM launches over all pairs of all n
// blocks, so the whole inverse costs ~2 log2(b/64) launches.  The parked
// blocks stay: every reader of W (the GEMV sweeps, G_i, Ga_i, W_i^T W_i and
// the merge levels themselves) masks the strict upper triangle element by
// element, so it is never cleared.
//
// Selected inversion then runs the Takahashi recurrence backwards in i with
// G_i = L_i W_i and Ga_i = B_i W_i:
//     Sigma_{n,i}   = -Sigma_{n,i+1} G_i - Sigma_T Ga_i
//     Sigma_{i+1,i} = -Sigma_{i+1,i+1} G_i - Sigma_{n,i+1}^T Ga_i
//     Sigma_ii      =  W_i^T W_i - G_i^T Sigma_{i+1,i} - Ga_i^T Sigma_{n,i}
// with Sigma_T = T^-T T^-1.  G_i, Ga_i and W_i^T W_i do not depend on the
// recurrence: one launch computes them for every block at once (full
// occupancy), leaving two dependent launches per block.  Sigma_ii is written
// full symmetric.
#include <rocblas/rocblas.h>

#include "bta_internal.h"
#include "../../include/pyinla_bta.h"

#ifndef BTA_SELINV_SPLIT
#define BTA_SELINV_SPLIT 1
#endif

namespace bta {

int potri_build(Workspace& ws, PotriPlan& pp, double* D, double* L, double* B, double* T, int n, int b, int a);

__global__ void __launch_bounds__(BTA_THREADS) k_leaf_copy(double* W, const double* winv, int b, int nT) {
  const int blk = blockIdx.x / nT, k = blockIdx.x % nT;
  const int mv = min(64, b - k * 64);
  const double* src = winv + ((int64_t)blk * nT + k) * 4096;
  double* dst = W + (int64_t)blk * b * b + (int64_t)(k * 64) * b + k * 64;
  const int c = threadIdx.x & 63;
  for (int r = threadIdx.x >> 6; r < mv; r += 4)
    if (c < mv) dst[(int64_t)r * b + c] = src[r * 64 + c];
}

size_t Workspace::plan_bytes() const {
  size_t t = 0;
  for (const auto& kv : trtri_plans) t += kv.second->bytes;
  for (const auto& kv : update_plans) t += kv.second->bytes;
  for (const auto& kv : potri_plans) t += kv.second->grouped.bytes + kv.second->with_blas.bytes;
  if (flow) {
    t += flow->buf_bytes;
    for (const auto& sc : flow->scheds) t += sc.bytes;
  }
  return t;
}

void Workspace::clear_plans() {
  trtri_plans.clear();
  update_plans.clear();
  potri_plans.clear();
  trtri_cur = update_cur = nullptr;
  potri_cur = nullptr;
}

int trtri_prepare(Workspace& ws, const double* D, int n, int b) {
  if (!ws.wfull) return BTA_E_NORES;  // W is reserved with the workspace (bta_ws_reserve)
  const PlanKey key{{D, nullptr, nullptr, nullptr}, 0, ws.tile_grouped, 0};
  int rc = 0;
  ws.trtri_cur = plan_lookup(ws, ws.trtri_plans, key, [&](DevPlan& plan) {
    const int nT = (b + 63) / 64;
    std::vector<GemmLaunch> ls;
    for (int s = 1; s < nT; s *= 2) {
      const int ts = choose_tile(ws.tile_grouped, s * 64, s * 64);
      GemmLaunch la(ts), lb(ts);
      for (int t0 = 0; t0 + s < nT; t0 += 2 * s) {
        const int r1 = t0 * 64, r2 = (t0 + s) * 64;
        const int m1 = s * 64;
        const int m2 = min((t0 + 2 * s) * 64, b) - r2;
        for (int i = 0; i < n; ++i) {
          double* W = ws.wfull + (int64_t)i * b * b;
          const double* Di = D + (int64_t)i * b * b;
          // T1^T = W11^T L21^T  -> upper block (r1, r2), m1 x m2
          GemmJob ja = job(W + (int64_t)r1 * b + r2, m1, m2, b, 0.0);
          add_seg(ja, seg(W + (int64_t)r1 * b + r1, b, 1, Di + (int64_t)r2 * b + r1, b, 1, m1, 1.0, TRI_A_UPPER));
          la.add(ja);
          // W21 = -W22 T1  -> lower block (r2, r1), m2 x m1
          GemmJob jb = job(W + (int64_t)r2 * b + r1, m2, m1, b, 0.0);
          add_seg(jb, seg(W + (int64_t)r2 * b + r2, b, 0, W + (int64_t)r1 * b + r2, b, 1, m2, -1.0, TRI_A_LOWER));
          lb.add(jb);
        }
      }
      ls.push_back(std::move(la));
      ls.push_back(std::move(lb));
    }
    return plan.build(ls);
  }, &rc);
  return rc;
}

int trtri_enqueue(Workspace& ws, const double* D, int n, int b, hipStream_t st) {
  const int nT = (b + 63) / 64;
  hipLaunchKernelGGL(k_leaf_copy, dim3(n * nT), dim3(BTA_THREADS), 0, st, ws.wfull, ws.winv, b, nT);
  for (size_t q = 0; q < ws.trtri_cur->launches.size(); ++q) {
    TimedScope ts(ws, KC_GROUPED, st);
    ws.trtri_cur->launch((int)q, st, ws.gemm_variant);
  }
  return (int)hipGetLastError();
}

int potri_prepare(Workspace& ws, double* D, double* L, double* B, double* T, int n, int b, int a) {
  int rc = trtri_prepare(ws, D, n, b);
  if (rc) return rc;
  // G_i = L_i W_i and Ga_i = B_i W_i for every block (they depend on the
  // factor only), kept for the whole recurrence in the reserved buffer
  // (n b^2 + n max(ws.a, 1) b doubles; a <= ws.a)
  if (!ws.gfull) return BTA_E_NORES;
  if (a > ws.a) return BTA_E_SHAPE;
  const PlanKey key{{D, L, B, T}, a, ws.tile_grouped, ws.use_blas};
  ws.potri_cur = plan_lookup(ws, ws.potri_plans, key, [&](PotriPlan& pp) {
    return potri_build(ws, pp, D, L, B, T, n, b, a);
  }, &rc);
  return rc;
}

int potri_build(Workspace& ws, PotriPlan& pp, double* D, double* L, double* B, double* T, int n, int b, int a) {
  int rc = 0;
  double* Gall = ws.gfull;
  double* Gaall = ws.gfull + (size_t)n * b * b;
  std::vector<GemmLaunch> ls;
  const int ts = choose_tile(ws.tile_grouped, b, b);
  // Sigma_ii's update (lower-triangle tiles, K = b) as split tiles: alone it
  // would fill only about half of the resident slots with workgroups that each
  // run the whole K loop
  const int split_sd = (BTA_SELINV_SPLIT && ts == 64 && b >= 1024) ? 2 : 1;
  // One launch of everything the recurrence does not depend on:
  //   G_i = L_i W_i, Ga_i = B_i W_i, Sigma_ii <- W_i^T W_i (the first term of
  //   Sigma_ii; D_i is free once W_i exists), Sigma_T = T^-T T^-1.
  {
    GemmLaunch l0(ts);
    for (int i = 0; i < n; ++i) {
      double* Di = D + (int64_t)i * b * b;
      double* Wi = ws.wfull + (int64_t)i * b * b;
      if (i < n - 1) {
        GemmJob j = job(Gall + (int64_t)i * b * b, b, b, b, 0.0);
        add_seg(j, seg(L + (int64_t)i * b * b, b, 0, Wi, b, 0, b, 1.0, TRI_B_LOWER));
        l0.add(j);
      }
      if (a > 0) {
        GemmJob j = job(Gaall + (int64_t)i * a * b, a, b, b, 0.0);
        add_seg(j, seg(B + (int64_t)i * a * b, b, 0, Wi, b, 0, b, 1.0, TRI_B_LOWER));
        l0.add(j);
      }
      GemmJob sd = job(Di, b, b, b, 0.0, /*lower=*/1, /*mirror=*/1);
      add_seg(sd, seg(Wi, b, 1, Wi, b, 0, b, 1.0, TRI_A_UPPER | TRI_B_LOWER));
      l0.add(sd);
    }
    if (a > 0) {
      GemmJob j = job(T, a, a, a, 0.0);
      add_seg(j, seg(ws.tinv, 64, 1, ws.tinv, 64, 0, a, 1.0));
      l0.add(j);
    }
    ls.push_back(std::move(l0));
  }
  // The recurrence, two dependent launches per block (i = n-1 .. 0):
  //   l2(i): Sigma_{i+1,i} = -Sigma_{i+1,i+1} G_i - Sigma_{n,i+1}^T Ga_i   (into L_i)
  //   l3(i): Sigma_ii     +=  - G_i^T Sigma_{i+1,i} - Ga_i^T Sigma_{n,i}  (into D_i)
  //          Sigma_{n,i-1} = -Sigma_{n,i} G_{i-1} - Sigma_T Ga_{i-1}     (into B_{i-1})
  // Sigma_{n,i-1} needs only Sigma_{n,i}, so it rides in l3(i), whose
  // lower-triangle tiles fill half of the resident slots, and l2 keeps its
  // b x b tiles (one full wave at b = 2000).  Sigma_{n,n-1} = -Sigma_T Ga_{n-1}
  // is l2(n-1).  Ga_i = B_i W_i was formed in l0, so B_{i-1} is free.
  auto arrow_job = [&](int i) {  // Sigma_{n,i}
    double* Bi = B + (int64_t)i * a * b;
    GemmJob s2 = job(Bi, a, b, b, 0.0);
    if (i < n - 1) add_seg(s2, seg(B + (int64_t)(i + 1) * a * b, b, 0, Gall + (int64_t)i * b * b, b, 0, b, -1.0));
    add_seg(s2, seg(T, a, 0, Gaall + (int64_t)i * a * b, b, 0, a, -1.0));
    return s2;
  };
  for (int i = n - 1; i >= 0; --i) {
    const bool hasL = i < n - 1;
    double* Di = D + (int64_t)i * b * b;
    double* Li = hasL ? L + (int64_t)i * b * b : nullptr;
    double* Bi = a > 0 ? B + (int64_t)i * a * b : nullptr;
    double* Dn = hasL ? D + (int64_t)(i + 1) * b * b : nullptr;
    double* Bn = (hasL && a > 0) ? B + (int64_t)(i + 1) * a * b : nullptr;
    double* G = hasL ? Gall + (int64_t)i * b * b : nullptr;
    double* Ga = a > 0 ? Gaall + (int64_t)i * a * b : nullptr;
    GemmLaunch l2(ts), l3(ts);
    if (a > 0 && !hasL) l2.add(arrow_job(i));
    if (hasL) {
      GemmJob s2 = job(Li, b, b, b, 0.0);  // Sigma_{i+1,i}
      add_seg(s2, seg(Dn, b, 0, G, b, 0, b, -1.0));
      if (a > 0) add_seg(s2, seg(Bn, b, 1, Ga, b, 0, a, -1.0));
      l2.add(s2);
    }
    // the few Sigma_{n,i-1} tiles go first: dispatched ahead of the b x b
    // tiles they start at once instead of trailing the launch
    if (a > 0 && i > 0) l3.add(arrow_job(i - 1));
    if (hasL || a > 0) {
      GemmJob sd = job(Di, b, b, b, 1.0, /*lower=*/1, /*mirror=*/1);  // Sigma_ii
      sd.split = split_sd;
      if (hasL) add_seg(sd, seg(G, b, 1, Li, b, 0, b, -1.0));
      if (a > 0) add_seg(sd, seg(Ga, b, 1, Bi, b, 0, a, -1.0));
      l3.add(sd);
    }
    if (l2.tiles() > 0) ls.push_back(std::move(l2));
    if (l3.tiles() > 0) ls.push_back(std::move(l3));
  }
  rc = pp.grouped.build(ls);
  if (rc) return rc;
  if (ws.use_blas && n > 1) {
    // The same recurrence with Sigma_{i+1,i} = -Sigma_{i+1,i+1} G_i as one
    // rocBLAS dgemm per block.  Its arrow term -Sigma_{n,i+1}^T Ga_i is written
    // into L_i one launch earlier, by l3(i+1) (Sigma_{n,i+1} is final there, L_i
    // is free once G_i exists), and the dgemm accumulates onto it.
    if (!ws.blas) {
      rocblas_handle h;
      if (rocblas_create_handle(&h) != rocblas_status_success) return BTA_E_LIB;
      ws.blas = h;
    }
    std::vector<GemmLaunch> lb;
    lb.push_back(ls.front());  // l0
    pp.order.push_back(0);
    for (int i = n - 1; i >= 0; --i) {
      const bool hasL = i < n - 1;
      double* Di = D + (int64_t)i * b * b;
      double* Li = hasL ? L + (int64_t)i * b * b : nullptr;
      double* Bi = a > 0 ? B + (int64_t)i * a * b : nullptr;
      double* Ga = a > 0 ? Gaall + (int64_t)i * a * b : nullptr;
      if (a > 0 && !hasL) {
        GemmLaunch l2(ts);
        l2.add(arrow_job(i));
        pp.order.push_back((int)lb.size());
        lb.push_back(std::move(l2));
      }
      if (hasL) {
        pp.order.push_back(-1 - (int)pp.gemms.size());
        pp.gemms.push_back({D + (int64_t)(i + 1) * b * b, Gall + (int64_t)i * b * b, Li, a > 0 ? 1.0 : 0.0});
      }
      GemmLaunch l3(ts);
      if (a > 0 && i > 0) {
        l3.add(arrow_job(i - 1));
        GemmJob s2 = job(L + (int64_t)(i - 1) * b * b, b, b, b, 0.0);  // arrow term of Sigma_{i,i-1}
        add_seg(s2, seg(Bi, b, 1, Gaall + (int64_t)(i - 1) * a * b, b, 0, a, -1.0));
        l3.add(s2);
      }
      if (hasL || a > 0) {
        GemmJob sd = job(Di, b, b, b, 1.0, /*lower=*/1, /*mirror=*/1);  // Sigma_ii
        sd.split = split_sd;
        if (hasL) add_seg(sd, seg(Gall + (int64_t)i * b * b, b, 1, Li, b, 0, b, -1.0));
        if (a > 0) add_seg(sd, seg(Ga, b, 1, Bi, b, 0, a, -1.0));
        l3.add(sd);
      }
      if (l3.tiles() > 0) {
        pp.order.push_back((int)lb.size());
        lb.push_back(std::move(l3));
      }
    }
    rc = pp.with_blas.build(lb);
    if (rc) return rc;
    pp.has_blas = true;
  }
  return 0;
}

int potri_enqueue(Workspace& ws, double* D, double* L, double* B, double* T, int n, int b, int a,
                  hipStream_t st) {
  (void)L;
  (void)B;
  (void)T;
  if (!ws.w_valid || ws.w_of != D) {
    const int rc = trtri_enqueue(ws, D, n, b, st);
    if (rc) return rc;
  }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cap) != hipSuccess) cap = hipStreamCaptureStatusNone;
  PotriPlan& pp = *ws.potri_cur;
  if (!pp.has_blas || cap != hipStreamCaptureStatusNone) {
    for (size_t q = 0; q < pp.grouped.launches.size(); ++q) {
      TimedScope ts(ws, KC_GROUPED, st);
      pp.grouped.launch((int)q, st, ws.gemm_variant);
    }
    return (int)hipGetLastError();
  }
  rocblas_handle h = static_cast<rocblas_handle>(ws.blas);
  if (rocblas_set_stream(h, st) != rocblas_status_success) return BTA_E_LIB;
  const double minus_one = -1.0;
  for (const int q : pp.order) {
    if (q >= 0) {
      TimedScope ts(ws, KC_GROUPED, st);
      pp.with_blas.launch(q, st, ws.gemm_variant);
      continue;
    }
    const PotriPlan::BlasGemm& g = pp.gemms[-1 - q];
    TimedScope ts(ws, KC_BLAS, st);
    // row-major C = A B is column-major C^T = B^T A^T
    if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, b, b, b, &minus_one, g.B, b, g.A, b,
                      &g.beta, g.C, b) != rocblas_status_success)
      return BTA_E_LIB;
  }
  return (int)hipGetLastError();
}

}  // namespace bta
