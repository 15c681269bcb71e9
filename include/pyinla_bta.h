/*
 * pyinla_bta.h -- C ABI of the MI355X (gfx950) BTA solver library
 * (libpyinla_bta.so).
 *
 * The library replaces the arithmetic that pyINLA's SerinvSolver hands to the
 * external Serinv package, on block-tridiagonal-arrowhead (BTA) matrices:
 *
 *   D  (n, b, b)   diagonal blocks          A_diagonal_blocks
 *   L  (n-1, b, b) sub-diagonal blocks      A_lower_diagonal_blocks
 *   B  (n, a, b)   arrow (bottom) blocks    A_arrow_bottom_blocks
 *   T  (a, a)      arrow tip                A_arrow_tip_block
 *
 * (src/pyinla/solvers/structured_solver.py:36-60).  All blocks are FP64,
 * row-major, contiguous, with leading dimension b (a for T).  a == 0 selects
 * the BT variant (no arrowhead: B and T may be NULL).
 *
 * Conventions
 *   - Pointers are device pointers on the workspace's device.
 *   - stream is a hipStream_t passed as void* (NULL = default stream).
 *   - Every function returns 0 on success, a negative BTA_E_* code on invalid
 *     arguments, or a positive hipError_t code from the runtime.
 *   - Device memory.  Every buffer the arithmetic touches is allocated by
 *     bta_ws_create or bta_ws_reserve and counted by bta_ws_bytes /
 *     bta_ws_allocated: the per-tile inverses, flags and scratch, the
 *     persistent panel's exchange buffers, and (reserved on request) the block
 *     inverses W_i (n b^2 doubles) and the selected inversion's G_i / Ga_i
 *     (n b^2 + n max(a,1) b doubles).  The only other device allocations are
 *     launch plans -- job tables and task lists of a few hundred KB -- built
 *     the first time a call sees a new set of block pointers (or ahead of time
 *     by bta_ws_prepare) and kept with the workspace (at most 16 per kind; a
 *     full cache is emptied, which synchronises the device and drops the
 *     recorded graphs).  bta_potri additionally creates a rocBLAS handle on
 *     its first call with BTA_OPT_BLAS on.  Repeated calls with the same
 *     pointers allocate nothing.
 *   - A solve or selected inversion on a workspace without the reservation it
 *     needs returns BTA_E_NORES (no hidden allocation).
 *   - Calls are asynchronous on the stream; results that a host needs
 *     (logdet, info) are read with bta_ws_status, which synchronises.
 */
#ifndef PYINLA_BTA_H
#define PYINLA_BTA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BTA_OK 0
#define BTA_E_ARG (-1)      /* invalid shape / pointer / option */
#define BTA_E_SHAPE (-2)    /* shape does not match the workspace */
#define BTA_E_TIMEOUT (-3)  /* a bounded in-kernel wait timed out */
#define BTA_E_LIB (-4)      /* a rocBLAS call failed */
#define BTA_E_NORES (-5)    /* the workspace lacks the reservation the call needs */
#define BTA_E_SCRATCH (-6)  /* the shape exceeds the fixed scratch of bta_potrf_noalloc */

typedef struct bta_ws* bta_ws_t;

/* Library version (major*10000 + minor*100 + patch). */
int bta_version(void);

/* Workspace flags (bta_ws_create, bta_ws_reserve, bta_ws_bytes). */
#define BTA_WS_GRAPH 1   /* capture each (operation, pointers) sequence once in a hipGraph and replay it */
#define BTA_WS_SOLVE 2   /* reserve the block inverses W_i used by bta_potrs (n b^2 doubles) */
#define BTA_WS_SELINV 4  /* reserve W_i and the G_i / Ga_i of bta_potri (2 n b^2 + n max(a,1) b doubles) */

/* Device workspace for one solver instance of shape (n, b, a), a <= 64.  Holds
 * the inverse of every 64x64 diagonal tile of the factor, readiness flags, the
 * log-determinant partials, the scratch of the solves, the persistent panel's
 * exchange buffers (b > 64) and the buffers `flags` reserves.  Replaces the
 * buffer ownership of SerinvSolver.__init__ (structured_solver.py:22-70),
 * which allocates everything once; the block buffers themselves are owned by
 * the caller. */
int bta_ws_create(int device, int n, int b, int a, int flags, bta_ws_t* out);
int bta_ws_destroy(bta_ws_t ws);
/* Reserve more buffers (BTA_WS_SOLVE / BTA_WS_SELINV) on an existing
 * workspace; a no-op for what it already holds.  Not a hot call. */
int bta_ws_reserve(bta_ws_t ws, int flags);
/* Bytes bta_ws_create allocates for (n, b, a, flags), plans excluded. */
size_t bta_ws_bytes(int n, int b, int a, int flags);
/* Bytes the workspace holds now: buffers, reservations and cached plans.
 * Replaces SerinvSolver.get_solver_memory's count of solver-owned memory
 * (structured_solver.py:540-550). */
size_t bta_ws_allocated(bta_ws_t ws);
/* Build ahead of time the launch plans of the operations in `ops` (1 =
 * bta_potrf, 2 = bta_potrs, 4 = bta_potri; bits may be combined) for these
 * block pointers, so that the calls that follow allocate nothing (e.g. before
 * graph capture).  No reference counterpart. */
int bta_ws_prepare(bta_ws_t ws, double* D, double* L, double* B, double* T, int n, int b, int a, int ops);

/* In-place BTA (a > 0) or BT (a == 0) Cholesky, Q = LL^T, with the
 * log-determinant accumulated on the fly.  Replaces serinv.pobtaf(D, L, B, T)
 * / serinv.pobtf(D, L) as called by SerinvSolver.cholesky
 * (structured_solver.py:83-93).  On exit D[i] holds the lower factor block with
 * zero strict upper triangle, L[i], B[i] the factor's sub-diagonal and arrow
 * blocks, T the tip factor.  A non-positive pivot is reported through
 * bta_ws_status (info), as LAPACK's potrf info. */
int bta_potrf(bta_ws_t ws, double* D, double* L, double* B, double* T, int n, int b, int a,
              void* stream);

/* bta_potrf without a workspace and without any device allocation.  No
 * counterpart in the reference: for callers that cannot allocate device memory
 * (code under hipGraph stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same factorization and outputs as bta_potrf, computed by
 * the per-tile-column panel path; every buffer beyond the caller's blocks and
 * outputs lives in a fixed device array of the library (3.1 MiB per GPU, held
 * until the process ends).  Limits: b <= 4096, a <= 64, n * ceil(b/64) + 1 <=
 * 131072 and n < 2^20; larger shapes return BTA_E_SCRATCH before anything is
 * enqueued.  Runs on the current HIP device (pointers must live there).
 * Outputs, device pointers, each may be NULL: logdet[0] as bta_ws_status
 * reports it, info[0] (0, or 1 + global row of the first non-positive pivot),
 * err[0] (BTA_E_TIMEOUT if a bounded wait expired, else 0).  All calls on one
 * GPU share the array: calls that overlap on one GPU (different streams, or
 * threads without ordering) race and must not be issued. */
int bta_potrf_noalloc(double* D, double* L, double* B, double* T, int n, int b, int a, double* logdet,
                      int64_t* info, int* err, void* stream);

/* Batched factorization of independent problems that may differ in (n, b, a)
 * (the 2 dim(theta) + 1 finite-difference evaluations, prior / conditional
 * pairs, mixed-size batches).  Problem q uses workspace ws[q] (created for its
 * n, b and an arrowhead >= a[q]) and blocks D[q], L[q], B[q], T[q]; a[q] == 0
 * factors the BT part only.  Every block step / tile column is one launch over
 * all problems that have it, and bta_batch_potrf replays the whole sequence as
 * one hipGraph (use_graph != 0).  Results (logdet, info) are read per problem
 * with bta_ws_status.  No reference counterpart in one process: the reference
 * spreads these evaluations over MPI ranks (pyinla.py:93-106,541-572). */
typedef struct bta_batch* bta_batch_t;
int bta_batch_create(int nprob, bta_ws_t* ws, double** D, double** L, double** B, double** T,
                     const int* a, int use_graph, bta_batch_t* out);
int bta_batch_potrf(bta_batch_t batch, void* stream);
int bta_batch_destroy(bta_batch_t batch);

/* Synchronise the stream and return the status of the last bta_potrf:
 *   info   : 0 if positive definite, else 1 + global row of the first failing
 *            pivot (LAPACK convention);
 *   logdet : 2 * sum log diag of the factor (the reference's
 *            SerinvSolver.logdet, structured_solver.py:151-166; for a == 0 the
 *            tip is excluded, as sparsity="bt" does);
 *   err    : BTA_E_TIMEOUT if a bounded wait expired, else 0.
 * Any output pointer may be NULL. */
int bta_ws_status(bta_ws_t ws, int64_t* info, double* logdet, int* err, void* stream);

/* Device address of the FP64 log-determinant written by the last bta_potrf
 * (valid after the stream reaches that point; no host synchronisation). */
double* bta_ws_logdet_ptr(bta_ws_t ws);

/* out[0] (device) = 2 sum_i sum_j log D_i[j][j] (+ 2 sum_j log T[j][j] if
 * a > 0) of the blocks as they are held now, summed in a fixed order.  This is
 * SerinvSolver.logdet(sparsity) (structured_solver.py:151-166): it reads the
 * current blocks, so after a BTA factorization a == 0 excludes the tip, as
 * logdet("bt") does, and it may follow any factorization (a == 0 or not). */
int bta_logdet(bta_ws_t ws, const double* D, const double* T, int n, int b, int a, double* out, void* stream);

/* Tuning options (no reference counterpart): the output-tile edge (64 or
 * 128; 0 = built-in heuristic) of the FP64-MFMA GEMMs used for the block
 * updates of bta_potrf (BTA_OPT_TILE_UPDATE) and for the block inverses and
 * selected inversion (BTA_OPT_TILE_GROUPED).  Set by the autotuner
 * (pyinla_amd/autotune.py); recorded graphs are dropped. */
#define BTA_OPT_TILE_UPDATE 1
#define BTA_OPT_TILE_GROUPED 2
/* Panel factorization of bta_potrf: 0 = automatic, 1 = one launch per 64-wide
 * tile column, 2 = one persistent dataflow launch per block (a chain
 * workgroup on the diagonal plus helper workgroups running a dependency-
 * levelled task list).  BTA_OPT_FLOW_HELPERS sets the helper count (0 =
 * automatic: half the device's resident capacity, less one) and
 * BTA_OPT_FLOW_DELTA the order of the helpers' task lists: 0 (default) = the
 * order of a host-side simulation of the launch, critical path first; d >= 1
 * = dependency-level order, a tile away from the diagonal receiving its
 * rank-64 updates in groups of d tile columns. */
#define BTA_OPT_PANEL 3
#define BTA_OPT_FLOW_HELPERS 4
#define BTA_OPT_FLOW_DELTA 5
/* Batched factorization in the persistent form: problems factored
 * concurrently per wave (taken from the first problem's workspace; 0 = 6,
 * at most 8). */
#define BTA_OPT_BATCH_LANES 6
/* Selected inversion: 1 (default) = the b x b x b products
 * Sigma_{i+1,i} = -Sigma_{i+1,i+1} G_i as rocBLAS dgemm calls on the caller's
 * stream (not while that stream is under graph capture), 0 = the grouped
 * FP64-MFMA kernel for every product. */
#define BTA_OPT_BLAS 7
/* Persistent panel: the near-diagonal tiles ((c,c), (c+1,c), (c+2,c)), whose
 * updates and solves the chain waits for, go to a pool of helpers / value
 * helpers (at least min(10, helpers / 8)) that own nothing else (default 16;
 * 1..64). */
#define BTA_OPT_FLOW_NEAR 8
/* 64x64 grouped GEMM main loop: 0 (default) = one LDS buffer, the next K
 * chunk loaded into registers while the current one is multiplied; 1 = two
 * LDS buffers with loads issued two K chunks ahead (slower on MI355X: 470 us
 * against 395 us for NN 2000^3; kept for comparison). */
#define BTA_OPT_GEMM_KERNEL 9
/* Persistent panel, block update D_{i+1} -= L_i L_i^T, B_{i+1} -= B_i L_i^T,
 * T -= B_i B_i^T: 1 = applied by each block's launch through helper tasks
 * that fill the gaps of the factorization; 2 = the launch only factors the
 * panel and the update follows as one grouped-GEMM launch; 0 (default) =
 * automatic (1 up to 48 tile columns, b <= 3072, else 2). */
#define BTA_OPT_FLOW_FOLD 10
int bta_ws_set_option(bta_ws_t ws, int option, int value);

/* Host-only diagnostic (no reference counterpart): the helper task lists of
 * the persistent panel launch for a panel of nT tile columns, nTL L tile rows
 * and nTa (0/1) arrow tile rows.  Task words are op << 30 | row << 20 |
 * col << 10 | k (op 0 = update with column k, 1 = solve, 2 = clear);
 * off[0..helpers] delimit the helpers' lists; near_div is BTA_OPT_FLOW_NEAR
 * (0 = the default).  With tasks == off == NULL only
 * *ntasks is returned.  Used by the CPU schedule simulator in the tests. */
int bta_flow_schedule(int nT, int nTL, int nTa, int helpers, int delta, int near_div, uint32_t* tasks, int64_t cap,
                      int* off, int64_t* ntasks);

/* Optional event timing (no reference counterpart; the reference times whole
 * calls with perf_counter, structured_solver.py:80-100).  While on, graphs are
 * bypassed and every launch of the classes 0 = panel factorization step,
 * 1 = block update GEMM, 2 = grouped GEMM (inverses / selected inversion),
 * 3 = GEMV sweep, 4 = rocBLAS dgemm of the selected inversion is bracketed
 * by a pair of hipEvents on its stream.  bta_ws_get_timing synchronises and returns the summed milliseconds and the
 * number of launches of a class since timing was switched on or reset. */
int bta_ws_set_timing(bta_ws_t ws, int on);
int bta_ws_reset_timing(bta_ws_t ws);
int bta_ws_get_timing(bta_ws_t ws, int kclass, double* total_ms, int64_t* count);

/* In-place triangular sweep with the factor held in D, L, B, T (as left by
 * bta_potrf).  trans == 0 solves LL y = rhs (serinv.pobtas(..., trans="N")),
 * trans == 1 solves LL^T x = rhs (serinv.pobtas(..., trans="C")); the pair
 * replaces SerinvSolver.solve (structured_solver.py:102-149).  a == 0 gives
 * pobts.  rhs is (n*b + a) x nrhs, row-major, leading dimension ldrhs
 * (ldrhs >= nrhs, any nrhs).  The first call after a factorization
 * also forms W_i = D_i^-1 for every block in the reserved buffer
 * (BTA_WS_SOLVE or BTA_WS_SELINV; BTA_E_NORES without it).  Launched
 * directly, never from a recorded graph. */
int bta_potrs(bta_ws_t ws, double* D, double* L, double* B, double* T, int n, int b, int a,
              double* rhs, int nrhs, int ldrhs, int trans, void* stream);

/* In-place selected inversion: the factor blocks are overwritten by the
 * blocks of Sigma = Q^-1 on the BTA pattern -- D[i] <- Sigma_ii (full
 * symmetric), L[i] <- Sigma_{i+1,i}, B[i] <- Sigma_{n,i}, T <- Sigma_T.
 * Replaces serinv.pobtasi(D, L, B, T) / pobtsi(D, L) as called by
 * SerinvSolver.selected_inversion (structured_solver.py:168-190).  With
 * BTA_OPT_BLAS on (the default) the Sigma_{i+1,i} products are rocBLAS dgemm
 * calls on `stream`; everything else runs in this library's kernels.  Returns
 * BTA_E_LIB if a rocBLAS call fails, BTA_E_NORES without BTA_WS_SELINV. */
int bta_potri(bta_ws_t ws, double* D, double* L, double* B, double* T, int n, int b, int a,
              void* stream);

/* Block inverses W_i = D_i^-1 of the diagonal factor blocks, enqueued on
 * `stream` and kept in the workspace until the next factorization: the
 * first bta_potrs / bta_potri after a factorization builds them, this lets a
 * caller build them earlier (e.g. while another solver's work leaves the
 * device idle).  No counterpart in the reference. */
int bta_block_inverses(bta_ws_t ws, double* D, int n, int b, void* stream);

/* ---------------------------------------------------------- assembly */

/* Dense BTA blocks of Q = sum_{v,w,m,k} coef[((v*nv+w)*nM+m)*nK+k] *
 * (M_m (x) E_vw (x) S_k) written straight into D (n,b,b) and L (n-1,b,b),
 * every entry (zeros included); nv fields (nv <= 4, b = nv*ns) are interleaved
 * per time block as the coregional model orders them
 * (coregional_model.py:754-786); nv = 1 is the single-field case.
 * Spatial S_k share one union CSR pattern (rowptr[b+1], colidx[nnz]) with
 * values vals[k*nnz + e] (nK <= 4); temporal M_m are tridiagonal, given by
 * Mdiag[m*n + t] and Moff[m*(n-1) + t] (nM*nK <= 16).  If a > 0 the arrow
 * blocks B are cleared and the tip T set to tip_diag * I (the fixed-effects
 * prior).  Replaces the sparse Kronecker assembly of
 * SpatioTemporalSubModel.construct_Q_prior (spatio_temporal.py:114-149), the
 * regression tip (regression.py:28-34), Model.construct_Q_prior
 * (model.py:306-382) and the sparse->BTA scatter that follows it
 * (structured_solver.py:192-250). */
int bta_assemble_kron(double* D, double* L, double* B, double* T, int n, int b, int a, int nv,
                      const int* rowptr, const int* colidx, const double* vals, int nnz, int nK,
                      const double* Mdiag, const double* Moff, int nM, const double* coef,
                      double tip_diag, void* stream);

/* Scatter COO values into the blocks: dst(code[e]) (+)= scale * val[src_idx[e]]
 * (src_idx NULL: val[e]); code = (buffer << 60) | element offset with buffer
 * 0..3 = D, L, B, T.  Destinations must be unique (no atomics; deterministic).
 * accumulate != 0 adds (e.g. Q_cond = Q_prior + exp(theta_o) A^T A,
 * model.py:384-423), else stores (the drop-in sparse scatter of
 * SerinvSolver._spmatrix_to_bta, structured_solver.py:252-390, whose
 * per-block index it precomputes once, as blockmapping.py:66-195 does). */
int bta_scatter_coo(double* D, double* L, double* B, double* T, const uint64_t* code,
                    const double* val, const int* src_idx, int64_t count, double scale,
                    int accumulate, void* stream);

/* out[e] = block entry at code[e]: the gather of the selected inverse on a
 * sparsity pattern (SerinvSolver._structured_to_spmatrix,
 * structured_solver.py:470-538). */
int bta_gather_coo(const double* D, const double* L, const double* B, const double* T,
                   const uint64_t* code, double* out, int64_t count, void* stream);

/* out[0] = sum_e wval[e] * G(e), G = block entry at code[e] (mode 0) or
 * x[grow[e]] * x[gcol[e]] (mode 1); deterministic two-pass reduction using
 * work[nwork].  No reference counterpart: it evaluates tr(Sigma A^T A) and
 * x^T A^T A x for the adjoint gradient of f(theta), whose finite-difference
 * form is PyINLA._objective_function (pyinla.py:512-600). */
/* out[j] = a_j^T S a_j for every row a_j of the CSR matrix (rowptr[m+1],
 * colidx, val; global column indices into the n*b + a unknowns), where S is
 * the symmetric matrix held on the BTA pattern by D (full symmetric blocks),
 * L, B and T.  After bta_potri this is diag(A Sigma A^T), the marginal
 * variances of the observations that PyINLA.get_marginal_variances_observations
 * forms as (a @ Sigma_sparse @ a.T).diagonal() (pyinla.py:960-1016).  A pair of
 * columns outside the BTA pattern yields NaN. */
int bta_quadform_rows(const double* D, const double* L, const double* B, const double* T, int n, int b, int a,
                      const int* rowptr, const int* colidx, const double* val, int64_t m, double* out,
                      void* stream);

int bta_bilinear_coo(const double* D, const double* L, const double* B, const double* T,
                     const uint64_t* code, const int* grow, const int* gcol, const double* x,
                     const double* wval, int64_t count, int mode, double* work, int nwork,
                     double* out, void* stream);

/* Kronecker contractions per block t, field pair (v, w) and spatial matrix k
 * (nK <= 4, nv <= 4, b = nv*ns): with q = ((t*nv + v)*nv + w)*nK + k,
 * mode 0: out[2q] = <D_t[v,w], S_k>_F, out[2q+1] = <L_t[v,w], S_k>_F;
 * mode 1: x_{t,v}^T S_k x_{t,w} and x_{t+1,v}^T S_k x_{t,w}.  work holds
 * n*nv*ceil(ns/64)*2*nv*nK doubles.  No reference counterpart (adjoint
 * traces tr(Sigma dQ/dtheta) restricted to the BTA pattern). */
int bta_kron_contract(const double* D, const double* L, const double* x, int n, int b, int nv,
                      const int* rowptr, const int* colidx, const double* vals, int nnz, int nK,
                      int mode, double* work, double* out, void* stream);

/* Diagnostic (no reference counterpart): C = beta C + alpha op(A) op(B) on row-
 * major FP64 matrices (op(X) = X, or X^T when ta / tb = 1; C m x n, K = k)
 * through the grouped-GEMM kernel the selected inversion uses: variant 0 =
 * the default 64x64 kernel, 1 = the 64x64 kernel with two LDS buffers, 2 =
 * 128x128 tiles.  Builds a one-job plan (allocates), launches on `stream` and
 * synchronises it; for tests and kernel measurements. */
int bta_gemm(const double* A, int lda, int ta, const double* B, int ldb, int tb, double* C, int ldc, int m, int n,
             int k, double alpha, double beta, int variant, void* stream);

/* Diagnostic (no reference counterpart): runs one v_mfma_f64_16x16x4f64 with
 * A[i][k] = i + 100k, B[k][j] = j + 1000k and writes each lane's four
 * accumulator values to out[4*lane + r] (256 doubles, device memory), so a
 * test can confirm the operand layout every kernel here relies on. */
int bta_selftest_mfma(double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PYINLA_BTA_H */
