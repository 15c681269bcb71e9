"""GPU parity of the BTA triangular sweeps (pobtas) and selected inversion (pobtasi).

Expected values: the CPU oracle's pobtas / pobtasi (pinned against dense
LAPACK in tests/test_oracle.py) on the same factor.  Bars: rel 1e-9 on
solutions and selected-inverse entries.
"""
import numpy as np
import pytest
import torch

from oracle import bta as ob
from pyinla_amd import _lib
from pyinla_amd.solver import SerinvSolver
from pyinla_amd.workloads import known_factor_bta

pytestmark = pytest.mark.gpu

RTOL = 1e-9


def _rel(x, y):
    return float(np.abs(np.asarray(x) - np.asarray(y)).max() / max(np.abs(y).max(), 1e-300))


CASES = [(1, 1, 0), (1, 40, 3), (2, 70, 2), (3, 17, 3), (3, 64, 4), (4, 92, 4), (2, 130, 0), (3, 200, 9),
         (50, 92, 4), (3, 100, 64), (3, 600, 5), (2, 640, 0),  # b >= 512 exercises the 128x128-tile GEMM
         (3, 1100, 5), (2, 1024, 0)]  # b >= 1024: Sigma_ii's update as split tiles (1100: ragged edge tiles)


def _kf(n, b, a, seed):
    # b >= 1024: off-diagonal draws scaled to the b = 92 block norms (unscaled,
    # Sigma spans ~7 orders of magnitude at n = 3 and the 1e-9 bar measures the
    # conditioning, not the kernels)
    return known_factor_bta(n, b, a, seed=seed, norm_b=92 if b >= 1024 else None)


@pytest.mark.parametrize("n,b,a", CASES)
@pytest.mark.parametrize("use_graph", [False, True])
def test_solve_matches_oracle(cuda, n, b, a, use_graph):
    kf = _kf(n, b, a, seed=100 + n * b + a)
    s = SerinvSolver(None, b, n, a, use_graph=use_graph)
    s.load_blocks(kf.Q_D, kf.Q_L, kf.Q_B if a else None, kf.Q_T if a else None)
    s.factorize("bta")
    rng = np.random.default_rng(7)
    R = rng.normal(size=(n * b + a, 3))
    xr = R.copy()
    ob.pobtas(kf.Lf_D, kf.Lf_L, kf.Lf_B if a else None, kf.Lf_T if a else None, xr, "N")
    ob.pobtas(kf.Lf_D, kf.Lf_L, kf.Lf_B if a else None, kf.Lf_T if a else None, xr, "C")
    for _ in range(2):
        x = torch.from_numpy(R.copy()).to(cuda)
        out = s.solve(x, "bta")
        assert out is x
        assert _rel(x.cpu().numpy(), xr) < RTOL
    # 1-D rhs, numpy in / out
    v = R[:, 0].copy()
    s.solve(v, "bta")
    assert _rel(v, xr[:, 0]) < RTOL


@pytest.mark.parametrize("n,b,a", CASES)
def test_selected_inversion_matches_oracle(cuda, n, b, a):
    kf = _kf(n, b, a, seed=200 + n * b + a)
    s = SerinvSolver(None, b, n, a, use_graph=False)
    s.load_blocks(kf.Q_D, kf.Q_L, kf.Q_B if a else None, kf.Q_T if a else None)
    s.factorize("bta")
    s.selected_inversion("bta")
    D, L = kf.Lf_D.copy(), kf.Lf_L.copy()
    B = kf.Lf_B.copy() if a else None
    T = kf.Lf_T.copy() if a else None
    ob.pobtasi(D, L, B, T)
    Dg = s.A_diagonal_blocks.cpu().numpy()
    assert _rel(Dg, D) < RTOL
    # the variances (diagonal of Sigma) per entry as well
    dvar, dvar_g = np.diagonal(D, axis1=1, axis2=2), np.diagonal(Dg, axis1=1, axis2=2)
    assert np.all(dvar > 0) and np.all(np.abs(dvar_g - dvar) <= RTOL * dvar)
    if n > 1:
        assert _rel(s.A_lower_diagonal_blocks.cpu().numpy(), L) < RTOL
    if a:
        assert _rel(s.A_arrow_bottom_blocks.cpu().numpy(), B) < RTOL
        assert _rel(s.A_arrow_tip_block.cpu().numpy(), T) < RTOL
        tvar = np.diag(T)
        assert np.all(np.abs(np.diag(s.A_arrow_tip_block.cpu().numpy()) - tvar) <= RTOL * tvar)


def test_bt_solve_and_selinv_on_bta_buffers(cuda):
    kf = known_factor_bta(4, 70, 5, seed=9)
    s = SerinvSolver(None, 70, 4, 5, use_graph=True)
    s.load_blocks(kf.Q_D, kf.Q_L, kf.Q_B, kf.Q_T)
    s.factorize("bt")
    D, L = kf.Q_D.copy(), kf.Q_L.copy()
    ob.pobtf(D, L)
    rhs = np.random.default_rng(1).normal(size=4 * 70)
    xr = rhs.copy()
    ob.pobts(D, L, xr, "N")
    ob.pobts(D, L, xr, "C")
    x = torch.from_numpy(rhs.copy()).to(cuda)
    s.solve(x, "bt")
    assert _rel(x.cpu().numpy(), xr) < RTOL
    s.selected_inversion("bt")
    ob.pobtsi(D, L)
    assert _rel(s.A_diagonal_blocks.cpu().numpy(), D) < RTOL
    assert _rel(s.A_lower_diagonal_blocks.cpu().numpy(), L) < RTOL


def test_batched_factorization_mixed_shapes(cuda):
    """Problems of different (n, b, a), BT and BTA, factored by one batched
    graph: each must equal its single-problem oracle."""
    from pyinla_amd.batch import BatchedFactorization
    shapes = [(3, 92, 4, "bta"), (5, 64, 6, "bta"), (2, 130, 0, "bt"), (4, 70, 5, "bt"), (1, 17, 3, "bta")]
    kfs, solvers, sps_ = [], [], []
    for q, (n, b, a, sp) in enumerate(shapes):
        kf = known_factor_bta(n, b, a, seed=300 + q)
        s = SerinvSolver(None, b, n, a)
        s.load_blocks(kf.Q_D, kf.Q_L, kf.Q_B if a else None, kf.Q_T if a else None)
        kfs.append(kf)
        solvers.append(s)
        sps_.append(sp)
    batch = BatchedFactorization(solvers, sps_)
    for rep in range(2):
        if rep:
            for kf, s in zip(kfs, solvers):
                s.load_blocks(kf.Q_D, kf.Q_L, kf.Q_B if kf.a else None, kf.Q_T if kf.a else None)
        batch.factorize()
        for (n, b, a, sp), kf, s in zip(shapes, kfs, solvers):
            D, L = kf.Q_D.copy(), kf.Q_L.copy()
            B = kf.Q_B.copy() if sp == "bta" and a else None
            T = kf.Q_T.copy() if sp == "bta" and a else None
            ob.pobtaf(D, L, B, T)
            info, ld = s.status()
            assert info == 0
            ld_o = ob.logdet(D, T)
            assert abs(ld - ld_o) <= 1e-10 * abs(ld_o)
            assert _rel(s.A_diagonal_blocks.cpu().numpy(), D) < RTOL
            if n > 1:
                assert _rel(s.A_lower_diagonal_blocks.cpu().numpy(), L) < RTOL
            if B is not None:
                assert _rel(s.A_arrow_bottom_blocks.cpu().numpy(), B) < RTOL


def test_autotuner_keeps_results_exact(cuda):
    """The autotuner times its candidates (panel variant, near pool, GEMM
    tiles, main loop, rocBLAS / grouped products) on the solver's matrix,
    restores the blocks, and the factorization / selected inversion with the
    chosen options still match the oracle."""
    from pyinla_amd import autotune
    kf = known_factor_bta(3, 600, 5, seed=77)
    s = SerinvSolver(None, 600, 3, 5)
    s.load_blocks(kf.Q_D, kf.Q_L, kf.Q_B, kf.Q_T)
    autotune._cache.clear()
    choice = autotune.autotune(s, "bta", reps=1)
    assert choice["update"] in (64, 128) and choice["grouped"] in (64, 128)
    assert choice["panel"] in (1, 2) and choice["near"] in (16, 8, 24)
    assert choice["gemm_kernel"] in (0, 1) and choice["blas"] in (0, 1)
    assert _rel(s.A_diagonal_blocks.cpu().numpy(), kf.Q_D) < 1e-15  # restored
    s.factorize("bta")
    assert _rel(s.A_diagonal_blocks.cpu().numpy(), kf.Lf_D) < RTOL
    s.selected_inversion("bta")
    D, L, B, T = kf.Lf_D.copy(), kf.Lf_L.copy(), kf.Lf_B.copy(), kf.Lf_T.copy()
    ob.pobtasi(D, L, B, T)
    assert _rel(s.A_diagonal_blocks.cpu().numpy(), D) < RTOL
    assert _rel(s.A_arrow_bottom_blocks.cpu().numpy(), B) < RTOL


@pytest.mark.parametrize("force_flow", [False, True])
def test_batched_flow_factorization(cuda, force_flow):
    """All problems wider than one tile: the batch runs as persistent dataflow
    launches (one chain + an equal share of helpers per problem); with
    force_flow the single-tile-column problems take that form too."""
    from pyinla_amd.batch import BatchedFactorization
    shapes = [(3, 200, 6, "bta"), (2, 130, 0, "bt"), (4, 92, 4, "bta"), (1, 300, 5, "bta"), (3, 150, 3, "bt")]
    if force_flow:
        shapes.append((3, 60, 2, "bta"))
    kfs, solvers, sps_ = [], [], []
    for q, (n, b, a, sp) in enumerate(shapes):
        kf = known_factor_bta(n, b, a, seed=500 + q)
        s = SerinvSolver(None, b, n, a)
        s.load_blocks(kf.Q_D, kf.Q_L, kf.Q_B if a else None, kf.Q_T if a else None)
        kfs.append(kf)
        solvers.append(s)
        sps_.append(sp)
    if force_flow:
        solvers[0].set_panel("flow")
    print(f"[batch-flow] force={force_flow} shapes={shapes}", flush=True)
    batch = BatchedFactorization(solvers, sps_)
    for rep in range(2):
        if rep:
            for kf, s in zip(kfs, solvers):
                s.load_blocks(kf.Q_D, kf.Q_L, kf.Q_B if kf.a else None, kf.Q_T if kf.a else None)
        batch.factorize()
        for (n, b, a, sp), kf, s in zip(shapes, kfs, solvers):
            D, L = kf.Q_D.copy(), kf.Q_L.copy()
            B = kf.Q_B.copy() if sp == "bta" and a else None
            T = kf.Q_T.copy() if sp == "bta" and a else None
            ob.pobtaf(D, L, B, T)
            info, ld = s.status()
            assert info == 0
            ld_o = ob.logdet(D, T)
            assert abs(ld - ld_o) <= 1e-10 * abs(ld_o)
            assert _rel(s.A_diagonal_blocks.cpu().numpy(), D) < RTOL
            if n > 1:
                assert _rel(s.A_lower_diagonal_blocks.cpu().numpy(), L) < RTOL
            if B is not None:
                assert _rel(s.A_arrow_bottom_blocks.cpu().numpy(), B) < RTOL
                assert _rel(s.A_arrow_tip_block.cpu().numpy(), T) < RTOL


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("blas", [False, True])
@pytest.mark.parametrize("n,b,a", [(2, 70, 2), (4, 92, 4), (3, 130, 0), (3, 300, 6), (3, 1100, 5)])
def test_selected_inversion_blas_and_grouped(cuda, n, b, a, blas, use_graph):
    """Both forms of the Sigma_{i+1,i} products (rocBLAS dgemm / grouped
    kernel) against the oracle, twice in a row on the same solver; with
    use_graph the second call replays the recorded graph, which holds the
    all-grouped plan (rocBLAS is not called under capture)."""
    kf = _kf(n, b, a, seed=300 + n * b + a)
    s = SerinvSolver(None, b, n, a, use_graph=use_graph)
    s.set_selinv_blas(blas)
    D, L = kf.Lf_D.copy(), kf.Lf_L.copy()
    B = kf.Lf_B.copy() if a else None
    T = kf.Lf_T.copy() if a else None
    ob.pobtasi(D, L, B, T)
    for _ in range(2):
        s.load_blocks(kf.Q_D, kf.Q_L, kf.Q_B if a else None, kf.Q_T if a else None)
        s.factorize("bta")
        s.selected_inversion("bta")
        assert _rel(s.A_diagonal_blocks.cpu().numpy(), D) < RTOL
        assert _rel(s.A_lower_diagonal_blocks.cpu().numpy(), L) < RTOL
        if a:
            assert _rel(s.A_arrow_bottom_blocks.cpu().numpy(), B) < RTOL
            assert _rel(s.A_arrow_tip_block.cpu().numpy(), T) < RTOL


@pytest.mark.parametrize("blas", [False, True])
@pytest.mark.parametrize("n,b,a", [(2, 70, 2), (3, 130, 0), (3, 300, 6), (2, 600, 5)])
def test_selected_inversion_double_buffered_gemm(cuda, n, b, a, blas):
    """The double-buffered 64x64 grouped GEMM (BTA_OPT_GEMM_KERNEL 1: paired
    clamped loads, masks at the LDS store, one pipeline per K segment) on the
    block inverses and the selected inversion -- several segments per job,
    triangular operands, ragged edges, lower / mirrored outputs -- against
    the oracle."""
    kf = known_factor_bta(n, b, a, seed=500 + n * b + a)
    s = SerinvSolver(None, b, n, a, use_graph=False)
    _lib.check(s.lib.bta_ws_set_option(s._ws, 9, 1), "set_option")
    s.set_selinv_blas(blas)
    D, L = kf.Lf_D.copy(), kf.Lf_L.copy()
    B = kf.Lf_B.copy() if a else None
    T = kf.Lf_T.copy() if a else None
    ob.pobtasi(D, L, B, T)
    s.load_blocks(kf.Q_D, kf.Q_L, kf.Q_B if a else None, kf.Q_T if a else None)
    s.factorize("bta")
    s.selected_inversion("bta")
    assert _rel(s.A_diagonal_blocks.cpu().numpy(), D) < RTOL
    assert _rel(s.A_lower_diagonal_blocks.cpu().numpy(), L) < RTOL
    if a:
        assert _rel(s.A_arrow_bottom_blocks.cpu().numpy(), B) < RTOL
        assert _rel(s.A_arrow_tip_block.cpu().numpy(), T) < RTOL


@pytest.mark.parametrize("n,b,a", [(3, 600, 5), (2, 130, 0), (4, 92, 4), (2, 300, 6)])
def test_selected_inversion_with_128_tiles(cuda, n, b, a):
    """The 128x128-tile grouped GEMM (block inverses + selected inversion,
    ragged edges, triangular operands, lower/mirror outputs) matches the
    oracle."""
    kf = known_factor_bta(n, b, a, seed=400 + n * b + a)
    s = SerinvSolver(None, b, n, a, use_graph=False)
    _lib.check(s.lib.bta_ws_set_option(s._ws, 2, 128), "set_option")
    s.load_blocks(kf.Q_D, kf.Q_L, kf.Q_B if a else None, kf.Q_T if a else None)
    s.factorize("bta")
    rng = np.random.default_rng(3)
    R = rng.normal(size=(n * b + a, 2))
    xr = R.copy()
    ob.pobtas(kf.Lf_D, kf.Lf_L, kf.Lf_B if a else None, kf.Lf_T if a else None, xr, "N")
    ob.pobtas(kf.Lf_D, kf.Lf_L, kf.Lf_B if a else None, kf.Lf_T if a else None, xr, "C")
    x = torch.from_numpy(R.copy()).to(cuda)
    s.solve(x, "bta")
    assert _rel(x.cpu().numpy(), xr) < RTOL
    s.selected_inversion("bta")
    D, L = kf.Lf_D.copy(), kf.Lf_L.copy()
    B = kf.Lf_B.copy() if a else None
    T = kf.Lf_T.copy() if a else None
    ob.pobtasi(D, L, B, T)
    assert _rel(s.A_diagonal_blocks.cpu().numpy(), D) < RTOL
    if n > 1:
        assert _rel(s.A_lower_diagonal_blocks.cpu().numpy(), L) < RTOL
    if a:
        assert _rel(s.A_arrow_bottom_blocks.cpu().numpy(), B) < RTOL
        assert _rel(s.A_arrow_tip_block.cpu().numpy(), T) < RTOL


@pytest.mark.parametrize("n,b,a,nrhs", [(2, 130, 3, 500), (3, 64, 0, 257), (2, 300, 6, 40)])
def test_solve_many_rhs(cuda, n, b, a, nrhs):
    """Many right-hand sides: the transposed sweep processes them in column
    groups sized to the workspace scratch; every column must match the oracle."""
    kf = known_factor_bta(n, b, a, seed=900 + n * b + a + nrhs)
    s = SerinvSolver(None, b, n, a, use_graph=False)
    s.load_blocks(kf.Q_D, kf.Q_L, kf.Q_B if a else None, kf.Q_T if a else None)
    s.factorize("bta")
    R = np.random.default_rng(5).normal(size=(n * b + a, nrhs))
    xr = R.copy()
    ob.pobtas(kf.Lf_D, kf.Lf_L, kf.Lf_B if a else None, kf.Lf_T if a else None, xr, "N")
    ob.pobtas(kf.Lf_D, kf.Lf_L, kf.Lf_B if a else None, kf.Lf_T if a else None, xr, "C")
    x = torch.from_numpy(R.copy()).to(cuda)
    s.solve(x, "bta")
    assert _rel(x.cpu().numpy(), xr) < RTOL


@pytest.mark.parametrize("use_graph", [False, True])
def test_block_inverses_built_ahead(cuda, use_graph):
    """bta_block_inverses before the solve / selected inversion gives the same
    results as letting them build the inverses, and a refactorization
    invalidates the prepared inverses."""
    n, b, a = 3, 200, 5
    kf = known_factor_bta(n, b, a, seed=321)
    s = SerinvSolver(None, b, n, a, use_graph=use_graph)
    rng = np.random.default_rng(9)
    R = rng.normal(size=(n * b + a, 2))
    xr = R.copy()
    ob.pobtas(kf.Lf_D, kf.Lf_L, kf.Lf_B, kf.Lf_T, xr, "N")
    ob.pobtas(kf.Lf_D, kf.Lf_L, kf.Lf_B, kf.Lf_T, xr, "C")
    for rep in range(2):
        # second round: new values (scaled matrix) so that stale inverses would show
        scale = 1.0 + rep
        s.load_blocks(kf.Q_D * scale, kf.Q_L * scale, kf.Q_B * scale, kf.Q_T * scale)
        s.factorize("bta")
        s.prepare_block_inverses("bta")
        x = torch.from_numpy(R.copy()).to(cuda)
        s.solve(x, "bta")
        assert _rel(x.cpu().numpy(), xr / scale) < RTOL
    s.selected_inversion("bta")
    D, L, B, T = kf.Lf_D.copy(), kf.Lf_L.copy(), kf.Lf_B.copy(), kf.Lf_T.copy()
    ob.pobtasi(D, L, B, T)
    assert _rel(s.A_diagonal_blocks.cpu().numpy(), D / 2.0) < RTOL
    assert _rel(s.A_lower_diagonal_blocks.cpu().numpy(), L / 2.0) < RTOL


def _oracle_all(kf, sparsity):
    """Oracle factor, logdet, solve and selected inverse of kf's matrix."""
    a = kf.a if sparsity == "bta" else 0
    D, L = kf.Q_D.copy(), kf.Q_L.copy()
    B = kf.Q_B.copy() if a else None
    T = kf.Q_T.copy() if a else None
    ob.pobtaf(D, L, B, T)
    ld = ob.logdet(D, T)
    n, b = kf.n, kf.b
    rhs = np.random.default_rng(5).normal(size=n * b + a)
    x = rhs.copy()
    ob.pobtas(D, L, B, T, x, "N")
    ob.pobtas(D, L, B, T, x, "C")
    Ds, Ls = D.copy(), L.copy()
    Bs = B.copy() if a else None
    Ts = T.copy() if a else None
    ob.pobtasi(Ds, Ls, Bs, Ts)
    return dict(ld=ld, rhs=rhs, x=x, S=(Ds, Ls, Bs, Ts), F=(D, L, B, T))


def _load(s, kf, sparsity):
    a = kf.a if sparsity == "bta" else 0
    s.load_blocks(kf.Q_D, kf.Q_L, kf.Q_B if a else None, kf.Q_T if a else None)


def _check_all(cuda, s, kf, sparsity, ref):
    """factorize -> logdet -> solve -> selected_inversion on one solver."""
    _load(s, kf, sparsity)
    s.factorize(sparsity)
    info, ld = s.status()
    assert info == 0
    assert abs(ld - ref["ld"]) <= 1e-10 * abs(ref["ld"])
    x = torch.from_numpy(ref["rhs"].copy()).to(cuda)
    s.solve(x, sparsity)
    assert _rel(x.cpu().numpy(), ref["x"]) < RTOL
    s.selected_inversion(sparsity)
    Ds, Ls, Bs, Ts = ref["S"]
    assert _rel(s.A_diagonal_blocks.cpu().numpy(), Ds) < RTOL
    assert _rel(s.A_lower_diagonal_blocks.cpu().numpy(), Ls) < RTOL
    if Bs is not None:
        assert _rel(s.A_arrow_bottom_blocks.cpu().numpy(), Bs) < RTOL
        assert _rel(s.A_arrow_tip_block.cpu().numpy(), Ts) < RTOL


@pytest.mark.parametrize("b", [70, 130])
def test_alternating_bta_and_bt_on_one_solver_with_graphs(cuda, b):
    """The reference's evaluation loop factors Q_cond ("bta") then Q_prior
    ("bt") on one solver, over and over (pyinla.py:655,1102), with
    SerinvSolver's default graph replay: bta -> bt -> bta -> bt, each
    followed by a solve and a selected inversion, against the oracle.  Every
    launch plan a recorded graph uses must still be alive on the replays."""
    n, a = 5, 5
    kf = known_factor_bta(n, b, a, seed=41 + b)
    refs = {sp: _oracle_all(kf, sp) for sp in ("bta", "bt")}
    s = SerinvSolver(None, b, n, a, use_graph=True)
    for _ in range(2):
        for sp in ("bta", "bt"):
            _check_all(cuda, s, kf, sp, refs[sp])


def test_bt_selected_inversion_then_bta_on_bta_buffers(cuda):
    """selected_inversion("bt") then a BTA factorization and
    selected_inversion("bta") on the same solver: the G_a buffer of the
    second call holds the full arrowhead (n a b doubles)."""
    n, b, a = 4, 130, 7
    kf = known_factor_bta(n, b, a, seed=77)
    s = SerinvSolver(None, b, n, a, use_graph=False)
    _check_all(cuda, s, kf, "bt", _oracle_all(kf, "bt"))
    _check_all(cuda, s, kf, "bta", _oracle_all(kf, "bta"))


def test_logdet_reads_the_held_blocks(cuda):
    """logdet(sparsity) as the reference computes it
    (structured_solver.py:151-166): 2 sum log diag of the blocks held now, so
    after a BTA factorization logdet("bt") leaves the tip out."""
    n, b, a = 4, 92, 4
    kf = known_factor_bta(n, b, a, seed=3)
    ref = _oracle_all(kf, "bta")
    D, L, B, T = ref["F"]
    s = SerinvSolver(None, b, n, a, use_graph=False)
    _load(s, kf, "bta")
    s.factorize("bta")
    ld_bt = 2.0 * sum(np.log(np.diag(D[i])).sum() for i in range(n))
    ld_tip = 2.0 * np.log(np.diag(T)).sum()
    assert abs(s.logdet("bta") - ref["ld"]) <= 1e-12 * abs(ref["ld"])
    assert abs(s.logdet("bt") - ld_bt) <= 1e-12 * abs(ld_bt)
    assert abs(s.logdet("bta") - s.logdet("bt") - ld_tip) <= 1e-12 * abs(ref["ld"])


def test_reservations_and_memory(cuda):
    """A solver reserved for factorizations only refuses a solve (no hidden
    allocation); the bytes it reports grow by exactly the reserved buffers."""
    n, b, a = 3, 130, 4
    kf = known_factor_bta(n, b, a, seed=12)
    sf = SerinvSolver(None, b, n, a, use_graph=False, reserve="factor")
    ss = SerinvSolver(None, b, n, a, use_graph=False, reserve="solve")
    si = SerinvSolver(None, b, n, a, use_graph=False, reserve="selinv")
    assert ss.memory_bytes() - sf.memory_bytes() == 8 * n * b * b
    assert si.memory_bytes() - ss.memory_bytes() == 8 * (n * b * b + n * a * b)
    _load(sf, kf, "bta")
    sf.factorize("bta")
    with pytest.raises(_lib.BTALibraryError, match="reservation"):
        sf.solve(torch.zeros(n * b + a, dtype=torch.float64, device=cuda), "bta")
    _load(ss, kf, "bta")
    ss.factorize("bta")
    ss.solve(torch.zeros(n * b + a, dtype=torch.float64, device=cuda), "bta")
    with pytest.raises(_lib.BTALibraryError, match="reservation"):
        ss.selected_inversion("bta")
    # get_solver_memory accumulates like the reference (structured_solver.py:548)
    m1 = si.get_solver_memory()
    assert si.get_solver_memory() == 2 * m1
