"""GPU parity of the persistent dataflow panel (k_panel_flow) against the
per-column panel launches and the known factor.

Both panel variants compute the same tile algorithm; only the order in which
the rank-64 updates of a tile are summed can differ (deferred updates), so the
factors agree to rounding.  The known-factor construction fixes the exact
factor (Cholesky uniqueness), so each variant is also checked against it at
the FP64 bars of the path (rel 1e-9 blocks, rel 1e-10 log-determinant).
"""
import numpy as np
import pytest

from pyinla_amd.solver import SerinvSolver
from pyinla_amd.workloads import known_factor_bta

pytestmark = pytest.mark.gpu


def _rel(x, y):
    return float(np.abs(np.asarray(x) - np.asarray(y)).max() / max(np.abs(y).max(), 1e-300))


def _factor(kf, mode, helpers=0, delta=None, use_graph=False, calls=1, fold=None):
    s = SerinvSolver(None, kf.b, kf.n, kf.a, use_graph=use_graph)
    s.set_panel(mode, helpers, delta, fold)
    out = []
    for _ in range(calls):
        s.load_blocks(kf.Q_D, kf.Q_L, kf.Q_B if kf.a else None, kf.Q_T if kf.a else None)
        s.factorize("bta")
        info, ld = s.status()
        r = dict(info=info, logdet=ld, D=s.A_diagonal_blocks.cpu().numpy(),
                 L=s.A_lower_diagonal_blocks.cpu().numpy())
        if kf.a:
            r["B"] = s.A_arrow_bottom_blocks.cpu().numpy()
            r["T"] = s.A_arrow_tip_block.cpu().numpy()
        out.append(r)
    return out


def _check_known(kf, r):
    assert r["info"] == 0
    assert _rel(r["D"], kf.Lf_D) < 1e-9
    if kf.n > 1:
        assert _rel(r["L"], kf.Lf_L) < 1e-9
    if kf.a:
        assert _rel(r["B"], kf.Lf_B) < 1e-9
        assert _rel(r["T"], kf.Lf_T) < 1e-9
    ld = 2 * (np.log(np.diagonal(kf.Lf_D, axis1=1, axis2=2)).sum()
              + (np.log(np.diag(kf.Lf_T)).sum() if kf.a else 0.0))
    assert abs(r["logdet"] - ld) <= 1e-10 * abs(ld)


@pytest.mark.parametrize("n,b,a,helpers,delta", [
    (3, 130, 6, 0, 2), (2, 300, 4, 5, 0), (3, 300, 0, 0, 3), (1, 200, 6, 0, 2),
    (2, 640, 6, 0, 2), (2, 1000, 9, 64, 1), (4, 92, 4, 2, 5), (2, 128, 64, 0, 2)])
def test_flow_matches_known_factor_and_columns(cuda, n, b, a, helpers, delta):
    kf = known_factor_bta(n, b, a, seed=100 + n + b + a)
    print(f"[flow] n={n} b={b} a={a} helpers={helpers} delta={delta}", flush=True)
    rf = _factor(kf, "flow", helpers, delta)[0]
    _check_known(kf, rf)
    rc = _factor(kf, "columns")[0]
    assert abs(rf["logdet"] - rc["logdet"]) <= 1e-12 * abs(rc["logdet"])
    assert _rel(rf["D"], rc["D"]) < 1e-12
    if n > 1:
        assert _rel(rf["L"], rc["L"]) < 1e-12


@pytest.mark.parametrize("n,b,a,helpers", [(3, 130, 6, 0), (2, 300, 4, 5), (3, 300, 0, 0), (1, 200, 6, 0),
                                             (2, 1000, 9, 64), (3, 3200, 6, 0)])
def test_flow_block_update_as_gemm(cuda, n, b, a, helpers):
    """The persistent panel without its next-block tasks, the block update
    then applied by one grouped-GEMM launch per block: the known factor, and
    the folded form to rounding (the update's sums run in another order).
    b = 3200 (50 tile columns) is where the automatic choice switches."""
    print(f"[flow] update as GEMM n={n} b={b} a={a} helpers={helpers}", flush=True)
    if b <= 2000:
        kf = known_factor_bta(n, b, a, seed=200 + n + b + a)
    else:  # off-diagonal draws scaled to the b = 92 block norms (else Q is singular in FP64)
        from types import SimpleNamespace
        from pyinla_amd.workloads import known_factor_bta_torch
        (Ld, Ll, Lb, Lt), (QD, QL, QB, QT) = known_factor_bta_torch(n, b, a, seed=7, device=cuda, norm_b=92)
        kf = SimpleNamespace(n=n, b=b, a=a, Q_D=QD.cpu().numpy(), Q_L=QL.cpu().numpy(), Q_B=QB.cpu().numpy(),
                             Q_T=QT.cpu().numpy(), Lf_D=Ld.cpu().numpy(), Lf_L=Ll.cpu().numpy(),
                             Lf_B=Lb.cpu().numpy(), Lf_T=Lt.cpu().numpy())
        del Ld, Ll, Lb, Lt, QD, QL, QB, QT
    rg = _factor(kf, "flow", helpers, fold="gemm")[0]
    _check_known(kf, rg)
    rp = _factor(kf, "flow", helpers, fold="panel")[0]
    assert abs(rg["logdet"] - rp["logdet"]) <= 1e-12 * abs(rp["logdet"])
    assert _rel(rg["D"], rp["D"]) < 1e-12
    if n > 1:
        assert _rel(rg["L"], rp["L"]) < 1e-12
    ra = _factor(kf, "flow", helpers, fold="auto")[0]
    assert _rel(ra["D"], rp["D"] if b <= 3072 else rg["D"]) == 0.0


@pytest.mark.parametrize("use_graph", [False, True])
def test_flow_repeated_calls(cuda, use_graph):
    """Flags carry a per-call, per-block tag: back-to-back calls (and graph
    replays) must not see the previous call's flags as ready."""
    kf = known_factor_bta(3, 260, 5, seed=9)
    print(f"[flow] repeated calls use_graph={use_graph}", flush=True)
    for r in _factor(kf, "flow", use_graph=use_graph, calls=3):
        _check_known(kf, r)


def test_flow_reports_non_spd(cuda):
    kf = known_factor_bta(3, 150, 2, seed=5)
    QD = kf.Q_D.copy()
    QD[1, 100, 100] = -5.0
    s = SerinvSolver(None, kf.b, kf.n, kf.a, use_graph=False)
    s.set_panel("flow")
    s.load_blocks(QD, kf.Q_L, kf.Q_B, kf.Q_T)
    s.factorize("bta")
    info, _ = s.status()
    assert info == 1 + 150 + 100
