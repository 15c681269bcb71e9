"""GPU parity of the BTA Cholesky (pobtaf / pobtf) and log-determinant.

Expected values come from the CPU oracle (oracle/bta.py), which is itself
pinned in tests/test_oracle.py; the known-factor construction additionally
fixes the exact factor (Cholesky uniqueness).  Tolerances are the FP64 bars of
the BTA path: rel 1e-10 on the log-determinant, rel 1e-9 on factor blocks.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import bta as ob
from pyinla_amd import _lib
from pyinla_amd.solver import SerinvSolver
from pyinla_amd.workloads import known_factor_bta

pytestmark = pytest.mark.gpu

RTOL_LOGDET = 1e-10
RTOL_BLOCK = 1e-9


def _rel(x, y):
    x = np.asarray(x)
    y = np.asarray(y)
    den = max(np.abs(y).max(), 1e-300)
    return float(np.abs(x - y).max() / den)


def test_mfma_operand_layout(cuda):
    lib = _lib.load()
    out = torch.zeros(256, dtype=torch.float64, device=cuda)
    _lib.check(lib.bta_selftest_mfma(C.c_void_p(out.data_ptr()), None), "selftest")
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(64, 4)
    l = np.arange(64)
    A = np.zeros((16, 4))
    B = np.zeros((4, 16))
    A[l % 16, l // 16] = (l % 16) + 100.0 * (l // 16)
    B[l // 16, l % 16] = (l % 16) + 1000.0 * (l // 16)
    Cm = A @ B
    exp = np.zeros((64, 4))
    for lane in range(64):
        for r in range(4):
            exp[lane, r] = Cm[4 * r + lane // 16, lane % 16]
    np.testing.assert_array_equal(got, exp)


def _run_factor(kf, sparsity="bta", use_graph=False):
    a = kf.a if sparsity == "bta" else 0
    s = SerinvSolver(None, kf.b, kf.n, kf.a, use_graph=use_graph)
    s.load_blocks(kf.Q_D, kf.Q_L, kf.Q_B if kf.a else None, kf.Q_T if kf.a else None)
    s.factorize(sparsity)
    info, ld = s.status()
    res = dict(info=info, logdet=ld,
               D=s.A_diagonal_blocks.cpu().numpy(), L=s.A_lower_diagonal_blocks.cpu().numpy())
    if a:
        res["B"] = s.A_arrow_bottom_blocks.cpu().numpy()
        res["T"] = s.A_arrow_tip_block.cpu().numpy()
    return res


@pytest.mark.parametrize("n,b,a", [(1, 1, 0), (3, 1, 0), (2, 17, 3), (3, 64, 4), (3, 65, 6),
                                   (4, 92, 4), (2, 130, 0), (3, 200, 9), (2, 128, 64)])
def test_potrf_known_factor(cuda, n, b, a):
    kf = known_factor_bta(n, b, a, seed=11 + n + b + a)
    r = _run_factor(kf)
    assert r["info"] == 0
    assert _rel(r["D"], kf.Lf_D) < RTOL_BLOCK
    if n > 1:
        assert _rel(r["L"], kf.Lf_L) < RTOL_BLOCK
    if a:
        assert _rel(r["B"], kf.Lf_B) < RTOL_BLOCK
        assert _rel(r["T"], kf.Lf_T) < RTOL_BLOCK
    ld_exact = 2 * (np.log(np.diagonal(kf.Lf_D, axis1=1, axis2=2)).sum()
                    + (np.log(np.diag(kf.Lf_T)).sum() if a else 0.0))
    assert abs(r["logdet"] - ld_exact) <= RTOL_LOGDET * abs(ld_exact) + 1e-12


@pytest.mark.parametrize("use_graph", [False, True])
def test_potrf_plain_bta_full_size(cuda, use_graph):
    """n=50, b=92, a=4: the plain FP64 BTA case, against the oracle."""
    kf = known_factor_bta()
    D, L, B, T = kf.Q_D.copy(), kf.Q_L.copy(), kf.Q_B.copy(), kf.Q_T.copy()
    ob.pobtaf(D, L, B, T)
    ld_o = ob.logdet(D, T)
    for _ in range(2):  # a second call replays the graph / reuses flags
        r = _run_factor(kf, use_graph=use_graph)
        assert r["info"] == 0
        assert abs(r["logdet"] - ld_o) <= RTOL_LOGDET * abs(ld_o)
        assert _rel(r["D"], D) < RTOL_BLOCK
        assert _rel(r["L"], L) < RTOL_BLOCK
        assert _rel(r["B"], B) < RTOL_BLOCK
        assert _rel(r["T"], T) < RTOL_BLOCK


def test_potrf_bt_on_bta_buffers(cuda):
    """sparsity='bt' factors only the BT part and excludes the tip from logdet."""
    kf = known_factor_bta(4, 70, 5, seed=3)
    D, L = kf.Q_D.copy(), kf.Q_L.copy()
    ob.pobtf(D, L)
    r = _run_factor(kf, sparsity="bt")
    assert r["info"] == 0
    assert _rel(r["D"], D) < RTOL_BLOCK and _rel(r["L"], L) < RTOL_BLOCK
    ld_o = ob.logdet(D)
    assert abs(r["logdet"] - ld_o) <= RTOL_LOGDET * abs(ld_o)


def test_potrf_reports_non_spd(cuda):
    kf = known_factor_bta(3, 70, 2, seed=5)
    QD = kf.Q_D.copy()
    QD[1, 66, 66] = -5.0  # row 70 + 66 = 136 of the global matrix
    s = SerinvSolver(None, kf.b, kf.n, kf.a, use_graph=False)
    s.load_blocks(QD, kf.Q_L, kf.Q_B, kf.Q_T)
    s.factorize("bta")
    info, _ = s.status()
    assert info == 1 + 70 + 66
