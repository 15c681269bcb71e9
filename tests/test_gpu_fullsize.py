"""Full-size parity through identities that hold at any size.

At the benchmark shape (n = 200 blocks of b = 2000) a dense oracle is out of
reach, so the HIP path is checked on a Kronecker matrix Q = M (x) S (M
tridiagonal SPD, S dense SPD), for which
    logdet Q = b logdet M + n logdet S,
    Sigma_{ij} = (M^-1)_{ij} S^-1   (selected inverse blocks),
    Q^-1 vec(R) = vec(M^-1 R S^-1)  (t-major vec),
all computable on the CPU from n x n and b x b matrices.  An arrowhead is
added through a Schur-complement construction whose logdet is known too.
"""
import numpy as np
import pytest
import torch

from pyinla_amd.solver import SerinvSolver

pytestmark = pytest.mark.gpu


def _kron_case(n, b, seed):
    rng = np.random.default_rng(seed)
    M = np.diag(2.0 + rng.random(n)) + np.diag(-0.9 * rng.random(n - 1), -1)
    M = np.tril(M) + np.tril(M, -1).T
    X = rng.normal(size=(b, b)) / np.sqrt(b)
    S = X @ X.T + 0.5 * np.eye(b)
    return M, S


@pytest.mark.parametrize("n,b", [(200, 2000)])
def test_kronecker_identities_at_benchmark_size(cuda, n, b):
    M, S = _kron_case(n, b, 11)
    s = SerinvSolver(None, b, n, 0)
    St = torch.from_numpy(S).to(cuda)
    for t in range(n):
        s.A_diagonal_blocks[t].copy_(St).mul_(M[t, t])
        if t < n - 1:
            s.A_lower_diagonal_blocks[t].copy_(St).mul_(M[t + 1, t])
    s.factorize("bt")
    info, ld = s.status()
    assert info == 0
    ld_exact = b * np.linalg.slogdet(M)[1] + n * np.linalg.slogdet(S)[1]
    assert abs(ld - ld_exact) <= 1e-10 * abs(ld_exact)
    # solve: X = M^-1 R S^-1
    rng = np.random.default_rng(3)
    R = rng.normal(size=(n, b))
    x = torch.from_numpy(R.reshape(-1).copy()).to(cuda)
    s.solve(x, "bt")
    Si = np.linalg.inv(S)
    Mi = np.linalg.inv(M)
    Xexp = Mi @ R @ Si
    Xg = x.cpu().numpy().reshape(n, b)
    assert np.abs(Xg - Xexp).max() <= 1e-9 * np.abs(Xexp).max()
    # selected inverse blocks
    s.selected_inversion("bt")
    for t in (0, 1, n // 2, n - 2, n - 1):
        Dg = s.A_diagonal_blocks[t].cpu().numpy()
        assert np.abs(Dg - Mi[t, t] * Si).max() <= 1e-9 * np.abs(Mi[t, t] * Si).max()
        if t < n - 1:
            Lg = s.A_lower_diagonal_blocks[t].cpu().numpy()
            assert np.abs(Lg - Mi[t + 1, t] * Si).max() <= 1e-9 * np.abs(Mi[t, t] * Si).max()
