"""GPU parity at the block sizes of the benchmark configurations.

Each configuration is run at its real block size b (the size that sets the
conditioning of the diagonal blocks and every kernel's tiling), with the
number of time blocks reduced where the CPU oracle would otherwise take
minutes:

  * Gaussian spatio-temporal model, 40 x 40 mesh (b = 1600) and 50 x 40 mesh
    (b = 2000), nt = 3: log-determinants, x*, f and the adjoint gradient
    against the oracle (oracle/model.py) and its central FD (step 1e-4);
  * the 40 x 40 model at its full nt = 100: f against the oracle, and the
    2 dim(theta) + 1 = 9 FD evaluations (18 factorizations) as one batched
    launch sequence against the one-by-one FD loop;
  * three interleaved fields of 1500 nodes (b = 4500), nt = 3: f parts
    against the oracle, the adjoint gradient against the GPU central FD of
    f and, on two components, the oracle's FD;
  * 24 known-factor BTA matrices with nt in {50, 100, 200} and b in
    {256, 512, 1024, 2048}, a = 6, as one batch: log-determinants against the
    exact 2 sum log diag of the known factor, first and last diagonal factor
    blocks against the known factor (Cholesky uniqueness);
  * n = 200, b = 2000, a = 6: a BTA matrix whose space-time part is a
    Kronecker product M (x) S and whose tip makes the Schur complement a
    chosen SPD C, so that logdet, the solution and the selected-inverse
    blocks (arrow and tip included) follow in closed form on the CPU;
  * n = 200, b = 2000: the adjoint gradient against the 9-point central FD
    of f with step 1e-4.

Bars: rel 1e-10 on log-determinants and f, rel 1e-9 on solutions and
selected-inverse entries (per block: max |error| over the block's largest
entry), rel 1e-6 between the adjoint gradient and finite differences.
"""
import numpy as np
import pytest
import torch

from oracle import model as om
from pyinla_amd.coefficients import CoefficientModel, CoregionalCoefficients
from pyinla_amd.solver import SerinvSolver
from pyinla_amd.st_model import STGaussianModel
from pyinla_amd.workloads import coregional_problem, st_problem

pytestmark = pytest.mark.gpu


def _rel(x, y):
    return float(np.abs(np.asarray(x) - np.asarray(y)).max() / max(np.abs(y).max(), 1e-300))


def _log(msg):
    print(f"[benchmark-sizes] {msg}", flush=True)


def _check_f_parts(parts, ref):
    for key in ("ld_c", "ld_p", "quad"):
        assert abs(parts[key] - ref[key]) <= 1e-10 * abs(ref[key]), key
    assert abs(parts["f"] - ref["f"]) <= 1e-10 * max(abs(ref["ld_c"]), abs(ref["f"]))


@pytest.mark.timeout(600)
@pytest.mark.parametrize("nx,ny", [(40, 40), (50, 40)])
def test_st_model_at_benchmark_block_size(cuda, nx, ny):
    _log(f"st model {nx}x{ny} (b={nx * ny}), nt=3")
    prob = st_problem(nt=3, nx=nx, ny=ny, nb=6, seed=3, hx=1.0 / (nx - 1), hy=1.0 / (ny - 1))
    cm = CoefficientModel("baseline")
    model = STGaussianModel(prob, coef=cm, device=cuda)
    parts = model.forward(prob.theta)
    ref = om.evaluate_f(prob, cm.table, prob.theta, return_parts=True)
    _check_f_parts(parts, ref)
    x = model.x.cpu().numpy()
    # per entry, with a floor at 1e-12 of the largest entry
    assert np.all(np.abs(x - ref["x"]) <= 1e-9 * np.abs(ref["x"]) + 1e-12 * np.abs(ref["x"]).max())
    f_adj, g_adj = model.value_and_grad(prob.theta)
    assert abs(f_adj - ref["f"]) <= 1e-10 * max(abs(ref["ld_c"]), abs(ref["f"]))
    _, g_fd = om.fd_gradient(prob, cm.table, prob.theta, eps=1e-4)
    assert _rel(g_adj, g_fd) < 1e-6, (g_adj, g_fd)
    # marginal variances per entry: diag(Sigma) of Q_cond through the explicit
    # block inverses at this block size (the oracle's agrees with a dense
    # inverse to ~2e-13 per entry on the 40 x 40 mesh, kappa(Q_cond) ~ 8e3)
    _log("marginal variances vs oracle, per entry")
    lat, obs = model.marginal_variances(prob.theta)
    lat_ref, obs_ref = om.marginal_variances(prob, cm.table, prob.theta)
    lat = lat.cpu().numpy()
    assert np.all(lat_ref > 0)
    assert np.all(np.abs(lat - lat_ref) <= 1e-9 * lat_ref), np.max(np.abs(lat - lat_ref) / lat_ref)
    for v, r in zip(obs, obs_ref):
        v = v.cpu().numpy()
        assert np.all(np.abs(v - r) <= 1e-9 * np.abs(r) + 1e-12 * np.abs(r).max())


@pytest.mark.timeout(900)
def test_st_model_full_size_nt100_and_batched_fd(cuda):
    """The 40 x 40 mesh at nt = 100 (160,006 unknowns): f against the
    oracle; the 9 FD evaluations with all 18 factorizations batched against
    the sequential FD loop."""
    _log("st model 40x40, nt=100: f vs oracle")
    prob = st_problem(nt=100, nx=40, ny=40, nb=6, seed=1)
    cm = CoefficientModel("baseline")
    model = STGaussianModel(prob, coef=cm, device=cuda)
    parts = model.forward(prob.theta)
    ref = om.evaluate_f(prob, cm.table, prob.theta, return_parts=True)
    _check_f_parts(parts, ref)
    _log("batched FD (wave 9) vs sequential FD")
    f1, g1 = model.fd_gradient(prob.theta, eps=1e-3)
    f2, g2 = model.fd_gradient_batched(prob.theta, eps=1e-3, wave=9)
    assert abs(f1 - f2) <= 1e-12 * abs(f1)
    assert abs(f1 - ref["f"]) <= 1e-10 * max(abs(ref["ld_c"]), abs(ref["f"]))
    assert _rel(g2, g1) < 1e-8


@pytest.mark.timeout(900)
def test_coregional_at_benchmark_block_size(cuda):
    _log("coregional nv=3, 50x30 mesh (b=4500), nt=3")
    prob = coregional_problem(nv=3, nt=3, nx=50, ny=30, nb_per=3, obs_per_node=2, seed=4)
    cm = CoregionalCoefficients(3)
    model = STGaussianModel(prob, coef=cm, device=cuda)
    parts = model.forward(prob.theta)
    ref = om.evaluate_f(prob, cm.table, prob.theta, return_parts=True)
    _check_f_parts(parts, ref)
    x = model.x.cpu().numpy()
    assert _rel(x, ref["x"]) < 1e-9
    _log("adjoint gradient vs GPU FD (31 evaluations)")
    f_adj, g_adj = model.value_and_grad(prob.theta)
    _, g_fd = model.fd_gradient(prob.theta, eps=1e-4)
    assert _rel(g_adj, g_fd) < 1e-6, (g_adj, g_fd)
    _log("two components against the oracle FD")
    eps = 1e-4
    for i in (2, 9):  # theta_e of field 0, the first coregionalisation weight
        e = np.zeros_like(prob.theta)
        e[i] = eps
        gp = om.evaluate_f(prob, cm.table, prob.theta + e)
        gm = om.evaluate_f(prob, cm.table, prob.theta - e)
        g_o = (gp - gm) / (2 * eps)
        assert abs(g_adj[i] - g_o) <= 1e-6 * np.abs(g_adj).max(), (i, g_adj[i], g_o)


@pytest.mark.timeout(900)
def test_mixed_batch_of_24_at_benchmark_sizes(cuda):
    """The 24 seeded members (nt in {50,100,200}, b in {256,...,2048}, a = 6)
    factored as one batch: logdet against the known factor's exact value and
    the first and last diagonal blocks against the known factor."""
    from pyinla_amd.batch import BatchedFactorization
    from pyinla_amd.workloads import known_factor_bta_torch
    rng = np.random.default_rng(2)
    shapes = [(int(rng.choice([50, 100, 200])), int(rng.choice([256, 512, 1024, 2048]))) for _ in range(24)]
    a = 6
    solvers, exact, first_last = [], [], []
    for q, (n, b) in enumerate(shapes):
        Lf, Q = known_factor_bta_torch(n, b, a, seed=100 + q, device=cuda, norm_b=92)
        exact.append(2.0 * (torch.log(torch.diagonal(Lf[0], dim1=1, dim2=2)).sum()
                            + torch.log(torch.diag(Lf[3])).sum()).item())
        first_last.append((Lf[0][0].cpu().numpy(), Lf[0][n - 1].cpu().numpy()))
        del Lf
        s = SerinvSolver(None, b, n, a, device=cuda, use_graph=False, reserve="factor")
        s.load_blocks(*Q)
        del Q
        solvers.append(s)
    _log(f"24 problems loaded ({sum(n * b * b for n, b in shapes) * 16 / 1e9:.0f} GB of blocks)")
    batch = BatchedFactorization(solvers, ["bta"] * len(solvers))
    batch.factorize()
    for (n, b), s, ld_ex, (d0, dl) in zip(shapes, solvers, exact, first_last):
        info, ld = s.status()
        assert info == 0
        assert abs(ld - ld_ex) <= 1e-10 * abs(ld_ex), (n, b, ld, ld_ex)
        assert _rel(s.A_diagonal_blocks[0].cpu().numpy(), d0) < 1e-9
        assert _rel(s.A_diagonal_blocks[n - 1].cpu().numpy(), dl) < 1e-9


def _kron_schur_case(n, b, a, seed):
    """Q = [[M (x) S, B^T], [B, T]] with T = C + B (M (x) S)^-1 B^T, so that
    the Schur complement of the arrowhead is the chosen SPD C."""
    rng = np.random.default_rng(seed)
    M = np.diag(2.0 + rng.random(n)) + np.diag(-0.9 * rng.random(n - 1), -1)
    M = np.tril(M) + np.tril(M, -1).T
    X = rng.normal(size=(b, b)) / np.sqrt(b)
    S = X @ X.T + 0.5 * np.eye(b)
    R = rng.normal(size=(a, n, b)) * 0.3 / np.sqrt(n * b)  # arrow rows, block i = R[:, i, :]
    G = rng.normal(size=(a, a)) / np.sqrt(a)
    C = G @ G.T + np.eye(a)
    Mi = np.linalg.inv(M)
    Si = np.linalg.inv(S)
    Y = np.einsum("ts,ksc->ktc", Mi, R) @ Si  # Y_k = M^-1 R_k S^-1 (t-major)
    X_ = np.einsum("ktc,ltc->kl", R, Y)      # B K^-1 B^T
    T = C + X_
    return M, S, R, C, T, Mi, Si, Y


@pytest.mark.timeout(900)
def test_full_size_bta_schur_identity(cuda):
    """n = 200, b = 2000, a = 6 with the Schur complement fixed to C:
    logdet Q = b logdet M + n logdet S + logdet C; Q^-1 r by block
    elimination; Sigma_T = C^-1, Sigma_{n,i} = -C^-1 Z_i,
    Sigma_ii = (M^-1)_ii S^-1 + Z_i^T C^-1 Z_i,
    Sigma_{i+1,i} = (M^-1)_{i+1,i} S^-1 + Z_{i+1}^T C^-1 Z_i, with Z_i the
    block-i columns of B K^-1 (K = M (x) S)."""
    n, b, a = 200, 2000, 6
    _log(f"Kronecker-Schur BTA case n={n} b={b} a={a}: CPU closed forms")
    M, S, R, C, T, Mi, Si, Y = _kron_schur_case(n, b, a, 21)
    s = SerinvSolver(None, b, n, a, device=cuda, use_graph=False)
    St = torch.from_numpy(S).to(cuda)
    for t in range(n):
        s.A_diagonal_blocks[t].copy_(St).mul_(M[t, t])
        if t < n - 1:
            s.A_lower_diagonal_blocks[t].copy_(St).mul_(M[t + 1, t])
    s.A_arrow_bottom_blocks.copy_(torch.from_numpy(np.ascontiguousarray(R.transpose(1, 0, 2))))
    s.A_arrow_tip_block.copy_(torch.from_numpy(T))
    _log("factorize")
    s.factorize("bta")
    info, ld = s.status()
    assert info == 0
    ld_exact = b * np.linalg.slogdet(M)[1] + n * np.linalg.slogdet(S)[1] + np.linalg.slogdet(C)[1]
    assert abs(ld - ld_exact) <= 1e-10 * abs(ld_exact)
    assert abs(s.logdet("bta") - ld_exact) <= 1e-10 * abs(ld_exact)
    # solve
    rng = np.random.default_rng(4)
    r_st = rng.normal(size=(n, b))
    r_a = rng.normal(size=a)
    Kinv_r = Mi @ r_st @ Si
    x_a = np.linalg.solve(C, r_a - np.einsum("ktc,tc->k", R, Kinv_r))
    x_st = Kinv_r - np.einsum("k,ktc->tc", x_a, Y)
    x = torch.from_numpy(np.concatenate([r_st.ravel(), r_a])).to(cuda)
    s.solve(x, "bta")
    xg = x.cpu().numpy()
    assert _rel(xg[: n * b].reshape(n, b), x_st) < 1e-9
    assert _rel(xg[n * b:], x_a) < 1e-9
    _log("selected inversion")
    s.selected_inversion("bta")
    Ci = np.linalg.inv(C)
    assert _rel(s.A_arrow_tip_block.cpu().numpy(), Ci) < 1e-9
    Bg = s.A_arrow_bottom_blocks
    for t in (0, 1, n // 2, n - 2, n - 1):
        Z = Y[:, t, :]
        assert _rel(Bg[t].cpu().numpy(), -Ci @ Z) < 1e-9, t
        Sd = Mi[t, t] * Si + Z.T @ Ci @ Z
        assert _rel(s.A_diagonal_blocks[t].cpu().numpy(), Sd) < 1e-9, t
        if t < n - 1:
            Sl = Mi[t + 1, t] * Si + Y[:, t + 1, :].T @ Ci @ Z
            assert np.abs(s.A_lower_diagonal_blocks[t].cpu().numpy() - Sl).max() <= 1e-9 * np.abs(Sd).max(), t


@pytest.mark.timeout(600)
def test_adjoint_gradient_full_size_vs_9_point_fd(cuda):
    """The benchmark workload (nt = 200, 50 x 40 mesh, nb = 6, dim theta =
    4): adjoint gradient against the 2 d + 1 = 9-point central FD of f with
    step 1e-4."""
    _log("adjoint vs 9-point FD at nt=200, ns=2000")
    prob = st_problem(nt=200, nx=50, ny=40, nb=6, seed=3, hx=1.0 / 49, hy=1.0 / 39)
    model = STGaussianModel(prob, device=cuda)
    f_adj, g_adj = model.value_and_grad(prob.theta)
    f_fd, g_fd = model.fd_gradient(prob.theta, eps=1e-4)
    assert abs(f_adj - f_fd) <= 1e-12 * abs(f_fd)
    assert _rel(g_adj, g_fd) < 1e-6, (g_adj, g_fd)
