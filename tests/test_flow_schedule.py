"""CPU simulation of the persistent panel schedule (bta_flow.hip).

The helper task lists come from the library itself (bta_flow_schedule, host
code only).  A simulator runs the chain workgroup and the helpers as
coroutines under random interleavings, with the device's wait/publish rules:
a helper or the chain only proceeds once the flag it waits for has been
published.  It checks that

* no interleaving deadlocks (both the dependency-level order and the order
  of a simulated launch are valid orders),
* every tile receives exactly its updates before it is solved or taken over
  by the chain (update counts are tracked per tile),
* every factor address is stored to by one agent only (the cache-coherence
  rule the kernel relies on),
* the result equals numpy's Cholesky factor of the panel [D; L; B], and the
  next-block updates D_{i+1} -= L L^T, B_{i+1} -= B L^T, T -= B B^T folded
  into the same launch are applied exactly once per column.
"""
import ctypes as C
import random

import numpy as np
import pytest

from pyinla_amd import _lib

UPD, TRSM, ZERO, NEXT = 0, 1, 2, 3
SPAN = 4


def next_end(k, nT):
    """End of the next-block task starting at column k (the library's rule:
    spans of SPAN columns, the last SPAN columns one each)."""
    tail = nT - SPAN
    return min(k + SPAN, tail) if k < tail else k + 1


def schedule(nT, nTL, nTa, helpers, delta, near_div=0):
    lib = _lib.load()
    n = C.c_int64()
    assert lib.bta_flow_schedule(nT, nTL, nTa, helpers, delta, near_div, None, 0, None, C.byref(n)) == 0
    tasks = (C.c_uint32 * max(1, n.value))()
    off = (C.c_int * (helpers + 1))()
    assert lib.bta_flow_schedule(nT, nTL, nTa, helpers, delta, near_div, tasks, n.value, off, None) == 0
    words = list(tasks)[: n.value]
    lists = []
    for h in range(helpers):
        lst = []
        for w in words[off[h]:off[h + 1]]:
            lst.append((w >> 30, (w >> 20) & 1023, (w >> 10) & 1023, w & 1023))
        lists.append(lst)
    return lists


class Sim:
    """Tile-level model of one k_panel_flow launch (64-wide tiles; rows and
    columns beyond the matrix are zero-padded as on the device)."""

    def __init__(self, D, L, B, ts=64, Dn=None, Bn=None, T=None):
        self.ts = ts
        b = D.shape[0]
        self.b = b
        self.nT = -(-b // ts)
        self.nTL = self.nT if L is not None else 0
        self.nTa = 1 if B is not None and B.shape[0] > 0 else 0
        self.R = self.nT + self.nTL + self.nTa
        pad = self.nT * ts
        self.P = np.zeros((self.R * ts, pad))
        self.P[:b, :b] = D
        if self.nTL:
            self.P[self.nT * ts:self.nT * ts + b, :b] = L
        if self.nTa:
            self.P[(self.nT + self.nTL) * ts:(self.nT + self.nTL) * ts + B.shape[0], :b] = B
        self.Dn = np.zeros((pad, pad)) if Dn is None else np.pad(Dn, ((0, pad - b), (0, pad - b)))
        self.Bn = np.zeros((ts, pad)) if Bn is None else np.pad(Bn, ((0, ts - Bn.shape[0]), (0, pad - b)))
        self.T = np.zeros((ts, ts)) if T is None else np.pad(T, ((0, ts - T.shape[0]), (0, ts - T.shape[0])))
        self.next_cols = {}  # next-block target -> columns applied
        self.xbuf = {}
        self.hacc = {}
        self.winv = {}
        self.flags = set()
        self.writer = {}   # factor tile -> agent that stored to it
        self.nupd = {}     # factor tile -> number of updates applied (helpers + chain)

    def tile(self, r, c):
        t = self.ts
        return self.P[r * t:(r + 1) * t, c * t:(c + 1) * t]

    def store(self, agent, r, c, val):
        key = (r, c)
        prev = self.writer.setdefault(key, agent)
        assert prev == agent, f"tile {key} stored by {prev} and {agent}"
        self.tile(r, c)[...] = val

    def valid(self, c):
        return min(self.ts, self.b - self.ts * c)

    def handoff(self, r, c):
        return r < self.nT and ((r == c and c >= 1) or r == c + 1)

    # -- agents (generators yielding flag keys to wait for)
    def chol(self, A, mv):
        A = A.copy()
        A[mv:, :] = 0.0
        A[:, mv:] = 0.0
        A[range(mv, self.ts), range(mv, self.ts)] = 1.0
        Lf = np.linalg.cholesky(np.tril(A) + np.tril(A, -1).T)
        return Lf, np.linalg.inv(Lf)

    def chain(self):
        """The chain workgroup with the device's wait and publish points
        (flow_chain): per column k it waits for BOTH hand-off tiles before
        the solve of (k+1, k); X(k+1, k) is published after the update of
        (k+1, k+1), and every inverse right after its factorization."""
        nT = self.nT
        Lf, W = self.chol(self.tile(0, 0), self.valid(0))
        self.store("chain", 0, 0, Lf)
        self.winv[0] = W
        self.flags.add(("w", 0))
        for k in range(nT - 1):
            if k == 0:
                Cs = self.tile(1, 0).copy()
                Cd = self.tile(1, 1).copy()
            else:
                yield ("h", k, 0)
                yield ("h", k, 1)
                Cs = self.hacc[("s", k)].copy()
                Cd = self.hacc[("d", k + 1)].copy()
            assert self.nupd.get((k + 1, k), 0) == k
            assert self.nupd.get((k + 1, k + 1), 0) == k
            X = Cs @ W.T
            self.store("chain", k + 1, k, X)
            self.xbuf[(k + 1, k)] = X.copy()
            Cd -= X @ X.T
            self.flags.add(("x", k + 1, k))
            self.nupd[(k + 1, k + 1)] = k + 1
            Lf, W = self.chol(Cd, self.valid(k + 1))
            self.store("chain", k + 1, k + 1, Lf)
            self.winv[k + 1] = W
            self.flags.add(("w", k + 1))

    def helper(self, h, tasks):
        nT = self.nT
        name = f"h{h}"
        for op, r, c, k in tasks:
            if op == NEXT:
                t = self.ts
                if r < nT:
                    tgt, ra, rb = self.Dn[r * t:(r + 1) * t, c * t:(c + 1) * t], nT + r, nT + c
                elif r == nT:
                    tgt, ra, rb = self.Bn[:, c * t:(c + 1) * t], nT + self.nTL, nT + c
                else:
                    tgt, ra, rb = self.T, nT + self.nTL, nT + self.nTL
                for kk in range(k, next_end(k, nT)):
                    yield ("x", ra, kk)
                    if rb != ra:
                        yield ("x", rb, kk)
                    tgt -= self.xbuf[(ra, kk)] @ self.xbuf[(rb, kk)].T
                    self.next_cols.setdefault((r, c), []).append(kk)
                continue
            if op == ZERO:
                self.store(name, r, c, np.zeros((self.ts, self.ts)))
                continue
            if op == UPD:
                yield ("x", r, k)
                if r != c:
                    yield ("x", c, k)
                A = self.xbuf[(r, k)]
                Bm = self.xbuf[(c, k)]
                hand = self.handoff(r, c)
                if hand and k > 0:
                    acc = self.hacc[("d" if r == c else "s", c)].copy()
                else:
                    acc = self.tile(r, c).copy()
                assert self.nupd.get((r, c), 0) == k or not hand
                acc -= A @ Bm.T
                self.nupd[(r, c)] = self.nupd.get((r, c), 0) + 1
                if hand:
                    self.hacc[("d" if r == c else "s", c)] = acc
                    if r == c + 1 and k == c - 1:
                        self.flags.add(("h", c, 0))
                    elif r == c and k == c - 2:
                        self.flags.add(("h", c - 1, 1))
                else:
                    self.store(name, r, c, acc)
            else:
                yield ("w", k)
                assert self.nupd.get((r, c), 0) == c, (r, c, self.nupd.get((r, c), 0))
                X = self.tile(r, c) @ self.winv[k].T
                self.store(name, r, c, X)
                self.xbuf[(r, c)] = X.copy()
                self.flags.add(("x", r, c))

    def run(self, lists, seed):
        rng = random.Random(seed)
        agents = {"chain": self.chain()}
        for h, lst in enumerate(lists):
            agents[f"h{h}"] = self.helper(h, lst)
        waiting = {a: None for a in agents}
        live = set(agents)
        while live:
            ready = [a for a in live if waiting[a] is None or waiting[a] in self.flags]
            assert ready, f"deadlock: waits {[(a, waiting[a]) for a in live]}"
            a = rng.choice(ready)
            try:
                waiting[a] = next(agents[a])
            except StopIteration:
                live.discard(a)


def _panel(b, a, with_l, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(b, b)) / np.sqrt(b)
    D = X @ X.T + np.eye(b)
    L = rng.normal(size=(b, b)) if with_l else None
    B = rng.normal(size=(a, b)) if a else None
    return D, L, B


@pytest.mark.parametrize("b,a,helpers,delta", [(300, 6, 5, 2), (256, 0, 3, 0), (200, 9, 7, 1), (130, 4, 2, 3)])
def test_schedule_simulation_next_block_updates(b, a, helpers, delta):
    D, L, B = _panel(b, a, True, 8)
    rng = np.random.default_rng(9)
    Dn = rng.normal(size=(b, b))
    Bn = rng.normal(size=(a, b)) if a else None
    T = rng.normal(size=(a, a)) if a else None
    nT = -(-b // 64)
    lists = schedule(nT, nT, 1 if a else 0, helpers, delta)
    s = Sim(D, L, B, Dn=Dn, Bn=Bn, T=T)
    s.run(lists, 1)
    for key, cols in s.next_cols.items():
        assert sorted(cols) == list(range(nT)), key
    Lf = np.linalg.cholesky(D)
    Li = np.linalg.solve(Lf, L.T).T
    got = s.Dn[:b, :b]
    exp = Dn - Li @ Li.T
    assert np.abs(np.tril(got) - np.tril(exp)).max() <= 1e-9 * np.abs(exp).max()
    if a:
        Bi = np.linalg.solve(Lf, B.T).T
        np.testing.assert_allclose(s.Bn[:a, :b], Bn - Bi @ Li.T, rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(s.T[:a, :a], T - Bi @ Bi.T, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("b,a,with_l,helpers,delta", [
    (300, 6, True, 5, 2), (300, 6, False, 3, 0), (256, 0, True, 7, 1), (200, 4, True, 2, 5),
    (130, 6, True, 4, 2), (64, 6, True, 3, 2), (450, 9, True, 11, 3)])
def test_schedule_simulation_matches_cholesky(b, a, with_l, helpers, delta):
    D, L, B = _panel(b, a, with_l, 7)
    nT = -(-b // 64)
    lists = schedule(nT, nT if with_l else 0, 1 if a else 0, helpers, delta)
    for seed in range(3):
        s = Sim(D, L, B)
        s.run(lists, seed)
        Lf = np.linalg.cholesky(D)
        got = s.P[:b, :b]
        assert np.abs(np.tril(got) - Lf).max() <= 1e-10 * np.abs(Lf).max()
        assert np.abs(np.triu(got, 1)).max() == 0.0
        if with_l:
            exp = np.linalg.solve(Lf, L.T).T
            np.testing.assert_allclose(s.P[nT * 64:nT * 64 + b, :b], exp, rtol=1e-9, atol=1e-9)
        if a:
            r0 = (2 * nT if with_l else nT) * 64
            exp = np.linalg.solve(Lf, B.T).T
            np.testing.assert_allclose(s.P[r0:r0 + a, :b], exp, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("delta", [2, 0])
def test_schedule_lists_are_well_formed(delta):
    """Solves of a tile follow all of its updates, updates of a tile come in
    column order, and every task appears exactly once."""
    nT, helpers = 9, 6
    lists = schedule(nT, nT, 1, helpers, delta)
    seen = set()
    for lst in lists:
        done_upd = {}
        for op, r, c, k in lst:
            key = (op, r, c, k)
            assert key not in seen
            seen.add(key)
            if op == UPD:
                j = done_upd.get((r, c), 0)
                assert k == j or (r == c and k == j), (r, c, k, j)
                done_upd[(r, c)] = j + 1
            elif op == TRSM:
                assert done_upd.get((r, c), 0) == c
    R = 2 * nT + 1
    n_upd = sum(1 for (op, *_rest) in seen if op == UPD)
    exp = 0
    for r in range(R):
        for c in range(nT):
            if r < nT and r < c:
                continue
            exp += c - (1 if (r == c and c >= 1) else 0)
    assert n_upd == exp


def test_simulator_catches_planted_schedule_mistakes():
    """A helper list run backwards deadlocks; a dropped update trips a count
    check, reads a hand-off tile never written, or deadlocks."""
    import copy
    D, L, B = _panel(300, 6, True, 7)
    lists = schedule(5, 5, 1, 5, 2)
    bad = copy.deepcopy(lists)
    bad[2] = bad[2][::-1]
    with pytest.raises(AssertionError):
        Sim(D, L, B).run(bad, 0)
    bad = copy.deepcopy(lists)
    h = next(h for h, lst in enumerate(bad) if any(t[0] == UPD and t[2] >= 2 for t in lst))
    q = next(q for q, t in enumerate(bad[h]) if t[0] == UPD and t[2] >= 2)
    del bad[h][q]
    with pytest.raises((AssertionError, KeyError)):
        Sim(D, L, B).run(bad, 0)


@pytest.mark.parametrize("nT,nTL,nTa,helpers,near", [(32, 32, 1, 255, 0), (32, 0, 0, 127, 0), (16, 16, 1, 40, 0),
                                                      (32, 0, 1, 255, 0), (71, 0, 0, 255, 0), (71, 0, 1, 255, 0),
                                                      (71, 71, 1, 255, 0), (71, 0, 0, 200, 0), (40, 0, 0, 511, 0),
                                                      (32, 32, 1, 255, 4), (32, 0, 0, 255, 2), (32, 32, 0, 255, 3)])
def test_list_order_at_production_shape(nT, nTL, nTa, helpers, near):
    """The list order (delta 0) for panels of b = 2048 and b = 4500 (71 tile
    columns) with full-device helper counts, blocks with and without L: no
    interleaving deadlocks and the factor is right (small tiles keep the
    numerics cheap; the schedule depends on the tile counts only).  The last
    block of the b = 4500 shape deadlocked on the device while TRSM tasks
    did not depend on the chain step that publishes inv(L_kk)."""
    ts = 4 if nT <= 32 else 2
    b = nT * ts
    D, L, B = _panel(b, min(3, ts) if nTa else 0, nTL > 0, 5)
    lists = schedule(nT, nTL, nTa, helpers, 0, near)
    s = Sim(D, L, B, ts=ts)
    s.run(lists, 3)
    Lf = np.linalg.cholesky(D)
    assert np.abs(np.tril(s.P[:b, :b]) - Lf).max() <= 1e-10 * np.abs(Lf).max()
