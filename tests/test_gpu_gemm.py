"""GPU numerics of the grouped FP64-MFMA GEMM (bta_gemm, the kernel of the
block inverses and the selected inversion) against torch float64 matmul on
the same inputs: every transpose combination, ragged edges, odd leading
dimensions (the element-wise load path), even-padded leading dimensions
with ragged rows and odd K (the paired-load path with its clamped addresses)
and beta / alpha.  Bar: max |error|
<= 1e-13 * max |C| (K <= 600, FP64 accumulation of O(1) entries)."""
import ctypes as C

import numpy as np
import pytest
import torch

from pyinla_amd import _lib

pytestmark = pytest.mark.gpu


def _ptr(t):
    return C.c_void_p(t.data_ptr())


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("m,n,k,pad", [(64, 64, 64, 0), (130, 70, 33, 1), (257, 190, 600, 0), (5, 300, 17, 3),
                                       (2000, 64, 250, 0), (67, 130, 33, -1), (5, 300, 17, -1), (257, 190, 601, -1),
                                       (129, 1, 3, -1)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_against_torch(cuda, variant, m, n, k, pad, ta, tb):
    lib = _lib.load()
    g = torch.Generator(device=cuda)
    g.manual_seed(m * 7 + n * 3 + k + 11 * ta + 5 * tb)
    f64 = dict(dtype=torch.float64, device=cuda)
    Ar, Ac = (k, m) if ta else (m, k)
    Br, Bc = (n, k) if tb else (k, n)
    # pad < 0: every leading dimension is the next even number above the width
    pa, pb, pc = ((2 - (x & 1) for x in (Ac, Bc, n)) if pad < 0 else (pad, pad, pad))
    A = torch.randn((Ar, Ac + pa), generator=g, **f64)
    B = torch.randn((Br, Bc + pb), generator=g, **f64)
    Cm = torch.randn((m, n + pc), generator=g, **f64)
    opA = A[:, :Ac].T if ta else A[:, :Ac]
    opB = B[:, :Bc].T if tb else B[:, :Bc]
    ref = 0.5 * Cm[:, :n] - 1.5 * (opA @ opB)
    out = Cm.clone()
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.bta_gemm(_ptr(A), Ac + pa, ta, _ptr(B), Bc + pb, tb, _ptr(out), n + pc, m, n, k, -1.5, 0.5,
                            variant, st), "bta_gemm")
    torch.cuda.synchronize()
    err = (out[:, :n] - ref).abs().max().item()
    assert err <= 1e-13 * ref.abs().max().item()
    if pc:  # columns past n untouched
        assert torch.equal(out[:, n:], Cm[:, n:])


def test_gemm_nn_2000_rate(cuda):
    """The product shape of the selected inversion (NN, 2000^3): the kernel's
    rate is printed (measured on MI355X), the result checked."""
    lib = _lib.load()
    b = 2000
    f64 = dict(dtype=torch.float64, device=cuda)
    A = torch.randn((b, b), **f64)
    B = torch.randn((b, b), **f64)
    out = torch.empty((b, b), **f64)
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for variant in (0, 1):
        times = []
        for _ in range(4):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            _lib.check(lib.bta_gemm(_ptr(A), b, 0, _ptr(B), b, 0, _ptr(out), b, b, b, b, 1.0, 0.0, variant, st),
                       "bta_gemm")
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1))
        print(f"[gemm] variant {variant}: NN 2000^3 in {min(times):.3f} ms "
              f"(includes plan upload) = {2 * b ** 3 / min(times) / 1e9:.1f} TFLOP/s", flush=True)
    ref = A @ B
    assert (out - ref).abs().max().item() <= 1e-12 * ref.abs().max().item()
