"""The allocation-free H_eff matvec (ct_heff_apply_noalloc) against ct_heff_apply.

Both entry points run the same kernels with the same launch shapes; only the
intermediates live elsewhere (the library's static scratch array instead of the
handle's workspaces), so the outputs must agree bit for bit.  Two calls in a
row, the second on smaller dimensions, check that nothing depends on what the
previous call left in the shared scratch array.
"""
import ctypes

import numpy as np
import pytest
import torch

import chemtensor_mojo_amd as ct
from chemtensor_mojo_amd import _lib
from chemtensor_mojo_amd import hamiltonians as ham

pytestmark = pytest.mark.gpu


def _problem(rng, Dl, Dlb, Dr, Drb, w, dt):
    W = ham.j1j2_tensors(6, 1.0, 0.5)[2] if w == 8 else ham.heisenberg_xxz_tensors(6, 1.0, 1.0, 0.0)[2]
    L = rng.standard_normal((w, Dl, Dlb)) / np.sqrt(Dl)
    R = rng.standard_normal((w, Dr, Drb)) / np.sqrt(Dr)
    th = rng.standard_normal((Dl, 2, 2, Dr))
    dev = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")
    return dev(L), dev(W), dev(R), dev(th), W


def _run(name, ctx, prec, L, W, R, th, out, dims):
    Dl, Dlb, Dr, Drb, w = dims
    dcode = _lib.F64 if L.dtype == torch.float64 else _lib.F32
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    _lib.call(name, ctx.handle, prec, dcode, p(L), p(W), p(W), p(R), p(th), p(out), Dl, Dlb, 2, Dr, Drb, w, w, w)


# first call: the step-4 GEMM splits K (64x64 tiles, 4 splits); second: smaller
CALLS = [(64, 64, 64, 64, 8), (40, 36, 24, 30, 5)]


@pytest.mark.parametrize("prec,dt", [(ct.PREC_FP64, torch.float64), (ct.PREC_FP32, torch.float32),
                                     (ct.PREC_FP32, torch.float64)])
def test_noalloc_matches_heff_apply(ctx, ora, prec, dt):
    rng = np.random.default_rng(11)
    for dims in CALLS:
        Dl, Dlb, Dr, Drb, w = dims
        L, W, R, th, Wn = _problem(rng, Dl, Dlb, Dr, Drb, w, dt)
        n = Dlb * 4 * Drb
        ref = torch.full((n,), float("nan"), dtype=dt, device="cuda")
        got = torch.full((n,), float("nan"), dtype=dt, device="cuda")
        _run("ct_heff_apply", ctx, prec, L, W, R, th, ref, dims)
        _run("ct_heff_apply_noalloc", ctx, prec, L, W, R, th, got, dims)
        torch.cuda.synchronize()
        assert torch.equal(ref, got), dims
        # and both against the FP64 oracle on the same inputs
        exp = ora.heff_apply(L.double().cpu().numpy(), Wn, Wn, R.double().cpu().numpy(),
                             th.double().cpu().numpy())
        err = np.linalg.norm(got.double().cpu().numpy().reshape(exp.shape) - exp) / np.linalg.norm(exp)
        assert err < (1e-13 if prec == ct.PREC_FP64 else 2e-6), (dims, err)


def test_noalloc_rejects_beyond_scratch(ctx):
    rng = np.random.default_rng(3)
    # D = 512, w = 8 in FP64 needs ~270 MiB of intermediates: rejected before any launch
    D, w = 512, 8
    L = torch.zeros(w * D * D, dtype=torch.float64, device="cuda")
    th = torch.zeros(D * 4 * D, dtype=torch.float64, device="cuda")
    W = torch.tensor(ham.j1j2_tensors(6, 1.0, 0.5)[2], device="cuda")
    out = torch.full((D * 4 * D,), 7.0, dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.CTError) as e:
        _run("ct_heff_apply_noalloc", ctx, ct.PREC_FP64, L, W, L, th, out, (D, D, D, D, w))
    assert e.value.status == 7 and "MiB" in str(e.value)
    torch.cuda.synchronize()
    assert bool((out == 7.0).all())
    # 3xBF16 is not offered allocation-free
    with pytest.raises(_lib.CTError) as e2:
        _run("ct_heff_apply_noalloc", ctx, ct.PREC_BF16X3, L, W, L, th, out, (D, D, D, D, w))
    assert e2.value.status == -2
    # a regular call still works afterwards
    Ls, Ws, Rs, ths, Wn = _problem(rng, 16, 16, 16, 16, 8, torch.float64)
    o1 = torch.empty(16 * 4 * 16, dtype=torch.float64, device="cuda")
    o2 = torch.empty_like(o1)
    _run("ct_heff_apply", ctx, ct.PREC_FP64, Ls, Ws, Rs, ths, o1, (16, 16, 16, 16, 8))
    _run("ct_heff_apply_noalloc", ctx, ct.PREC_FP64, Ls, Ws, Rs, ths, o2, (16, 16, 16, 16, 8))
    torch.cuda.synchronize()
    assert torch.equal(o1, o2)
