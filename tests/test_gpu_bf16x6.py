"""GPU parity of the 6xBF16 form (CT_PREC_BF16X6): FP32-accurate contractions
on the BF16 matrix cores.

Each operand x is split into three bf16 values h + m + l that carry its FP32
significand; six products (all terms down to 2^-16 |ab|) are accumulated in
FP32.  The claim under test is FP32 accuracy: every result is held to the
FP32 form's tolerance against the FP64 oracle, and its error is compared with
the FP32 MFMA path's on the same inputs.
"""
import ctypes

import numpy as np
import pytest
import torch

import chemtensor_mojo_amd as ct
from chemtensor_mojo_amd import hamiltonians as ham
from chemtensor_mojo_amd._lib import call

pytestmark = pytest.mark.gpu

X6 = ct.PREC_BF16X6


def dev(ctx, a, dtype=torch.float64):
    return ct.create_dense_tensor_from_data(ctx, a, a.shape, dtype)


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _bf16_rne(x32):
    """float32 array -> bf16 bit patterns, round to nearest even."""
    u = x32.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _bf16_val(b):
    return (b.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("rows,K,s_row_is_1,dt", [(300, 70, True, np.float64), (257, 33, True, np.float32),
                                                  (64, 1024, True, np.float64), (50, 90, False, np.float64)])
def test_split_bf16x6_bit_exact(ctx, rows, K, s_row_is_1, dt):
    """ct_split_bf16x6 against a numpy restatement of the three-way split
    h = bf16(x), m = bf16(x - h), l = bf16(x - h - m) of the FP32 value, with
    zero padding to 32; h + m + l reproduces the FP32 value exactly.  Both
    split kernels: rows contiguous (the transposing one) and k contiguous."""
    rng = np.random.default_rng(rows * 3 + K)
    K32 = (K + 31) // 32 * 32
    if s_row_is_1:
        X = rng.standard_normal((K, rows)).astype(dt)  # element (r, k) at X[k][r]
        s_row, s_k, x = 1, rows, X.T
    else:
        X = rng.standard_normal((rows, K)).astype(dt)
        s_row, s_k, x = K, 1, X
    ld, ldl = 2 * K32, K32 + 8
    xt = torch.tensor(X, device="cuda")
    out = torch.full((rows * ld,), 0x7FC0, dtype=torch.int16, device="cuda")
    outl = torch.full((rows * ldl,), 0x7FC0, dtype=torch.int16, device="cuda")
    call("ct_split_bf16x6", ctx.handle, 0 if dt == np.float64 else 1, vp(xt), rows, K, s_row, s_k, vp(out), ld,
         vp(outl), ldl)
    got = out.cpu().numpy().view(np.uint16).reshape(rows, K32 // 32, 2, 32)
    gotl = outl.cpu().numpy().view(np.uint16).reshape(rows, ldl)[:, :K32]
    x32 = np.zeros((rows, K32), np.float32)
    x32[:, :K] = x.astype(np.float32)
    h = _bf16_rne(x32)
    r = x32 - _bf16_val(h)
    m = _bf16_rne(r)
    l3 = _bf16_rne(r - _bf16_val(m))
    assert np.array_equal(got[:, :, 0, :], h.reshape(rows, K32 // 32, 32))
    assert np.array_equal(got[:, :, 1, :], m.reshape(rows, K32 // 32, 32))
    assert np.array_equal(gotl, l3)
    recon = _bf16_val(h).astype(np.float64) + _bf16_val(m) + _bf16_val(l3)
    assert np.array_equal(recon, x32.astype(np.float64))


@pytest.mark.parametrize("M,N,K,dt", [
    # 256x128 tiles
    (4096, 1024, 8192, torch.float64),  # the last matvec GEMM at D = 1024, w = 8: 128 tiles, split-K = 2
    (1000, 1100, 3000, torch.float32),  # ragged M / N / K, split-K = 4 over a padded K
    (2048, 2048, 256, torch.float64),   # 128 tiles, single K range
    # 128x128 tiles (fewer than 128 tiles of 256x128 after splitting K)
    (64, 96, 4096, torch.float64),      # one tile: split-K = 32
    (200, 300, 1000, torch.float32),    # ragged, split-K = 8 over a padded K
    (300, 190, 77, torch.float64),      # ragged, single K range
    (513, 129, 32, torch.float32),      # one K step
])
def test_gemm_bf16x6_nt_fp32_accuracy(ctx, M, N, K, dt):
    """ct_gemm_bf16x6_nt against FP64 products on the device: the six-product
    sum of the split values (FP32 accumulation error only) and the exact
    product, whose error must not exceed the FP32 MFMA GEMM's (ct_tensordot,
    CT_PREC_FP32) on the same inputs.  Split buffers start as BF16 NaN
    patterns, so an unwritten pad shows up as NaN in C."""
    g = torch.Generator(device="cuda").manual_seed(M + 7 * N + K)
    A = torch.randn(M, K, dtype=torch.float64, device="cuda", generator=g)
    B = torch.randn(N, K, dtype=torch.float64, device="cuda", generator=g)
    K32 = (K + 31) // 32 * 32
    ld, ldl = 2 * K32, K32
    sa = torch.full((M * ld,), 0x7FC0, dtype=torch.int16, device="cuda")
    sb = torch.full((N * ld,), 0x7FC0, dtype=torch.int16, device="cuda")
    la = torch.full((M * ldl,), 0x7FC0, dtype=torch.int16, device="cuda")
    lb = torch.full((N * ldl,), 0x7FC0, dtype=torch.int16, device="cuda")
    call("ct_split_bf16x6", ctx.handle, 0, vp(A), M, K, K, 1, vp(sa), ld, vp(la), ldl)
    call("ct_split_bf16x6", ctx.handle, 0, vp(B), N, K, K, 1, vp(sb), ld, vp(lb), ldl)
    C = torch.full((M, N), float("nan"), dtype=dt, device="cuda")
    call("ct_gemm_bf16x6_nt", ctx.handle, vp(sa), ld, vp(la), ldl, vp(sb), ld, vp(lb), ldl,
         0 if dt == torch.float64 else 1, vp(C), M, N, K, N)
    torch.cuda.synchronize()

    def planes(t, tl, r):
        u = t.view(r, ld).reshape(r, K32 // 32, 2, 32).to(torch.int32) & 0xFFFF
        f = (u << 16).view(torch.float32).to(torch.float64)
        ul = tl.view(r, ldl).to(torch.int32) & 0xFFFF
        fl = (ul << 16).view(torch.float32).to(torch.float64)
        return f[:, :, 0, :].reshape(r, K32)[:, :K], f[:, :, 1, :].reshape(r, K32)[:, :K], fl[:, :K]

    (ah, am, al), (bh, bm, bl) = planes(sa, la, M), planes(sb, lb, N)
    mm = lambda x, y: x @ y.T.contiguous()
    six = mm(al, bh) + mm(ah, bl) + mm(am, bm) + mm(am, bh) + mm(ah, bm) + mm(ah, bh)
    exact = mm(A, B)
    Cd = C.to(torch.float64)
    assert not torch.isnan(Cd).any()
    r6 = float(torch.linalg.norm(Cd - six) / torch.linalg.norm(six))
    re = float(torch.linalg.norm(Cd - exact) / torch.linalg.norm(exact))
    assert r6 < (2e-6 if dt == torch.float64 else 4e-6), r6
    # the FP32 MFMA GEMM of the same product
    C32 = torch.empty((M, N), dtype=torch.float64, device="cuda")
    ct.dense_tensor_dot(ct.DenseTensor(C32.view(-1), [M, N]), ct.DenseTensor(A.reshape(-1), [M, K]),
                        ct.DenseTensor(B.reshape(-1), [N, K]), ctx, 1, False, False, prec=ct.PREC_FP32)
    r32 = float(torch.linalg.norm(C32 - exact) / torch.linalg.norm(exact))
    if dt == torch.float64:
        assert re <= 1.5 * r32 + 1e-8, (re, r32)
    assert re < 3e-6, re


def test_tensordot_bf16x6_vs_oracle(ctx, ora):
    """dense_tensor_dot in the 6xBF16 form (operands split in workspace) on the
    tile-edge, split-K and axis-inference cases of the FP32 test, to the FP32
    tolerance."""
    cases = [((129, 70), (70, 131), 1, False, True), ((300, 257), (300, 190), 1, True, True),
             ((190, 33), (77, 33), 1, False, False), ((64, 256), (128, 64), 1, True, False),
             ((64, 4096), (4096, 96), 1, False, True), ((3, 5000), (7, 5000), 1, False, False),
             ((2, 3, 1000), (3, 1000, 130), 2, False, True), ((1024, 4096), (4096, 2048), 1, False, True)]
    rng = np.random.default_rng(5)
    for sA, sB, nm, aA, aB in cases:
        A = rng.standard_normal(sA)
        B = rng.standard_normal(sB)
        exp = ora.tensordot(A, B, nm, aA, aB)
        for dt in (torch.float32, torch.float64):
            C = ct.create_dense_tensor(ctx, list(exp.shape), dtype=dt)
            ct.dense_tensor_dot(C, dev(ctx, A, dt), dev(ctx, B, dt), ctx, nm, aA, aB, prec=X6)
            assert rel(C.to_numpy(), exp) < 2e-6, (sA, sB, dt)


def _rand_problem(rng, Dl, Dr, d, wL, wM, wR):
    L = rng.standard_normal((wL, Dl, Dl)) / np.sqrt(Dl)
    R = rng.standard_normal((wR, Dr, Dr)) / np.sqrt(Dr)
    Wi = rng.standard_normal((wL, d, d, wM))
    Wj = rng.standard_normal((wM, d, d, wR))
    th = rng.standard_normal((Dl, d, d, Dr))
    return L, R, Wi, Wj, th


@pytest.mark.parametrize("shape", [(1, 2, 2, 1, 5, 1), (7, 9, 2, 5, 5, 5), (64, 64, 2, 5, 5, 5),
                                   (130, 96, 2, 8, 8, 8), (256, 256, 2, 8, 8, 8), (40, 40, 3, 4, 4, 4),
                                   # w*d^2 > 64: the mixing as a batched six-product GEMM (quantum-chemistry
                                   # MPO bonds), with the 8-byte and the per-element unpack
                                   (64, 48, 2, 40, 40, 40), (24, 20, 2, 501, 562, 501), (20, 9, 2, 40, 40, 40),
                                   (64, 24, 2, 501, 562, 501)])
def test_heff_bf16x6_vs_oracle(ctx, ora, shape):
    """H_eff in the 6xBF16 form against the oracle, FP64 and FP32 storage, to
    the FP32 form's tolerance."""
    Dl, Dr, d, wL, wM, wR = shape
    rng = np.random.default_rng(7 + Dl)
    L, R, Wi, Wj, th = _rand_problem(rng, Dl, Dr, d, wL, wM, wR)
    exp = ora.heff_apply(L, Wi, Wj, R, th)
    for dt in (torch.float64, torch.float32):
        out = ct.apply_two_site_heff(dev(ctx, th, dt), dev(ctx, L, dt), dev(ctx, R, dt), ct.MPOSite(dev(ctx, Wi, dt)),
                                     ct.MPOSite(dev(ctx, Wj, dt)), ctx, prec=X6)
        assert rel(out.to_numpy(), exp) < 2e-6, dt


def test_heff_bf16x6_D1024_matches_fp32_accuracy(ctx):
    """The benchmark matvec (D = 1024, J1-J2 bulk MPO, w = 8): the 6xBF16 result
    is at least as close to the FP64 one as the FP32 MFMA result is."""
    rng = np.random.default_rng(11)
    D, w = 1024, 8
    L = rng.standard_normal((w, D, D)) / np.sqrt(D)
    R = rng.standard_normal((w, D, D)) / np.sqrt(D)
    W = ham.j1j2_tensors(6, 1.0, 0.5)[2]
    th = rng.standard_normal((D, 2, 2, D))
    tL, tR, tth = dev(ctx, L), dev(ctx, R), dev(ctx, th)
    tW = ct.MPOSite(dev(ctx, W))
    o64 = ct.apply_two_site_heff(tth, tL, tR, tW, tW, ctx, prec=ct.PREC_FP64).to_numpy()
    o32 = ct.apply_two_site_heff(tth, tL, tR, tW, tW, ctx, prec=ct.PREC_FP32).to_numpy()
    o6 = ct.apply_two_site_heff(tth, tL, tR, tW, tW, ctx, prec=X6).to_numpy()
    assert np.isfinite(o6).all()
    e32, e6 = rel(o32, o64), rel(o6, o64)
    assert e6 < 2e-6
    assert e6 <= 1.2 * e32, (e6, e32)


def test_heff_bf16x6_prepared_matches_one_shot(ctx):
    """The prepared operator (split environments kept with the operator) gives
    the one-shot call's results bit for bit, also after calls of another shape
    on the same handle."""
    rng = np.random.default_rng(21)
    Dl, Dr, w = 96, 80, 8
    L = rng.standard_normal((w, Dl, Dl))
    R = rng.standard_normal((w, Dr, Dr))
    W = ham.j1j2_tensors(6, 1.0, 0.5)[2]
    tL, tR, tW = dev(ctx, L), dev(ctx, R), ct.MPOSite(dev(ctx, W))
    op = ct.HeffOperator(tL, tR, tW, tW, ctx, prec=X6)
    for seed in (1, 2):
        th = dev(ctx, np.random.default_rng(seed).standard_normal((Dl, 2, 2, Dr)))
        ref = ct.apply_two_site_heff(th, tL, tR, tW, tW, ctx, prec=X6).to_numpy()
        L2, R2 = dev(ctx, rng.standard_normal((w, 40, 40))), dev(ctx, rng.standard_normal((w, 56, 56)))
        ct.apply_two_site_heff(dev(ctx, rng.standard_normal((40, 2, 2, 56))), L2, R2, tW, tW, ctx, prec=X6)
        assert np.array_equal(op.apply(th).to_numpy(), ref)
    op.close()


@pytest.mark.parametrize("shape", [(96, 2, 128, 8, 8), (33, 3, 20, 4, 4), (20, 2, 24, 40, 40)])
def test_env_updates_bf16x6(ctx, ora, shape):
    Dl, d, Dr, wL, wR = shape
    rng = np.random.default_rng(3 + Dl)
    A = rng.standard_normal((Dl, d, Dr))
    W = rng.standard_normal((wL, d, d, wR))
    L = rng.standard_normal((wL, Dl, Dl))
    R = rng.standard_normal((wR, Dr, Dr))
    gotL = ct.update_left_environment(dev(ctx, L), ct.MPSSite(dev(ctx, A)), ct.MPOSite(dev(ctx, W)), ctx, prec=X6)
    gotR = ct.update_right_environment(dev(ctx, R), ct.MPSSite(dev(ctx, A)), ct.MPOSite(dev(ctx, W)), ctx, prec=X6)
    assert rel(gotL.to_numpy(), ora.env_update_left(L, A, W)) < 2e-6
    assert rel(gotR.to_numpy(), ora.env_update_right(R, A, W)) < 2e-6


def test_lanczos_bf16x6_refined(ctx, ora):
    """6xBF16 Krylov matvecs with the FP64 refinement: the energy matches the
    FP64 Lanczos of the oracle like the FP32 form's does."""
    # a symmetric H_eff: L, R symmetric per channel, W symmetric in (s, s')
    rng = np.random.default_rng(21)
    Dl = Dr = 48
    w, d = 5, 2
    L = rng.standard_normal((w, Dl, Dl))
    L = 0.5 * (L + L.transpose(0, 2, 1))
    R = rng.standard_normal((w, Dr, Dr))
    R = 0.5 * (R + R.transpose(0, 2, 1))
    Wi = rng.standard_normal((w, d, d, w))
    Wi = 0.5 * (Wi + Wi.transpose(0, 2, 1, 3))
    Wj = rng.standard_normal((w, d, d, w))
    Wj = 0.5 * (Wj + Wj.transpose(0, 2, 1, 3))
    th = rng.standard_normal((Dl, d, d, Dr))
    E_o = ora.lanczos_two_site(th, L, R, Wi, Wj, max_iter=20, tol=1e-12)[0]
    E, _, info = ct.lanczos_two_site_optimize(dev(ctx, th), dev(ctx, L), dev(ctx, R), ct.MPOSite(dev(ctx, Wi)),
                                              ct.MPOSite(dev(ctx, Wj)), ctx, 20, 1e-12, prec=X6, refine=1,
                                              return_info=True)
    assert abs(E - E_o) / abs(E_o) < 1e-8


def test_dmrg_bf16x6_refined_vs_fp64(ctx):
    """A whole refined 6xBF16 run (L = 20 antiferromagnetic Heisenberg chain,
    chi = 64) ends within 1e-9 of the FP64 run, whose energy is checked
    against the oracle elsewhere."""
    mpo_t = ham.heisenberg_tensors(20, -1.0)
    mps_t = ct.random_mps_c_compatible_tensors(20, 2, 16, 42)
    out = {}
    for prec, refine in ((ct.PREC_FP64, 0), (X6, 2)):
        p = ct.DMRGParams(num_sweeps=6, chi_max=64, eps_trunc=1e-12, max_krylov_iter=20, krylov_tol=1e-12,
                          energy_tol=0.0, jacobi_max_rot=-1, prec=prec, refine=refine)
        E, _ = ct.dmrg_two_site(ctx, ct.mpo_from_tensors(ctx, mpo_t), ct.mps_from_tensors(ctx, mps_t), p)
        out[prec] = E
    assert abs(out[X6] - out[ct.PREC_FP64]) <= 1e-9 * abs(out[ct.PREC_FP64]), out
