"""The reference's call surface around the kernels, on the GPU through the C ABI:

- DenseTensor views (reshape / flatten_dims, dense_tensor.mojo:547-635);
- build_two_site_theta (dmrg.mojo:93-140) and build_right_environments
  (environments.mojo:420-472) against independent numpy contractions, the
  oracle's environment updates and the dense <psi|H|psi>;
- the float32 boundary: lanczos_two_site_optimize, dense_tensor_svd_trunc,
  split_two_site_theta_svd and dmrg_two_site on the reference's float32
  tensors.  Inside, these run in FP64 on the widened values, so a float32 call
  must give exactly the FP64 call's results on the same (float32-representable)
  inputs, rounded to float32 on the way out; the FP64 call is itself checked
  against the oracle;
- repeated DMRGRun.run calls that stop early on energy_tol, and the two split
  modes at a truncation threshold near the Gram form's rounding level.
"""
import numpy as np
import pytest
import torch

import chemtensor_mojo_amd as ct
from chemtensor_mojo_amd import hamiltonians as ham

pytestmark = pytest.mark.gpu


def dev(ctx, a, dtype=torch.float64):
    return ct.create_dense_tensor_from_data(ctx, a, a.shape, dtype)


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def f32(a):
    """The float32-representable neighbour of a (as float64)."""
    return np.asarray(a, dtype=np.float32).astype(np.float64)


# ------------------------------------------------------------------ views
def test_reshape_and_flatten_dims_are_views(ctx):
    rng = np.random.default_rng(0)
    X = rng.standard_normal((2, 3, 4, 5))
    t = dev(ctx, X)
    # the reference docstring's examples (dense_tensor.mojo:573-590)
    f = t.flatten_dims(1, 3, ctx)
    assert f.shape == [2, 12, 5] and f.stride == [60, 5, 1]
    assert f.storage.data_ptr() == t.storage.data_ptr()
    assert np.array_equal(f.to_numpy(), X.reshape(2, 12, 5))
    v = dev(ctx, rng.standard_normal((2, 3, 4))).flatten_dims(0, 3, ctx)
    assert v.shape == [24]
    b = dev(ctx, rng.standard_normal((8, 5, 10))).flatten_dims(1, 3, ctx)
    assert b.shape == [8, 50]
    r = t.reshape([6, 20])
    assert r.shape == [6, 20] and r.stride == [20, 1] and r.storage.data_ptr() == t.storage.data_ptr()
    assert np.array_equal(r.to_numpy(), X.reshape(6, 20))
    # a view shares storage: writing through it is seen by the original
    r.storage[7] = 42.0
    assert t.to_numpy().reshape(-1)[7] == 42.0
    # errors as the reference raises them (:597-598, :616-630)
    for bad in [(-1, 2), (2, 2), (3, 1), (0, 5)]:
        with pytest.raises(ValueError):
            t.flatten_dims(*bad, ctx)
    with pytest.raises(ValueError):
        t.reshape([])
    with pytest.raises(ValueError):
        t.reshape([120, 0])
    with pytest.raises(ValueError):
        t.reshape([7, 17])


# ----------------------------------------------------------- theta / envs
@pytest.mark.parametrize("dims,dt,tol", [((7, 2, 9, 5), torch.float64, 1e-14), ((33, 2, 64, 40), torch.float64, 1e-14),
                                         ((33, 2, 64, 40), torch.float32, 2e-6), ((16, 3, 24, 16), torch.float64, 1e-14)])
def test_build_two_site_theta(ctx, dims, dt, tol):
    Dl, d, Dm, Dr = dims
    rng = np.random.default_rng(Dl + Dm)
    Ai = rng.standard_normal((Dl, d, Dm))
    Aj = rng.standard_normal((Dm, d, Dr))
    th = ct.build_two_site_theta(ct.MPSSite(dev(ctx, Ai, dt)), ct.MPSSite(dev(ctx, Aj, dt)), ctx)
    assert th.shape == [Dl, d, d, Dr] and th.dtype == dt
    exp = np.einsum("asm,mtb->astb", Ai, Aj)
    assert rel(th.to_numpy(), exp) < tol
    with pytest.raises(ValueError):
        ct.build_two_site_theta(ct.MPSSite(dev(ctx, Ai, dt)), ct.MPSSite(dev(ctx, Aj[:-1], dt)), ctx)


@pytest.mark.parametrize("builder", ["xxz", "j1j2"])
def test_build_right_environments(ctx, ora, builder):
    L = 8
    mpo_t = ham.heisenberg_xxz_tensors(L, 1.0, 0.7, 0.2) if builder == "xxz" else ham.j1j2_tensors(L, 1.0, 0.5)
    mps_t = ct.random_mps_c_compatible_tensors(L, 2, 6, 11)
    mps = ct.mps_from_tensors(ctx, mps_t)
    mpo = ct.mpo_from_tensors(ctx, mpo_t)
    envs = ct.build_right_environments(mps, mpo, ctx)
    assert len(envs) == L + 1
    # R[N] = identity [w_N, D_N, D_N] (environments.mojo:441-461)
    assert np.array_equal(envs[L].to_numpy(), np.ones((1, 1, 1)))
    R = np.ones((1, 1, 1))
    for i in range(L - 1, -1, -1):
        R = ora.env_update_right(R, mps_t[i], mpo_t[i])
        assert list(envs[i].shape) == list(R.shape)
        assert rel(envs[i].to_numpy(), R) < 1e-13, i
    # the full chain is <psi|H|psi> of the (unnormalised) state vector
    psi = mps_t[0]
    for A in mps_t[1:]:
        psi = np.tensordot(psi, A, axes=([-1], [0]))
    psi = psi.reshape(-1)
    e = float(psi @ (ham.mpo_to_dense(mpo_t) @ psi))
    assert abs(envs[0].to_numpy().reshape(-1)[0] - e) <= 1e-12 * max(1.0, abs(e))


# --------------------------------------------------------- float32 boundary
def _sym_problem32(rng, Dl, Dr, w=5):
    d = 2
    L = rng.standard_normal((w, Dl, Dl))
    L = f32(0.5 * (L + L.transpose(0, 2, 1)))
    R = rng.standard_normal((w, Dr, Dr))
    R = f32(0.5 * (R + R.transpose(0, 2, 1)))
    Wi = rng.standard_normal((w, d, d, w))
    Wi = f32(0.5 * (Wi + Wi.transpose(0, 2, 1, 3)))
    Wj = rng.standard_normal((w, d, d, w))
    Wj = f32(0.5 * (Wj + Wj.transpose(0, 2, 1, 3)))
    th = f32(rng.standard_normal((Dl, d, d, Dr)))
    return L, R, Wi, Wj, th


@pytest.mark.parametrize("prec,refine", [(ct.PREC_FP64, 0), (ct.PREC_FP32, 2), (ct.PREC_BF16X3, 2)])
def test_lanczos_float32_boundary(ctx, ora, prec, refine):
    rng = np.random.default_rng(4)
    L, R, Wi, Wj, th = _sym_problem32(rng, 24, 30)
    res = {}
    for dt in (torch.float64, torch.float32):
        E, v, info = ct.lanczos_two_site_optimize(dev(ctx, th, dt), dev(ctx, L, dt), dev(ctx, R, dt),
                                                  ct.MPOSite(dev(ctx, Wi, dt)), ct.MPOSite(dev(ctx, Wj, dt)), ctx, 20,
                                                  1e-10, prec=prec, refine=refine, return_info=True)
        assert v.dtype == dt
        res[dt] = (E, v.storage.cpu().numpy(), info)
    (E64, v64, i64), (E32, v32, i32) = res[torch.float64], res[torch.float32]
    assert E32 == E64 and i32 == i64
    assert np.array_equal(v32, v64.astype(np.float32))
    if prec == ct.PREC_FP64:
        E_o, v_o, nK_o, _, _ = ora.lanczos_two_site(th, L, R, Wi, Wj, max_iter=20, tol=1e-10)
        assert i64["num_iter"] == nK_o
        assert abs(E64 - E_o) <= 1e-10 * max(1.0, abs(E_o))
        assert min(rel(v32, v_o.ravel()), rel(-v32, v_o.ravel())) < 1e-6


def test_svd_trunc_float32_boundary(ctx, ora):
    rng = np.random.default_rng(12)
    A = f32(rng.standard_normal((90, 70)) @ np.diag(np.exp(-0.12 * np.arange(70))))
    U64, S64, V64, c64, dw64 = ct.dense_tensor_svd_trunc(dev(ctx, A), ctx, 50, 1e-8, return_discarded=True)
    U32, S32, V32, c32, dw32 = ct.dense_tensor_svd_trunc(dev(ctx, A, torch.float32), ctx, 50, 1e-8,
                                                         return_discarded=True)
    assert c32 == c64 and dw32 == dw64 and U32.dtype == torch.float32
    for a, b in [(U32, U64), (S32, S64), (V32, V64)]:
        assert np.array_equal(a.storage.cpu().numpy(), b.storage.cpu().numpy().astype(np.float32))
    U_o, S_o, Vt_o, chi_o = ora.svd_trunc(A, 50, 1e-8)
    assert c64 == chi_o and rel(S64.to_numpy(), S_o) < 1e-12


@pytest.mark.parametrize("mode", [ct.SVD_GESVD, ct.SVD_GRAM])
@pytest.mark.parametrize("l2r", [True, False])
def test_theta_split_float32_boundary(ctx, mode, l2r):
    rng = np.random.default_rng(3 + mode)
    th = f32(rng.standard_normal((20, 2, 2, 18)) * np.exp(-0.2 * np.arange(18))[None, None, None, :])
    out = {}
    for dt in (torch.float64, torch.float32):
        Ai, Aj, dw = ct.split_two_site_theta_svd(dev(ctx, th, dt), 30, 1e-9, ctx, left_to_right=l2r, svd_mode=mode)
        assert Ai.tensor.dtype == dt and Aj.tensor.dtype == dt
        out[dt] = (Ai.tensor.storage.cpu().numpy(), Aj.tensor.storage.cpu().numpy(), dw, Ai.tensor.shape)
    a64, b64, dw64, s64 = out[torch.float64]
    a32, b32, dw32, s32 = out[torch.float32]
    assert s32 == s64 and dw32 == dw64
    assert np.array_equal(a32, a64.astype(np.float32)) and np.array_equal(b32, b64.astype(np.float32))


def test_dmrg_float32_boundary(ctx, ora):
    """dmrg_two_site on float32 MPS/MPO tensors (the reference's storage) runs
    the FP64 sweep on the widened values: the same sweep energies as the FP64
    call, bit for bit, and float32 sites equal to the FP64 sites rounded."""
    mpo_t = [f32(w) for w in ham.ising_1d_tensors(10, 1.0, 0.3, 0.9)]
    mps_t = [f32(a) for a in ct.random_mps_c_compatible_tensors(10, 2, 4, 7)]
    res = {}
    for dt in (torch.float64, torch.float32):
        p = ct.DMRGParams(num_sweeps=3, chi_max=6, eps_trunc=1e-12, max_krylov_iter=4, krylov_tol=1e-12,
                          energy_tol=0.0)
        E, out, st = ct.dmrg_two_site(ctx, ct.mpo_from_tensors(ctx, mpo_t, dt), ct.mps_from_tensors(ctx, mps_t, dt),
                                      p, return_stats=True)
        res[dt] = (E, st.sweep_energies, [s.tensor.storage.cpu().numpy() for s in out.sites],
                   [s.tensor.dtype for s in out.sites])
    E64, se64, s64, _ = res[torch.float64]
    E32, se32, s32, dts = res[torch.float32]
    assert E32 == E64 and se32 == se64
    assert all(d == torch.float32 for d in dts)
    assert all(np.array_equal(a, b.astype(np.float32)) for a, b in zip(s32, s64))
    ref = ora.dmrg_two_site(mps_t, mpo_t, 3, 6, 1e-12, 4, 1e-12, 0.0)
    assert np.allclose(se64, ref["sweep_E"], rtol=1e-10, atol=0)
    with pytest.raises(TypeError):
        ct.DMRGRun(ctx, ct.mpo_from_tensors(ctx, mpo_t, torch.float64), ct.mps_from_tensors(ctx, mps_t, torch.float32),
                   ct.DMRGParams(num_sweeps=1, chi_max=6, eps_trunc=1e-12, max_krylov_iter=4, krylov_tol=1e-12,
                                 energy_tol=0.0))


# ------------------------------------------------------ repeated runs / modes
def test_dmrg_run_twice_stops_early(ctx):
    """DMRGRun.run may be called again; a call that stops on energy_tol returns
    exactly the sweeps it ran (dmrg.mojo:466-474)."""
    mpo_t = ham.heisenberg_tensors(10, 1.0)
    mps_t = ct.random_mps_c_compatible_tensors(10, 2, 8, 43)
    p = ct.DMRGParams(num_sweeps=3, chi_max=32, eps_trunc=1e-12, max_krylov_iter=10, krylov_tol=1e-12,
                      energy_tol=1e-9, jacobi_max_rot=-1)
    run = ct.DMRGRun(ctx, ct.mpo_from_tensors(ctx, mpo_t), ct.mps_from_tensors(ctx, mps_t), p)
    try:
        E1, se1 = run.run(3)
        E2, se2 = run.run(6)
        st = run.stats()
    finally:
        run.close()
    assert 1 <= len(se1) <= 3 and se1[-1] == E1
    assert 2 <= len(se2) < 6 and se2[-1] == E2        # converged: stops after two sweeps of the second call
    assert all(np.isfinite(se2)) and all(e < -10.0 for e in se2)
    assert abs(se2[-1] - se2[-2]) < 1e-9
    assert st.sweeps == len(se1) + len(se2)


def test_split_modes_at_tiny_eps(ctx):
    """eps_trunc = 1e-13 with chi_max not binding: the Gram split's eigenvalues
    sigma^2 carry an absolute error ~ k * 1e-16 * sigma_max^2, so its chi may
    differ from the full SVD's by the few singular values inside that band,
    while both discarded weights stay within that band of each other and of
    the threshold."""
    rng = np.random.default_rng(5)
    Dl, d, Dr = 32, 2, 32
    k = Dl * d
    # sigma_i = 10^(-i/5): the discarded weight keeping c is ~10^(-0.4 c), so the rule
    # keeps 33 (6.3e-14 <= 1e-13 < 1.6e-13); the Gram noise level is ~1e-14
    U, _ = np.linalg.qr(rng.standard_normal((k, k)))
    V, _ = np.linalg.qr(rng.standard_normal((k, k)))
    sig = 10.0 ** (-np.arange(k) / 5.0)
    th = ((U * sig) @ V.T).reshape(Dl, d, d, Dr)
    out = {}
    for mode in (ct.SVD_GESVD, ct.SVD_GRAM):
        Ai, Aj, dw = ct.split_two_site_theta_svd(dev(ctx, th), 1000, 1e-13, ctx, svd_mode=mode)
        out[mode] = (Ai.right_bond_dim(), dw)
    (c_s, dw_s), (c_g, dw_g) = out[ct.SVD_GESVD], out[ct.SVD_GRAM]
    band = 4 * k * 1.1e-16 * sig[0] ** 2 / np.sum(sig ** 2)
    tail = np.cumsum((sig ** 2)[::-1])[::-1] / np.sum(sig ** 2)  # tail[c] = exact discarded weight keeping c
    assert c_s == int(np.argmax(tail <= 1e-13))                  # the full SVD resolves the rule exactly
    assert abs(dw_s - tail[c_s]) <= band
    assert abs(dw_g - dw_s) <= band + abs(tail[c_g] - tail[c_s])
    assert abs(c_g - c_s) <= 3
    assert dw_g <= 1e-13 + band
