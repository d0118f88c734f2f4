"""Adjoint contractions and dE/dJ (new capability, no reference counterpart).

Pinned by numpy.einsum of the same networks, by the Hellmann-Feynman
derivative <psi|dH/dJ_b|psi> from exact diagonalisation, and by central
finite differences of DMRG energies (h = 1e-5, FP64)."""
import numpy as np
import pytest
import torch

import chemtensor_mojo_amd as ct
from chemtensor_mojo_amd import adjoint as adj
from chemtensor_mojo_amd import hamiltonians as ham

pytestmark = pytest.mark.gpu


def dev(ctx, a):
    return ct.create_dense_tensor_from_data(ctx, a, a.shape)


def test_env_adjoint_vs_einsum(ctx):
    rng = np.random.default_rng(0)
    Dl, d, Dr, wL, wR = 9, 2, 13, 5, 4
    L = rng.standard_normal((wL, Dl, Dl))
    A = rng.standard_normal((Dl, d, Dr))
    R = rng.standard_normal((wR, Dr, Dr))
    g = adj.env_adjoint(dev(ctx, L), dev(ctx, A), dev(ctx, R), wL, wR, ctx).cpu().numpy()
    ref = np.einsum("lab,asr,btc,mrc->lstm", L, A, A, R)
    assert np.allclose(g, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_heff_adjoint_vs_einsum(ctx):
    import ctypes
    from chemtensor_mojo_amd._lib import call
    rng = np.random.default_rng(1)
    Dl, d, Dr, wL, wM, wR = 6, 2, 7, 5, 3, 4
    L = rng.standard_normal((wL, Dl, Dl))
    R = rng.standard_normal((wR, Dr, Dr))
    Wi = rng.standard_normal((wL, d, d, wM))
    Wj = rng.standard_normal((wM, d, d, wR))
    th = rng.standard_normal((Dl, d, d, Dr))
    gi = torch.empty(wL * d * d * wM, dtype=torch.float64, device="cuda")
    gj = torch.empty(wM * d * d * wR, dtype=torch.float64, device="cuda")
    tL, tR, tWi, tWj, tth = (dev(ctx, x) for x in (L, R, Wi, Wj, th))
    call("ct_heff_adjoint", ctx.handle, tL.ptr(), tWi.ptr(), tWj.ptr(), tR.ptr(), tth.ptr(),
         ctypes.c_void_p(gi.data_ptr()), ctypes.c_void_p(gj.data_ptr()), Dl, d, Dr, wL, wM, wR)
    refi = np.einsum("lab,asrc,mrvn,ncd,buvd->lsum", L, th, Wj, R, th)
    refj = np.einsum("lab,asrc,lsum,ncd,buvd->mrvn", L, th, Wi, R, th)
    assert np.allclose(gi.cpu().numpy().reshape(refi.shape), refi, rtol=1e-11, atol=1e-11 * np.abs(refi).max())
    assert np.allclose(gj.cpu().numpy().reshape(refj.shape), refj, rtol=1e-11, atol=1e-11 * np.abs(refj).max())


def test_two_site_energy_autograd_vs_cpu_einsum(ctx):
    """TwoSiteEnergy: E and dE/d(theta, W_i, W_ip1) against torch autograd of the
    same network written as an einsum on the CPU (non-symmetric L, R, W, so the
    H^T theta term is exercised)."""
    rng = np.random.default_rng(11)
    Dl, d, Dr, wL, wM, wR = 5, 2, 6, 3, 4, 2
    L = rng.standard_normal((wL, Dl, Dl))
    R = rng.standard_normal((wR, Dr, Dr))
    Wi = rng.standard_normal((wL, d, d, wM))
    Wj = rng.standard_normal((wM, d, d, wR))
    th = rng.standard_normal((Dl, d, d, Dr))
    cpu = [torch.tensor(x, requires_grad=True) for x in (th, Wi, Wj)]
    Lc, Rc = torch.tensor(L), torch.tensor(R)
    num = torch.einsum("lab,asrc,lsum,mrvn,ncd,buvd->", Lc, cpu[0], cpu[1], cpu[2], Rc, cpu[0])
    E_ref = num / torch.sum(cpu[0] * cpu[0])
    E_ref.backward()
    gpu = [torch.tensor(x, device="cuda", requires_grad=True) for x in (th, Wi, Wj)]
    E = adj.two_site_energy(gpu[0], gpu[1], gpu[2], torch.tensor(L, device="cuda"), torch.tensor(R, device="cuda"), ctx)
    E.backward()
    assert abs(E.item() - E_ref.item()) <= 1e-12 * abs(E_ref.item())
    for g, c in zip(gpu, cpu):
        gg, gc = g.grad.cpu().numpy(), c.grad.numpy()
        assert np.allclose(gg, gc, rtol=1e-11, atol=1e-11 * np.abs(gc).max())
    with pytest.raises(ValueError):
        adj.two_site_energy(gpu[0], gpu[1], gpu[2], torch.tensor(L, device="cuda", requires_grad=True),
                            torch.tensor(R, device="cuda"), ctx)


def statevector(mps_np):
    psi = mps_np[0]
    for A in mps_np[1:]:
        psi = np.tensordot(psi, A, axes=([psi.ndim - 1], [0]))
    return psi.reshape(-1)


def solver_factory(ctx, L, chi, sweeps=6, prec=ct.PREC_FP64, refine=0):
    mps0 = ct.random_mps_c_compatible_tensors(L, 2, 8, 44)

    def solve(J):
        p = ct.DMRGParams(num_sweeps=sweeps, chi_max=chi, eps_trunc=1e-14, max_krylov_iter=20, krylov_tol=1e-14,
                          energy_tol=0.0, prec=prec, refine=refine, jacobi_max_rot=20000)
        E, mps = ct.dmrg_two_site(ctx, ct.mpo_from_tensors(ctx, ham.random_bond_heisenberg_tensors(J)),
                                  ct.mps_from_tensors(ctx, mps0), p)
        return E, mps

    return solve


def test_dEdJ_vs_hellmann_feynman_and_ed(ctx):
    L = 8
    rng = np.random.default_rng(5)
    J = rng.uniform(0.5, 1.5, L - 1)
    solve = solver_factory(ctx, L, 16)
    E, mps = solve(J)
    model = adj.random_bond_heisenberg_model(L)
    E_adj, dE = adj.coupling_gradient(ctx, mps, model, J)
    # ED ground state and exact derivative <psi0|dH/dJ_b|psi0>
    H = ham.mpo_to_dense(ham.random_bond_heisenberg_tensors(J))
    w, V = np.linalg.eigh(H)
    psi0 = V[:, 0]
    assert abs(E - w[0]) < 1e-9 * abs(w[0])
    assert abs(E_adj - w[0]) < 1e-9 * abs(w[0])
    psi = statevector(mps.to_numpy())
    for b in range(L - 1):
        e = np.zeros(L - 1)
        e[b] = 1.0
        dH = ham.mpo_to_dense(ham.random_bond_heisenberg_tensors(e))
        hf_dmrg = psi @ dH @ psi / (psi @ psi)
        hf_ed = psi0 @ dH @ psi0
        assert abs(dE[b] - hf_dmrg) < 1e-10 * max(1.0, abs(hf_dmrg))
        assert abs(dE[b] - hf_ed) < 1e-7 * max(1.0, abs(hf_ed))


def test_dEdJ_vs_central_differences(ctx):
    L = 8
    rng = np.random.default_rng(6)
    J = rng.uniform(0.5, 1.5, L - 1)
    solve = solver_factory(ctx, L, 16)
    E, mps = solve(J)
    _, dE = adj.coupling_gradient(ctx, mps, adj.random_bond_heisenberg_model(L), J)
    h = 1e-5
    for b in (0, 3, 6):
        Jp, Jm = J.copy(), J.copy()
        Jp[b] += h
        Jm[b] -= h
        fd = (solve(Jp)[0] - solve(Jm)[0]) / (2 * h)
        assert abs(fd - dE[b]) < 1e-6 * max(1.0, abs(dE[b])), (b, fd, dE[b])


def test_autograd_function(ctx):
    L = 6
    J = torch.tensor(np.random.default_rng(7).uniform(0.5, 1.5, L - 1), dtype=torch.float64, requires_grad=True)
    solve = solver_factory(ctx, L, 8)
    model = adj.random_bond_heisenberg_model(L)
    E = adj.dmrg_energy(J, solve, model, ctx)
    E.backward()
    _, mps = solve(J.detach().numpy())
    _, dE = adj.coupling_gradient(ctx, mps, model, J.detach().numpy())
    assert torch.allclose(J.grad, torch.from_numpy(dE), rtol=1e-9, atol=1e-12)
    # the energy is the ED ground state
    w0 = np.linalg.eigvalsh(ham.mpo_to_dense(ham.random_bond_heisenberg_tensors(J.detach().numpy())))[0]
    assert abs(E.item() - w0) < 1e-9 * abs(w0)
