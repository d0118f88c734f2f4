"""Two-site DMRG sweeps with local dimension d > 2 (spin-S Heisenberg chains,
d = 2S + 1): the path reads d from the MPS and MPO tensors (dmrg.mojo:310-372
contracts theta[Dl, d, d, Dr] with W[wL, d, d, wR]; the split reshapes theta to
(Dl d) x (d Dr), dmrg.mojo:143-251), and the reference's chains are all d = 2,
so these runs hold the d > 2 shapes against the FP64 oracle and exact
diagonalisation.

H = J sum_i S_i . S_{i+1} as an MPO of bond dimension 5 built in the reference's
lower-triangular layout (the one hamiltonians.heisenberg_tensors uses for
S = 1/2).  Its normalisation is pinned by the two-site spectrum,
E = J/2 [S_tot(S_tot+1) - 2 S(S+1)] for S_tot = 0..2S, checked on the CPU.
"""
import numpy as np
import pytest
import torch

import chemtensor_mojo_amd as ct
from chemtensor_mojo_amd import hamiltonians as ham


def spin_ops(S):
    """S^z, S^+ in the basis m = S, S-1, ..., -S (index k = S - m)."""
    d = int(round(2 * S)) + 1
    m = S - np.arange(d)
    Sz = np.diag(m)
    Sp = np.zeros((d, d))
    for k in range(1, d):  # S^+ |m_k> = sqrt(S(S+1) - m_k(m_k+1)) |m_k + 1>, and m_k + 1 = m_{k-1}
        Sp[k - 1, k] = np.sqrt(S * (S + 1) - m[k] * (m[k] + 1))
    return Sz, Sp


def spin_heisenberg_tensors(num_sites, S, J=1.0):
    Sz, Sp = spin_ops(S)
    Sm, I = Sp.T.copy(), np.eye(Sz.shape[0])
    grid = [[I, Sp, Sm, Sz, None],
            [None, None, None, None, 0.5 * J * Sm],
            [None, None, None, None, 0.5 * J * Sp],
            [None, None, None, None, J * Sz],
            [None, None, None, None, I]]
    d = Sz.shape[0]
    out = []
    for i in range(num_sites):
        rows = [0] if i == 0 else list(range(5))
        cols = [4] if i == num_sites - 1 else list(range(5))
        out.append(ham._site_from_grid(grid, rows, cols, d=d))
    return out


@pytest.mark.parametrize("S", [1.0, 1.5])
def test_spin_operators_and_two_site_spectrum(S):
    Sz, Sp = spin_ops(S)
    Sm = Sp.T
    d = Sz.shape[0]
    # [S+, S-] = 2 Sz, S^2 = S(S+1)
    assert np.allclose(Sp @ Sm - Sm @ Sp, 2 * Sz, atol=1e-14)
    Sx, Sy2 = (Sp + Sm) / 2, -((Sp - Sm) / 2) @ ((Sp - Sm) / 2)
    assert np.allclose(Sx @ Sx + Sy2 + Sz @ Sz, S * (S + 1) * np.eye(d), atol=1e-13)
    H2 = ham.mpo_to_dense(spin_heisenberg_tensors(2, S, J=1.0))
    ev = np.linalg.eigvalsh(H2)
    expect = np.sort(np.concatenate([np.full(2 * St + 1, 0.5 * (St * (St + 1) - 2 * S * (S + 1)))
                                     for St in range(int(round(2 * S)) + 1)]))
    assert np.allclose(ev, expect, atol=1e-12)


CASES = {
    # S: (L, start bond dim, chi_max = the exact bond dimension d^(L/2))
    1.0: (8, 9, 81),
    1.5: (6, 8, 64),
}


@pytest.fixture(scope="module")
def chains(ora):
    out = {}
    for S, (L, D0, chi) in CASES.items():
        d = int(round(2 * S)) + 1
        mpo_t = spin_heisenberg_tensors(L, S)
        mps_t = ct.random_mps_c_compatible_tensors(L, d, D0, 42)
        E0 = float(np.linalg.eigvalsh(ham.mpo_to_dense(mpo_t))[0])
        E_o = ora.dmrg_two_site(mps_t, mpo_t, 4, chi, 1e-12, 20, 1e-12, 0.0, jacobi_max_rot=20000)["E"]
        out[S] = (L, d, chi, mpo_t, mps_t, E0, E_o)
    return out


def _run(ctx, mpo_t, mps_t, dtype, **kw):
    p = ct.DMRGParams(num_sweeps=4, eps_trunc=1e-12, max_krylov_iter=20, krylov_tol=1e-12, energy_tol=0.0,
                      jacobi_max_rot=-1, **kw)
    return ct.dmrg_two_site(ctx, ct.mpo_from_tensors(ctx, mpo_t, dtype=dtype), ct.mps_from_tensors(ctx, mps_t,
                                                                                                   dtype=dtype), p)


@pytest.mark.gpu
@pytest.mark.parametrize("S", [1.0, 1.5])
@pytest.mark.parametrize("svd_mode", [ct.SVD_GRAM, ct.SVD_GESVD])
def test_spin_chain_fp64_vs_oracle_and_ed(ctx, chains, S, svd_mode):
    L, d, chi, mpo_t, mps_t, E0, E_o = chains[S]
    E, mps = _run(ctx, mpo_t, mps_t, torch.float64, chi_max=chi, svd_mode=svd_mode)
    dims = [s.tensor.shape for s in mps.sites]
    print(f"spin-{S} L={L} (d={d}) svd_mode {svd_mode}: E {E:.14f} oracle {E_o:.14f} ED {E0:.14f} "
          f"bonds {[sh[2] for sh in dims[:-1]]}")
    assert all(sh[1] == d for sh in dims)
    assert abs(E - E_o) <= 1e-10 * abs(E_o)
    assert abs(E - E0) <= 1e-10 * abs(E0)


def _ed(mpo_t, dtype):
    """Exact E0 of the MPO as stored: float32 storage rounds sqrt(2) (spin 1)
    and sqrt(3) (spin 3/2), which moves E0 by 2.3e-8 and 1.4e-8 relative."""
    if dtype == torch.float32:
        mpo_t = [w.astype(np.float32).astype(np.float64) for w in mpo_t]
    return float(np.linalg.eigvalsh(ham.mpo_to_dense(mpo_t))[0])


@pytest.mark.gpu
@pytest.mark.parametrize("S", [1.0, 1.5])
@pytest.mark.parametrize("dtype,prec,refine", [(torch.float32, ct.PREC_FP64, 0), (torch.float32, ct.PREC_FP32, 1),
                                               (torch.float64, ct.PREC_BF16X3, 1)])
def test_spin_chain_mixed_forms(ctx, chains, S, dtype, prec, refine):
    """float32 storage and the FP32 / 3xBF16 matvecs with the FP64 Rayleigh
    quotient of each Ritz vector, against the exact E0 of the stored MPO to
    1e-9 (observed at spin 1: 1.1e-12 with the FP64 matvec, 2.9e-11 with the
    FP32 and 3xBF16 matvecs; at spin 3/2 below 5e-15)."""
    L, d, chi, mpo_t, mps_t, _, _ = chains[S]
    E0 = _ed(mpo_t, dtype)
    E, _ = _run(ctx, mpo_t, mps_t, dtype, chi_max=chi, prec=prec, refine=refine)
    print(f"spin-{S} L={L} ({dtype}, prec {prec}, refine {refine}): E {E:.14f} ED of the stored MPO {E0:.14f} "
          f"rel {abs(E - E0) / abs(E0):.1e}")
    assert abs(E - E0) <= 1e-9 * abs(E0)
