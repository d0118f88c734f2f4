"""Quantum-chemistry Hamiltonian and its MPO (no reference counterpart).

The Jordan-Wigner construction is checked against an independent
occupation-number implementation that applies each creation / annihilation
operator with the explicit fermionic sign (-1)^(number of occupied modes
before it); the MPO against the dense operator it was built from; the
Hamiltonian against particle-number conservation and the integral symmetry.
"""
import itertools

import numpy as np
import pytest

from chemtensor_mojo_amd import hamiltonians as ham


def _apply(op, j, state, M):
    """op 'a' or 'c' (annihilate / create) on mode j of basis index state
    (site 0 = most significant bit); returns (sign, new_state) or None."""
    bit = 1 << (M - 1 - j)
    occ = (state & bit) != 0
    if (op == "a") != occ:
        return None
    before = bin(state >> (M - j)).count("1") if j > 0 else 0
    return (-1) ** before, state ^ bit


def _term_matrix(ops, M):
    """Matrix of a product of ladder operators, ops = [(kind, mode), ...]
    applied right to left."""
    dim = 2 ** M
    out = np.zeros((dim, dim))
    for s in range(dim):
        amp, cur = 1.0, s
        for kind, j in reversed(ops):
            r = _apply(kind, j, cur, M)
            if r is None:
                amp = 0.0
                break
            amp *= r[0]
            cur = r[1]
        if amp != 0.0:
            out[cur, s] += amp
    return out


def _brute_force_h(h, v):
    n = h.shape[0]
    M = 2 * n
    H = np.zeros((2 ** M, 2 ** M))
    for p, q in itertools.product(range(n), repeat=2):
        for s in (0, 1):
            H += h[p, q] * _term_matrix([("c", 2 * p + s), ("a", 2 * q + s)], M)
    for p, q, r, t in itertools.product(range(n), repeat=4):
        if v[p, q, r, t] == 0.0:
            continue
        for s1, s2 in itertools.product((0, 1), repeat=2):
            # a+_{p s1} a+_{r s2} a_{t s2} a_{q s1}
            H += 0.5 * v[p, q, r, t] * _term_matrix(
                [("c", 2 * p + s1), ("c", 2 * r + s2), ("a", 2 * t + s2), ("a", 2 * q + s1)], M)
    return H


def test_integrals_symmetry():
    h, v = ham.qc_integrals(4, seed=3)
    assert np.allclose(h, h.T)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.allclose(v, v.transpose(perm))


@pytest.mark.parametrize("n_orb,seed", [(1, 5), (2, 11), (2, 12)])
def test_jordan_wigner_matches_occupation_number_rules(n_orb, seed):
    h, v = ham.qc_integrals(n_orb, seed=seed)
    H = ham.qc_hamiltonian_sparse(h, v).toarray()
    assert np.abs(H - _brute_force_h(h, v)).max() < 1e-12
    assert np.abs(H - H.T).max() < 1e-12


def test_particle_number_conserved_and_penalty():
    h, v = ham.qc_integrals(3, seed=8)
    H = ham.qc_hamiltonian_sparse(h, v).toarray()
    M = 6
    N = np.diag([bin(s).count("1") for s in range(2 ** M)]).astype(float)
    assert np.abs(H @ N - N @ H).max() < 1e-12
    Hp = ham.qc_hamiltonian_sparse(h, v, penalty=2.0, n_elec=3).toarray()
    assert np.abs(Hp - (H + 2.0 * (N - 3 * np.eye(2 ** M)) @ (N - 3 * np.eye(2 ** M)))).max() < 1e-12


@pytest.mark.parametrize("n_orb", [2, 3, 4])
def test_qc_mpo_round_trip_and_fci(n_orb):
    sites, info = ham.qc_tensors(n_orb, seed=13)
    H = ham.qc_hamiltonian_sparse(info["h"], info["v"], 10.0, n_orb).toarray()
    assert np.abs(ham.mpo_to_dense(sites) - H).max() < 1e-10 * max(1.0, np.abs(H).max())
    assert sites[0].shape[0] == 1 and sites[-1].shape[3] == 1
    # the penalised ground state is the half-filling ground state
    assert abs(info["E_penalised"] - info["E_fci"]) < 1e-10 * max(1.0, abs(info["E_fci"]))


def _mpo_apply(sites, x):
    """H x for an MPO and a state vector (site 0 the most significant bit)."""
    T = x.reshape([1, -1])  # [bond, remaining input bits then finished output bits]
    for k, W in enumerate(sites):
        w = T.shape[0]
        T = T.reshape(w, 2, -1)  # [a, din_k, rest]
        T = np.einsum("air,aijb->bjr", T, W)  # [b, dout_k, rest]
        T = T.reshape(W.shape[3], -1)
        # move the finished output bit to the back so the next site leads
        T = T.reshape(W.shape[3], 2, -1).transpose(0, 2, 1).reshape(W.shape[3], -1)
    return T.reshape(-1)


@pytest.mark.parametrize("n_orb,penalty", [(1, 0.0), (2, 0.0), (3, 1.5), (4, 0.7)])
def test_string_mpo_matches_sparse_operator(n_orb, penalty):
    h, v = ham.qc_integrals(n_orb, seed=30 + n_orb)
    sites = ham.qc_mpo_tensors(h, v, penalty, n_orb)
    H = ham.qc_hamiltonian_sparse(h, v, penalty, n_orb).toarray()
    assert np.abs(ham.mpo_to_dense(sites) - H).max() < 1e-12 * max(1.0, np.abs(H).max())


def test_string_mpo_has_optimal_bond_dimensions():
    """The sparse tensor-train SVD finds the operator-Schmidt ranks, i.e. the
    bond dimensions of the SVD of the dense operator."""
    h, v = ham.qc_integrals(5, seed=4)
    sites = ham.qc_mpo_tensors(h, v, 1.0, 5)
    dense = ham.mpo_from_dense(ham.qc_hamiltonian_sparse(h, v, 1.0, 5).toarray(), 10)
    assert [W.shape[3] for W in sites] == [W.shape[3] for W in dense]


def test_string_mpo_apply_matches_sparse_six_orbitals():
    h, v = ham.qc_integrals(6, seed=9)
    sites = ham.qc_mpo_tensors(h, v, 1.0, 6)
    H = ham.qc_hamiltonian_sparse(h, v, 1.0, 6)
    rng = np.random.default_rng(0)
    for _ in range(2):
        x = rng.standard_normal(4 ** 6)
        ref = H @ x
        assert np.abs(_mpo_apply(sites, x) - ref).max() < 1e-11 * np.abs(ref).max()


def test_string_mpo_center_rank_is_complementary_operator_count():
    """For generic integrals the central operator-Schmidt rank over k = 2n/2
    spin-orbitals per side is 1 (identity) + 1 (complete left block)
    + 2k (single ladder operators) + 2k (their right complements)
    + k(k-1) (pairs a+a+ and aa) + k^2 (pairs a+a) (Chan et al.,
    J. Chem. Phys. 145, 014102 (2016))."""
    n = 8
    h, v = ham.qc_integrals(n, seed=2)
    sites = ham.qc_mpo_tensors(h, v, 1.0, n)
    k = n
    assert sites[k - 1].shape[3] == 2 + 4 * k + k * (k - 1) + k * k
    assert [W.shape[3] for W in sites[:-1]] == [W.shape[3] for W in sites[:-1]][::-1]


def test_penalty_only_mpo_rank():
    """mu (N - n)^2 alone: left operators I, N_L, N_L^2 -> rank <= 3."""
    n = 4
    sites = ham.qc_mpo_tensors(np.zeros((n, n)), np.zeros((n,) * 4), 2.0, 3)
    assert max(W.shape[3] for W in sites[:-1]) <= 3
    M = 2 * n
    N = np.diag([bin(s).count("1") for s in range(2 ** M)]).astype(float)
    ref = 2.0 * (N - 3 * np.eye(2 ** M)) @ (N - 3 * np.eye(2 ** M))
    assert np.abs(ham.mpo_to_dense(sites) - ref).max() < 1e-12 * np.abs(ref).max()


def test_number_sector_random_mps():
    L, ne = 10, 5
    sites = ham.number_sector_random_mps_tensors(L, ne, 8, seed=3)
    psi = sites[0]
    for A in sites[1:]:
        psi = np.tensordot(psi, A, axes=([-1], [0]))
    psi = psi.reshape(-1)
    occ = np.array([bin(i).count("1") for i in range(2 ** L)])
    assert np.linalg.norm(psi) > 0
    assert np.abs(psi[occ != ne]).max() == 0.0
    assert max(A.shape[2] for A in sites) == 8
