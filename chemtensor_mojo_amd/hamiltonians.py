"""MPO builders (host-side input construction).

Site tensors have the reference layout W[wL, d_in, d_out, wR] with operator
element op[d_in][d_out] stored at d_in*d + d_out (src/state/mpo_state.mojo:14-19,
src/state/hamiltonians.mojo:454-458).  The builders produce FP64 numpy arrays
(``*_tensors``) and device MPOs (``create_*``).

Reference builders restated: Ising (hamiltonians.mojo:55-174), Heisenberg XXX
(:177-341), Heisenberg XXZ (:344-521).  Builders without a reference
counterpart (J1-J2 chain, random-bond Heisenberg) follow the same
operator-automaton construction and the same Pauli-unit conventions.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np

I2 = np.eye(2)
X = np.array([[0.0, 1.0], [1.0, 0.0]])
Z = np.array([[1.0, 0.0], [0.0, -1.0]])
SP = np.array([[0.0, 1.0], [0.0, 0.0]])  # S+ (hamiltonians.mojo:212-216)
SM = np.array([[0.0, 0.0], [1.0, 0.0]])  # S-


def _site_from_grid(grid, rows: Sequence[int], cols: Sequence[int], d: int = 2) -> np.ndarray:
    """W[wl, din, dout, wr] = grid[rows[wl]][cols[wr]][din, dout] (None = 0)."""
    W = np.zeros((len(rows), d, d, len(cols)))
    for a, r in enumerate(rows):
        for b, c in enumerate(cols):
            op = grid[r][c]
            if op is not None:
                W[a, :, :, b] = op
    return W


def _chain(grid, w: int, first_row: int, last_col: int, num_sites: int, bulk=None) -> list[np.ndarray]:
    """First site = row first_row, last site = column last_col, bulk = full grid
    (the reference's row-vector / matrix / column-vector layout).  ``bulk``
    optionally gives a per-site grid."""
    sites = []
    for i in range(num_sites):
        g = bulk(i) if bulk is not None else grid
        rows = [first_row] if i == 0 else list(range(w))
        cols = [last_col] if i == num_sites - 1 else list(range(w))
        sites.append(_site_from_grid(g, rows, cols))
    return sites


def ising_1d_tensors(num_sites: int, J: float = 1.0, h_longitudinal: float = 0.0,
                     g_transverse: float = 1.0) -> list[np.ndarray]:
    """H = -J sum Z_i Z_{i+1} - g sum X_i - h sum Z_i (hamiltonians.mojo:55-174), w = 3."""
    if num_sites < 2:
        raise ValueError("Ising model requires at least 2 sites")
    on = -h_longitudinal * Z - g_transverse * X
    grid = [[I2, None, None],
            [-J * Z, None, None],
            [on, Z, I2]]
    return _chain(grid, 3, 2, 0, num_sites)


def heisenberg_tensors(num_sites: int, J: float = 1.0) -> list[np.ndarray]:
    """H = J sum (XX + YY + ZZ) (antiferromagnet for J > 0), hamiltonians.mojo:177-341, w = 5."""
    if num_sites < 2:
        raise ValueError("Heisenberg model requires at least 2 sites")
    grid = [[I2, SP, SM, Z, None],
            [None, None, None, None, 2.0 * J * SM],
            [None, None, None, None, 2.0 * J * SP],
            [None, None, None, None, J * Z],
            [None, None, None, None, I2]]
    return _chain(grid, 5, 0, 4, num_sites)


def heisenberg_xxz_tensors(num_sites: int, J: float = 1.0, D: float = 1.0, h: float = 0.0) -> list[np.ndarray]:
    """H = -J sum (XX + YY + D ZZ) - h sum Z (hamiltonians.mojo:344-521), w = 5.

    Note the sign: J = 1, D = 1 is the ferromagnet with E0 = -(L-1)."""
    if num_sites < 2:
        raise ValueError("Heisenberg model requires at least 2 sites")
    grid = [[I2, None, None, None, None],
            [-2.0 * J * SM, None, None, None, None],
            [-2.0 * J * SP, None, None, None, None],
            [-J * Z, None, None, None, None],
            [-h * Z, SP, SM, D * Z, I2]]
    return _chain(grid, 5, 4, 0, num_sites)


def random_bond_heisenberg_tensors(couplings: Sequence[float]) -> list[np.ndarray]:
    """H = sum_i J_i (X_i X_{i+1} + Y_i Y_{i+1} + Z_i Z_{i+1}), one coupling per bond.

    Same w = 5 automaton as heisenberg_tensors; bond i's coupling sits in the
    receiving tensor W_{i+1} (rows 1-3 of its last column), so
    dW_{i+1}/dJ_i is the J = 1 pattern of those rows."""
    Js = [float(x) for x in couplings]
    n = len(Js) + 1
    if n < 2:
        raise ValueError("need at least one bond")

    def grid_at(i):
        Jb = Js[i - 1] if i >= 1 else 0.0
        return [[I2, SP, SM, Z, None],
                [None, None, None, None, 2.0 * Jb * SM],
                [None, None, None, None, 2.0 * Jb * SP],
                [None, None, None, None, Jb * Z],
                [None, None, None, None, I2]]

    return _chain(None, 5, 0, 4, n, bulk=grid_at)


def random_bond_coupling_derivative(num_sites: int, bond: int) -> tuple[int, np.ndarray]:
    """(site, dW_site/dJ_bond) for random_bond_heisenberg_tensors."""
    W = random_bond_heisenberg_tensors([0.0] * (num_sites - 1))[bond + 1]
    G = np.zeros_like(W)
    last = W.shape[3] - 1
    G[1, :, :, last] = 2.0 * SM
    G[2, :, :, last] = 2.0 * SP
    G[3, :, :, last] = Z
    return bond + 1, G


def j1j2_tensors(num_sites: int, J1: float = 1.0, J2: float = 0.5) -> list[np.ndarray]:
    """H = J1 sum S_i.S_{i+1} + J2 sum S_i.S_{i+2} in the Pauli units of
    heisenberg_tensors (S.S -> XX + YY + ZZ), open chain, w = 8.

    Channels: 0 start, 1-3 = S+,S-,Z placed, 4-6 = S+,S-,Z delayed by one
    site, 7 done.  No reference counterpart."""
    if num_sites < 3:
        raise ValueError("J1-J2 chain requires at least 3 sites")
    n = 8
    grid = [[None] * n for _ in range(n)]
    grid[0][0] = I2
    grid[0][1], grid[0][2], grid[0][3] = SP, SM, Z
    grid[1][4] = grid[2][5] = grid[3][6] = I2
    grid[1][7], grid[2][7], grid[3][7] = 2.0 * J1 * SM, 2.0 * J1 * SP, J1 * Z
    grid[4][7], grid[5][7], grid[6][7] = 2.0 * J2 * SM, 2.0 * J2 * SP, J2 * Z
    grid[7][7] = I2
    return _chain(grid, n, 0, 7, num_sites)


def mpo_to_dense(sites: Sequence[np.ndarray]) -> np.ndarray:
    """Full 2^L x 2^L matrix <out|H|in> of an MPO (test helper, small L);
    restates mpo_to_full_matrix (src/tests/state/mpo/mpo_test_helpers.mojo:23-119)
    with rows indexed by the bra (d_out) and columns by the ket (d_in)."""
    T = sites[0]  # [1, din, dout, w]
    M = T[0].transpose(1, 0, 2)  # [dout, din, w]
    for W in sites[1:]:
        # M[out, in, a] x W[a, din, dout, b] -> [out*dout, in*din, b]
        M = np.einsum("xya,aijb->xjyib", M, W)
        s = M.shape
        M = M.reshape(s[0] * s[1], s[2] * s[3], s[4])
    return M[:, :, 0]


# ------------------------------------------------ quantum chemistry (no reference counterpart)
def qc_integrals(n_orb: int, seed: int = 42, v_scale: float = 0.1) -> tuple[np.ndarray, np.ndarray]:
    """Random ab initio-shaped integrals over n_orb spatial orbitals: h_pq real
    symmetric N(0,1); v_pqrs = (pq|rs) N(0, v_scale) with the 8-fold
    permutational symmetry of real orbitals imposed by averaging."""
    rng = np.random.default_rng(seed)
    h = rng.standard_normal((n_orb, n_orb))
    h = 0.5 * (h + h.T)
    v = rng.standard_normal((n_orb,) * 4) * v_scale
    perms = [(0, 1, 2, 3), (1, 0, 2, 3), (0, 1, 3, 2), (1, 0, 3, 2),
             (2, 3, 0, 1), (3, 2, 0, 1), (2, 3, 1, 0), (3, 2, 1, 0)]
    v = sum(v.transpose(p) for p in perms) / 8.0
    return h, v


def _jw_annihilators(num_sites: int):
    """Sparse a_j = Z x .. x Z x sigma^- x I x .. x I (site 0 the most
    significant bit of the basis index; Z|1> = -|1>, sigma^- = |0><1|)."""
    import scipy.sparse as sp
    Zs = sp.csr_matrix(Z)
    Is = sp.identity(2, format="csr")
    Ms = sp.csr_matrix(np.array([[0.0, 1.0], [0.0, 0.0]]))
    ops = []
    for j in range(num_sites):
        m = sp.identity(1, format="csr")
        for i in range(num_sites):
            m = sp.kron(m, Zs if i < j else (Ms if i == j else Is), format="csr")
        ops.append(m)
    return ops


def qc_hamiltonian_sparse(h: np.ndarray, v: np.ndarray, penalty: float = 0.0, n_elec: int | None = None):
    """Second-quantised H = sum_pq h_pq E_pq + 1/2 sum_pqrs (pq|rs) (E_pq E_rs - delta_qr E_ps)
    (E_pq = sum_sigma a+_{p sigma} a_{q sigma}) on 2*n_orb spin-orbital sites
    (site 2p = p up, 2p+1 = p down) under the Jordan-Wigner mapping, plus
    penalty * (N - n_elec)^2 when penalty > 0 (n_elec defaults to half filling).
    Returns a scipy sparse matrix <out|H|in> of size 4^n_orb (test / small-system use)."""
    import scipy.sparse as sp
    n = h.shape[0]
    M = 2 * n
    a = _jw_annihilators(M)
    ad = [x.T.tocsr() for x in a]
    E = [[sum(ad[2 * p + s] @ a[2 * q + s] for s in (0, 1)) for q in range(n)] for p in range(n)]
    H = sp.csr_matrix((4 ** n, 4 ** n))
    for p in range(n):
        for q in range(n):
            H = H + h[p, q] * E[p][q]
    for p in range(n):
        for q in range(n):
            F = sp.csr_matrix((4 ** n, 4 ** n))
            for r in range(n):
                for s in range(n):
                    if v[p, q, r, s] != 0.0:
                        F = F + v[p, q, r, s] * E[r][s]
            H = H + 0.5 * (E[p][q] @ F)
    for p in range(n):
        for s in range(n):
            H = H - 0.5 * sum(v[p, q, q, s] for q in range(n)) * E[p][s]
    if penalty > 0.0:
        ne = n if n_elec is None else n_elec
        N = sum(ad[j] @ a[j] for j in range(M))
        Dn = N - ne * sp.identity(4 ** n, format="csr")
        H = H + penalty * (Dn @ Dn)
    return H.tocsr()


def mpo_from_dense(H: np.ndarray, num_sites: int, tol: float = 1e-12) -> list[np.ndarray]:
    """Exact MPO of a dense 2^L x 2^L operator <out|H|in> by successive SVDs
    (tensor-train decomposition of the operator with the site pair index
    d_in*2 + d_out), dropping singular values below tol * the largest at each
    cut.  Inverse of mpo_to_dense; small systems only."""
    L = num_sites
    T = np.asarray(H, dtype=np.float64).reshape([2] * (2 * L))
    order = []
    for k in range(L):
        order += [L + k, k]  # (in_k, out_k)
    C = T.transpose(order).reshape(4, -1)
    sites, r_prev = [], 1
    for k in range(L - 1):
        C = C.reshape(r_prev * 4, -1)
        U, S, Vt = np.linalg.svd(C, full_matrices=False)
        r = max(1, int((S > tol * S[0]).sum())) if S[0] > 0 else 1
        sites.append(U[:, :r].reshape(r_prev, 2, 2, r))
        C = S[:r, None] * Vt[:r]
        r_prev = r
    sites.append(C.reshape(r_prev, 2, 2, 1))
    return sites


def qc_tensors(n_orb: int, seed: int = 42, penalty: float = 10.0, n_elec: int | None = None,
               v_scale: float = 0.1) -> tuple[list[np.ndarray], dict]:
    """MPO tensors of the penalised quantum-chemistry Hamiltonian on 2*n_orb
    sites (exact, by SVD of the dense operator; n_orb <= 6) and its exact
    ground state: {"E_fci", "E_penalised", "n_elec"} with E_fci the lowest
    eigenvalue of H in the n_elec sector (half filling by default)."""
    import scipy.sparse as sp
    h, v = qc_integrals(n_orb, seed, v_scale)
    ne = n_orb if n_elec is None else n_elec
    Hs = qc_hamiltonian_sparse(h, v, penalty, ne)
    Hd = Hs.toarray()
    M = 2 * n_orb
    occ = np.array([bin(i).count("1") for i in range(2 ** M)])
    sector = np.where(occ == ne)[0]
    H0 = qc_hamiltonian_sparse(h, v, 0.0).toarray()
    e_fci = float(np.linalg.eigvalsh(H0[np.ix_(sector, sector)])[0])
    e_pen = float(np.linalg.eigvalsh(Hd)[0])
    return mpo_from_dense(Hd, M), {"E_fci": e_fci, "E_penalised": e_pen, "n_elec": ne, "h": h, "v": v}
