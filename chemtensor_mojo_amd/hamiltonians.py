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
