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
    flat = [E[p][q].tocoo() for p in range(n) for q in range(n)]

    def lincomb(c):  # sum_pq c[p, q] E_pq, assembled once from coordinates
        c = np.asarray(c).ravel()
        nz = [k for k in range(c.size) if c[k] != 0.0]
        if not nz:
            return sp.csr_matrix((4 ** n, 4 ** n))
        return sp.coo_matrix((np.concatenate([c[k] * flat[k].data for k in nz]),
                              (np.concatenate([flat[k].row for k in nz]), np.concatenate([flat[k].col for k in nz]))),
                             shape=(4 ** n, 4 ** n)).tocsr()

    H = lincomb(h - 0.5 * np.einsum("pqqs->ps", v))
    for p in range(n):
        for q in range(n):
            H = H + 0.5 * (E[p][q] @ lincomb(v[p, q]))
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


# Local operator basis of the operator strings: I, Z, c+ = |1><0|, c = |0><1|
# (letters 0..3; |1> = occupied).  Every 2x2 matrix is a unique combination.
_ANN = np.array([[0.0, 1.0], [0.0, 0.0]])
_LETTERS = np.stack([I2, Z, _ANN.T, _ANN])


def _qc_ladder_terms(h: np.ndarray, v: np.ndarray, penalty: float, n_elec: int):
    """The Hamiltonian of qc_hamiltonian_sparse as a list of products of at
    most four ladder operators: (modes [T, 4], creator [T, 4], coef [T]);
    mode -1 pads shorter products.  Two-body part in normal order:
    1/2 sum (pq|rs) a+_{p s} a+_{r t} a_{s t} a_{q s}; penalty
    mu (N - n)^2 = mu [sum_i (1 - 2n) n_i + sum_{i != j} n_i n_j + n^2]."""
    n = h.shape[0]
    M = 2 * n
    modes, cre, coefs = [], [], []

    def add(mo, cr, c):
        modes.append(mo)
        cre.append(cr)
        coefs.append(c)

    p, q, s = (x.ravel() for x in np.meshgrid(np.arange(n), np.arange(n), np.arange(2), indexing="ij"))
    mo = np.full((p.size, 4), -1)
    mo[:, 0], mo[:, 1] = 2 * p + s, 2 * q + s
    cr = np.zeros((p.size, 4), bool)
    cr[:, 0] = True
    add(mo, cr, h[p, q])
    p, q, r, t, s1, s2 = (x.ravel() for x in np.meshgrid(*([np.arange(n)] * 4 + [np.arange(2)] * 2),
                                                        indexing="ij"))
    mo = np.stack([2 * p + s1, 2 * r + s2, 2 * t + s2, 2 * q + s1], 1)
    c = 0.5 * v[p, q, r, t]
    keep = (mo[:, 0] != mo[:, 1]) & (mo[:, 2] != mo[:, 3]) & (c != 0.0)
    cr = np.zeros((int(keep.sum()), 4), bool)
    cr[:, :2] = True
    add(mo[keep], cr, c[keep])
    if penalty > 0.0:
        i, j = (x.ravel() for x in np.meshgrid(np.arange(M), np.arange(M), indexing="ij"))
        off = i != j
        mo = np.stack([i[off], i[off], j[off], j[off]], 1)
        cr = np.zeros(mo.shape, bool)
        cr[:, 0] = cr[:, 2] = True
        add(mo, cr, np.full(mo.shape[0], penalty))
        mo = np.full((M, 4), -1)
        mo[:, 0] = mo[:, 1] = np.arange(M)
        cr = np.zeros(mo.shape, bool)
        cr[:, 0] = True
        add(mo, cr, np.full(M, penalty * (1.0 - 2.0 * n_elec)))
        add(np.full((1, 4), -1), np.zeros((1, 4), bool), np.array([penalty * n_elec * n_elec]))
    return np.concatenate(modes), np.concatenate(cre), np.concatenate(coefs)


def _jw_strings(modes: np.ndarray, cre: np.ndarray, coefs: np.ndarray, num_sites: int):
    """Jordan-Wigner images of ladder-operator products as operator strings.
    a_j = Z x .. x Z x c x I x .. x I, so a product is a tensor product whose
    site-i factor is the ordered product of Z (operator mode > i), c or c+
    (mode == i) and I (mode < i).  Each factor is expanded in the letter basis
    (n and 1-n take two letters, so a term yields up to 2^k strings).  A
    string is a uint64 key, 2 bits per site, site 0 most significant.
    Returns (sorted unique keys, summed coefficients), zeros dropped."""
    L = num_sites
    if L > 32:
        raise ValueError("operator strings hold at most 32 sites")
    T = coefs.size
    lead = np.zeros((T, L), np.int64)
    lead_c = np.zeros((T, L))
    alt = np.full((T, L), -1, np.int64)
    alt_c = np.zeros((T, L))
    rows = np.arange(T)
    for i in range(L):
        m = np.broadcast_to(I2, (T, 2, 2)).copy()
        for k in range(modes.shape[1]):
            ladder = np.where(cre[:, k][:, None, None], _ANN.T, _ANN)
            f = np.where((modes[:, k] > i)[:, None, None], Z,
                         np.where((modes[:, k] == i)[:, None, None], ladder, I2))
            m = m @ f
        C = np.stack([0.5 * (m[:, 0, 0] + m[:, 1, 1]), 0.5 * (m[:, 0, 0] - m[:, 1, 1]), m[:, 1, 0], m[:, 0, 1]], 1)
        nz = C != 0.0
        cnt = nz.sum(1)
        if cnt.max() > 2:
            raise AssertionError("a product of ladder operators on one site spans more than two letters")
        first = np.argmax(nz, 1)
        lead[:, i], lead_c[:, i] = first, C[rows, first]
        two = cnt == 2
        second = 3 - np.argmax(nz[:, ::-1], 1)
        alt[two, i], alt_c[two, i] = second[two], C[rows[two], second[two]]
    alive = np.all(lead_c != 0.0, 1)
    lead, lead_c, alt, alt_c, coefs = lead[alive], lead_c[alive], alt[alive], alt_c[alive], coefs[alive]
    shifts = (2 * (L - 1 - np.arange(L))).astype(np.uint64)
    ndouble = (alt >= 0).sum(1)
    keys_out, c_out = [], []
    for nd in range(int(ndouble.max(initial=0)) + 1):
        sel = ndouble == nd
        if not sel.any():
            continue
        ld, lc, al, ac, co = lead[sel], lead_c[sel], alt[sel], alt_c[sel], coefs[sel]
        r = np.arange(co.size)
        dpos = np.argsort(al < 0, axis=1, kind="stable")[:, :nd]
        for combo in range(1 << nd):
            let, cc = ld.copy(), lc.copy()
            for b in range(nd):
                if combo >> b & 1:
                    col = dpos[:, b]
                    let[r, col], cc[r, col] = al[r, col], ac[r, col]
            keys_out.append(np.bitwise_or.reduce(let.astype(np.uint64) << shifts[None, :], axis=1))
            c_out.append(co * np.prod(cc, 1))
    keys, inv = np.unique(np.concatenate(keys_out), return_inverse=True)
    cs = np.bincount(inv.ravel(), weights=np.concatenate(c_out), minlength=keys.size)
    nz = cs != 0.0
    return keys[nz], cs[nz]


def mpo_from_strings(keys: np.ndarray, coefs: np.ndarray, num_sites: int, tol: float = 1e-12) -> list[np.ndarray]:
    """Exact MPO of H = sum_s coefs[s] * string_s (keys as from _jw_strings)
    by the tensor-train SVD of H over the letter basis, run on the sparse
    representation: at site k the remaining operator is a dense
    [w * 4] x [distinct suffixes] matrix (row = left bond state, letter at
    site k), factored by QR of its transpose and an SVD of the triangular
    factor; singular values below tol * the largest are dropped.  The bond
    dimensions are the operator-Schmidt ranks of H, as for mpo_from_dense,
    without forming the 4^L operator."""
    L = num_sites
    R = coefs[None, :].astype(np.float64)
    suffix = keys.astype(np.uint64)
    factors = []
    for k in range(L):
        sh = np.uint64(2 * (L - 1 - k))
        letter = ((suffix >> sh) & np.uint64(3)).astype(np.int64)
        rest = suffix & ((np.uint64(1) << sh) - np.uint64(1)) if k < L - 1 else np.zeros_like(suffix)
        urest, idx = np.unique(rest, return_inverse=True)
        w = R.shape[0]
        Mk = np.zeros((w, 4, urest.size))
        Mk[:, letter, idx.ravel()] = R
        if k == L - 1:
            factors.append(Mk.reshape(w, 4, 1))
            break
        Mk = Mk.reshape(w * 4, urest.size)
        if Mk.shape[1] > Mk.shape[0]:
            U, S, _ = np.linalg.svd(np.linalg.qr(Mk.T, mode="r").T, full_matrices=False)
        else:
            U, S, _ = np.linalg.svd(Mk, full_matrices=False)
        rank = max(1, int((S > tol * S[0]).sum())) if S[0] > 0 else 1
        U = U[:, :rank]
        factors.append(U.reshape(w, 4, rank))
        R = U.T @ Mk
        suffix = urest
    # W[a, din, dout, b] = sum_letter F[a, letter, b] * op_letter[dout, din]
    return [np.ascontiguousarray(np.einsum("aob,oij->ajib", F, _LETTERS)) for F in factors]


def qc_mpo_tensors(h: np.ndarray, v: np.ndarray, penalty: float = 0.0, n_elec: int | None = None,
                   tol: float = 1e-12) -> list[np.ndarray]:
    """MPO of the operator of qc_hamiltonian_sparse on 2*n_orb sites (n_orb <= 16)
    from its Jordan-Wigner operator strings, without the dense operator."""
    n = h.shape[0]
    ne = n if n_elec is None else n_elec
    keys, cs = _jw_strings(*_qc_ladder_terms(h, v, penalty, ne), 2 * n)
    return mpo_from_strings(keys, cs, 2 * n, tol)


def qc_tensors(n_orb: int, seed: int = 42, penalty: float = 10.0, n_elec: int | None = None,
               v_scale: float = 0.1, exact: bool | None = None) -> tuple[list[np.ndarray], dict]:
    """MPO tensors of the penalised quantum-chemistry Hamiltonian on 2*n_orb
    sites (qc_mpo_tensors) and, when exact (default n_orb <= 6), full CI:
    "E_fci", the lowest eigenvalue of H in the n_elec sector (half filling by
    default), and for n_orb <= 6 "E_penalised", the lowest eigenvalue of the
    penalised operator over all sectors."""
    h, v = qc_integrals(n_orb, seed, v_scale)
    ne = n_orb if n_elec is None else n_elec
    sites = qc_mpo_tensors(h, v, penalty, ne)
    info = {"n_elec": ne, "h": h, "v": v}
    if exact is None:
        exact = n_orb <= 6
    if exact:
        M = 2 * n_orb
        occ = np.array([bin(i).count("1") for i in range(2 ** M)])
        sector = np.where(occ == ne)[0]
        Hs = qc_hamiltonian_sparse(h, v, 0.0).tocsr()[sector][:, sector]
        if sector.size <= 4000:
            info["E_fci"] = float(np.linalg.eigvalsh(Hs.toarray())[0])
        else:
            from scipy.sparse.linalg import eigsh
            info["E_fci"] = float(eigsh(Hs, k=1, which="SA", tol=1e-14)[0][0])
        if n_orb <= 6:
            info["E_penalised"] = float(np.linalg.eigvalsh(qc_hamiltonian_sparse(h, v, penalty, ne).toarray())[0])
    return sites, info


def number_sector_random_mps_tensors(num_sites: int, n_elec: int, D: int, seed: int = 42) -> list[np.ndarray]:
    """Random MPS [Dl, 2, Dr] of definite particle number n_elec: every bond
    index carries a left electron count q (0 <= q <= n_elec, reachable from both
    ends), the bond's indices cycle through the reachable counts nearest the
    mean filling, and A[l, s, r] ~ N(0,1) only where q_r = q_l + s.  The state
    lies entirely in the n_elec sector; N-conserving Hamiltonians keep it there."""
    rng = np.random.default_rng(seed)
    L = num_sites
    charges = []
    for k in range(L + 1):
        lo, hi = max(0, n_elec - (L - k)), min(k, n_elec)
        feasible = sorted(range(lo, hi + 1), key=lambda q: (abs(q - k * n_elec / L), q))
        Dk = min(D, 2 ** min(k, L - k))
        charges.append(np.array([feasible[i % len(feasible)] for i in range(Dk)]))
    sites = []
    for k in range(L):
        ql, qr = charges[k], charges[k + 1]
        A = np.zeros((ql.size, 2, qr.size))
        for s in (0, 1):
            mask = (ql[:, None] + s) == qr[None, :]
            A[:, s, :] = np.where(mask, rng.standard_normal(mask.shape), 0.0)
        sites.append(A / max(np.linalg.norm(A), 1e-300))
    return sites
