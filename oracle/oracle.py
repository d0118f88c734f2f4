"""numpy/ctypes front end of the FP64 CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this module; the product package never does.  Every
routine forwards to oracle/liboracle.so (dmrg_oracle.c), which restates the
reference algorithm in FP64 and cites the reference file:line of each step.

Parity pins (checked by tests/test_oracle_*.py):
  * PCG32 / seed_rng_state / randnf  - published pcg32 demo outputs, and the
    Box-Muller moments; rng_c_compat.mojo:14-231.
  * tensordot                        - the reference's brute-force test oracle
    (test_dense_tensor_dot.mojo:25-106) on the DT-DOT cases, integer KATs of
    test_c_replica_dot.mojo, numpy.tensordot.
  * heff_apply / env updates         - numpy.einsum of the tensor network, the
    projection identity H_eff = P^T H P against a dense Hamiltonian, identity
    MPO invariants.
  * jacobi_smallest / lanczos        - numpy.linalg.eigh on the same matrices.
  * svd_trunc                        - sigma=[3,2] KAT of test_svd_lapack.mojo,
    reconstruction, truncation rule brute force.
  * dmrg_two_site                    - exact diagonalisation and closed forms
    (E0=-3 for the 4-site reference XXZ, E0=-(L-1) ferromagnet).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i64p = ctypes.POINTER(ctypes.c_int64)
_dp = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.ora_pcg32_random.restype = ctypes.c_uint32
        L.ora_rand_uint32.restype = ctypes.c_uint32
        L.ora_rand_uint64.restype = ctypes.c_uint64
        L.ora_rand_interval.restype = ctypes.c_uint64
        L.ora_rand_interval.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        L.ora_randnf.restype = ctypes.c_float
        L.ora_pcg32_seed.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64]
        L.ora_seed_rng_state.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        L.ora_random_mps_dims.restype = ctypes.c_int64
        L.ora_random_mps_dims.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, _i64p]
        L.ora_random_mps_fill.restype = ctypes.c_int64
        L.ora_random_mps_fill.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, _dp]
        L.ora_jacobi_smallest.restype = ctypes.c_double
        L.ora_jacobi_smallest.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]
        L.ora_svd_trunc.argtypes = [_dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
                                    _dp, _dp, _dp]
        L.ora_set_num_threads.argtypes = [ctypes.c_int]
        L.ora_get_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _i64(seq) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(seq, dtype=np.int64))


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def set_num_threads(n: int) -> None:
    lib().ora_set_num_threads(int(n))


def get_num_threads() -> int:
    return int(lib().ora_get_num_threads())


# ---------------------------------------------------------------- RNG
class _Pcg32(ctypes.Structure):
    _fields_ = [("state", ctypes.c_uint64), ("inc", ctypes.c_uint64)]


class _Rng(ctypes.Structure):
    _fields_ = [("g0", _Pcg32), ("g1", _Pcg32)]


def pcg32_sequence(initstate: int, initseq: int, n: int) -> list[int]:
    r = _Pcg32()
    lib().ora_pcg32_seed(ctypes.byref(r), initstate, initseq)
    return [lib().ora_pcg32_random(ctypes.byref(r)) for _ in range(n)]


class Rng:
    """seed_rng_state + draws, rng_c_compat.mojo:59-149."""

    def __init__(self, seed: int):
        self._s = _Rng()
        lib().ora_seed_rng_state(ctypes.byref(self._s), seed)

    def rand_uint32(self) -> int:
        return lib().ora_rand_uint32(ctypes.byref(self._s))

    def rand_uint64(self) -> int:
        return lib().ora_rand_uint64(ctypes.byref(self._s))

    def rand_interval(self, bound: int) -> int:
        return lib().ora_rand_interval(ctypes.byref(self._s), bound)

    def randnf(self) -> float:
        return lib().ora_randnf(ctypes.byref(self._s))


def random_mps(nsites: int, d: int, chi_max: int, seed: int) -> list[np.ndarray]:
    """create_random_mps_c_compatible (rng_c_compat.mojo:174-231), FP64 copies of float32 draws."""
    dims = np.zeros(nsites + 1, dtype=np.int64)
    total = lib().ora_random_mps_dims(nsites, d, chi_max, seed, dims.ctypes.data_as(_i64p))
    buf = np.zeros(total, dtype=np.float64)
    lib().ora_random_mps_fill(nsites, d, chi_max, seed, _p(buf))
    out, off = [], 0
    for i in range(nsites):
        n = int(dims[i] * d * dims[i + 1])
        out.append(buf[off:off + n].reshape(int(dims[i]), d, int(dims[i + 1])).copy())
        off += n
    return out


# ------------------------------------------------------------ tensors
def permute(a, perm) -> np.ndarray:
    a = _f64(a)
    out = np.empty([a.shape[p] for p in perm], dtype=np.float64)
    sh = _i64(a.shape)
    pm = np.ascontiguousarray(np.asarray(perm, dtype=np.int32))
    lib().ora_permute(_p(a), ctypes.c_int(a.ndim), sh.ctypes.data_as(_i64p),
                      pm.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(out))
    return out


def tensordot_shape(sA, sB, ndim_mult=1, axrange_A=False, axrange_B=False):
    sA, sB = _i64(sA), _i64(sB)
    sC = np.zeros(16, dtype=np.int64)
    r = lib().ora_tensordot_shape(len(sA), sA.ctypes.data_as(_i64p), len(sB), sB.ctypes.data_as(_i64p),
                                  ndim_mult, int(axrange_A), int(axrange_B), sC.ctypes.data_as(_i64p))
    if r < 0:
        raise ValueError("Dim mismatch")
    return tuple(int(x) for x in sC[:r])


def _dot(fn, A, B, ndim_mult, axrange_A, axrange_B):
    A, B = _f64(A), _f64(B)
    shape = tensordot_shape(A.shape, B.shape, ndim_mult, axrange_A, axrange_B)
    C = np.zeros(shape if shape else (1,), dtype=np.float64)
    sA, sB = _i64(A.shape), _i64(B.shape)
    rc = fn(_p(A), A.ndim, sA.ctypes.data_as(_i64p), _p(B), B.ndim, sB.ctypes.data_as(_i64p),
            ndim_mult, int(axrange_A), int(axrange_B), _p(C))
    if rc != 0:
        raise ValueError("Dim mismatch")
    return C


def tensordot(A, B, ndim_mult=1, axrange_A=False, axrange_B=False) -> np.ndarray:
    """dense_tensor_dot semantics (dense_tensor.mojo:963-1331), FP64 + dgemm."""
    return _dot(lib().ora_tensordot, A, B, ndim_mult, axrange_A, axrange_B)


def tensordot_brute(A, B, ndim_mult=1, axrange_A=False, axrange_B=False) -> np.ndarray:
    """Brute-force loops of test_dense_tensor_dot.mojo:25-106 (no axis inference)."""
    return _dot(lib().ora_tensordot_brute, A, B, ndim_mult, axrange_A, axrange_B)


# ------------------------------------------------------ DMRG pieces
def heff_apply(L, Wi, Wj, R, theta) -> np.ndarray:
    """apply_two_site_heff (dmrg.mojo:254-374)."""
    L, Wi, Wj, R, theta = map(_f64, (L, Wi, Wj, R, theta))
    wL, Dl, Dlb = L.shape
    wR, Dr, Drb = R.shape
    d = theta.shape[1]
    wM = Wi.shape[3]
    out = np.zeros((Dlb, d, d, Drb), dtype=np.float64)
    lib().ora_heff_apply(_p(L), _p(Wi), _p(Wj), _p(R), _p(theta), _p(out),
                         *[ctypes.c_int64(int(x)) for x in (Dl, Dlb, d, Dr, Drb, wL, wM, wR)])
    return out


def env_update_left(L, A, W) -> np.ndarray:
    """update_left_environment (environments.mojo:99-201)."""
    L, A, W = map(_f64, (L, A, W))
    Dl, d, Dr = A.shape
    wL, wR = W.shape[0], W.shape[3]
    out = np.zeros((wR, Dr, Dr), dtype=np.float64)
    lib().ora_env_update_left(_p(L), _p(A), _p(W), _p(out),
                              *[ctypes.c_int64(int(x)) for x in (Dl, d, Dr, wL, wR)])
    return out


def env_update_right(R, A, W) -> np.ndarray:
    """update_right_environment (environments.mojo:309-417)."""
    R, A, W = map(_f64, (R, A, W))
    Dl, d, Dr = A.shape
    wL, wR = W.shape[0], W.shape[3]
    out = np.zeros((wL, Dl, Dl), dtype=np.float64)
    lib().ora_env_update_right(_p(R), _p(A), _p(W), _p(out),
                               *[ctypes.c_int64(int(x)) for x in (Dl, d, Dr, wL, wR)])
    return out


def jacobi_smallest(T, tol=1e-10, max_rot=200):
    """_jacobi_smallest_eigpair (dmrg.mojo:507-598)."""
    A = _f64(T).copy()
    n = A.shape[0]
    vec = np.zeros(n, dtype=np.float64)
    e = lib().ora_jacobi_smallest(_p(A), n, tol, max_rot, _p(vec))
    return e, vec


def lanczos_two_site(theta, L, R, Wi, Wj, max_iter=20, tol=1e-10, max_rot=200):
    """lanczos_two_site_optimize (dmrg.mojo:486-756) -> (E0, theta*, nK, alpha, beta).
    max_rot: Jacobi rotation cap (200 in the reference, dmrg.mojo:510)."""
    L, Wi, Wj, R = map(_f64, (L, Wi, Wj, R))
    th = _f64(theta).copy()
    Dl, d, _, Dr = th.shape
    wL, wM, wR = L.shape[0], Wi.shape[3], R.shape[0]
    E = ctypes.c_double()
    nK = ctypes.c_int()
    alpha = np.zeros(max_iter)
    beta = np.zeros(max_iter)
    rc = lib().ora_lanczos_two_site(
        _p(L), _p(Wi), _p(Wj), _p(R), _p(th),
        *[ctypes.c_int64(int(x)) for x in (Dl, d, Dr, wL, wM, wR)],
        ctypes.c_int(max_iter), ctypes.c_double(tol), ctypes.c_int(max_rot), ctypes.byref(E), ctypes.byref(nK),
        _p(alpha), _p(beta))
    if rc != 0:
        raise ValueError("Initial vector has zero norm")
    k = nK.value
    return E.value, th, k, alpha[:k].copy(), beta[:max(k - 1, 0)].copy()


def svd_trunc(A, chi_max, eps_trunc=1e-12):
    """dense_tensor_svd_trunc_lapack_f64 (dense_tensor.mojo:1483-1592) -> (U, S, Vt, chi)."""
    A = _f64(A)
    m, n = A.shape
    k = min(m, n, chi_max)
    U = np.zeros((m, max(k, 1)))
    S = np.zeros(max(k, 1))
    Vt = np.zeros((max(k, 1), n))
    chi = lib().ora_svd_trunc(_p(A), ctypes.c_int64(m), ctypes.c_int64(n), ctypes.c_int64(chi_max),
                              ctypes.c_double(eps_trunc), _p(U), _p(S), _p(Vt))
    if chi < 0:
        raise RuntimeError(f"dgesdd failed, info={-chi - 100}")
    U = U.reshape(-1)[: m * chi].reshape(m, chi)
    Vt = Vt.reshape(-1)[: chi * n].reshape(chi, n)
    return U, S[:chi].copy(), Vt, chi


class _Params(ctypes.Structure):
    _fields_ = [("num_sweeps", ctypes.c_int), ("chi_max", ctypes.c_int64), ("eps_trunc", ctypes.c_double),
                ("max_krylov_iter", ctypes.c_int), ("krylov_tol", ctypes.c_double),
                ("energy_tol", ctypes.c_double), ("jacobi_max_rot", ctypes.c_int)]


def dmrg_two_site(mps_sites, mpo_sites, num_sweeps, chi_max, eps_trunc=1e-10, max_krylov_iter=20,
                  krylov_tol=1e-10, energy_tol=0.0, jacobi_max_rot=200):
    """dmrg_two_site (dmrg.mojo:377-483).

    Returns dict(E, sweep_E, trunc_weight, sweeps, mps)."""
    N = len(mps_sites)
    d = mps_sites[0].shape[1]
    dims = _i64([s.shape[0] for s in mps_sites] + [mps_sites[-1].shape[2]])
    wd = _i64([w.shape[0] for w in mpo_sites] + [mpo_sites[-1].shape[3]])
    mps_in = _f64(np.concatenate([_f64(s).ravel() for s in mps_sites]))
    mpo = _f64(np.concatenate([_f64(w).ravel() for w in mpo_sites]))
    # capacity: every bond at min(d^l, d^(N-l), chi_max)
    cap_dims = [min(d ** min(l, N - l), chi_max) if l not in (0, N) else int(dims[l]) for l in range(N + 1)]
    cap = sum(max(cap_dims[i], dims[i]) * d * max(cap_dims[i + 1], dims[i + 1]) for i in range(N))
    out = np.zeros(int(cap) + 1)
    sweep_E = np.zeros(num_sweeps)
    tw = ctypes.c_double()
    Ef = ctypes.c_double()
    prm = _Params(num_sweeps, chi_max, eps_trunc, max_krylov_iter, krylov_tol, energy_tol, jacobi_max_rot)
    rc = lib().ora_dmrg_two_site(ctypes.c_int(N), ctypes.c_int64(d), dims.ctypes.data_as(_i64p), _p(mps_in),
                                 wd.ctypes.data_as(_i64p), _p(mpo), ctypes.byref(prm), _p(out),
                                 ctypes.c_int64(out.size), _p(sweep_E), ctypes.byref(tw), ctypes.byref(Ef))
    if rc < 0:
        raise RuntimeError(f"oracle dmrg failed rc={rc}")
    sites, off = [], 0
    for i in range(N):
        n = int(dims[i] * d * dims[i + 1])
        sites.append(out[off:off + n].reshape(int(dims[i]), d, int(dims[i + 1])).copy())
        off += n
    return {"E": Ef.value, "sweep_E": sweep_E[:rc].copy(), "trunc_weight": tw.value, "sweeps": rc,
            "mps": sites}
