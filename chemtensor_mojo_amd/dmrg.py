"""Two-site DMRG call surface, mirroring src/algorithms/dmrg.mojo.

``dmrg_two_site`` runs the whole sweep schedule inside libdmrg_hip.so (C++
driver, every tensor resident in HBM); the per-step functions
(``apply_two_site_heff``, ``lanczos_two_site_optimize``,
``split_two_site_theta_svd``, ...) expose the same kernels one call at a time
with the reference's signatures.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._lib import PREC_BF16X3, PREC_BF16X6, PREC_FP32, PREC_FP64, SVD_GESVD, SVD_GRAM, call
from .dense_tensor import DenseTensor, DeviceContext, default_prec
from .state import MatrixProductOperator, MatrixProductState, MPOSite, MPSSite


@dataclass
class DMRGParams:
    """DMRGParams (dmrg.mojo:22-36) plus precision controls without a reference counterpart.

    prec            precision of the Krylov matvecs (PREC_FP64 / PREC_FP32 / PREC_BF16X3 /
                    PREC_BF16X6: FP32 accuracy from six BF16 products)
    refine          0: reference energies; 1: FP64 Rayleigh quotient of each local
                    Ritz vector; 2: plus one residual-correction step
    env_prec        precision of the environment updates
    jacobi_max_rot  rotation cap of the Jacobi eigensolve of T; 200 is the
                    reference's value (dmrg.mojo:510), which leaves 20x20 T
                    matrices unconverged late in a run; -1 solves T exactly
                    (Sturm multisection + inverse iteration: the converged
                    Jacobi result to rounding, ~200x faster than converging
                    the Jacobi)
    svd_mode        SVD_GRAM (default: FP64 eigendecomposition of theta theta^T or
                    theta^T theta, 10-25x faster than a full SVD at D >= 512) or
                    SVD_GESVD (full FP64 SVD, the reference's factorisation).  The
                    Gram eigenvalues sigma^2 carry an absolute error of about
                    k * 1e-16 * sigma_max^2 (k = min(Dl d, d Dr)), so singular
                    values below ~sqrt(k) * 1e-8 sigma_max are resolved only
                    roughly.  For eps_trunc well above k * 1e-16 both modes keep
                    the same chi and discarded weight; for smaller eps_trunc (and
                    chi_max not binding) the Gram split can keep a few more or
                    fewer vectors than the full SVD, a difference of order
                    k * 1e-16 in the discarded weight; use SVD_GESVD when that
                    matters.
    graphs          replay each bond's Lanczos solve from a HIP graph cached per
                    bond once its shapes repeat (identical results; removes the
                    per-kernel launch cost that dominates at small bond dimensions)
    """
    num_sweeps: int
    chi_max: int
    eps_trunc: float
    max_krylov_iter: int
    krylov_tol: float
    energy_tol: float
    two_site: bool = True
    verbose: bool = False
    prec: int = PREC_FP64
    refine: int = 0
    env_prec: int = PREC_FP64
    jacobi_max_rot: int = 200
    svd_mode: int = SVD_GRAM
    graphs: bool = True


@dataclass
class DMRGStats:
    energy: float = 0.0
    sweeps: int = 0
    local_updates: int = 0
    trunc_weight_sum: float = 0.0
    trunc_weight_max: float = 0.0
    max_bond: int = 0
    t_total: float = 0.0
    t_lanczos: float = 0.0
    t_svd: float = 0.0
    t_env: float = 0.0
    t_theta: float = 0.0
    matvecs: int = 0
    matvec_flops: float = 0.0
    env_flops: float = 0.0
    graph_replays: int = 0
    kept_start: int = 0
    sweep_energies: list = field(default_factory=list)


def build_two_site_theta(A_i: MPSSite, A_ip1: MPSSite, ctx: DeviceContext) -> DenseTensor:
    """theta[Dl,d,d,Dr] = A_i[Dl*d,Dm] @ A_{i+1}[Dm,d*Dr] (dmrg.mojo:93-140)."""
    Dl, d, Dm = A_i.tensor.shape
    Dm2, d2, Dr = A_ip1.tensor.shape
    if Dm2 != Dm:
        raise ValueError("Bond dimension mismatch in build_two_site_theta")
    if d2 != d:
        raise ValueError("Physical dimensions must match")
    t = A_i.tensor
    out = torch.empty(Dl * d * d * Dr, dtype=t.dtype, device=t.storage.device)
    call("ct_two_site_theta", ctx.handle, default_prec(t.dtype), t.dtype_code, t.ptr(), A_ip1.tensor.ptr(),
         ctypes.c_void_p(out.data_ptr()), Dl, d, Dm, Dr)
    return DenseTensor(out, [Dl, d, d, Dr])


def apply_two_site_heff(theta: DenseTensor, L_env: DenseTensor, R_env: DenseTensor, W_i: MPOSite, W_ip1: MPOSite,
                        ctx: DeviceContext, prec: int | None = None) -> DenseTensor:
    """H_eff|theta> without forming H_eff (dmrg.mojo:254-374)."""
    Dl, d, _, Dr = theta.shape
    wL, Dl2, Dlb = L_env.shape
    wR, Dr2, Drb = R_env.shape
    wM = W_i.tensor.shape[3]
    if Dl2 != Dl or Dr2 != Dr or W_i.tensor.shape[0] != wL or W_ip1.tensor.shape[3] != wR \
            or W_ip1.tensor.shape[0] != wM:
        raise ValueError("apply_two_site_heff: shape mismatch")
    dt = theta.dtype
    out = torch.empty(Dlb * d * d * Drb, dtype=dt, device=theta.storage.device)
    p = default_prec(dt) if prec is None else prec
    call("ct_heff_apply", ctx.handle, p, theta.dtype_code, L_env.ptr(), W_i.tensor.ptr(), W_ip1.tensor.ptr(),
         R_env.ptr(), theta.ptr(), ctypes.c_void_p(out.data_ptr()), Dl, Dlb, d, Dr, Drb, wL, wM, wR)
    return DenseTensor(out, [Dlb, d, d, Drb])


class HeffOperator:
    """Prepared two-site H_eff (ct_heff_create): the environments and MPO tensors
    are fixed, ``apply(theta)`` is apply_two_site_heff (dmrg.mojo:254-374)
    without re-deriving W_i W_{i+1} and the precision-specific copies of L and R
    on every call - the form a Lanczos-type eigensolver uses.  With FP64 the
    operator reads L and R in place; this object keeps them alive."""

    def __init__(self, L_env: DenseTensor, R_env: DenseTensor, W_i: MPOSite, W_ip1: MPOSite, ctx: DeviceContext,
                 prec: int | None = None):
        wL, Dl, Dlb = L_env.shape
        wR, Dr, Drb = R_env.shape
        wM = W_i.tensor.shape[3]
        d = W_i.tensor.shape[1]
        if W_i.tensor.shape[0] != wL or W_ip1.tensor.shape[3] != wR or W_ip1.tensor.shape[0] != wM \
                or W_ip1.tensor.shape[1] != d:
            raise ValueError("HeffOperator: shape mismatch")
        dt = L_env.dtype
        for t in (R_env, W_i.tensor, W_ip1.tensor):
            if t.dtype != dt:
                raise TypeError("HeffOperator: L, R and W must share one dtype")
        self.ctx = ctx
        self.dtype = dt
        self.dims = (Dl, Dlb, d, Dr, Drb)
        self._keep = (L_env, R_env, W_i, W_ip1)
        p = default_prec(dt) if prec is None else prec
        op = ctypes.c_void_p()
        call("ct_heff_create", ctx.handle, p, L_env.dtype_code, L_env.ptr(), W_i.tensor.ptr(), W_ip1.tensor.ptr(),
             R_env.ptr(), Dl, Dlb, d, Dr, Drb, wL, wM, wR, ctypes.byref(op))
        self.op = op
        _lib.track(self)

    def apply(self, theta: DenseTensor, out: torch.Tensor | None = None) -> DenseTensor:
        Dl, Dlb, d, Dr, Drb = self.dims
        if list(theta.shape) != [Dl, d, d, Dr] or theta.dtype != self.dtype:
            raise ValueError("HeffOperator.apply: theta must be [Dl, d, d, Dr] of the operator's dtype")
        if out is None:
            out = torch.empty(Dlb * d * d * Drb, dtype=self.dtype, device=theta.storage.device)
        call("ct_heff_apply_prepared", self.op, theta.ptr(), ctypes.c_void_p(out.data_ptr()))
        return DenseTensor(out, [Dlb, d, d, Drb])

    def close(self) -> None:
        if getattr(self, "op", None) is not None and self.op.value:
            _lib.load().ct_heff_destroy(self.op)
            self.op = None
        _lib.untrack(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _common_dtype(*tensors: DenseTensor) -> torch.dtype:
    dt = tensors[0].dtype
    for t in tensors[1:]:
        if t.dtype != dt:
            raise TypeError("all tensors must share one dtype (float64 or float32)")
    return dt


def lanczos_two_site_optimize(initial_theta: DenseTensor, L_env: DenseTensor, R_env: DenseTensor, W_i: MPOSite,
                              W_ip1: MPOSite, ctx: DeviceContext, max_iter: int, tol: float,
                              prec: int = PREC_FP64, refine: int = 0, return_info: bool = False,
                              max_rotations: int = 200):
    """Lowest eigenpair of the local H_eff (dmrg.mojo:486-756), FP64 Krylov basis on
    the GPU.  Returns (E0, theta*) like the reference, theta* in the tensors' dtype
    (float64, or the reference's float32, widened to FP64 inside the solver); with
    return_info also a dict with the Krylov dimension, the raw Lanczos value and
    the residual norm.  prec selects the matvec precision."""
    dt = _common_dtype(initial_theta, L_env, R_env, W_i.tensor, W_ip1.tensor)
    Dl, d, _, Dr = initial_theta.shape
    wL, wM, wR = L_env.shape[0], W_i.tensor.shape[3], R_env.shape[0]
    th = initial_theta.storage.clone()
    E = ctypes.c_double()
    n = ctypes.c_int()
    El = ctypes.c_double()
    rn = ctypes.c_double()
    call("ct_lanczos_two_site", ctx.handle, int(prec), initial_theta.dtype_code, L_env.ptr(), W_i.tensor.ptr(),
         W_ip1.tensor.ptr(), R_env.ptr(), ctypes.c_void_p(th.data_ptr()), Dl, d, Dr, wL, wM, wR, int(max_iter),
         float(tol), int(max_rotations), int(refine), ctypes.byref(E), ctypes.byref(n), ctypes.byref(El),
         ctypes.byref(rn))
    out = DenseTensor(th, [Dl, d, d, Dr])
    if return_info:
        return E.value, out, {"num_iter": n.value, "E_lanczos": El.value, "residual_norm": rn.value}
    return E.value, out


def dense_tensor_svd_trunc(tensor: DenseTensor, ctx: DeviceContext, chi_max: int, eps_trunc: float = 1e-12,
                           return_discarded: bool = False):
    """Truncated SVD of a [m, n] matrix on the GPU (dense_tensor.mojo:1483-1623), in FP64
    (float32 input is widened and the factors rounded back).  Returns (U [m,chi],
    S [chi], Vt [chi,n], chi) like the reference; with return_discarded also the
    discarded weight."""
    if tensor.rank() != 2:
        raise ValueError("dense_tensor_svd_trunc expects a matrix")
    dt = tensor.dtype
    m, n = tensor.shape
    k = min(m, n, chi_max)
    dev = tensor.storage.device
    U = torch.empty(max(m * k, 1), dtype=dt, device=dev)
    S = torch.empty(max(k, 1), dtype=dt, device=dev)
    Vt = torch.empty(max(k * n, 1), dtype=dt, device=dev)
    chi = ctypes.c_int64()
    dw = ctypes.c_double()
    call("ct_svd_trunc", ctx.handle, tensor.dtype_code, tensor.ptr(), m, n, chi_max, float(eps_trunc),
         ctypes.c_void_p(U.data_ptr()), ctypes.c_void_p(S.data_ptr()), ctypes.c_void_p(Vt.data_ptr()),
         ctypes.byref(chi), ctypes.byref(dw))
    c = chi.value
    if c == 0:
        res = (None, None, None, 0)
    else:
        res = (DenseTensor(U[: m * c], [m, c]), DenseTensor(S[:c], [c]), DenseTensor(Vt[: c * n], [c, n]), c)
    return res + (dw.value,) if return_discarded else res


def split_two_site_theta_svd(theta: DenseTensor, chi_max: int, eps_trunc: float, ctx: DeviceContext,
                             left_to_right: bool = True, svd_mode: int = SVD_GESVD):
    """theta[Dl,d,d,Dr] -> (A_i[Dl,d,chi], A_{i+1}[chi,d,Dr], discarded_weight)
    (dmrg.mojo:143-252): S goes right on L->R, left on R->L.  The split runs in
    FP64 on the GPU; float32 theta (the reference's dtype) is widened on entry and
    the site tensors are returned in float32."""
    dt = theta.dtype
    Dl, d, _, Dr = theta.shape
    k = min(Dl * d, d * Dr, chi_max)
    dev = theta.storage.device
    Ai = torch.empty(Dl * d * k, dtype=dt, device=dev)
    Aj = torch.empty(k * d * Dr, dtype=dt, device=dev)
    chi = ctypes.c_int64()
    dw = ctypes.c_double()
    call("ct_theta_split", ctx.handle, int(svd_mode), theta.dtype_code, theta.ptr(), Dl, d, Dr, chi_max,
         float(eps_trunc), int(bool(left_to_right)), ctypes.c_void_p(Ai.data_ptr()), ctypes.c_void_p(Aj.data_ptr()),
         ctypes.byref(chi), ctypes.byref(dw))
    c = chi.value
    if c == 0:
        raise RuntimeError("truncation kept no singular value")
    return (MPSSite(DenseTensor(Ai[: Dl * d * c], [Dl, d, c])), MPSSite(DenseTensor(Aj[: c * d * Dr], [c, d, Dr])),
            dw.value)


class DMRGRun:
    """Owns a library-side DMRG state (MPS, environments, workspaces on the device)."""

    def __init__(self, ctx: DeviceContext, mpo: MatrixProductOperator, mps: MatrixProductState, params: DMRGParams):
        if not params.two_site:
            raise ValueError("Only two-site DMRG is currently implemented")
        N = mps.num_sites()
        if N < 2:
            raise ValueError("DMRG requires at least 2 sites")
        if mpo.num_sites() != N:
            raise ValueError("MPS and MPO must have the same number of sites")
        self.ctx = ctx
        self.N = N
        self.d = mps.physical_dim
        self.params = params
        # The driver copies the sites into its own FP64 storage at create time;
        # float32 sites (the reference's dtype) are widened there and mps()
        # returns float32 again.
        self.dtype = _common_dtype(*[s.tensor for s in mps.sites], *[s.tensor for s in mpo.sites])
        self._mpo = [s.tensor.storage.contiguous() for s in mpo.sites]
        self._mps = [s.tensor.storage.contiguous() for s in mps.sites]
        p = _lib.DMRGParamsC(params.num_sweeps, params.chi_max, params.eps_trunc, params.max_krylov_iter,
                             params.krylov_tol, params.energy_tol, params.jacobi_max_rot, params.svd_mode, params.prec,
                             params.refine, params.env_prec, int(params.verbose), int(params.graphs))
        wd = _lib.i64_array(mpo.bond_dims)
        md = _lib.i64_array(mps.bond_dims)
        wp = (ctypes.c_void_p * N)(*[t.data_ptr() for t in self._mpo])
        mp = (ctypes.c_void_p * N)(*[t.data_ptr() for t in self._mps])
        st = ctypes.c_void_p()
        call("ct_dmrg_create", ctx.handle, ctypes.byref(p), N, self.d, _lib.F64 if self.dtype == torch.float64
             else _lib.F32, wd, wp, md, mp, ctypes.byref(st))
        self.state = st
        _lib.track(self)

    def run(self, num_sweeps: int = 0) -> tuple[float, list[float]]:
        n = num_sweeps if num_sweeps > 0 else self.params.num_sweeps
        E = ctypes.c_double()
        se = (ctypes.c_double * max(n, 1))()
        ran = ctypes.c_int()
        call("ct_dmrg_run", self.state, n, ctypes.byref(E), se, ctypes.byref(ran))
        return E.value, list(se)[: ran.value]

    def stats(self) -> DMRGStats:
        c = _lib.DMRGStatsC()
        call("ct_dmrg_get_stats", self.state, ctypes.byref(c))
        return DMRGStats(**{k: getattr(c, k) for k, _ in c._fields_})

    def dims(self) -> list[int]:
        d = (ctypes.c_int64 * (self.N + 1))()
        call("ct_dmrg_get_dims", self.state, d)
        return list(d)

    def mps(self) -> MatrixProductState:
        dims = self.dims()
        sites = []
        for i in range(self.N):
            n = dims[i] * self.d * dims[i + 1]
            t = torch.empty(n, dtype=self.dtype, device=self.ctx.device)
            call("ct_dmrg_get_site", self.state, i, _lib.F64 if self.dtype == torch.float64 else _lib.F32,
                 ctypes.c_void_p(t.data_ptr()))
            sites.append(MPSSite(DenseTensor(t, [dims[i], self.d, dims[i + 1]])))
        return MatrixProductState(sites)

    def close(self) -> None:
        if self.state is not None and self.state.value:
            _lib.load().ct_dmrg_destroy(self.state)
            self.state = None
        _lib.untrack(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dmrg_two_site(ctx: DeviceContext, mpo: MatrixProductOperator, mps: MatrixProductState, params: DMRGParams,
                  return_stats: bool = False):
    """Two-site DMRG (dmrg.mojo:377-483).  Returns (energy, optimised MPS) like the
    reference; the energy is the last local Lanczos eigenvalue (:458-463), or its
    FP64-refined value when params.refine != 0."""
    run = DMRGRun(ctx, mpo, mps, params)
    try:
        E, sweep_E = run.run()
        out = run.mps()
        st = run.stats()
        st.sweep_energies = sweep_E
    finally:
        run.close()
    if return_stats:
        return E, out, st
    return E, out


__all__ = [
    "DMRGParams", "DMRGStats", "DMRGRun", "dmrg_two_site", "apply_two_site_heff", "lanczos_two_site_optimize",
    "build_two_site_theta", "split_two_site_theta_svd", "dense_tensor_svd_trunc", "PREC_FP64", "PREC_FP32",
    "PREC_BF16X3", "PREC_BF16X6",
]
