"""Adjoint of the environment contractions and dE/dJ behind torch.autograd.

No reference counterpart (the reference has no differentiation).  For an MPS
|psi> (e.g. the DMRG result) and an MPO H(J) whose tensors depend on
couplings J,

    E(J) = <psi|H(J)|psi> / <psi|psi>
    dE/dW_i = L_i . A_i (x) A_i . R_{i+1} / <psi|psi>      (ct_env_adjoint)
    dE/dJ_b = sum_i < dE/dW_i , dW_i/dJ_b >

At a variational optimum (converged DMRG, negligible truncation) this is the
Hellmann-Feynman derivative <psi|dH/dJ|psi>, i.e. the derivative of the
ground-state energy.  The backward pass recomputes the environments from the
stored MPS: the right environments are built once and the left ones are
carried along a single left-to-right pass, so only one side's environments
are held in memory.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import call
from .dense_tensor import DenseTensor, DeviceContext
from .state import (MatrixProductOperator, MatrixProductState, MPOSite, mpo_from_tensors, update_left_environment,
                    update_right_environment)


def _ident(ctx: DeviceContext, w: int, D: int) -> DenseTensor:
    t = torch.zeros((w, D, D), dtype=torch.float64, device=ctx.device)
    t[:, torch.arange(D), torch.arange(D)] = 1.0
    return DenseTensor(t)


def env_adjoint(L: DenseTensor, A: DenseTensor, R: DenseTensor, wL: int, wR: int, ctx: DeviceContext) -> torch.Tensor:
    """gW[wl,s,s',wr] = sum L[wl,a,a'] A[a,s,b] A[a',s',b'] R[wr,b,b'] (FP64, device)."""
    Dl, d, Dr = A.shape
    g = torch.empty(wL * d * d * wR, dtype=torch.float64, device=ctx.device)
    call("ct_env_adjoint", ctx.handle, L.ptr(), A.ptr(), R.ptr(), ctypes.c_void_p(g.data_ptr()), Dl, d, Dr, wL, wR)
    return g.view(wL, d, d, wR)


def energy_and_mpo_gradient(ctx: DeviceContext, mps: MatrixProductState, mpo: MatrixProductOperator):
    """(E, [dE/dW_i]) for E = <psi|H|psi>/<psi|psi>, all on the device in FP64."""
    N = mps.num_sites()
    d = mps.physical_dim
    Wid = MPOSite(DenseTensor(torch.eye(d, dtype=torch.float64, device=ctx.device).reshape(1, d, d, 1)))
    # right environments of H and of the identity (norm), from the right end
    R = [None] * (N + 1)
    Rn = [None] * (N + 1)
    R[N] = _ident(ctx, mpo.bond_dimension(N), mps.bond_dimension(N))
    Rn[N] = _ident(ctx, 1, mps.bond_dimension(N))
    for i in range(N - 1, -1, -1):
        R[i] = update_right_environment(R[i + 1], mps.sites[i], mpo.sites[i], ctx)
        Rn[i] = update_right_environment(Rn[i + 1], mps.sites[i], Wid, ctx)
    E_num = float(R[0].to_numpy().ravel()[0])
    norm = float(Rn[0].to_numpy().ravel()[0])
    grads = []
    L = _ident(ctx, mpo.bond_dimension(0), mps.bond_dimension(0))
    for i in range(N):
        g = env_adjoint(L, mps.sites[i].tensor, R[i + 1], mpo.bond_dimension(i), mpo.bond_dimension(i + 1), ctx)
        grads.append(g / norm)
        if i + 1 < N:
            L = update_left_environment(L, mps.sites[i], mpo.sites[i], ctx)
        R[i + 1] = None  # released as the pass moves right
    return E_num / norm, grads


def expectation(ctx: DeviceContext, mps: MatrixProductState, mpo_tensors: Sequence[np.ndarray]) -> float:
    """<psi|O|psi>/<psi|psi> for an MPO given as host tensors (forward contraction only)."""
    O = mpo_from_tensors(ctx, mpo_tensors)
    d = mps.physical_dim
    N = mps.num_sites()
    L = _ident(ctx, O.bond_dimension(0), mps.bond_dimension(0))
    Ln = _ident(ctx, 1, mps.bond_dimension(0))
    Wid = MPOSite(DenseTensor(torch.eye(d, dtype=torch.float64, device=ctx.device).reshape(1, d, d, 1)))
    for i in range(N):
        L = update_left_environment(L, mps.sites[i], O.sites[i], ctx)
        Ln = update_left_environment(Ln, mps.sites[i], Wid, ctx)
    return float(L.to_numpy().ravel()[0]) / float(Ln.to_numpy().ravel()[0])


class CouplingModel:
    """An MPO family H(J): ``build(J)`` gives host tensors, ``derivative(J, b)``
    gives [(site, dW_site/dJ_b)] (host arrays)."""

    def __init__(self, build: Callable[[np.ndarray], list], derivative: Callable[[np.ndarray, int], list]):
        self.build = build
        self.derivative = derivative


def random_bond_heisenberg_model(num_sites: int) -> CouplingModel:
    from . import hamiltonians as ham

    def build(J):
        return ham.random_bond_heisenberg_tensors(np.asarray(J, dtype=np.float64))

    def derivative(J, b):
        site, G = ham.random_bond_coupling_derivative(num_sites, b)
        return [(site, G)]

    return CouplingModel(build, derivative)


def coupling_gradient(ctx: DeviceContext, mps: MatrixProductState, model: CouplingModel, J: np.ndarray):
    """(E, dE/dJ) at fixed psi through the adjoint kernels."""
    J = np.asarray(J, dtype=np.float64)
    mpo = mpo_from_tensors(ctx, model.build(J))
    E, grads = energy_and_mpo_gradient(ctx, mps, mpo)
    gh = [g.cpu().numpy() for g in grads]
    dE = np.zeros(J.size)
    for b in range(J.size):
        for site, dW in model.derivative(J, b):
            dE[b] += float(np.sum(gh[site] * dW))
    return E, dE


def heff_adjoint(L: DenseTensor, R: DenseTensor, W_i: DenseTensor, W_ip1: DenseTensor, theta: DenseTensor,
                 ctx: DeviceContext):
    """Gradients of the unnormalised two-site energy <theta|H_eff(W_i, W_ip1)|theta> with respect to the two MPO
    tensors (ct_heff_adjoint, FP64):
        gWi[wl,s1,s1',wm] = sum L[wl,a,a'] theta[a,s1,s2,b] W_ip1[wm,s2,s2',wr] R[wr,b,b'] theta[a',s1',s2',b']
        gWj[wm,s2,s2',wr] = the same network with W_ip1 removed and W_i kept."""
    Dl, d, _, Dr = theta.shape
    wL, wM, wR = W_i.shape[0], W_i.shape[3], W_ip1.shape[3]
    gi = torch.empty(wL * d * d * wM, dtype=torch.float64, device=ctx.device)
    gj = torch.empty(wM * d * d * wR, dtype=torch.float64, device=ctx.device)
    call("ct_heff_adjoint", ctx.handle, L.ptr(), W_i.ptr(), W_ip1.ptr(), R.ptr(), theta.ptr(),
         ctypes.c_void_p(gi.data_ptr()), ctypes.c_void_p(gj.data_ptr()), Dl, d, Dr, wL, wM, wR)
    return gi.view(wL, d, d, wM), gj.view(wM, d, d, wR)


class TwoSiteEnergy(torch.autograd.Function):
    """E = <theta|H_eff|theta> / <theta|theta> of the two-site local problem
    (H_eff of apply_two_site_heff, dmrg.mojo:254-374), differentiable in
    theta, W_i and W_ip1 (FP64 device tensors; L and R are constants):

        dE/dW_i, dE/dW_ip1 = ct_heff_adjoint / <theta|theta>
        dE/dtheta          = (H_eff + H_eff^T) theta / <theta|theta> - 2 E theta / <theta|theta>

    H_eff^T theta is the matvec with every operand transposed (L[w]^T,
    W with its two physical indices swapped, R[w]^T), so no symmetry of H is
    assumed.

    forward(theta [Dl,d,d,Dr], W_i [wL,d,d,wM], W_ip1 [wM,d,d,wR],
            L [wL,Dl,Dl], R [wR,Dr,Dr], ctx)"""

    @staticmethod
    def forward(fctx, theta, W_i, W_ip1, L, R, ctx: DeviceContext):
        if L.requires_grad or R.requires_grad:
            raise ValueError("TwoSiteEnergy: the environments are constants (no gradient with respect to L or R)")
        for t in (theta, W_i, W_ip1, L, R):
            if t.dtype != torch.float64 or t.device.type != "cuda":
                raise ValueError("TwoSiteEnergy: FP64 device tensors required")
        Dl, d, _, Dr = theta.shape
        if L.shape[1:] != (Dl, Dl) or R.shape[1:] != (Dr, Dr) or W_i.shape[0] != L.shape[0] \
                or W_ip1.shape[3] != R.shape[0] or W_i.shape[3] != W_ip1.shape[0]:
            raise ValueError("TwoSiteEnergy: shape mismatch")
        th, Wi, Wj, Lc, Rc = (DenseTensor(t.detach().contiguous()) for t in (theta, W_i, W_ip1, L, R))
        Hth = _heff(th, Lc, Rc, Wi, Wj, ctx)
        nrm = float(torch.dot(th.storage, th.storage))
        E = float(torch.dot(th.storage, Hth)) / nrm
        fctx.save_for_backward(theta.detach(), W_i.detach(), W_ip1.detach(), L.detach(), R.detach(), Hth)
        fctx.ctx, fctx.E, fctx.nrm = ctx, E, nrm
        return torch.tensor(E, dtype=torch.float64, device=theta.device)

    @staticmethod
    def backward(fctx, gE):
        theta, W_i, W_ip1, L, R, Hth = fctx.saved_tensors
        ctx, E, nrm = fctx.ctx, fctx.E, fctx.nrm
        gth = gWi = gWj = None
        th = DenseTensor(theta.contiguous())
        if fctx.needs_input_grad[1] or fctx.needs_input_grad[2]:
            gi, gj = heff_adjoint(DenseTensor(L.contiguous()), DenseTensor(R.contiguous()),
                                  DenseTensor(W_i.contiguous()), DenseTensor(W_ip1.contiguous()), th, ctx)
            gWi = gE * gi / nrm if fctx.needs_input_grad[1] else None
            gWj = gE * gj / nrm if fctx.needs_input_grad[2] else None
        if fctx.needs_input_grad[0]:
            # H^T theta: the same matvec on transposed operands
            Lt = DenseTensor(L.transpose(1, 2).contiguous())
            Rt = DenseTensor(R.transpose(1, 2).contiguous())
            Wit = DenseTensor(W_i.transpose(1, 2).contiguous())
            Wjt = DenseTensor(W_ip1.transpose(1, 2).contiguous())
            HTth = _heff(th, Lt, Rt, Wit, Wjt, ctx)
            g = (Hth + HTth - 2.0 * E * th.storage) / nrm
            gth = gE * g.view(theta.shape)
        return gth, gWi, gWj, None, None, None


def _heff(th: DenseTensor, L: DenseTensor, R: DenseTensor, Wi: DenseTensor, Wj: DenseTensor,
          ctx: DeviceContext) -> torch.Tensor:
    from .dmrg import apply_two_site_heff
    return apply_two_site_heff(th, L, R, MPOSite(Wi), MPOSite(Wj), ctx, prec=_lib.PREC_FP64).storage


def two_site_energy(theta: torch.Tensor, W_i: torch.Tensor, W_ip1: torch.Tensor, L: torch.Tensor, R: torch.Tensor,
                    ctx: DeviceContext) -> torch.Tensor:
    """Differentiable two-site local energy (see TwoSiteEnergy)."""
    return TwoSiteEnergy.apply(theta, W_i, W_ip1, L, R, ctx)


class DMRGEnergy(torch.autograd.Function):
    """E(J) = ground-state energy of H(J) by two-site DMRG (forward), with
    dE/dJ from the adjoint environment contractions (backward).

    forward(J, solver) : J a 1-D float64 tensor; ``solver(J_numpy) -> (E, MPS)``
    backward           : gE * dE/dJ, environments recomputed from the saved MPS.
    """

    @staticmethod
    def forward(fctx, J: torch.Tensor, solver, model: CouplingModel, ctx: DeviceContext):
        Jn = J.detach().cpu().numpy().astype(np.float64)
        E, mps = solver(Jn)
        fctx.mps = mps
        fctx.model = model
        fctx.ctx = ctx
        fctx.Jn = Jn
        return torch.tensor(E, dtype=torch.float64)

    @staticmethod
    def backward(fctx, gE):
        _, dE = coupling_gradient(fctx.ctx, fctx.mps, fctx.model, fctx.Jn)
        return gE * torch.from_numpy(dE), None, None, None


def dmrg_energy(J: torch.Tensor, solver, model: CouplingModel, ctx: DeviceContext) -> torch.Tensor:
    """Differentiable DMRG ground-state energy (see DMRGEnergy)."""
    return DMRGEnergy.apply(J, solver, model, ctx)


__all__ = ["env_adjoint", "heff_adjoint", "TwoSiteEnergy", "two_site_energy", "energy_and_mpo_gradient", "expectation", "CouplingModel", "random_bond_heisenberg_model",
           "coupling_gradient", "DMRGEnergy", "dmrg_energy"]
