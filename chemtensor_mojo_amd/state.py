"""MPS / MPO containers and the environment updates, mirroring src/state.

Layouts (reference conventions): MPS site [Dl, d, Dr] (mps_state.mojo:13-19),
MPO site [wL, d_in, d_out, wR] (mpo_state.mojo:14-19), environments
[w, D_ket, D_bra] with identity boundaries (dmrg.mojo:70-90,
environments.mojo:441-461).
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
import torch

from . import _lib
from ._lib import PREC_FP64, call
from .dense_tensor import DenseTensor, DeviceContext, create_dense_tensor_from_data, default_prec


class MPSSite:
    """mps_state.mojo:13-49."""

    def __init__(self, tensor: DenseTensor):
        if tensor.rank() != 3:
            raise ValueError("MPS site tensor must have rank 3 [Dl, d, Dr]")
        self.tensor = tensor

    def rank(self) -> int:
        return 3

    def shape(self) -> list[int]:
        return list(self.tensor.shape)

    def physical_dim(self) -> int:
        return self.tensor.shape[1]

    def left_bond_dim(self) -> int:
        return self.tensor.shape[0]

    def right_bond_dim(self) -> int:
        return self.tensor.shape[2]


class MatrixProductState:
    """mps_state.mojo:52-217 (container and call surface)."""

    def __init__(self, sites: Sequence[MPSSite]):
        if len(sites) < 1:
            raise ValueError("MPS needs at least one site")
        self.sites = list(sites)
        self.physical_dim = self.sites[0].physical_dim()
        for a, b in zip(self.sites[:-1], self.sites[1:]):
            if a.right_bond_dim() != b.left_bond_dim():
                raise ValueError("MPS bond dimension mismatch")
        self.bond_dims = [s.left_bond_dim() for s in self.sites] + [self.sites[-1].right_bond_dim()]

    def num_sites(self) -> int:
        return len(self.sites)

    def bond_dimension(self, index: int) -> int:
        return self.bond_dims[index]

    def site_shape(self, index: int) -> list[int]:
        return self.sites[index].shape()

    def describe(self) -> str:
        s = f"MatrixProductState(num_sites={self.num_sites()}, d={self.physical_dim}, bonds={self.bond_dims})"
        print(s)
        return s

    def to_numpy(self) -> list[np.ndarray]:
        return [s.tensor.to_numpy() for s in self.sites]


class MPOSite:
    """mpo_state.mojo:10-54."""

    def __init__(self, tensor: DenseTensor):
        if tensor.rank() != 4:
            raise ValueError("MPO site tensor must have rank 4 [wL, d_in, d_out, wR]")
        self.tensor = tensor

    def shape(self) -> list[int]:
        return list(self.tensor.shape)

    def left_bond_dim(self) -> int:
        return self.tensor.shape[0]

    def physical_in_dim(self) -> int:
        return self.tensor.shape[1]

    def physical_out_dim(self) -> int:
        return self.tensor.shape[2]

    def right_bond_dim(self) -> int:
        return self.tensor.shape[3]


class MatrixProductOperator:
    """mpo_state.mojo:57-193."""

    def __init__(self, sites: Sequence[MPOSite]):
        self.sites = list(sites)
        for a, b in zip(self.sites[:-1], self.sites[1:]):
            if a.right_bond_dim() != b.left_bond_dim():
                raise ValueError("MPO bond dimension mismatch")
        self.bond_dims = [s.left_bond_dim() for s in self.sites] + [self.sites[-1].right_bond_dim()]

    def num_sites(self) -> int:
        return len(self.sites)

    def bond_dimension(self, index: int) -> int:
        return self.bond_dims[index]

    def site_shape(self, index: int) -> list[int]:
        return self.sites[index].shape()

    def to_numpy(self) -> list[np.ndarray]:
        return [s.tensor.to_numpy() for s in self.sites]


def mpo_from_tensors(ctx: DeviceContext, tensors: Sequence[np.ndarray],
                     dtype: torch.dtype = torch.float64) -> MatrixProductOperator:
    return MatrixProductOperator([MPOSite(create_dense_tensor_from_data(ctx, t, t.shape, dtype)) for t in tensors])


def mps_from_tensors(ctx: DeviceContext, tensors: Sequence[np.ndarray],
                     dtype: torch.dtype = torch.float64) -> MatrixProductState:
    return MatrixProductState([MPSSite(create_dense_tensor_from_data(ctx, t, t.shape, dtype)) for t in tensors])


def create_product_mps(ctx: DeviceContext, physical_dim: int, basis: Sequence[int],
                       dtype: torch.dtype = torch.float64) -> MatrixProductState:
    """Product state |basis[0] basis[1] ...> with all bonds 1 (mps_state.mojo:258-304)."""
    if len(basis) < 1:
        raise ValueError("basis must contain at least one element")
    if physical_dim < 1:
        raise ValueError("physical_dim must be >= 1")
    sites = []
    for i, b in enumerate(basis):
        if not 0 <= b < physical_dim:
            raise ValueError(f"Invalid basis index {b} at site {i} (must be in [0, physical_dim))")
        v = np.zeros((1, physical_dim, 1))
        v[0, b, 0] = 1.0
        sites.append(v)
    return mps_from_tensors(ctx, sites, dtype)


def random_mps_c_compatible_tensors(nsites: int, d: int, chi_max: int, seed: int) -> list[np.ndarray]:
    """Host tensors of create_random_mps_c_compatible (rng_c_compat.mojo:174-231),
    generated by libdmrg_hip.so's PCG32 restatement (host code)."""
    lib = _lib.load()
    dims = (ctypes.c_int64 * (nsites + 1))()
    total = ctypes.c_int64()
    rc = lib.ct_random_mps_c_compatible(nsites, d, chi_max, ctypes.c_uint64(seed), dims, None, ctypes.byref(total))
    if rc != 0:
        raise ValueError(_lib.last_error())
    buf = np.zeros(max(total.value, 1), dtype=np.float64)
    rc = lib.ct_random_mps_c_compatible(nsites, d, chi_max, ctypes.c_uint64(seed), dims,
                                        buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(total))
    if rc != 0:
        raise ValueError(_lib.last_error())
    out, off = [], 0
    for i in range(nsites):
        n = dims[i] * d * dims[i + 1]
        out.append(buf[off:off + n].reshape(dims[i], d, dims[i + 1]).copy())
        off += n
    return out


def create_random_mps_c_compatible(ctx: DeviceContext, nsites: int, d: int, chi_max: int, seed: int,
                                   dtype: torch.dtype = torch.float64) -> MatrixProductState:
    return mps_from_tensors(ctx, random_mps_c_compatible_tensors(nsites, d, chi_max, seed), dtype)


def _identity_env(ctx: DeviceContext, w: int, D: int, dtype) -> DenseTensor:
    t = torch.zeros((w, D, D), dtype=dtype, device=ctx.device)
    t[:, torch.arange(D), torch.arange(D)] = 1.0
    return DenseTensor(t)


def update_left_environment(L_prev: DenseTensor, A_site: MPSSite, W_site: MPOSite, ctx: DeviceContext,
                            prec: int | None = None) -> DenseTensor:
    """L'[wR,Dr,Dr'] = sum L[wL,Dl,Dl'] A[Dl,s,Dr] W[wL,s,s',wR] A[Dl',s',Dr']
    (environments.mojo:99-201)."""
    A, W = A_site.tensor, W_site.tensor
    wL, Dl, Dlb = L_prev.shape
    Dl2, d, Dr = A.shape
    if Dl2 != Dl or Dlb != Dl or W.shape[0] != wL or W.shape[1] != d:
        raise ValueError("update_left_environment: shape mismatch")
    wR = W.shape[3]
    out = torch.empty(wR * Dr * Dr, dtype=A.dtype, device=A.storage.device)
    p = default_prec(A.dtype) if prec is None else prec
    call("ct_env_update_left", ctx.handle, p, A.dtype_code, L_prev.ptr(), A.ptr(), W.ptr(),
         ctypes.c_void_p(out.data_ptr()), Dl, d, Dr, wL, wR)
    return DenseTensor(out, [wR, Dr, Dr])


def update_right_environment(R_next: DenseTensor, A_site: MPSSite, W_site: MPOSite, ctx: DeviceContext,
                             prec: int | None = None) -> DenseTensor:
    """R'[wL,Dl,Dl'] = sum R[wR,Dr,Dr'] A[Dl,s,Dr] W[wL,s,s',wR] A[Dl',s',Dr']
    (environments.mojo:309-417)."""
    A, W = A_site.tensor, W_site.tensor
    wR, Dr, Drb = R_next.shape
    Dl, d, Dr2 = A.shape
    if Dr2 != Dr or Drb != Dr or W.shape[3] != wR or W.shape[1] != d:
        raise ValueError("update_right_environment: shape mismatch")
    wL = W.shape[0]
    out = torch.empty(wL * Dl * Dl, dtype=A.dtype, device=A.storage.device)
    p = default_prec(A.dtype) if prec is None else prec
    call("ct_env_update_right", ctx.handle, p, A.dtype_code, R_next.ptr(), A.ptr(), W.ptr(),
         ctypes.c_void_p(out.data_ptr()), Dl, d, Dr, wL, wR)
    return DenseTensor(out, [wL, Dl, Dl])


def build_right_environments(mps: MatrixProductState, mpo: MatrixProductOperator,
                             ctx: DeviceContext) -> list[DenseTensor]:
    """R[N] = identity, R[i] = update_right_environment(R[i+1], A_i, W_i)
    (environments.mojo:420-472)."""
    N = mps.num_sites()
    dt = mps.sites[0].tensor.dtype
    envs: list[DenseTensor | None] = [None] * (N + 1)
    envs[N] = _identity_env(ctx, mpo.bond_dimension(N), mps.bond_dimension(N), dt)
    for i in range(N - 1, -1, -1):
        envs[i] = update_right_environment(envs[i + 1], mps.sites[i], mpo.sites[i], ctx)
    return envs  # type: ignore[return-value]


def expectation_value(mps: MatrixProductState, mpo: MatrixProductOperator, ctx: DeviceContext) -> float:
    """<mps|mpo|mps> by left-environment sweeps (environments.mojo:475-518)."""
    dt = mps.sites[0].tensor.dtype
    L = _identity_env(ctx, mpo.bond_dimension(0), mps.bond_dimension(0), dt)
    for i in range(mps.num_sites()):
        L = update_left_environment(L, mps.sites[i], mpo.sites[i], ctx)
    if L.size != 1:
        raise ValueError(f"Unexpected final environment size in expectation_value: {L.size}")
    return float(L.to_numpy().reshape(-1)[0])


__all__ = [
    "MPSSite", "MatrixProductState", "MPOSite", "MatrixProductOperator", "mpo_from_tensors", "mps_from_tensors",
    "create_product_mps", "create_random_mps_c_compatible", "random_mps_c_compatible_tensors",
    "update_left_environment", "update_right_environment", "build_right_environments", "expectation_value",
    "PREC_FP64",
]
