"""Dense tensor core on the GPU, mirroring the reference's DenseTensor API.

Reference: src/m_tensor/dense_tensor.mojo.  A ``DenseTensor`` is a row-major
tensor whose storage is a torch CUDA tensor (used only as device memory); all
arithmetic runs in libdmrg_hip.so through the C ABI.  ``DeviceContext``
replaces MAX's DeviceContext for this path: one GPU, one stream, one library
handle.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
import torch

from . import _lib
from ._lib import F32, F64, PREC_BF16X3, PREC_BF16X6, PREC_FP32, PREC_FP64, call


def _dtype_code(t: torch.dtype) -> int:
    if t == torch.float64:
        return F64
    if t == torch.float32:
        return F32
    raise TypeError(f"unsupported dtype {t} (float64 or float32)")


def default_prec(t: torch.dtype) -> int:
    return PREC_FP64 if t == torch.float64 else PREC_FP32


class DeviceContext:
    """One GPU + the torch current stream + a libdmrg_hip handle.

    Mirrors the reference's use of MAX ``DeviceContext`` (e.g.
    src/algorithms/dmrg.mojo:378): every call is ordered on this stream.
    """

    def __init__(self, device: int = 0):
        if not torch.cuda.is_available():
            raise RuntimeError("chemtensor_mojo_amd needs a GPU (torch.cuda.is_available() is False)")
        _lib.load()
        torch.cuda.set_device(device)
        self.device_index = device
        self.device = torch.device("cuda", device)
        self.stream = torch.cuda.current_stream(self.device)
        h = ctypes.c_void_p()
        call("ct_handle_create", device, ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h))
        self.handle = h
        _lib.track(self)

    def synchronize(self) -> None:
        call("ct_synchronize", self.handle)

    def close(self) -> None:
        if self.handle is not None and self.handle.value:
            _lib.load().ct_handle_destroy(self.handle)
            self.handle = None
        _lib.untrack(self)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def compute_row_major_strides(shape: Sequence[int], ndim: int | None = None) -> list[int]:
    """dense_tensor.mojo:767-788."""
    n = len(shape) if ndim is None else ndim
    st = [1] * n
    for i in range(n - 2, -1, -1):
        st[i] = st[i + 1] * int(shape[i + 1])
    return st


class DenseTensor:
    """Row-major tensor on the device (dense_tensor.mojo:263-764).

    ``reshape``/``flatten_dims`` are views sharing storage (as :547-635);
    ``transpose`` materialises a permuted copy (as :460-545)."""

    def __init__(self, storage: torch.Tensor, shape: Sequence[int] | None = None):
        if not storage.is_cuda:
            raise ValueError("DenseTensor storage must be a CUDA tensor")
        self.storage = storage.reshape(-1) if storage.is_contiguous() else storage.contiguous().reshape(-1)
        self.shape = [int(s) for s in (storage.shape if shape is None else shape)]
        if int(np.prod(self.shape)) != self.storage.numel() and not (len(self.shape) == 0 and self.storage.numel() == 1):
            raise ValueError(f"shape {self.shape} does not match storage of {self.storage.numel()} elements")
        self.stride = compute_row_major_strides(self.shape)
        self.size = int(np.prod(self.shape)) if self.shape else 1

    # ---- metadata
    @property
    def dtype(self) -> torch.dtype:
        return self.storage.dtype

    @property
    def dtype_code(self) -> int:
        return _dtype_code(self.storage.dtype)

    def ptr(self) -> ctypes.c_void_p:
        return ctypes.c_void_p(self.storage.data_ptr())

    def rank(self) -> int:
        return len(self.shape)

    def reshape(self, shape: Sequence[int]) -> "DenseTensor":
        """View with a new shape (dense_tensor.mojo:614-635); dims must be >= 1."""
        shape = [int(s) for s in shape]
        if len(shape) == 0:
            raise ValueError("reshape requires rank >= 1")
        if any(s < 1 for s in shape):
            raise ValueError("reshape dimensions must be >= 1")
        if int(np.prod(shape)) != self.size:
            raise ValueError(f"cannot reshape {self.shape} to {shape}")
        return DenseTensor(self.storage, shape)

    def flatten_dims(self, i_start: int, i_end: int, ctx: DeviceContext | None = None) -> "DenseTensor":
        """Merge axes [i_start, i_end) into one (dense_tensor.mojo:547-612)."""
        if not (0 <= i_start < i_end <= len(self.shape)):
            raise ValueError("invalid flatten range")
        new = self.shape[:i_start] + [int(np.prod(self.shape[i_start:i_end]))] + self.shape[i_end:]
        return DenseTensor(self.storage, new)

    def transpose(self, perm: Sequence[int], ctx: DeviceContext) -> "DenseTensor":
        """Permuted copy: output axis i = input axis perm[i] (dense_tensor.mojo:460-545)."""
        perm = [int(p) for p in perm]
        out = torch.empty(self.size, dtype=self.dtype, device=self.storage.device)
        shape = _lib.i64_array(self.shape)
        cperm = (ctypes.c_int * len(perm))(*perm)
        call("ct_permute", ctx.handle, self.dtype_code, self.ptr(), len(self.shape), shape, cperm,
             ctypes.c_void_p(out.data_ptr()))
        return DenseTensor(out, [self.shape[p] for p in perm])

    # ---- BLAS-1 (dense_tensor.mojo:637-754), FP64 accumulation
    def norm_sq(self, ctx: DeviceContext) -> float:
        r = ctypes.c_double()
        call("ct_norm_sq", ctx.handle, self.dtype_code, self.ptr(), self.size, ctypes.byref(r))
        return r.value

    def norm(self, ctx: DeviceContext) -> float:
        return float(np.sqrt(self.norm_sq(ctx)))

    def scale_in_place(self, alpha: float, ctx: DeviceContext) -> None:
        call("ct_scale", ctx.handle, self.dtype_code, float(alpha), self.ptr(), self.size)

    def to_numpy(self) -> np.ndarray:
        torch.cuda.current_stream(self.storage.device).synchronize()
        return self.storage.detach().cpu().numpy().reshape(self.shape if self.shape else (1,)).copy()

    def copy(self) -> "DenseTensor":
        return DenseTensor(self.storage.clone(), self.shape)

    def __repr__(self) -> str:
        return f"DenseTensor(shape={self.shape}, dtype={self.dtype})"


def dot_product(a: DenseTensor, b: DenseTensor, ctx: DeviceContext) -> float:
    """dense_tensor.mojo:706-729."""
    if a.size != b.size or a.dtype != b.dtype:
        raise ValueError("dot_product: size/dtype mismatch")
    r = ctypes.c_double()
    call("ct_dot", ctx.handle, a.dtype_code, a.ptr(), b.ptr(), a.size, ctypes.byref(r))
    return r.value


def axpy_in_place(y: DenseTensor, alpha: float, x: DenseTensor, ctx: DeviceContext) -> None:
    """y += alpha x (dense_tensor.mojo:731-754)."""
    if x.size != y.size or x.dtype != y.dtype:
        raise ValueError("axpy: size/dtype mismatch")
    call("ct_axpy", ctx.handle, y.dtype_code, float(alpha), x.ptr(), y.ptr(), y.size)


# ---------------------------------------------------------------- creation
def create_dense_tensor_uninitialized(ctx: DeviceContext, shape: Sequence[int],
                                      dtype: torch.dtype = torch.float64) -> DenseTensor:
    """dense_tensor.mojo:874-908."""
    n = int(np.prod([int(s) for s in shape])) if len(shape) else 1
    return DenseTensor(torch.empty(n, dtype=dtype, device=ctx.device), shape)


def create_dense_tensor(ctx: DeviceContext, shape: Sequence[int], init_value: float = 0.0,
                        dtype: torch.dtype = torch.float64) -> DenseTensor:
    """dense_tensor.mojo:815-872 (fill with init_value)."""
    n = int(np.prod([int(s) for s in shape])) if len(shape) else 1
    return DenseTensor(torch.full((n,), float(init_value), dtype=dtype, device=ctx.device), shape)


def create_dense_tensor_from_data(ctx: DeviceContext, data, shape: Sequence[int],
                                  dtype: torch.dtype = torch.float64) -> DenseTensor:
    """dense_tensor.mojo:910-961 (host data copied to the device)."""
    arr = np.ascontiguousarray(np.asarray(data, dtype=np.float64).reshape(-1))
    t = torch.from_numpy(arr).to(device=ctx.device, dtype=dtype)
    return DenseTensor(t, shape)


# --------------------------------------------------------------- tensordot
def dense_tensor_dot(C: DenseTensor, A: DenseTensor, B: DenseTensor, ctx: DeviceContext, ndim_mult: int = 1,
                     axrange_A: bool = False, axrange_B: bool = False, prec: int | None = None) -> None:
    """Generalised tensordot into the preallocated C (dense_tensor.mojo:963-1331).

    Contracts ndim_mult leading (axrange=True) or trailing (False) axes of A and
    B, aligned; with both flags False the B side is inferred as the reference
    does (:1141-1159).  C's shape must be exactly A's free axes then B's;
    a scalar result accepts [] or [1].  Raises ValueError on a dimension
    mismatch like the reference's "Dim mismatch"."""
    if not (A.dtype == B.dtype == C.dtype):
        raise TypeError("dense_tensor_dot: A, B and C must share a dtype")
    p = default_prec(A.dtype) if prec is None else prec
    sA, sB, sC = _lib.i64_array(A.shape), _lib.i64_array(B.shape), _lib.i64_array(C.shape)
    try:
        call("ct_tensordot", ctx.handle, p, A.dtype_code, A.ptr(), len(A.shape), sA, B.ptr(), len(B.shape), sB,
             int(ndim_mult), int(bool(axrange_A)), int(bool(axrange_B)), C.ptr(), len(C.shape), sC)
    except _lib.CTError as e:
        if e.status < 0 or e.status == 6:
            raise ValueError(str(e)) from None
        raise


__all__ = [
    "DeviceContext", "DenseTensor", "dense_tensor_dot", "dot_product", "axpy_in_place",
    "create_dense_tensor", "create_dense_tensor_uninitialized", "create_dense_tensor_from_data",
    "compute_row_major_strides", "PREC_FP64", "PREC_FP32", "PREC_BF16X3", "PREC_BF16X6",
]
