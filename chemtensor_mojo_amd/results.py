"""DMRG result and benchmark records in the reference's wire formats.

- ``dmrg_results_to_json_string`` / ``save_dmrg_results_to_json`` write the
  results JSON of src/algorithms/dmrg_results_json.mojo:12-102 (same keys, key
  order and layout: ``params`` with ``tol_split`` omitted when negative,
  ``entropy`` omitted when empty), so files from either implementation can be
  compared with the same tools.
- ``load_dmrg_reference`` reads that schema back
  (src/tests/algorithms/dmrg/dmrg_json_loader.mojo:13-105).
- ``append_bench_record`` appends one JSONL record of the benchmark format of
  src/tests/benchmarks/bench_contractions.mojo:221-253 with backend ``"hip"``.
- ``mps_norm`` and ``bond_entropies`` compute the ``norm`` and ``entropy``
  fields on the GPU.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np
import torch

from .dense_tensor import DenseTensor, DeviceContext
from .state import MatrixProductState, MPOSite, _identity_env, update_left_environment


@dataclass
class DMRGJsonParams:
    """The ``params`` object (dmrg_results_json.mojo:12-24); tol_split < 0 omits the key."""
    nsites: int
    d: int
    J: float
    D: float
    h: float
    num_sweeps: int
    maxiter_lanczos: int
    chi_max: int
    tol_split: float = -1.0


def _fmt_float(x: float) -> str:
    # shortest round-trip decimal form; JSON readers parse it exactly
    return repr(float(x))


def dmrg_results_to_json_string(impl: str, model: str, params: DMRGJsonParams, energy_final: float,
                                en_sweeps: Sequence[float], entropy: Sequence[float], bond_dims: Sequence[int],
                                norm: float) -> str:
    """Results JSON text (dmrg_results_json.mojo:35-79)."""
    out = "{\n"
    out += f'  "impl": "{impl}",\n'
    out += f'  "model": "{model}",\n'
    out += '  "params": {\n'
    out += f'    "nsites": {int(params.nsites)},\n'
    out += f'    "d": {int(params.d)},\n'
    out += f'    "J": {_fmt_float(params.J)},\n'
    out += f'    "D": {_fmt_float(params.D)},\n'
    out += f'    "h": {_fmt_float(params.h)},\n'
    out += f'    "num_sweeps": {int(params.num_sweeps)},\n'
    out += f'    "maxiter_lanczos": {int(params.maxiter_lanczos)},\n'
    out += f'    "chi_max": {int(params.chi_max)}'
    if params.tol_split >= 0.0:
        out += f',\n    "tol_split": {_fmt_float(params.tol_split)}'
    out += "\n  },\n"
    out += f'  "energy_final": {_fmt_float(energy_final)},\n'
    out += '  "en_sweeps": [' + ", ".join(_fmt_float(e) for e in en_sweeps) + "],\n"
    if len(entropy) > 0:
        out += '  "entropy": [' + ", ".join(_fmt_float(e) for e in entropy) + "],\n"
    out += '  "bond_dims": [' + ", ".join(str(int(b)) for b in bond_dims) + "],\n"
    out += f'  "norm": {_fmt_float(norm)}\n'
    out += "}\n"
    return out


def save_dmrg_results_to_json(path: str, impl: str, model: str, params: DMRGJsonParams, energy_final: float,
                              en_sweeps: Sequence[float], entropy: Sequence[float], bond_dims: Sequence[int],
                              norm: float) -> None:
    """Writes the results JSON file (dmrg_results_json.mojo:82-102)."""
    text = dmrg_results_to_json_string(impl, model, params, energy_final, en_sweeps, entropy, bond_dims, norm)
    with open(path, "w") as f:
        f.write(text)


@dataclass
class DMRGReference:
    """Fields of a results JSON file (dmrg_json_loader.mojo:13-34); tol_split = -1 when absent."""
    impl: str
    model: str
    nsites: int
    d: int
    J: float
    D: float
    h: float
    num_sweeps: int
    maxiter_lanczos: int
    chi_max: int
    tol_split: float
    energy_final: float
    en_sweeps: list = field(default_factory=list)
    bond_dims: list = field(default_factory=list)
    norm: float = 1.0
    entropy: list = field(default_factory=list)


def load_dmrg_reference(path: str) -> DMRGReference:
    """Reads a results JSON file written by either implementation (dmrg_json_loader.mojo:48-105)."""
    with open(path) as f:
        data = json.load(f)
    p = data["params"]
    return DMRGReference(
        impl=str(data["impl"]), model=str(data["model"]), nsites=int(p["nsites"]), d=int(p["d"]),
        J=float(p["J"]), D=float(p["D"]), h=float(p["h"]), num_sweeps=int(p["num_sweeps"]),
        maxiter_lanczos=int(p["maxiter_lanczos"]), chi_max=int(p["chi_max"]),
        tol_split=float(p.get("tol_split", -1.0)), energy_final=float(data["energy_final"]),
        en_sweeps=[float(x) for x in data["en_sweeps"]], bond_dims=[int(x) for x in data["bond_dims"]],
        norm=float(data["norm"]), entropy=[float(x) for x in data.get("entropy", [])])


def append_bench_record(path: str, operation: str, nsites: int, d: int, chi_max: int, time_seconds: float,
                        result: float, runs: int, backend: str = "hip") -> None:
    """Appends one benchmark JSONL line (bench_contractions.mojo:221-253), creating the folder."""
    rec = {"backend": backend, "operation": operation, "params": {"nsites": int(nsites), "d": int(d),
                                                                   "chi_max": int(chi_max)},
           "time_seconds": float(time_seconds), "result": float(result), "runs": int(runs)}
    parent = os.path.dirname(path)
    if parent:
        os.makedirs(parent, exist_ok=True)
    with open(path, "a") as f:
        f.write(json.dumps(rec, separators=(",", ":")) + "\n")


def mps_norm(mps: MatrixProductState, ctx: DeviceContext) -> float:
    """sqrt(<psi|psi>) through identity-MPO left environments (mps_state.mojo:258-304)."""
    d = mps.physical_dim
    dt = mps.sites[0].tensor.dtype
    Wid = MPOSite(DenseTensor(torch.eye(d, dtype=dt, device=ctx.device).reshape(1, d, d, 1)))
    L = _identity_env(ctx, 1, mps.bond_dimension(0), dt)
    for s in mps.sites:
        L = update_left_environment(L, s, Wid, ctx)
    return math.sqrt(max(float(L.to_numpy().reshape(-1)[0]), 0.0))


def bond_entropies(mps: MatrixProductState, ctx: DeviceContext) -> list[float]:
    """Von Neumann entanglement entropy -sum p ln p, p = sigma^2 / sum sigma^2, at each of the
    N-1 internal bonds.  A right-to-left SVD pass brings a copy of the state to right-canonical
    form, then a left-to-right pass reads the singular values of each bond (FP64 SVDs and
    contractions through the library; only the diagonal scalings are elementwise tensor ops)."""
    from .dense_tensor import create_dense_tensor, dense_tensor_dot
    from .dmrg import dense_tensor_svd_trunc
    sites = [DenseTensor(s.tensor.storage.to(torch.float64).clone(), list(s.tensor.shape)) for s in mps.sites]
    N = len(sites)
    for i in range(N - 1, 0, -1):  # right-canonical sites 1..N-1, centre on site 0
        Dl, d, Dr = sites[i].shape
        U, S, Vt, c = dense_tensor_svd_trunc(DenseTensor(sites[i].storage, [Dl, d * Dr]), ctx, min(Dl, d * Dr), 0.0)
        sites[i] = DenseTensor(Vt.storage, [c, d, Dr])
        US = DenseTensor((U.storage.view(Dl, c) * S.storage.view(1, c)).reshape(-1), [Dl, c])
        Dp, dp, _ = sites[i - 1].shape
        C = create_dense_tensor(ctx, [Dp, dp, c], dtype=torch.float64)
        dense_tensor_dot(C, sites[i - 1], US, ctx)
        sites[i - 1] = C
    ent = []
    for i in range(N - 1):  # move the centre right, recording each bond's spectrum
        Dl, d, Dr = sites[i].shape
        U, S, Vt, c = dense_tensor_svd_trunc(DenseTensor(sites[i].storage, [Dl * d, Dr]), ctx, min(Dl * d, Dr), 0.0)
        s2 = S.storage.double() ** 2
        p = (s2 / s2.sum()).cpu().numpy()
        p = p[p > 0]
        ent.append(float(-(p * np.log(p)).sum()))
        SV = DenseTensor((S.storage.view(c, 1) * Vt.storage.view(c, Dr)).reshape(-1), [c, Dr])
        _, dn, Dn = sites[i + 1].shape
        C = create_dense_tensor(ctx, [c, dn, Dn], dtype=torch.float64)
        dense_tensor_dot(C, SV, sites[i + 1], ctx)
        sites[i + 1] = C
    return ent


__all__ = ["DMRGJsonParams", "dmrg_results_to_json_string", "save_dmrg_results_to_json", "DMRGReference",
           "load_dmrg_reference", "append_bench_record", "mps_norm", "bond_entropies"]
