"""bench.py's accounting, on the CPU: the algorithmic FLOP counts of the
matvec and the environment updates (the reference contraction order,
dmrg.mojo:310-372 and environments.mojo:99-201), the HBM-traffic figures it
reads from the committed counter summary, and the one-GPU contract."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _matvec_flops_by_steps(Dl, d, Dr, wL, wM, wR):
    """Step by step, as the reference contracts: T1 = L.theta, then W_i, W_{i+1}, R."""
    s1 = 2 * wL * Dl * Dl * (d * d * Dr)        # L[wL, Dl, Dl'] . theta[Dl, d, d, Dr]
    s2 = 2 * (Dl * d * Dr) * (wL * d) * (d * wM)  # contract wL, s1 with W_i
    s3 = 2 * (Dl * Dr * d) * (wM * d) * (d * wR)  # contract wM, s2 with W_{i+1}
    s4 = 2 * (Dl * d * d) * (wR * Dr) * Dr        # contract wR, Dr with R
    return s1 + s2 + s3 + s4


@pytest.mark.parametrize("w,gflop", [(5, 86.74), (8, 139.59)])
def test_matvec_flops(w, gflop):
    F = bench.matvec_flops(1024, 2, 1024, w, w, w)
    assert F == _matvec_flops_by_steps(1024, 2, 1024, w, w, w)
    assert abs(F / 1e9 - gflop) < 0.005
    # closed form with all bonds equal: 4 w d^2 D^3 + 4 w^2 d^3 D^2
    D, d = 1024, 2
    assert F == 4 * w * d * d * D ** 3 + 4 * w * w * d ** 3 * D ** 2


@pytest.mark.parametrize("w,gflop", [(5, 43.16), (8, 69.26)])
def test_env_flops(w, gflop):
    D, d = 1024, 2
    by_steps = 2 * w * D * D * d * D + 2 * D * D * w * d * d * w + 2 * D * w * D * d * D
    assert bench.env_flops(D, d, w) == by_steps
    assert abs(bench.env_flops(D, d, w) / 1e9 - gflop) < 0.005


def test_profiled_traffic_reads_the_committed_summary():
    keys = ("fp32_stage4_gemm", "fp64_stage4_gemm", "bf16x3_stage4_gemm")
    vals = [bench.profiled_traffic(k, 1024, 8) for k in keys]
    assert all(v is not None and v > 0 for v in vals)
    # every figure is at least the algorithmic bytes of its product
    D, w = 1024, 8
    M, K, N = 4 * D, w * D, D
    alg = {"fp32_stage4_gemm": 4 * (M * K + K * N + M * N), "fp64_stage4_gemm": 8 * (M * K + K * N + M * N),
           "bf16x3_stage4_gemm": 4 * (M * K + K * N) + 8 * M * N}
    for k, v in zip(keys, vals):
        assert v >= alg[k]
    assert bench.profiled_traffic("fp32_stage4_gemm", 512, 8) is None
    assert bench.profiled_traffic("no_such_kernel", 1024, 8) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
