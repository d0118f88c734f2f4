"""Benchmark of the two-site DMRG hot path on one MI355X.

Workload (J1-J2 chain, MPO bond dimension w = 8, physical dim d = 2, bond
dimension D = 1024 - the precision configuration of the reference's
benchmark set): one "step" is one two-site H_eff matvec
(apply_two_site_heff, reference src/algorithms/dmrg.mojo:254-374) on
HBM-resident L, R, W and theta.  The matvec is timed in FP64 (the
oracle-equivalent path), FP32, 3xBF16 and 6xBF16 (FP32-accurate products on
the BF16 matrix cores); ``value`` is the FP32 matvec rate
(the reference computes in float32) in algorithmic TFLOP/s (reference
contraction order, 139.59 GFLOP per matvec) of the prepared operator the
Lanczos uses (W_i W_{i+1} and the L/R copies built once per local problem);
the one-shot rate, which rebuilds them every call, is reported beside it.

Also reported: the roofline of the dominant kernel of each form (the last
GEMM of the matvec, U[(Dl d d)][(wR Dr)] R[(wR Dr)][Dr']) timed live with HIP
events on the stream it runs on; fixed-D D=1024 DMRG sweeps in FP64 and in
the refined FP32 / 3xBF16 / 6xBF16 forms on two chains - one whose local problems are
converged after the first sweep (Majumdar-Ghosh point) and one whose Lanczos
runs its whole Krylov space (antiferromagnetic Heisenberg); the symmetric
eigensolver of the truncation; the matvec grid over D and w; the environment
updates and the theta build at D=1024; whole DMRG runs
of two benchmark configurations with per-stage times (XXZ antiferromagnet L=64,
chi <= 256, FP64; J1-J2 L=100, chi <= 1024, all four forms); and the FP64 CPU
oracle timed on the host cores.

Prints one JSON line.  N > 1 GPUs is not supported (the path runs on one GPU).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DMRG sweep s + H_eff matvec TFLOP/s at D=1024; |ΔE/E| vs FP64 oracle"
PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}  # TFLOP/s, MI355X dense MFMA
HBM_PEAK = 8000.0  # GB/s
WARM_S = 0.25  # untimed sustained work before each dominant-GEMM timing


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def matvec_flops(Dl, d, Dr, wL, wM, wR):
    return (2.0 * wL * Dl * Dl * d * d * Dr + 2.0 * Dl * d * Dr * wL * d * d * wM
            + 2.0 * Dl * Dr * d * wM * d * d * wR + 2.0 * Dl * d * d * wR * Dr * Dr)


def make_problem(D, w, seed=42):
    from chemtensor_mojo_amd import hamiltonians as ham
    rng = np.random.default_rng(seed)
    L = rng.standard_normal((w, D, D)) / np.sqrt(D)
    L = 0.5 * (L + L.transpose(0, 2, 1))
    R = rng.standard_normal((w, D, D)) / np.sqrt(D)
    R = 0.5 * (R + R.transpose(0, 2, 1))
    # bulk tensors: J1-J2 [8, 2, 2, 8] or XXZ (J=1, Delta=1) [5, 2, 2, 5]
    W = ham.j1j2_tensors(6, 1.0, 0.5)[2] if w == 8 else ham.heisenberg_xxz_tensors(6, 1.0, 1.0, 0.0)[2]
    assert W.shape[0] == w
    th = rng.standard_normal((D, 2, 2, D))
    th /= np.linalg.norm(th)
    return L, R, W, th


def time_events(fn, steps, warmup, min_warm_s=0.0):
    """Mean seconds per call of fn over `steps` calls (HIP events on the current stream) after `warmup` untimed
    calls, continued until at least min_warm_s of them have run (the GPU clock settles over the first tens of
    milliseconds of sustained MFMA work)."""
    import torch
    t_w = time.perf_counter()
    n = 0
    while n < warmup or time.perf_counter() - t_w < min_warm_s:
        fn()
        n += 1
        if min_warm_s > 0.0 and n % 8 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(steps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / 1e3 / steps, wall


def bench_matvec(ctx, D, w, steps, warmup):
    import torch
    import chemtensor_mojo_amd as ct
    L, R, W, th = make_problem(D, w)
    dev = lambda a: ct.create_dense_tensor_from_data(ctx, a, a.shape)
    tL, tR, tW, tth = dev(L), dev(R), ct.MPOSite(dev(W)), dev(th)
    F = matvec_flops(D, 2, D, w, w, w)
    res = {}
    outs = {}
    for name, prec in (("fp64", ct.PREC_FP64), ("fp32", ct.PREC_FP32), ("bf16x3", ct.PREC_BF16X3),
                       ("bf16x6", ct.PREC_BF16X6)):
        out_holder = {}

        # one-shot apply_two_site_heff: W_i W_{i+1} and the L/R copies are rebuilt every call
        def step():
            out_holder["o"] = ct.apply_two_site_heff(tth, tL, tR, tW, tW, ctx, prec=prec)
        t1, _ = time_events(step, steps, warmup)
        # prepared operator, as inside the Lanczos: environments prepared once per local problem
        op = ct.HeffOperator(tL, tR, tW, tW, ctx, prec=prec)
        buf = torch.empty(D * 4 * D, dtype=torch.float64, device="cuda")
        t, _ = time_events(lambda: op.apply(tth, buf), steps, warmup)
        op.close()
        outs[name] = out_holder["o"].to_numpy()
        res[name] = {"s_per_matvec": t, "tflops": F / t / 1e12, "s_per_matvec_one_shot": t1,
                     "tflops_one_shot": F / t1 / 1e12}
        log(f"matvec {name}: {t * 1e3:.3f} ms  {F / t / 1e12:.2f} TFLOP/s (one-shot incl. environment "
            f"preparation: {t1 * 1e3:.3f} ms, {F / t1 / 1e12:.2f} TFLOP/s)")
    ref = outs["fp64"]
    for name in ("fp32", "bf16x3", "bf16x6"):
        res[name]["rel_err_vs_fp64"] = float(np.linalg.norm(outs[name] - ref) / np.linalg.norm(ref))
    del tL, tR, tth
    torch.cuda.empty_cache()
    return res, F


def bench_matvec_mpo(ctx, D, Wi, Wj, steps, warmup, seed=7):
    """Prepared-operator matvec of given MPO tensors at bond dimension D (random symmetric environments), four
    forms; returns {form: {"s_per_matvec", "tflops"}} and the relative errors against FP64."""
    import torch
    import chemtensor_mojo_amd as ct
    rng = np.random.default_rng(seed)
    wL, wR = Wi.shape[0], Wj.shape[3]
    L = rng.standard_normal((wL, D, D)) / np.sqrt(D)
    L = 0.5 * (L + L.transpose(0, 2, 1))
    R = rng.standard_normal((wR, D, D)) / np.sqrt(D)
    R = 0.5 * (R + R.transpose(0, 2, 1))
    th = rng.standard_normal((D, 2, 2, D))
    th /= np.linalg.norm(th)
    dev = lambda a: ct.create_dense_tensor_from_data(ctx, a, a.shape)
    tL, tR, tth = dev(L), dev(R), dev(th)
    F = matvec_flops(D, 2, D, wL, Wi.shape[3], wR)
    res, outs = {}, {}
    for name, prec in (("fp64", ct.PREC_FP64), ("fp32", ct.PREC_FP32), ("bf16x3", ct.PREC_BF16X3),
                       ("bf16x6", ct.PREC_BF16X6)):
        op = ct.HeffOperator(tL, tR, ct.MPOSite(dev(Wi)), ct.MPOSite(dev(Wj)), ctx, prec=prec)
        buf = torch.empty(D * 4 * D, dtype=torch.float64, device="cuda")
        t, _ = time_events(lambda: op.apply(tth, buf), steps, warmup)
        outs[name] = buf.cpu().numpy()
        op.close()
        res[name] = {"s_per_matvec": t, "tflops": F / t / 1e12}
    for name in ("fp32", "bf16x3", "bf16x6"):
        res[name]["rel_err_vs_fp64"] = float(np.linalg.norm(outs[name] - outs["fp64"]) / np.linalg.norm(outs["fp64"]))
    del tL, tR, tth
    torch.cuda.empty_cache()
    return res, F


def env_flops(D, d, w):
    """Algorithmic FLOPs of update_left/right_environment with all bonds D (environments.mojo:99-201, 309-417):
    2 wL Dl Dl' d Dr + 2 Dl' Dr wL d^2 wR + 2 Dr wR Dl' d Dr' = 4 w d D^3 + 2 w^2 d^2 D^2."""
    return 4.0 * w * d * D ** 3 + 2.0 * w * w * d * d * D ** 2


def bench_env_theta(ctx, steps, warmup, D=1024, ws=(5, 8)):
    """Left / right environment updates (FP64, FP32, 3xBF16, 6xBF16 contractions on FP64 storage) and the two-site theta
    build (FP64) at bond dimension D, seeded random site tensor and environment, XXZ (w=5) / J1-J2 (w=8) bulk
    MPO; mean of `steps` calls timed with HIP events."""
    import torch
    import chemtensor_mojo_amd as ct
    from chemtensor_mojo_amd import hamiltonians as ham
    rng = np.random.default_rng(11)
    dev = lambda a: ct.create_dense_tensor_from_data(ctx, a, a.shape)
    A = dev(rng.standard_normal((D, 2, D)) / np.sqrt(2 * D))
    out = {}
    for w in ws:
        W = ham.j1j2_tensors(6, 1.0, 0.5)[2] if w == 8 else ham.heisenberg_xxz_tensors(6, 1.0, 1.0, 0.0)[2]
        E = rng.standard_normal((w, D, D)) / np.sqrt(D)
        E = dev(0.5 * (E + E.transpose(0, 2, 1)))
        site, mpo = ct.MPSSite(A), ct.MPOSite(dev(W))
        F = env_flops(D, 2, w)
        row = {"gflop": round(F / 1e9, 3)}
        for name, prec in (("fp64", ct.PREC_FP64), ("fp32", ct.PREC_FP32), ("bf16x3", ct.PREC_BF16X3),
                           ("bf16x6", ct.PREC_BF16X6)):
            tl, _ = time_events(lambda: ct.update_left_environment(E, site, mpo, ctx, prec=prec), steps, warmup)
            tr, _ = time_events(lambda: ct.update_right_environment(E, site, mpo, ctx, prec=prec), steps, warmup)
            row[name] = {"left_ms": round(tl * 1e3, 4), "left_tflops": round(F / tl / 1e12, 2),
                         "right_ms": round(tr * 1e3, 4), "right_tflops": round(F / tr / 1e12, 2)}
            log(f"env update w={w} D={D} {name}: left {tl * 1e3:.3f} ms ({F / tl / 1e12:.1f} TFLOP/s), "
                f"right {tr * 1e3:.3f} ms ({F / tr / 1e12:.1f} TFLOP/s)")
        out[f"w{w}"] = row
    Ft = 2.0 * (2 * D) * D * (2 * D)
    tt, _ = time_events(lambda: ct.build_two_site_theta(ct.MPSSite(A), ct.MPSSite(A), ctx), steps, warmup)
    out["theta_build_fp64"] = {"gflop": round(Ft / 1e9, 3), "ms": round(tt * 1e3, 4), "tflops": round(Ft / tt / 1e12, 2)}
    log(f"theta build D={D} fp64: {tt * 1e3:.3f} ms ({Ft / tt / 1e12:.1f} TFLOP/s)")
    torch.cuda.empty_cache()
    return out


def bench_matvec_grid(ctx, steps, warmup, Ds=(256, 512, 1024, 2048), ws=(5, 8)):
    """Prepared-operator matvec rate (TFLOP/s, mean over `steps` calls timed with HIP events) over D and w: the
    microbenchmark grid (XXZ bulk MPO for w=5, J1-J2 bulk MPO for w=8)."""
    import torch
    import chemtensor_mojo_amd as ct
    grid = {}
    for w in ws:
        for D in Ds:
            L, R, W, th = make_problem(D, w)
            dev = lambda a: ct.create_dense_tensor_from_data(ctx, a, a.shape)
            tL, tR, tW, tth = dev(L), dev(R), ct.MPOSite(dev(W)), dev(th)
            F = matvec_flops(D, 2, D, w, w, w)
            row = {"gflop": round(F / 1e9, 3)}
            for name, prec in (("fp64", ct.PREC_FP64), ("fp32", ct.PREC_FP32), ("bf16x3", ct.PREC_BF16X3),
                               ("bf16x6", ct.PREC_BF16X6)):
                op = ct.HeffOperator(tL, tR, tW, tW, ctx, prec=prec)
                buf = torch.empty(D * 4 * D, dtype=torch.float64, device="cuda")
                t, _ = time_events(lambda: op.apply(tth, buf), steps, warmup)
                op.close()
                row[name + "_ms"] = round(t * 1e3, 4)
                row[name + "_tflops"] = round(F / t / 1e12, 2)
            log(f"matvec grid w={w} D={D}: fp64 {row['fp64_tflops']} fp32 {row['fp32_tflops']} bf16x3 "
                f"{row['bf16x3_tflops']} bf16x6 {row['bf16x6_tflops']} TFLOP/s")
            grid[f"w{w}_D{D}"] = row
            del tL, tR, tth
            torch.cuda.empty_cache()
    return grid


def bench_dominant_gemm(ctx, D, w, steps, warmup):
    """Stage-4 GEMM of the matvec alone: [4D x wD] @ [wD x D], the same kernels."""
    import torch
    import chemtensor_mojo_amd as ct
    from chemtensor_mojo_amd import _lib
    M, K, N = 4 * D, w * D, D
    F = 2.0 * M * N * K
    g = torch.Generator(device="cuda").manual_seed(1)
    A64 = torch.randn(M * K, dtype=torch.float64, device="cuda", generator=g)
    B64 = torch.randn(K * N, dtype=torch.float64, device="cuda", generator=g)
    out = {}
    # FP64: ct_tensordot of A[M,K] . B[K,N] (A k-contiguous, B n-contiguous: the heff stage-4 kernel)
    tA = ct.DenseTensor(A64, [M, K])
    tB = ct.DenseTensor(B64, [K, N])
    tC = ct.create_dense_tensor(ctx, [M, N])
    t, _ = time_events(lambda: ct.dense_tensor_dot(tC, tA, tB, ctx), steps, warmup, WARM_S)
    out["fp64"] = {"s": t, "tflops": F / t / 1e12, "bytes": 8.0 * (M * K + K * N + M * N)}
    # FP32 (the reference's precision): the same product on float32 copies
    tA32 = ct.DenseTensor(A64.float(), [M, K])
    tB32 = ct.DenseTensor(B64.float(), [K, N])
    tC32 = ct.create_dense_tensor(ctx, [M, N], dtype=torch.float32)
    t, _ = time_events(lambda: ct.dense_tensor_dot(tC32, tA32, tB32, ctx), steps, warmup, WARM_S)
    out["fp32"] = {"s": t, "tflops": F / t / 1e12, "bytes": 4.0 * (M * K + K * N + M * N)}
    # 3xBF16: pre-split operands (split rows), time only the NT MFMA kernels
    ld = 2 * ((K + 31) // 32 * 32)
    As = torch.empty(M * ld, dtype=torch.int16, device="cuda")
    Bs = torch.empty(N * ld, dtype=torch.int16, device="cuda")
    vp = lambda x: ctypes.c_void_p(x.data_ptr())
    _lib.call("ct_split_bf16", ctx.handle, 0, vp(A64), M, K, K, 1, vp(As), ld)
    _lib.call("ct_split_bf16", ctx.handle, 0, vp(B64), N, K, 1, N, vp(Bs), ld)
    C32 = torch.empty(M * N, dtype=torch.float64, device="cuda")
    t, _ = time_events(lambda: _lib.call("ct_gemm_bf16x3_nt", ctx.handle, vp(As), ld, vp(Bs), ld, 0, vp(C32), M, N, K,
                                         N), steps, warmup, WARM_S)
    out["bf16x3"] = {"s": t, "tflops": F / t / 1e12, "mfma_tflops": 3 * F / t / 1e12,
                     "bytes": 2.0 * ld * (M + N) + 8.0 * M * N}
    # 6xBF16: the same split rows plus the l planes, time only the NT MFMA kernels
    ldl = (K + 31) // 32 * 32
    Al = torch.empty(M * ldl, dtype=torch.int16, device="cuda")
    Bl = torch.empty(N * ldl, dtype=torch.int16, device="cuda")
    _lib.call("ct_split_bf16x6", ctx.handle, 0, vp(A64), M, K, K, 1, vp(As), ld, vp(Al), ldl)
    _lib.call("ct_split_bf16x6", ctx.handle, 0, vp(B64), N, K, 1, N, vp(Bs), ld, vp(Bl), ldl)
    C6 = torch.empty(M * N, dtype=torch.float64, device="cuda")
    t, _ = time_events(lambda: _lib.call("ct_gemm_bf16x6_nt", ctx.handle, vp(As), ld, vp(Al), ldl, vp(Bs), ld, vp(Bl),
                                         ldl, 0, vp(C6), M, N, K, N), steps, warmup, WARM_S)
    out["bf16x6"] = {"s": t, "tflops": F / t / 1e12, "mfma_tflops": 6 * F / t / 1e12,
                     "bytes": 3.0 * ldl * (M + N) + 8.0 * M * N}
    ref = (A64.view(M, K) @ B64.view(K, N)).reshape(-1)
    out["bf16x3"]["rel_err"] = float(torch.linalg.norm(C32 - ref) / torch.linalg.norm(ref))
    out["bf16x6"]["rel_err"] = float(torch.linalg.norm(C6 - ref) / torch.linalg.norm(ref))
    out["fp64"]["rel_err"] = float(torch.linalg.norm(tC.storage - ref) / torch.linalg.norm(ref))
    out["fp32"]["rel_err"] = float(torch.linalg.norm(tC32.storage.double() - ref) / torch.linalg.norm(ref))
    for k, v in out.items():
        log(f"dominant GEMM {k}: {v['s'] * 1e3:.3f} ms {v['tflops']:.2f} TFLOP/s (err {v['rel_err']:.1e})")
    # Calibration stream for hardware byte counters: ct_scale in place over
    # 64 MiB of FP64 reads and writes exactly 64 MiB each per launch.
    cal = torch.ones(8 << 20, dtype=torch.float64, device="cuda")
    for _ in range(3):
        _lib.call("ct_scale", ctx.handle, 0, 1.0, vp(cal), cal.numel())
    torch.cuda.synchronize()
    del cal
    del A64, B64, As, Bs, Al, Bl, C32, C6, tC, tA32, tB32, tC32
    torch.cuda.empty_cache()
    return out, F


def bench_sweep(ctx, L_sites, D, prec, refine, J2=0.5, seed=42, sweeps=2, max_krylov=20, model="j1j2"):
    """Fixed-D two-site sweeps (eps_trunc = 0, chi_max = D) of the J1-J2 chain
    (model "j1j2") or the antiferromagnetic Heisenberg chain ("heisenberg",
    +sum(XX+YY+ZZ), w=5) from a seeded PCG32 random MPS whose bonds already sit
    at min(2^l, 2^(L-l), D); per-sweep wall and stage times."""
    import chemtensor_mojo_amd as ct
    from chemtensor_mojo_amd import hamiltonians as ham
    mpo_t = ham.j1j2_tensors(L_sites, 1.0, J2) if model == "j1j2" else ham.heisenberg_tensors(L_sites, 1.0)
    mpo = ct.mpo_from_tensors(ctx, mpo_t)
    mps = ct.create_random_mps_c_compatible(ctx, L_sites, 2, D, seed)
    p = ct.DMRGParams(num_sweeps=1, chi_max=D, eps_trunc=0.0, max_krylov_iter=max_krylov, krylov_tol=1e-12,
                      energy_tol=0.0, prec=prec, refine=refine, jacobi_max_rot=-1, svd_mode=ct.SVD_GRAM)
    run = ct.DMRGRun(ctx, mpo, mps, p)
    keys = ("t_lanczos", "t_svd", "t_env", "t_theta", "matvecs", "matvec_flops")
    try:
        Es, times, stages = [], [], []
        prev = {k: 0.0 for k in keys}
        for s in range(sweeps):
            t0 = time.perf_counter()
            E, _ = run.run(1)
            times.append(time.perf_counter() - t0)
            Es.append(E)
            st = run.stats()
            cur = {k: float(getattr(st, k)) for k in keys}
            stages.append({k: cur[k] - prev[k] for k in keys})
            prev = cur
            log(f"sweep {s} prec={prec} refine={refine}: {times[-1]:.3f} s E={E:.14f} maxD={st.max_bond} "
                f"lanczos={stages[-1]['t_lanczos']:.3f} svd={stages[-1]['t_svd']:.3f} env={stages[-1]['t_env']:.3f} "
                f"matvecs={int(stages[-1]['matvecs'])}")
        st = run.stats()
    finally:
        run.close()
    per = [{"s": times[i], "lanczos_s": g["t_lanczos"], "svd_s": g["t_svd"], "env_s": g["t_env"],
            "theta_s": g["t_theta"], "matvecs": int(g["matvecs"]),
            "matvec_tflops_in_sweep": (g["matvec_flops"] / g["t_lanczos"] / 1e12 if g["t_lanczos"] > 0 else None)}
           for i, g in enumerate(stages)]
    return {"energies": Es, "sweep_s": times, "per_sweep": per, "last_sweep": per[-1],
            "max_bond": st.max_bond, "trunc_weight_sum": st.trunc_weight_sum, "kept_start": st.kept_start}


# ------------------------------------------------------------------ studies
# Larger workloads run on request (--study NAME); each prints one JSON line.

def _sweep_deltas(run, nsweeps, tag):
    keys = ("t_lanczos", "t_svd", "t_env", "t_theta", "matvecs", "trunc_weight_sum")
    prev = {k: 0.0 for k in keys}
    out = []
    for s in range(nsweeps):
        t0 = time.perf_counter()
        E, _ = run.run(1)
        wall = time.perf_counter() - t0
        st = run.stats()
        cur = {k: float(getattr(st, k)) for k in keys}
        row = {"sweep": s, "E": E, "s": wall, "max_bond": st.max_bond}
        row.update({k: cur[k] - prev[k] for k in keys})
        prev = cur
        out.append(row)
        log(f"{tag} sweep {s}: {wall:.3f} s E={E:.14f} maxD={st.max_bond} lanczos={row['t_lanczos']:.3f} "
            f"svd={row['t_svd']:.3f} env={row['t_env']:.3f} matvecs={int(row['matvecs'])} "
            f"tw={row['trunc_weight_sum']:.3e}")
    return out


def config2_sweeps(ctx, modes):
    """Heisenberg XXZ chain L=64 (antiferromagnetic sign J=-1 of
    H = -J sum(XX+YY+Delta ZZ), so the bond dimension grows), Delta=1, w=5,
    PCG32 random start D=16, chi_max=256, eps_trunc=1e-10, 6 FP64 sweeps, 20
    Krylov vectors with converged tridiagonal solves (the reference's
    200-rotation cap stalls 20-vector runs at rounding-dependent energies);
    per sweep: wall time, Lanczos (H_eff matvecs and the tridiagonal solves),
    truncated split and environment update times."""
    import chemtensor_mojo_amd as ct
    from chemtensor_mojo_amd import hamiltonians as ham
    L, chi, sweeps = 64, 256, 6
    mpo_t = ham.heisenberg_xxz_tensors(L, -1.0, 1.0, 0.0)
    mps_t = ct.random_mps_c_compatible_tensors(L, 2, 16, 42)
    res = {}
    for mode, name in modes:
        p = ct.DMRGParams(num_sweeps=1, chi_max=chi, eps_trunc=1e-10, max_krylov_iter=20, krylov_tol=1e-10,
                          energy_tol=0.0, svd_mode=mode, jacobi_max_rot=-1)
        run = ct.DMRGRun(ctx, ct.mpo_from_tensors(ctx, mpo_t), ct.mps_from_tensors(ctx, mps_t), p)
        try:
            res[name] = _sweep_deltas(run, sweeps, name)
        finally:
            run.close()
    return mpo_t, mps_t, res


def config3_sweeps(ctx):
    """J1-J2 chain (J1=1, J2=0.5, w=8) L=100, PCG32 random start D=32,
    chi_max=1024, eps_trunc=1e-12, 8 sweeps, 20 Krylov vectors, converged
    tridiagonal solves; FP64 vs FP32, 3xBF16 and 6xBF16 matvecs with FP64
    refinement.  J2=0.5 is the Majumdar-Ghosh point, E0 = -3L/2 = -150 exactly."""
    import chemtensor_mojo_amd as ct
    from chemtensor_mojo_amd import hamiltonians as ham
    L, chi, sweeps = 100, 1024, 8
    mpo_t = ham.j1j2_tensors(L, 1.0, 0.5)
    mps_t = ct.random_mps_c_compatible_tensors(L, 2, 32, 42)
    res = {}
    for name, prec, refine in (("fp64", ct.PREC_FP64, 0), ("fp32_refined", ct.PREC_FP32, 2),
                               ("bf16x3_refined", ct.PREC_BF16X3, 2), ("bf16x6_refined", ct.PREC_BF16X6, 2)):
        p = ct.DMRGParams(num_sweeps=1, chi_max=chi, eps_trunc=1e-12, max_krylov_iter=20, krylov_tol=1e-12,
                          energy_tol=0.0, prec=prec, refine=refine, jacobi_max_rot=-1)
        run = ct.DMRGRun(ctx, ct.mpo_from_tensors(ctx, mpo_t), ct.mps_from_tensors(ctx, mps_t), p)
        try:
            res[name] = _sweep_deltas(run, sweeps, name)
        finally:
            run.close()
    E64 = res["fp64"][-1]["E"]
    tw64 = sum(r["trunc_weight_sum"] for r in res["fp64"])
    summary = {}
    for k, v in res.items():
        tw = sum(r["trunc_weight_sum"] for r in v)
        rise = max([0.0] + [v[i + 1]["E"] - v[i]["E"] for i in range(len(v) - 1)])
        summary[k] = {"E": v[-1]["E"], "dE_rel_vs_fp64": abs(v[-1]["E"] - E64) / abs(E64),
                      "dE_rel_vs_exact": abs(v[-1]["E"] + 1.5 * L) / (1.5 * L),
                      "max_sweep_rise_rel": rise / (1.5 * L),
                      "trunc_weight_sum": tw, "d_trunc_weight_vs_fp64": tw - tw64,
                      "max_bond": v[-1]["max_bond"], "total_s": sum(r["s"] for r in v),
                      "lanczos_s": sum(r["t_lanczos"] for r in v), "svd_s": sum(r["t_svd"] for r in v)}
        log(f"{k}: E={v[-1]['E']:.14f} dE/E vs fp64 {summary[k]['dE_rel_vs_fp64']:.2e} "
            f"vs exact {summary[k]['dE_rel_vs_exact']:.2e} tw={tw:.3e} time={summary[k]['total_s']:.2f} s")
    return summary, res


def study_stages(ctx, args):
    """Plain FP64 DMRG on the GPU next to the FP64 CPU oracle (both split
    modes on the GPU); the oracle runs the same sweeps on all host cores."""
    import chemtensor_mojo_amd as ct
    sweeps = 6
    modes = {"gram": (ct.SVD_GRAM, "gpu_fp64_gram"), "gesvd": (ct.SVD_GESVD, "gpu_fp64_gesvd")}
    mpo_t, mps_t, res = config2_sweeps(ctx, tuple(modes[m] for m in args.svd_modes.split(",")))
    chi = 256
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as ora
        ora.build()
        ora.set_num_threads(os.cpu_count() or 1)
        cur, rows = mps_t, []
        for sw in range(sweeps):  # one sweep per call, each restarting from the previous result
            t0 = time.perf_counter()
            ref = ora.dmrg_two_site(cur, mpo_t, 1, chi, 1e-10, 20, 1e-10, 0.0, jacobi_max_rot=20000)
            tc = time.perf_counter() - t0
            cur = ref["mps"]
            rows.append({"sweep": sw, "s": tc, "E": float(ref["sweep_E"][0]),
                         "max_bond": int(max(a.shape[2] for a in cur))})
            log(f"cpu oracle sweep {sw}: {tc:.2f} s E={rows[-1]['E']:.12f} maxD={rows[-1]['max_bond']}")
        cpu = {"value": rows[-1]["s"], "unit": "s (last sweep)", "cores": ora.get_num_threads(), "kind": "port",
               "sample": f"{sweeps} sweeps of the same run, FP64 oracle (OpenBLAS dgemm + LAPACKE dgesdd), "
                         "one oracle call per sweep", "sweeps": rows}
    g = res.get("gpu_fp64_gram") or next(iter(res.values()))
    line = {"study": "fp64_stages", "workload": " ".join((config2_sweeps.__doc__ + " " + study_stages.__doc__).split()),
            "sweeps": res, "cpu_oracle": cpu,
            "speedup_last_sweep_vs_cpu": (cpu["value"] / g[-1]["s"]) if cpu else None}
    print(json.dumps(line), flush=True)


def study_precision(ctx, args):
    """Precision study (the same runs as the default line's config-3 entry)."""
    summary, res = config3_sweeps(ctx)
    print(json.dumps({"study": "precision", "workload": " ".join(config3_sweeps.__doc__.split()),
                      "summary": summary, "sweeps": res}), flush=True)


def study_adjoint(ctx, args):
    """dE/dJ_b for a random-bond Heisenberg chain: L=64, J_b ~ U[0.5, 1.5]
    (numpy default_rng(64)), w=5, chi_max=512, FP32-refined forward (6 sweeps
    from a PCG32 random D=16 start), gradient of all 63 couplings through the
    adjoint kernels behind torch.autograd; checked against Hellmann-Feynman
    <psi|dH/dJ_b|psi> on every bond and against central differences (h=1e-5,
    FP64 sweeps warm-started from the converged MPS) on three bonds."""
    import torch
    import chemtensor_mojo_amd as ct
    from chemtensor_mojo_amd import adjoint as adj
    from chemtensor_mojo_amd import hamiltonians as ham
    L, chi = 64, 512
    J0 = np.random.default_rng(64).uniform(0.5, 1.5, L - 1)
    mps0 = ct.random_mps_c_compatible_tensors(L, 2, 16, 42)
    model = adj.random_bond_heisenberg_model(L)
    times = {}
    last = {}

    def solve(J):
        p = ct.DMRGParams(num_sweeps=6, chi_max=chi, eps_trunc=1e-12, max_krylov_iter=20, krylov_tol=1e-12,
                          energy_tol=0.0, prec=ct.PREC_FP32, refine=2, jacobi_max_rot=-1)
        E, mps = ct.dmrg_two_site(ctx, ct.mpo_from_tensors(ctx, ham.random_bond_heisenberg_tensors(J)),
                                  ct.mps_from_tensors(ctx, mps0), p)
        last["mps"] = mps
        return E, mps

    J = torch.tensor(J0, dtype=torch.float64, requires_grad=True)
    t0 = time.perf_counter()
    E = adj.dmrg_energy(J, solve, model, ctx)
    times["forward_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    E.backward()
    times["backward_s"] = time.perf_counter() - t0
    g = J.grad.numpy().copy()
    log(f"forward {times['forward_s']:.1f} s E={E.item():.12f}; backward {times['backward_s']:.2f} s")
    # Hellmann-Feynman on the same state: H is linear in J, so dH/dJ_b = H(e_b)
    mps = last["mps"]
    t0 = time.perf_counter()
    hf = np.array([adj.expectation(ctx, mps, ham.random_bond_heisenberg_tensors(np.eye(L - 1)[b]))
                   for b in range(L - 1)])
    times["hellmann_feynman_s"] = time.perf_counter() - t0
    hf_rel = float(np.max(np.abs(g - hf) / np.maximum(1.0, np.abs(hf))))
    log(f"max |adjoint - HF| / max(1,|HF|) over 63 bonds: {hf_rel:.2e}")
    # central differences in FP64, warm-started from the converged state
    fd = {}
    h = 1e-5
    for b in (0, 31, 62):
        vals = []
        for sgn in (1.0, -1.0):
            Jb = J0.copy()
            Jb[b] += sgn * h
            p = ct.DMRGParams(num_sweeps=4, chi_max=chi, eps_trunc=1e-12, max_krylov_iter=20, krylov_tol=1e-12,
                              energy_tol=0.0, jacobi_max_rot=-1)
            Eb, _ = ct.dmrg_two_site(ctx, ct.mpo_from_tensors(ctx, ham.random_bond_heisenberg_tensors(Jb)), mps, p)
            vals.append(Eb)
        fd[b] = (vals[0] - vals[1]) / (2 * h)
        log(f"bond {b}: adjoint {g[b]:.10f} HF {hf[b]:.10f} FD {fd[b]:.10f} rel {abs(fd[b] - g[b]) / abs(g[b]):.2e}")
    print(json.dumps({"study": "adjoint", "workload": study_adjoint.__doc__.replace("\n", " "), "E": E.item(),
                      "times": times, "max_rel_adjoint_vs_hf": hf_rel,
                      "fd": {str(b): {"fd": v, "adjoint": float(g[b]), "rel": abs(v - g[b]) / abs(g[b])}
                             for b, v in fd.items()},
                      "dEdJ": g.tolist()}), flush=True)


def bench_eigensolver(ctx, sizes=(1024, 2048), reps=3):
    """Symmetric eigensolver of the Gram split: ct_syevd (cooperative tridiagonalization + rocSOLVER dstedc +
    blocked back-transformation) against rocSOLVER dsyevd on Gram matrices theta theta^T, theta [n x n] with
    column scales 2^(-i/20) (decaying spectrum, as in a two-site split)."""
    import ctypes
    import torch
    from chemtensor_mojo_amd._lib import call
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    out = {}
    for n in sizes:
        g = torch.Generator(device="cuda").manual_seed(n)
        th = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
        th = th * (2.0 ** (-torch.arange(n, device="cuda", dtype=torch.float64) / 20.0))
        G = th @ th.T
        row = {}
        res = {}
        for name, method in (("ours", 0), ("rocsolver_dsyevd", 1)):
            w = torch.zeros(n, dtype=torch.float64, device="cuda")
            Z = torch.zeros(n, n, dtype=torch.float64, device="cuda")
            best = 1e9
            for _ in range(reps):
                A = G.clone()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                call("ct_syevd", ctx.handle, n, vp(A), vp(w), vp(Z), method)
                best = min(best, time.perf_counter() - t0)
            row[name + "_ms"] = round(best * 1e3, 3)
            res[name] = (w.clone(), Z.clone())
        w, Z = res["ours"]
        scale = float(w.abs().max())
        row["speedup"] = round(row["rocsolver_dsyevd_ms"] / row["ours_ms"], 3)
        row["max_eig_diff_rel"] = float((w - res["rocsolver_dsyevd"][0]).abs().max()) / scale
        row["residual_rel"] = float((G @ Z.T - Z.T * w).abs().max()) / scale
        row["orthogonality"] = float((Z @ Z.T - torch.eye(n, dtype=torch.float64, device="cuda")).abs().max())
        log(f"eigensolver n={n}: ours {row['ours_ms']:.2f} ms, rocSOLVER dsyevd {row['rocsolver_dsyevd_ms']:.2f} ms, "
            f"eig diff {row['max_eig_diff_rel']:.1e}, residual {row['residual_rel']:.1e}, orth {row['orthogonality']:.1e}")
        out[str(n)] = row
    return out


def study_qc(ctx, args):
    """Quantum-chemistry-shaped DMRG at reduced sizes: n_orb = 6 and 8 spatial orbitals (12 and 16 spin-orbital sites,
    Jordan-Wigner), h_pq symmetric N(0,1), (pq|rs) N(0,0.1) with 8-fold symmetry (seed 42), half filling through a
    particle-number penalty 1.0 (N - n_orb)^2 (2.4x the smallest penalty that makes the half-filled sector lowest at
    n_orb = 6: a larger one inflates |H| and with it the mixed-precision matvec error), exact MPO from the
    Jordan-Wigner operator strings; chi_max = 64 (exact on 12 sites) and 32 (truncated) at n_orb = 6, chi_max = 256
    (exact on 16 sites) at n_orb = 8; 6 sweeps, 20 Krylov vectors, exact tridiagonal solves; FP64 against FP32 and
    3xBF16 with FP64 refinement, all against full CI of the half-filled sector."""
    import chemtensor_mojo_amd as ct
    from chemtensor_mojo_amd import hamiltonians as ham
    out, problems = {}, {}
    for n_orb, D in ((6, 64), (6, 32), (8, 256)):
        if n_orb not in problems:
            t0 = time.perf_counter()
            mpo_t, info = ham.qc_tensors(n_orb, seed=42, penalty=1.0, exact=True)
            log(f"QC MPO n_orb={n_orb} and full CI in {time.perf_counter() - t0:.1f} s; bond dims "
                f"{[W.shape[3] for W in mpo_t[:-1]]}; E_fci = {info['E_fci']:.12f}")
            problems[n_orb] = (ct.mpo_from_tensors(ctx, mpo_t), info, [int(W.shape[3]) for W in mpo_t[:-1]])
        mpo, info, _ = problems[n_orb]
        L = 2 * n_orb
        for form, prec, refine in (("fp64", ct.PREC_FP64, 0), ("fp32_refined", ct.PREC_FP32, 2),
                                   ("bf16x3_refined", ct.PREC_BF16X3, 2)):
            name = f"n{n_orb}_D{D}_{form}"
            mps = ct.create_random_mps_c_compatible(ctx, L, 2, D, 42)
            p = ct.DMRGParams(num_sweeps=1, chi_max=D, eps_trunc=0.0, max_krylov_iter=20, krylov_tol=1e-12,
                              energy_tol=0.0, jacobi_max_rot=-1, prec=prec, refine=refine)
            run = ct.DMRGRun(ctx, mpo, mps, p)
            rows = _sweep_deltas(run, 6, name)
            run.close()
            E = rows[-1]["E"]
            out[name] = {"E": E, "E_fci": info["E_fci"], "dE_rel_vs_fci": abs(E - info["E_fci"]) / abs(info["E_fci"]),
                         "trunc_weight_sum": sum(r["trunc_weight_sum"] for r in rows),
                         "total_s": sum(r["s"] for r in rows), "sweeps": rows}
        E64 = out[f"n{n_orb}_D{D}_fp64"]["E"]
        for form in ("fp64", "fp32_refined", "bf16x3_refined"):
            o = out[f"n{n_orb}_D{D}_{form}"]
            o["dE_rel_vs_fp64"] = abs(o["E"] - E64) / abs(E64)
            log(f"n_orb={n_orb} D={D} {form}: E={o['E']:.14f} dE/E vs FCI {o['dE_rel_vs_fci']:.2e} vs fp64 "
                f"{o['dE_rel_vs_fp64']:.2e} trunc weight {o['trunc_weight_sum']:.2e} time={o['total_s']:.2f} s")
    print(json.dumps({"study": "qc", "workload": study_qc.__doc__.replace("\n", " "),
                      "mpo_bond_dims": {str(k): v[2] for k, v in problems.items()},
                      "summary": {k: {kk: vv for kk, vv in v.items() if kk != "sweeps"} for k, v in out.items()},
                      "sweeps": {k: v["sweeps"] for k, v in out.items()}}), flush=True)


def _mpo_expectation(mps_sites, mpo_sites):
    """<psi|O|psi> / <psi|psi> of an MPO over host MPS tensors [Dl, d, Dr]."""
    E = np.ones((1, 1, 1))
    N = np.ones((1, 1))
    for A, W in zip(mps_sites, mpo_sites):
        T = np.tensordot(E, A, axes=([1], [0]))          # [a, l', i, r]
        T = np.tensordot(T, W, axes=([0, 2], [0, 1]))    # [l', r, o, b]
        E = np.tensordot(T, A, axes=([0, 2], [0, 1]))    # [r, b, r']
        E = E.transpose(1, 0, 2)
        N = np.tensordot(np.tensordot(N, A, axes=([0], [0])), A, axes=([0, 1], [0, 1]))
    return float(E[0, 0, 0] / N[0, 0])


def study_qc16(ctx, args):
    """Quantum-chemistry-shaped DMRG at full size: n_orb = 16 spatial orbitals (32 spin-orbital sites, Jordan-Wigner),
    h_pq symmetric N(0,1), (pq|rs) N(0,0.1) with 8-fold symmetry (seed 42), half filling (16 electrons) through the
    particle-number penalty 1.0 (N - 16)^2, exact MPO from the operator strings (central bond dimension 562), seeded
    random start D=16 inside the 16-electron sector (hamiltonians.number_sector_random_mps_tensors; H conserves N, so
    the sweeps stay there), chi_max = 512, eps_trunc = 1e-12, 6 sweeps, 20 Krylov vectors, exact tridiagonal solves;
    FP64 against FP32, 3xBF16 and 6xBF16 with FP64 refinement.  Then one more sweep in each form, and in FP64 with the same
    refinement, from the same FP64 state (isolates the arithmetic from the different trajectories of truncated runs),
    and <(N - 16)^2> of each result."""
    import chemtensor_mojo_amd as ct
    from chemtensor_mojo_amd import hamiltonians as ham
    n_orb, D, sweeps, mu = 16, 512, int(args.qc_sweeps), float(args.qc_penalty)
    t0 = time.perf_counter()
    h, v = ham.qc_integrals(n_orb, 42)
    mpo_t = ham.qc_mpo_tensors(h, v, mu, n_orb)
    pen_t = ham.qc_mpo_tensors(np.zeros_like(h), np.zeros_like(v), 1.0, n_orb)
    log(f"QC MPO built in {time.perf_counter() - t0:.1f} s; bond dims {[W.shape[3] for W in mpo_t[:-1]]}")
    L = 2 * n_orb
    center, Fc = bench_matvec_mpo(ctx, D, mpo_t[L // 2 - 1], mpo_t[L // 2], 5, 2)
    for k, r in center.items():
        log(f"center-bond matvec {k}: {r['s_per_matvec'] * 1e3:.2f} ms {r['tflops']:.1f} TFLOP/s "
            f"(err vs fp64 {r.get('rel_err_vs_fp64', 0.0):.1e})")
    mpo = ct.mpo_from_tensors(ctx, mpo_t)
    forms = (("fp64", ct.PREC_FP64, 0), ("fp32_refined", ct.PREC_FP32, 2), ("bf16x3_refined", ct.PREC_BF16X3, 2),
             ("bf16x6_refined", ct.PREC_BF16X6, 2))

    def params(prec, refine, krylov=20):
        return ct.DMRGParams(num_sweeps=1, chi_max=D, eps_trunc=1e-12, max_krylov_iter=krylov, krylov_tol=1e-12,
                             energy_tol=0.0, jacobi_max_rot=-1, prec=prec, refine=refine)

    out, warm, fp64_state = {}, {}, None
    for form, prec, refine in forms:
        start = ct.mps_from_tensors(ctx, ham.number_sector_random_mps_tensors(L, n_orb, 16, 42))
        run = ct.DMRGRun(ctx, mpo, start, params(prec, refine))
        rows = _sweep_deltas(run, sweeps, form)
        state = run.mps()
        run.close()
        host = state.to_numpy()
        o = {"E": rows[-1]["E"], "trunc_weight_sum": sum(r["trunc_weight_sum"] for r in rows),
             "total_s": sum(r["s"] for r in rows), "max_bond": rows[-1]["max_bond"],
             "penalty_expectation": _mpo_expectation(host, pen_t), "sweeps": rows}
        out[form] = o
        log(f"{form}: E={o['E']:.13f} <(N-16)^2>={o['penalty_expectation']:.2e} "
            f"trunc weight {o['trunc_weight_sum']:.2e} time={o['total_s']:.1f} s")
        if form == "fp64":
            fp64_state = state
        else:
            del state
    E64 = out["fp64"]["E"]
    for form in out:
        out[form]["dE_rel_vs_fp64"] = abs(out[form]["E"] - E64) / abs(E64)
    warm_forms = (forms[0], ("fp64_refined", ct.PREC_FP64, 2)) + forms[1:]
    for form, prec, refine in warm_forms:
        run = ct.DMRGRun(ctx, mpo, fp64_state, params(prec, refine))
        rows = _sweep_deltas(run, 1, f"warm {form}")
        run.close()
        warm[form] = {"E": rows[-1]["E"], "s": rows[-1]["s"], "matvecs": rows[-1]["matvecs"],
                      "trunc_weight_sum": rows[-1]["trunc_weight_sum"]}
    W64, W64r = warm["fp64"]["E"], warm["fp64_refined"]["E"]
    for form in warm:
        warm[form]["dE_rel_vs_fp64"] = abs(warm[form]["E"] - W64) / abs(W64)
        warm[form]["dE_rel_vs_fp64_refined"] = abs(warm[form]["E"] - W64r) / abs(W64r)
        log(f"warm {form}: E={warm[form]['E']:.13f} dE/E vs fp64 {warm[form]['dE_rel_vs_fp64']:.2e} "
            f"vs fp64 refined {warm[form]['dE_rel_vs_fp64_refined']:.2e} {warm[form]['s']:.1f} s")
    print(json.dumps({"study": "qc16", "workload": study_qc16.__doc__.replace("\n", " "),
                      "mpo_bond_dims": [int(W.shape[3]) for W in mpo_t[:-1]],
                      "center_bond_matvec": {"D": D, "w": [int(mpo_t[L // 2 - 1].shape[0]), int(mpo_t[L // 2].shape[0]),
                                                           int(mpo_t[L // 2].shape[3])], "flops": Fc, **center},
                      "summary": {k: {kk: vv for kk, vv in v.items() if kk != "sweeps"} for k, v in out.items()},
                      "warm_sweep": warm, "sweeps": {k: v["sweeps"] for k, v in out.items()}}), flush=True)


def study_qc16_matvec(ctx, args):
    """The H_eff matvec at the centre bond of the 16-orbital quantum-chemistry MPO (w = 501, 562, 501; d = 2), D = 512,
    random symmetric environments, prepared operator, FP64 / FP32 / 3xBF16 / 6xBF16 (`--steps` timed calls each); used for
    the kernel profile of the large-w path."""
    from chemtensor_mojo_amd import hamiltonians as ham
    h, v = ham.qc_integrals(16, 42)
    mpo_t = ham.qc_mpo_tensors(h, v, 1.0, 16)
    res, F = bench_matvec_mpo(ctx, 512, mpo_t[15], mpo_t[16], args.steps, args.warmup)
    for k, r in res.items():
        log(f"centre-bond matvec {k}: {r['s_per_matvec'] * 1e3:.2f} ms {r['tflops']:.1f} TFLOP/s")
    print(json.dumps({"study": "qc16_matvec", "workload": study_qc16_matvec.__doc__.replace("\n", " "), "flops": F,
                      "by_form": res}), flush=True)


def study_eigensolver(ctx, args):
    """The Gram-split eigensolver alone at n = 2048 (the two-site split at D = 1024, d = 2), ours against rocSOLVER
    dsyevd; used for the kernel profile of the eigensolver."""
    print(json.dumps({"study": "eigensolver", "by_n": bench_eigensolver(ctx, sizes=(2048,))}), flush=True)


def bench_sweeps(ctx, args):
    """Fixed-D sweeps in FP64 and the refined FP32 / 3xBF16 / 6xBF16 forms on the two chains (see the module
    docstring)."""
    import chemtensor_mojo_amd as ct
    sweep = {}
    for key, model, L_sites, nsw in (("j1j2_mg", "j1j2", args.sweep_sites, 2),
                                     ("heisenberg_afm", "heisenberg", args.afm_sites, 2)):
        runs = {name: bench_sweep(ctx, L_sites, args.sweep_D, prec, refine, model=model, sweeps=nsw)
                for name, prec, refine in (("fp64", ct.PREC_FP64, 0), ("fp32_refined", ct.PREC_FP32, 2),
                                           ("bf16x3_refined", ct.PREC_BF16X3, 2),
                                           ("bf16x6_refined", ct.PREC_BF16X6, 2))}
        E64 = runs["fp64"]["energies"][-1]
        desc = ("J1-J2 chain at the Majumdar-Ghosh point (J2=0.5, w=8): after the first sweep every local "
                "problem is converged, so the sweep times the truncation and a one-vector Lanczos"
                if model == "j1j2" else
                "antiferromagnetic Heisenberg chain (w=5), far from converged at this bond dimension: every "
                "local Lanczos runs its full Krylov space")
        sweep[key] = {
            "workload": (f"{desc}; L={L_sites}, fixed D={args.sweep_D} (eps_trunc=0), PCG32 random start at "
                         f"full D, 20 Krylov vectors, exact tridiagonal solves, Gram-matrix FP64 truncation, "
                         f"{nsw} sweeps"),
            **runs,
            "dE_rel_bf16x3_vs_fp64": abs(runs["bf16x3_refined"]["energies"][-1] - E64) / abs(E64),
            "dE_rel_fp32_vs_fp64": abs(runs["fp32_refined"]["energies"][-1] - E64) / abs(E64),
            "speedup_sweep_bf16x3": [a / b for a, b in zip(runs["fp64"]["sweep_s"],
                                                           runs["bf16x3_refined"]["sweep_s"])],
            "speedup_sweep_fp32": [a / b for a, b in zip(runs["fp64"]["sweep_s"], runs["fp32_refined"]["sweep_s"])],
            "dE_rel_bf16x6_vs_fp64": abs(runs["bf16x6_refined"]["energies"][-1] - E64) / abs(E64),
            "speedup_sweep_bf16x6": [a / b for a, b in zip(runs["fp64"]["sweep_s"],
                                                           runs["bf16x6_refined"]["sweep_s"])],
        }
    return sweep


def cpu_baseline(D, w, budget_s):
    from oracle import oracle as ora
    ora.build()
    n = os.cpu_count() or 1
    ora.set_num_threads(n)
    L, R, W, th = make_problem(D, w)
    F = matvec_flops(D, 2, D, w, w, w)
    t0 = time.perf_counter()
    reps = 0
    while True:
        ora.heff_apply(L, W, W, R, th)
        reps += 1
        if time.perf_counter() - t0 > budget_s or reps >= 20:
            break
    t = (time.perf_counter() - t0) / reps
    return {"value": F / t / 1e12, "unit": "TFLOP/s", "cores": ora.get_num_threads(), "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{reps} FP64 H_eff matvecs at D={D}, w={w} (oracle, OpenBLAS dgemm, {t:.2f} s each)"}


def cpu_model() -> str:
    """The host CPU as /proc/cpuinfo names it, with its family / model / stepping
    numbers (the model name alone can be generic on virtualised hosts) and the
    logical CPUs the process may use."""
    fields = {}
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if not ln.strip():
                    break  # the first processor's block is enough
                k, _, v = ln.partition(":")
                fields[k.strip()] = v.strip()
    except OSError:
        pass
    name = fields.get("model name", "unknown")
    ids = ", ".join(f"{k} {fields[k]}" for k in ("cpu family", "model", "stepping") if k in fields)
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return f"{name} ({ids}; {ncpu} logical CPUs available)" if ids else f"{name} ({ncpu} logical CPUs available)"


def profiled_traffic(key, D, w):
    """HBM bytes per call of a dominant kernel at D=1024, w=8, from the counter
    summary committed in profiles/ (hardware counters cannot be read from inside
    this process); None when the summary has no row for the key."""
    if (D, w) != (1024, 8):
        return None
    path = os.path.join(ROOT, "profiles", "hbm_traffic_D1024_w8.md")
    try:
        with open(path) as f:
            for ln in f:
                cells = [c.strip() for c in ln.strip().strip("|").split("|")]
                if len(cells) >= 5 and cells[0] == key:
                    return float(cells[4])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--D", type=int, default=1024)
    ap.add_argument("--w", type=int, default=8)
    ap.add_argument("--sweep-sites", type=int, default=24)
    ap.add_argument("--sweep-D", type=int, default=1024)
    ap.add_argument("--afm-sites", type=int, default=40)
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--sweep-only", action="store_true",
                    help="run only the fixed-D sweeps and print them as one JSON line (for kernel traces of the sweeps)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the eigensolver and matvec-grid sections")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the L=64 XXZ (chi 256) and L=100 J1-J2 (chi 1024) sweep sections")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--qc-sweeps", type=int, default=6)
    ap.add_argument("--qc-penalty", type=float, default=1.0)
    ap.add_argument("--svd-modes", default="gram,gesvd",
                    help="split modes of the stages study, comma-separated: gram, gesvd")
    ap.add_argument("--study", choices=("stages", "precision", "adjoint", "eigensolver", "qc", "qc16",
                                            "qc16_matvec"), default=None,
                    help="run one of the larger studies instead of the default benchmark")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this path runs on one GPU; --gpus must be 1 (multi-GPU sharding is a design note only)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import chemtensor_mojo_amd as ct
    ctx = ct.DeviceContext(0)
    D, w = args.D, args.w
    log(f"device: {torch.cuda.get_device_name(0)}")
    if args.study:
        {"stages": study_stages, "precision": study_precision, "adjoint": study_adjoint,
         "eigensolver": study_eigensolver, "qc": study_qc, "qc16": study_qc16,
         "qc16_matvec": study_qc16_matvec}[args.study](ctx, args)
        ctx.close()
        return
    if args.sweep_only:
        print(json.dumps({"sweep": bench_sweeps(ctx, args)}), flush=True)
        ctx.close()
        return
    mv, F = bench_matvec(ctx, D, w, args.steps, args.warmup)
    gm, Fg = bench_dominant_gemm(ctx, D, w, args.steps, args.warmup)
    eig = None if args.no_extras else bench_eigensolver(ctx)
    mgrid = None if args.no_extras else bench_matvec_grid(ctx, max(5, args.steps // 2), args.warmup)
    envs = None if args.no_extras else bench_env_theta(ctx, max(5, args.steps // 2), args.warmup)
    sweep = None if args.no_sweep else bench_sweeps(ctx, args)
    configs = None
    if not args.no_configs:
        _, _, c2 = config2_sweeps(ctx, ((ct.SVD_GRAM, "fp64_gram"),))
        c3_summary, c3 = config3_sweeps(ctx)
        stage_keys = ("s", "t_lanczos", "t_svd", "t_env", "matvecs", "max_bond", "E")
        configs = {
            "xxz_afm_L64_chi256": {"workload": " ".join(config2_sweeps.__doc__.split()),
                                   "per_sweep": [{k: r[k] for k in stage_keys} for r in c2["fp64_gram"]],
                                   "total_s": sum(r["s"] for r in c2["fp64_gram"])},
            "j1j2_L100_chi1024": {"workload": " ".join(config3_sweeps.__doc__.split()), "summary": c3_summary,
                                  "per_sweep": {k: [{kk: r[kk] for kk in stage_keys} for r in v]
                                                for k, v in c3.items()}},
        }
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(D, w, args.cpu_budget)
        log(f"cpu baseline: {cpu['value']:.3f} TFLOP/s on {cpu['cores']} threads ({cpu['cpu_model']})")
    head = mv["fp32"]
    dom = gm["fp32"]
    dom_kernel = (f"gemm_mfma_kernel<float> + gemm_splitk_reduce<float> (stage-4 GEMM of the FP32 matvec, "
                  f"[{4 * D}x{w * D}].[{w * D}x{D}], 128x128 tiles, split-K 2)")
    line = {
        "metric": METRIC,
        "value": round(head["tflops"], 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["s_per_matvec"] * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded random environments/theta, J1-J2 bulk MPO)",
        "config": {"workload": f"two-site H_eff matvec, J1-J2 chain, D={D}, w={w}, d=2",
                   "matvec_gflop": round(F / 1e9, 3),
                   "value_basis": "FP32 matvec (the reference's float32 arithmetic) of the prepared H_eff operator "
                                  "(W_i W_i+1 and the L/R copies built once per local problem, as inside the "
                                  "Lanczos); FP64, 3xBF16 and 6xBF16 forms and the one-shot figures in matvec"},
        "matvec": {k: {kk: (float(f"{vv:.4g}") if isinstance(vv, float) else vv) for kk, vv in v.items()}
                   for k, v in mv.items()},
        "roofline": {"bound": "mfma", "kernel": dom_kernel,
                     "achieved": round(dom["tflops"], 3), "peak": PEAK["fp32"], "unit": "TFLOP/s",
                     "frac": round(dom["tflops"] / PEAK["fp32"], 4),
                     "traffic": profiled_traffic("fp32_stage4_gemm", D, w),
                     "note": "achieved = 2*M*N*K / (GEMM + split-K reduction time), HIP events on the library's "
                             "stream, mean of `steps` calls after >= 0.25 s of untimed calls; traffic = HBM bytes per call (FETCH_SIZE x 2 + WRITE_SIZE x 1, scales from a "
                             "64 MiB streaming kernel) from the counter runs in profiles/"},
        "roofline_fp64": {"bound": "mfma", "achieved": round(gm["fp64"]["tflops"], 3), "peak": PEAK["fp64"],
                          "unit": "TFLOP/s", "frac": round(gm["fp64"]["tflops"] / PEAK["fp64"], 4),
                          "traffic": profiled_traffic("fp64_stage4_gemm", D, w)},
        "roofline_bf16x3": {"bound": "mfma", "achieved": round(gm["bf16x3"]["tflops"], 3), "peak": PEAK["bf16"],
                            "unit": "TFLOP/s", "frac": round(gm["bf16x3"]["tflops"] / PEAK["bf16"], 4),
                            "mfma_issue_frac": round(3.0 * gm["bf16x3"]["tflops"] / PEAK["bf16"], 4),
                            "traffic": profiled_traffic("bf16x3_stage4_gemm", D, w),
                            "note": "achieved counts the algorithmic 2*M*N*K once although three BF16 products "
                                    "run per multiply-add; mfma_issue_frac counts all three against the dense "
                                    "BF16 peak"},
        "roofline_bf16x6": {"bound": "mfma", "kernel": "gemm_bf16x6_dma_kernel + splitk_reduce4_kernel (same GEMM, "
                                                       "256x128 tiles, split-K 2)",
                            "achieved": round(gm["bf16x6"]["tflops"], 3), "peak": round(PEAK["bf16"] / 6.0, 1),
                            "unit": "TFLOP/s", "frac": round(6.0 * gm["bf16x6"]["tflops"] / PEAK["bf16"], 4),
                            "rel_err": float(f"{gm['bf16x6']['rel_err']:.3g}"),
                            "rel_err_fp32": float(f"{gm['fp32']['rel_err']:.3g}"),
                            "vs_fp32_peak": round(gm["bf16x6"]["tflops"] / PEAK["fp32"], 4),
                            "traffic": profiled_traffic("bf16x6_stage4_gemm", D, w),
                            "note": "FP32-accurate GEMM on the BF16 matrix cores: six BF16 products per "
                                    "multiply-add; peak = dense BF16 peak / 6, so frac is the MFMA issue "
                                    "fraction; rel_err against the FP64 product, beside the FP32 MFMA GEMM's"},
        "sweep": sweep,
        "configs": configs,
        "eigensolver": {"workload": bench_eigensolver.__doc__.replace("\n", " ").replace("    ", " "), "by_n": eig},
        "matvec_grid": {"workload": bench_matvec_grid.__doc__.replace("\n", " ").replace("    ", " "), "by_w_D": mgrid},
        "env_update_theta": {"workload": " ".join(bench_env_theta.__doc__.split()), "by_w": envs},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
