#!/usr/bin/env python3
"""Benchmark of the periodic stiffness operator on one MI355X.

Workload: 3D H8 elasticity on a 256^3 periodic unit cube, fp32, mu=1,
nu=0.25, force f ~ N(0,1) with zero mean (synthetic, generated on the
device).  One step = one K.u apply (R2C plane sweep -> fused axis-0 pencil
with the modal block -> C2R plane sweep) with the input resident in HBM.

Prints ONE JSON line.  Besides the headline fp32 apply rate it reports the
plain fp64 path, the Green solve, fp32-vs-fp64 relative errors, the roofline
of the dominant kernel (HIP events on the launch stream) and the CPU
baseline (the oracle's literal restatement of the reference apply, timed on
this host).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "256³ K·u applies/s, % HBM roofline (fp32 vs plain fp64); p50/p99 served latency"
HBM_PEAK_GBS = 8000.0
KERNELS = ("plane_r2c", "pencil_block", "plane_c2r")


def _kernel_bytes(shape, esz):
    """Algorithmic HBM bytes per launch of each kernel: read its input once,
    write its output once.  Real field: DIM*|N|*esz; half-spectrum
    workspace: DIM*(|N|/NL)*(NL/2+1)*2*esz."""
    D = len(shape)
    N = 1
    for n in shape:
        N *= n
    field = D * N * esz
    half = D * (N // shape[-1]) * (shape[-1] // 2 + 1) * 2 * esz
    return {"plane_r2c": field + half, "pencil_block": 2 * half, "plane_c2r": half + field}


def _timed(fn, steps, sync):
    sync()
    t0 = time.perf_counter()
    for i in range(steps):
        fn(i)
    sync()
    return time.perf_counter() - t0


def _profile(op, fn, steps, sync):
    from bri17_amd import _lib
    L = _lib.lib()
    _lib.check(L.bri17_operator_profile(op.handle, 1))
    for i in range(steps):
        fn(i)
    sync()
    calls = ctypes.c_int()
    ms = (ctypes.c_double * 3)()
    _lib.check(L.bri17_operator_profile_read(op.handle, ctypes.byref(calls), ms))
    _lib.check(L.bri17_operator_profile(op.handle, 0))
    return {k: ms[i] / max(calls.value, 1) for i, k in enumerate(KERNELS)}


def _traffic(name):
    """HBM bytes per launch of `name` measured with rocprofv3 PMC counters
    (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), if a committed
    summary exists."""
    p = os.path.join(ROOT, "profiles", "traffic_256cube_f32.json")
    try:
        with open(p) as f:
            d = json.load(f)
        return d["kernels"][name]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(f_host, shape, mu, nu, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    import numpy as np
    pyoracle.lib()
    u = np.ascontiguousarray(f_host, dtype=np.float64)
    L = (1.0,) * len(shape)
    t0 = time.perf_counter()
    pyoracle.apply(shape, L, mu, nu, u, nthreads=threads)
    t_all = time.perf_counter() - t0
    out = {"value": 1.0 / t_all, "unit": "applies/s", "cores": threads, "kind": "port",
           "sample": "one full %s apply (fp64 c2c, literal restatement of tests/test_bri17.cpp:56-107) "
                     "on %d host threads" % ("x".join(map(str, shape)), threads)}
    if threads > 1 and os.environ.get("BRI17_CPU_SINGLE", "1") == "1":
        t0 = time.perf_counter()
        pyoracle.apply(shape, L, mu, nu, u, nthreads=1)
        out["single_thread_applies_per_s"] = 1.0 / (time.perf_counter() - t0)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain fp64 path")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from bri17_amd import StiffnessMatrix

    dev = torch.device("cuda:0")
    n = args.size
    shape = (n, n, n)
    mu, nu = 1.0, 0.25
    K, W = args.steps, args.warmup
    sync = torch.cuda.synchronize

    g = torch.Generator(device=dev).manual_seed(2026)
    f = torch.randn((3,) + shape, generator=g, device=dev, dtype=torch.float32)
    f -= f.mean(dim=(1, 2, 3), keepdim=True)
    op = StiffnessMatrix(shape, (1.0, 1.0, 1.0), mu, nu, "float32")
    outs = [torch.empty_like(f), torch.empty_like(f)]

    def step(i):
        op.apply(f, out=outs[i & 1])

    for i in range(W):
        step(i)
    t = _timed(step, K, sync)
    ms_step = 1e3 * t / K
    value = K / t
    per_kernel = _profile(op, step, K, sync)

    def solve_step(i):
        op.solve(f, out=outs[i & 1])

    for i in range(max(2, W // 4)):
        solve_step(i)
    t_solve = _timed(solve_step, K, sync)

    # fp32 error against the fp64 path (itself within 1e-12 of the oracle)
    result = {}
    if not args.no_fp64:
        f64 = f.double()
        op64 = StiffnessMatrix(shape, (1.0, 1.0, 1.0), mu, nu, "float64")
        o64 = [torch.empty_like(f64), torch.empty_like(f64)]

        def step64(i):
            op64.apply(f64, out=o64[i & 1])

        for i in range(max(2, W // 4)):
            step64(i)
        K64 = max(10, K // 4)
        t64 = _timed(step64, K64, sync)
        a32 = op.apply(f)
        a64 = op64.apply(f64)
        s32 = op.solve(f)
        s64 = op64.solve(f64)
        sync()
        e_apply = float((a32.double() - a64).norm() / a64.norm())
        e_solve = float((s32.double() - s64).norm() / s64.norm())
        result["fp64_plain"] = {"applies_per_s": K64 / t64, "ms_per_apply": 1e3 * t64 / K64,
                                "per_kernel_ms": _profile(op64, step64, K64, sync)}
        result["fp32_speedup_vs_fp64"] = value / (K64 / t64)
        result["rel_err_fp32_vs_fp64"] = {"apply": e_apply, "green_solve": e_solve}
        del op64, o64, f64, a64, s64

    dom = max(per_kernel, key=per_kernel.get)
    kb = _kernel_bytes(shape, 4)
    achieved = kb[dom] / (per_kernel[dom] * 1e-3) / 1e9
    b_alg = 2 * 3 * n ** 3 * 4
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "applies/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: f ~ N(0,1) per component, zero mean, generated on device",
        "config": {"workload": "3D H8 periodic stiffness K.u, %d^3 unit cube, mu=1, nu=0.25" % n,
                   "grid": list(shape), "components": 3, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": _traffic(dom)},
        "per_kernel_ms": per_kernel,
        "apply_roofline": {"algorithmic_bytes": b_alg, "achieved_GBs": b_alg / (ms_step * 1e-3) / 1e9,
                           "frac": b_alg / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "green_solves_per_s": K / t_solve,
    }
    line.update(result)
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(f.double().cpu().numpy(), shape, mu, nu, threads)
        line["cpu_baseline"]["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
