#!/usr/bin/env python3
"""Benchmark of the periodic stiffness operator on one MI355X.

Workload: 3D H8 elasticity on a 256^3 periodic unit cube, fp32, mu=1,
nu=0.25, force f ~ N(0,1) with zero mean (synthetic, generated on the
device).  One step = one K.u apply (R2C plane sweep -> fused axis-0 pencil
with the modal block -> C2R plane sweep) with the input resident in HBM.

Prints ONE JSON line.  Besides the headline fp32 apply rate it reports the
plain fp64 path, the Green solve, the fp32 errors against the fp64 oracle
(broadband input) and against the exact answer on the smoothest single
modes, the roofline of the dominant kernel (HIP events on the launch
stream), the batched and served workloads, and the CPU baseline (the
oracle's literal restatement of the reference apply and Green solve, timed
on this host).
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
    workspace: the R2C output of the real field with its two real columns
    (k = 0 and k = NL/2) packed into one complex column, DIM*|N|/2 complex
    values, the same bytes as the field."""
    D = len(shape)
    N = 1
    for n in shape:
        N *= n
    field = D * N * esz
    half = field
    return {"plane_r2c": field + half, "pencil_block": 2 * half, "plane_c2r": half + field}


def _settle(fn, sync, seconds):
    """Run `fn` back to back for `seconds` (at least 20 calls).  From idle,
    the device's power management runs the first ~100 applies of this
    workload at varying clocks (measured per step with events: ~255 us for
    the first few, ~300 us over steps 5-25, ~246 us from step ~100 on), so a
    short timed window measures that transient, not the kernels.  Returns
    the number of calls made."""
    sync()
    t0 = time.perf_counter()
    i = 0
    while i < 20 or time.perf_counter() - t0 < seconds:
        fn(i)
        i += 1
        if i % 20 == 0:
            sync()
    sync()
    return i


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


def _valu(name, kernel_ms):
    """VALU issue time of kernel `name` per launch, from the committed SQ
    counter summary (SQ_INSTS_VALU x the mean issue interval of the kernel's
    instruction mix / 1024 SIMDs; the intervals are measured by
    bench_issue_rate.hip), as a fraction of its measured duration."""
    p = os.path.join(ROOT, "profiles", "sq_256cube_f32.json")
    try:
        with open(p) as f:
            d = json.load(f)
        k = d["kernels"][name]
    except (OSError, KeyError, ValueError):
        return None
    return {"insts_per_launch": k["SQ_INSTS_VALU"], "issue_us": k["valu_issue_us"],
            "kernel_us": 1e3 * kernel_ms, "frac": k["valu_issue_us"] / (1e3 * kernel_ms),
            "model": d["model"]}


def _host_cpu():
    """CPU of this host: model name, logical CPUs, and the CPUs this process
    may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "affinity": allowed,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def _errs(got, ref):
    """2-norm relative error and max-abs error over max |ref|."""
    import numpy as np
    d = got - ref
    return {"l2": float(np.linalg.norm(d.ravel()) / np.linalg.norm(ref.ravel())),
            "max": float(np.abs(d).max() / np.abs(ref).max())}


def cpu_baseline(f_host, shape, mu, nu, threads, gpu_apply=None, gpu_solve=None):
    """The oracle (the C fp64 restatement of the reference apply,
    tests/test_bri17.cpp:56-107: full-spectrum c2c, per-k modal_stiffness
    recomputation, and of the Green solve with the per-k inverse and the
    k = 0 branch of bri17.hpp:336-341) timed on this host, on one 256^3 apply
    and one solve each (the faster of two runs), plus one single-thread
    apply, the 64-instance batched workload one instance after the other,
    and a 24-request sample of the serving mix one request after the other.
    Its results are also the checker of the GPU's fp32 results on the same
    (fp32-rounded) input."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.lib()
    u = f_host.astype("float64")
    L = (1.0,) * len(shape)

    def best_of(fn, reps):
        t, out = float("inf"), None
        for _ in range(reps):
            t0 = time.perf_counter()
            out = fn()
            t = min(t, time.perf_counter() - t0)
        return t, out

    t_apply, ref_apply = best_of(lambda: pyoracle.apply(shape, L, mu, nu, u, nthreads=threads), 2)
    t_solve, ref_solve = best_of(lambda: pyoracle.solve(shape, L, mu, nu, u, nthreads=threads), 2)
    cpu = _host_cpu()
    out = {"value": 1.0 / t_apply, "unit": "applies/s", "cores": threads, "kind": "port",
           "sample": "one full %s apply (fp64 c2c, literal restatement of tests/test_bri17.cpp:56-107) "
                     "on %d host threads, faster of 2 runs" % ("x".join(map(str, shape)), threads),
           "green_solves_per_s": 1.0 / t_solve, "host_cpu": cpu}
    if threads > 1 and os.environ.get("BRI17_CPU_SINGLE", "1") == "1":
        t0 = time.perf_counter()
        pyoracle.apply(shape, L, mu, nu, u, nthreads=1)
        out["single_thread_applies_per_s"] = 1.0 / (time.perf_counter() - t0)
    # the 64-instance batched workload, one instance after the other
    import numpy as np
    t0 = time.perf_counter()
    nodes = 0
    for i, (sh, bmu, bnu, m) in enumerate(batch_spec()):
        x = pyoracle.fill_uniform(900 + i, len(sh) * int(np.prod(sh))).reshape((len(sh),) + sh)
        (pyoracle.apply if m == "apply" else pyoracle.solve)(sh, (1.0,) * len(sh), bmu, bnu, x,
                                                            nthreads=threads)
        nodes += int(np.prod(sh))
    t_b = time.perf_counter() - t0
    out["batched_64"] = {"instances_per_s": 64 / t_b, "ms_per_batch": 1e3 * t_b, "nodes": nodes}
    # a sample of the serving mix, served one request after the other
    rng = np.random.default_rng(55)
    sample = serving_draw(rng, 24)
    lat = []
    for i, (side, m, smu, snu) in enumerate(sample):
        sh = (int(side),) * 3
        x = pyoracle.fill_uniform(950 + i, 3 * side ** 3).reshape((3,) + sh)
        t0 = time.perf_counter()
        (pyoracle.apply if m == "apply" else pyoracle.solve)(sh, (1.0, 1.0, 1.0), float(smu), float(snu),
                                                            x, nthreads=threads)
        lat.append(time.perf_counter() - t0)
    out["serving_sample"] = {"requests": len(sample), "requests_per_s": len(sample) / sum(lat),
                             "p50_ms": 1e3 * float(np.percentile(lat, 50)),
                             "p99_ms": 1e3 * float(np.percentile(lat, 99)),
                             "note": "compute time of each request alone (no queueing), same mix "
                                     "as the served workload"}
    errs = {}
    if gpu_apply is not None:
        errs["apply"] = _errs(gpu_apply, ref_apply)
    if gpu_solve is not None:
        errs["green_solve"] = _errs(gpu_solve, ref_solve)
    return out, errs


def smooth_mode_errors(op, dev, ks=((1, 0, 0), (3, 1, 2))):
    """fp32 error on the smoothest single Fourier modes, u = Re(a exp(i k.x)):
    K u = |h| Re(K^(k) a exp(i k.x)) and the Green solve
    Re(K^(k)^-1 a exp(i k.x)) / |h| exactly, with K^(k) from the host
    per-frequency function of the C ABI (bri17.hpp:247-292).  The max-norm
    error over max |exact| is reported: such fields have their energy where
    K^ is smallest, so the fp32 rounding of the transforms (spread over all
    frequencies and multiplied there by K^, up to N^2/10 times larger)
    dominates."""
    import numpy as np
    import torch
    from bri17_amd import pybri17
    shape, L = op.shape, op.L
    h = np.array(L) / np.array(shape)
    hv = float(np.prod(h))
    grid = pybri17.CartesianGrid3f64(shape, L)
    hooke = pybri17.Hooke3f64(op.mu, op.nu, grid)
    rng = np.random.default_rng(17)
    ax = [torch.arange(n, device=dev, dtype=torch.float64) for n in shape]
    res = {}
    for k in ks:
        K = np.zeros(9, dtype=np.complex128)
        hooke.modal_stiffness_matrix(np.array(k, dtype=np.intc), K)
        K = K.reshape(3, 3).real
        a = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        th = 2 * np.pi * (k[0] * ax[0][:, None, None] / shape[0] + k[1] * ax[1][None, :, None] / shape[1]
                          + k[2] * ax[2][None, None, :] / shape[2])
        c, s = torch.cos(th), torch.sin(th)

        def field(v):
            return torch.stack([float(v[d].real) * c - float(v[d].imag) * s for d in range(3)])

        u = field(a).float()
        e = {}
        for mode, exact in (("apply", hv * (K @ a)), ("green_solve", np.linalg.solve(K, a) / hv)):
            got = (op.apply(u) if mode == "apply" else op.solve(u)).double()
            ex = field(exact)
            e[mode] = float((got - ex).abs().max() / ex.abs().max())
            del got, ex
        res["k=%d,%d,%d" % k] = e
        del u, c, s, th
    torch.cuda.synchronize()
    return res


SIDES3_BATCH = (32, 48, 64, 96, 128)
SIDES2_BATCH = (64, 72, 81, 96, 108, 128, 144, 162, 192, 216, 243, 256, 288, 324, 384, 432, 486, 512)
SIDES_SERVE = (32, 36, 48, 54, 64, 72, 81, 96, 108, 128, 144, 162, 192, 216, 243, 256)


def batch_spec(seed=4):
    """(shape, mu, nu, mode) of the 64 instances of the batched workload."""
    import numpy as np
    rng = np.random.default_rng(seed)
    spec = []
    for i in range(64):
        shape = tuple(int(rng.choice(SIDES2_BATCH if i < 16 else SIDES3_BATCH)) for _ in range(2 if i < 16 else 3))
        spec.append((shape, float(rng.uniform(0.5, 2.0)), float(rng.uniform(0.0, 0.45)),
                     "solve" if rng.uniform() < 0.3 else "apply"))
    return spec


def serving_draw(rng, n):
    """n requests of the serving mix: (side, mode, mu, nu), cubic fp32 grids
    with side log-uniform in [32, 256] rounded in log space to 2^a 3^b, 70%
    K.u / 30% Green solve, mu ~ U[0.5, 2], nu ~ U[0, 0.45]."""
    import numpy as np
    logs = np.log(np.array(SIDES_SERVE, dtype=float))
    x = rng.uniform(np.log(32), np.log(256), size=n)
    side = [SIDES_SERVE[int(np.argmin(np.abs(logs - v)))] for v in x]
    mode = ["solve" if u < 0.3 else "apply" for u in rng.uniform(size=n)]
    return list(zip(side, mode, rng.uniform(0.5, 2.0, size=n), rng.uniform(0.0, 0.45, size=n)))


def batched(dev, steps, warmup, seed=4):
    """64 instances in one persistent launch, fp32: 48 3D grids with sides
    from {32,48,64,96,128}, 16 2D grids with sides 2^a 3^b in [64,512],
    mu ~ U[0.5,2], nu ~ U[0,0.45], 30% Green solves, u ~ U[-1,1]."""
    import numpy as np
    import torch
    from bri17_amd import Batch, StiffnessMatrix
    spec = batch_spec(seed)
    g = torch.Generator(device=dev).manual_seed(seed)
    ops = [StiffnessMatrix(sh, None, mu, nu, "float32") for sh, mu, nu, _ in spec]
    xs = [torch.rand(op.field_shape, generator=g, device=dev) * 2 - 1 for op in ops]
    outs = [torch.empty_like(x) for x in xs]
    items = [(op, m, x, o) for op, (_, _, _, m), x, o in zip(ops, spec, xs, outs)]
    b = Batch()
    for _ in range(warmup):
        b.run(items)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run(items)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    assert b.wait() == 0
    # the same with a new queue every launch (outputs alternate between two
    # buffer sets, so no launch repeats the previous one's items)
    outs2 = [torch.empty_like(x) for x in xs]
    items2 = [(op, m, x, o) for op, (_, _, _, m), x, o in zip(ops, spec, xs, outs2)]
    b.run(items2)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for k in range(steps):
        b.run(items if k % 2 == 0 else items2)
    torch.cuda.synchronize()
    t_new = (time.perf_counter() - t2) / steps
    assert b.wait() == 0
    nodes = sum(int(np.prod(sh)) for sh, *_ in spec)
    b_alg = sum(2 * len(sh) * int(np.prod(sh)) * 4 for sh, *_ in spec)
    # the same 64 instances one operator call after another, for comparison
    def seq():
        for op, m, x, o in items:
            (op.apply if m == "apply" else op.solve)(x, out=o)
    for _ in range(2):
        seq()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        seq()
    torch.cuda.synchronize()
    t_seq = (time.perf_counter() - t1) / steps
    return {"instances": 64, "nodes": nodes, "ms_per_launch": 1e3 * t, "instances_per_s": 64 / t,
            "queue": "the same items every step: each launch after the first reuses the device work "
                     "queue and is enqueued without waiting for the previous one",
            "ms_per_launch_new_queue": 1e3 * t_new,
            "achieved_GBs": b_alg / t / 1e9, "algorithmic_bytes": b_alg,
            "ms_sequential_calls": 1e3 * t_seq, "speedup_vs_sequential": t_seq / t, "launch": b.info()}


def serving(dev, n_requests=2000, n_capacity=400, load=0.7, seed=5, server_kw=None, rate=None):
    """Poisson arrivals of fp32 3D requests, side log-uniform in [32,256]
    rounded (in log space) to 2^a 3^b, 70% K.u / 30% Green solve, mu ~
    U[0.5,2], nu ~ U[0,0.45].  The arrival rate is `load` times the capacity
    measured by a closed-loop burst (or `rate`, requests/s, when given);
    `server_kw` are Server options.  Inputs stay resident in HBM (one per
    side, read-only); outputs come from a ring per side whose slots are
    reused only after their previous request finished."""
    import numpy as np
    import torch
    from bri17_amd import Server
    from bri17_amd.server import poisson_arrivals, sleep_until
    rng = np.random.default_rng(seed)

    def draw(n):
        return serving_draw(rng, n)

    g = torch.Generator(device=dev).manual_seed(seed)
    inputs = {n: torch.rand((3, n, n, n), generator=g, device=dev) * 2 - 1 for n in SIDES_SERVE}
    ring = 16
    outs = {n: [torch.empty_like(inputs[n]) for _ in range(ring)] for n in SIDES_SERVE}
    torch.cuda.synchronize()
    srv = Server(**(server_kw or {}))
    holder = {n: [None] * ring for n in SIDES_SERVE}
    count = {n: 0 for n in SIDES_SERVE}

    def submit(req):
        n, mode, mu, nu = req
        k = count[n] % ring
        count[n] += 1
        prev = holder[n][k]
        if prev is not None:
            prev.wait()
        r = srv.submit(inputs[n], outs[n][k], mode, float(mu), float(nu))
        holder[n][k] = r
        return r

    # warm-up: every side and mode once (plans, pools)
    for n in SIDES_SERVE:
        for m in ("apply", "solve"):
            submit((n, m, 1.0, 0.3))
    srv.drain()
    # capacity: closed-loop bursts (the faster of two: the first also grows
    # the server's buffer pool)
    cap = 0.0
    for _ in range(2):
        burst = draw(n_capacity)
        t0 = time.perf_counter()
        reqs = [submit(q) for q in burst]
        for r in reqs:
            r.wait()
        cap = max(cap, n_capacity / (time.perf_counter() - t0))
    # open loop: Poisson arrivals at load * capacity
    rate = rate or load * cap
    arr = poisson_arrivals(rate, n_requests, seed=seed + 1)
    mix = draw(n_requests)
    st0 = srv.stats()
    reqs = []
    t0 = time.perf_counter()
    for a, q in zip(arr, mix):
        sleep_until(t0 + a)
        reqs.append(submit(q))
    res = [r.wait() for r in reqs]
    t_end = time.perf_counter()
    st = srv.stats()
    srv.close()
    lat = np.array([r["latency_s"] for r in res]) * 1e3
    qd = np.array([r["queued_s"] for r in res]) * 1e3
    ok = sum(r["state"] == "done" for r in res)
    sides = np.array([q[0] for q in mix])
    by_side = {}
    for n in SIDES_SERVE:
        sel = sides == n
        if sel.sum() >= 10:
            by_side[str(n)] = {"requests": int(sel.sum()), "p50_ms": float(np.percentile(lat[sel], 50)),
                               "p99_ms": float(np.percentile(lat[sel], 99)),
                               "mean_queued_ms": float(qd[sel].mean())}
    waves = st["waves"] - st0["waves"]
    nodes = sum(n ** 3 for n, *_ in mix)
    return {"requests": n_requests, "completed": ok, "arrival_rate_hz": rate, "load": load,
            "capacity_rps": cap, "throughput_rps": n_requests / (t_end - t0),
            "throughput_Mnodes_per_s": nodes / (t_end - t0) / 1e6,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
            "mean_queued_ms": float(qd.mean()), "by_side": by_side,
            "waves": waves, "mean_wave_size": n_requests / max(waves, 1),
            "peak_waves_in_flight": st["peak_waves_in_flight"], "hbm_budget": st["hbm_budget"],
            "hbm_peak": st["hbm_peak"], "plans_resident": st["plans_resident"],
            "mix": "3D fp32, side log-uniform [32,256] -> 2^a 3^b, 70% apply / 30% solve"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain fp64 path")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-instance batched launch")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving run")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--settle", type=float, default=0.3, metavar="S",
                    help="seconds of back-to-back applies before the warmup steps, so that the "
                         "timed steps run at the device's steady clock (0: none)")
    ap.add_argument("--force-generic", action="store_true",
                    help="run the headline grid on the runtime-size kernels (profiling, A/B)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run 3 plain device copies of MIB MiB each (reference kernel for "
                         "calibrating rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from bri17_amd import StiffnessMatrix

    dev = torch.device("cuda:0")
    if args.copy_calibration:
        src = torch.ones(args.copy_calibration << 18, device=dev)  # MiB of float32
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    n = args.size
    shape = (n, n, n)
    mu, nu = 1.0, 0.25
    K, W = args.steps, args.warmup
    sync = torch.cuda.synchronize

    g = torch.Generator(device=dev).manual_seed(2026)
    f = torch.randn((3,) + shape, generator=g, device=dev, dtype=torch.float32)
    f -= f.mean(dim=(1, 2, 3), keepdim=True)
    op = StiffnessMatrix(shape, (1.0, 1.0, 1.0), mu, nu, "float32",
                         tuning=dict(force_generic=1) if args.force_generic else None)
    outs = [torch.empty_like(f), torch.empty_like(f)]

    def step(i):
        op.apply(f, out=outs[i & 1])

    settle_calls = _settle(step, sync, args.settle) if args.settle > 0 else 0
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

    # fp32 results on the headline input, checked against the fp64 oracle in
    # the CPU-baseline leg below
    a32_host = op.apply(f).double().cpu().numpy()
    s32_host = op.solve(f).double().cpu().numpy()
    smooth = smooth_mode_errors(op, dev) if n == 256 else None

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
        "device_settle": {"seconds": args.settle, "applies": settle_calls},
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: f ~ N(0,1) per component, zero mean, generated on device",
        "config": {"workload": "3D H8 periodic stiffness K.u, %d^3 unit cube, mu=1, nu=0.25" % n,
                   "grid": list(shape), "components": 3, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": _traffic(dom),
                     "limiter": "latency: LDS, memory and barriers at one workgroup per CU "
                                "(VALU issue 34-43%, HBM 55-70% of each kernel's time)",
                     "valu": _valu(dom, per_kernel[dom]) if n == 256 else None},
        "per_kernel_ms": per_kernel,
        "apply_roofline": {"algorithmic_bytes": b_alg, "achieved_GBs": b_alg / (ms_step * 1e-3) / 1e9,
                           "frac": b_alg / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "green_solves_per_s": K / t_solve,
    }
    line.update(result)
    if smooth is not None:
        line["smooth_mode_max_err_fp32"] = smooth
    # host-pointer entry point (copy in over PCIe, apply, copy out): not the
    # headline, which starts from inputs resident in HBM
    fh = f.cpu().numpy()
    oh = fh.copy()
    op.apply(fh, out=oh)
    t0 = time.perf_counter()
    for _ in range(3):
        op.apply(fh, out=oh)
    line["host_pointer_applies_per_s"] = 3 / (time.perf_counter() - t0)
    del fh, oh
    if not args.no_batch:
        line["batched_64"] = batched(dev, max(5, K // 10), 3)
    if not args.no_serve:
        line["serving"] = serving(dev, n_requests=args.serve_requests)
        # the same workload at a fixed arrival rate: comparable across builds
        # whose capacity (and hence the rate of the line above) differs
        fixed = serving(dev, n_requests=args.serve_requests, rate=5000.0)
        line["serving_fixed_rate"] = {k: fixed[k] for k in (
            "arrival_rate_hz", "throughput_rps", "p50_ms", "p99_ms", "mean_ms", "max_ms", "completed")}
    if not args.no_cpu:
        # every CPU this process may use, within the host's OpenMP thread
        # setting (the CPU share of one GPU on a shared host)
        cpu = _host_cpu()
        threads = cpu["affinity"]
        if cpu["omp_num_threads"] and cpu["omp_num_threads"].isdigit():
            threads = min(threads, int(cpu["omp_num_threads"]))
        base, errs = cpu_baseline(f.cpu().numpy(), shape, mu, nu, threads, a32_host, s32_host)
        base["gpu_over_cpu"] = value / base["value"]
        base["green_solve_gpu_over_cpu"] = (K / t_solve) / base["green_solves_per_s"]
        line["cpu_baseline"] = base
        line["rel_err_fp32_vs_oracle"] = errs
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
