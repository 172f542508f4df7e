"""Serving loop (bri17_server_*): continuous batching of mixed requests.

The central check is alone-vs-mix: each request's result served alone (a
wave of one on an idle server) equals, bitwise, its result served in a burst
of mixed grids, dtypes, materials and modes, where requests share waves and
waves overlap on several streams -- and equals the operator's own apply /
solve and the fp64 oracle within the stated tolerance.  Failure handling is
forced from the host: a footprint above the budget, an unsupported grid, a
NaN input, a tiny dependency-wait bound, cancellation."""
import os

import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

NT = min(16, os.cpu_count() or 1)
# cubic sides 2^a 3^b within [32, 128] and a few 2D grids
SIDES = (32, 36, 48, 54, 64, 72, 81, 96, 108, 128)
# every side of the serving workload: log-uniform in [32, 256] rounded to 2^a 3^b
SIDES_SERVE = (32, 36, 48, 54, 64, 72, 81, 96, 108, 128, 144, 162, 192, 216, 243, 256)


def _spec(seed=5, n=28):
    rng = np.random.default_rng(seed)
    spec = []
    for i in range(n):
        if i % 7 == 3:
            shape = (int(rng.choice((64, 96, 128, 192, 256))), int(rng.choice((72, 81, 128, 256))))
        else:
            shape = tuple([int(rng.choice(SIDES))] * 3) if i % 2 else \
                tuple(int(rng.choice(SIDES)) for _ in range(3))
        dtype = "float64" if i % 5 == 2 else "float32"
        mode = "solve" if rng.uniform() < 0.3 else "apply"
        L = tuple(float(v) for v in rng.uniform(0.5, 2.0, size=len(shape)))
        spec.append(dict(shape=shape, dtype=dtype, mode=mode, L=L,
                         mu=float(rng.uniform(0.5, 2.0)), nu=float(rng.uniform(0.0, 0.45))))
    return spec


def _inputs(device, spec, seed=11):
    import torch
    g = torch.Generator(device=device).manual_seed(seed)
    xs = []
    for s in spec:
        dt = getattr(torch, s["dtype"])
        xs.append((torch.rand((len(s["shape"]),) + s["shape"], generator=g, device=device, dtype=dt) * 2 - 1))
    torch.cuda.synchronize()
    return xs


def _serve(server, spec, xs, burst=True):
    """Submit every request (in one call when `burst`), wait for all."""
    import torch
    if burst:
        reqs = server.submit_many([dict(x=x, out=torch.empty_like(x), mode=s["mode"], mu=s["mu"], nu=s["nu"],
                                        L=s["L"]) for s, x in zip(spec, xs)])
    else:
        reqs = [server.submit(x, torch.empty_like(x), s["mode"], s["mu"], s["nu"], s["L"])
                for s, x in zip(spec, xs)]
    res = [r.wait(timeout=120) for r in reqs]
    return [r.out for r in reqs], res


def _spec_serving(seed=8, n=24):
    """The serving workload's mix: fp32 cubes with side log-uniform in
    [32, 256] rounded (in log space) to 2^a 3^b, 70% K u / 30% Green solve,
    mu ~ U[0.5, 2], nu ~ U[0, 0.45]; every side of the set appears at least
    once, the large runtime-size sides 144..243 and the fixed-size 256 among
    them."""
    rng = np.random.default_rng(seed)
    logs = np.log(np.array(SIDES_SERVE, dtype=float))
    sides = list(SIDES_SERVE)
    while len(sides) < n:
        v = rng.uniform(np.log(32), np.log(256))
        sides.append(SIDES_SERVE[int(np.argmin(np.abs(logs - v)))])
    rng.shuffle(sides)
    return [dict(shape=(int(n_),) * 3, dtype="float32", mode="solve" if rng.uniform() < 0.3 else "apply",
                 L=(1.0, 1.0, 1.0), mu=float(rng.uniform(0.5, 2.0)), nu=float(rng.uniform(0.0, 0.45)))
            for n_ in sides]


@pytest.mark.parametrize("mix", ["mixed_grids", "serving_sides"])
def test_alone_vs_mix_bitwise(device, oracle, mix):
    import torch
    from bri17_amd import Server, StiffnessMatrix
    spec = _spec() if mix == "mixed_grids" else _spec_serving()
    xs = _inputs(device, spec)
    print("requests:", [(s["shape"], s["dtype"][-2:], s["mode"]) for s in spec], flush=True)
    # alone: one request at a time on an idle server
    alone = []
    with Server(max_batch=8, max_waves=3) as srv:
        for s, x in zip(spec, xs):
            out, res = _serve(srv, [s], [x], burst=False)
            assert res[0]["state"] == "done" and res[0]["wave_size"] == 1, res[0]
            alone.append(out[0])
    # mix: everything at once -- shared waves, overlapping streams
    with Server(max_batch=8, max_waves=3) as srv:
        mix, res = _serve(srv, spec, xs)
        st = srv.stats()
    print("stats:", st, flush=True)
    assert all(r["state"] == "done" for r in res), res
    assert max(r["wave_size"] for r in res) > 1, "no request shared a wave"
    assert st["peak_waves_in_flight"] >= 2, "waves never overlapped"
    assert st["completed"] == len(spec) and st["hbm_admitted"] == 0 and st["running"] == 0
    assert st["hbm_peak"] <= st["hbm_budget"]
    for i, (s, x) in enumerate(zip(spec, xs)):
        assert torch.equal(alone[i], mix[i]), "request %d differs alone vs mix: %s" % (i, s)
        op = StiffnessMatrix(s["shape"], s["L"], s["mu"], s["nu"], s["dtype"])
        solo = op.apply(x) if s["mode"] == "apply" else op.solve(x)
        torch.cuda.synchronize()
        assert torch.equal(solo, mix[i]), "request %d differs from the operator's own %s" % (i, s["mode"])
    # oracle on a subset (fp64 oracle on the CPU)
    for i in range(0, len(spec), 4):
        s = spec[i]
        u = xs[i].double().cpu().numpy()
        fn = oracle.apply if s["mode"] == "apply" else oracle.solve
        ref = fn(s["shape"], s["L"], s["mu"], s["nu"], u, nthreads=NT)
        err = rel_err(mix[i].double().cpu().numpy(), ref)
        tol = 1e-12 if s["dtype"] == "float64" else 1e-5
        print("request", i, s["shape"], s["dtype"], s["mode"], "rel err %.2e" % err, flush=True)
        assert err <= tol


def test_heavy_requests_let_small_ones_pass(device):
    """Grids of at least heavy_nodes (2^22) nodes run one at a time by
    default; while one does, the small requests queued behind the other
    heavy ones start first.  With no limit the waves start in arrival
    order.  Either way every result is the same, bitwise."""
    import torch
    from bri17_amd import Server
    spec = [dict(shape=(n,) * 3, dtype="float32", mode=m, L=(1.0, 1.0, 1.0), mu=1.1, nu=0.3)
            for n, m in ((162, "apply"), (192, "solve"), (162, "apply"), (48, "apply"), (36, "solve"))]
    xs = _inputs(device, spec, seed=21)
    with Server(max_batch=8, max_waves=4) as srv:
        out, res = _serve(srv, spec, xs)
    with Server(max_batch=8, max_waves=4, max_heavy_waves=-1) as srv:
        out_fifo, res_fifo = _serve(srv, spec, xs)
    waves = [r["wave"] for r in res]
    waves_fifo = [r["wave"] for r in res_fifo]
    print("waves (heavy limit 1):", waves, " (no limit):", waves_fifo, flush=True)
    assert all(r["state"] == "done" for r in res + res_fifo)
    assert waves[0] < waves[1] < waves[2]
    assert max(waves[3:]) < waves[1], "the small requests waited behind a heavy one"
    assert waves_fifo[0] < waves_fifo[1] < waves_fifo[2] < min(waves_fifo[3:])
    for a, b in zip(out, out_fifo):
        assert torch.equal(a, b)


def test_host_memory_requests_match_device_requests(device):
    import torch
    from bri17_amd import Server
    spec = _spec(seed=6, n=10)
    xs = _inputs(device, spec, seed=12)
    with Server(max_batch=4, max_waves=2) as srv:
        dev_out, _ = _serve(srv, spec, xs)
        hreqs = [srv.submit(x.cpu().numpy(), None, s["mode"], s["mu"], s["nu"], s["L"]) for s, x in zip(spec, xs)]
        hres = [r.wait(timeout=120) for r in hreqs]
    assert all(r["state"] == "done" for r in hres)
    for i, r in enumerate(hreqs):
        assert np.array_equal(r.out, dev_out[i].cpu().numpy()), i


def test_budget_admission_queues_and_rejects(device):
    import torch
    from bri17_amd import Server, _lib
    spec = [dict(shape=(64, 64, 64), dtype="float32", mode="apply", L=(1, 1, 1), mu=1.0, nu=0.3),
            dict(shape=(48, 54, 72), dtype="float32", mode="solve", L=(1, 1, 1), mu=1.2, nu=0.2),
            dict(shape=(96, 96, 96), dtype="float64", mode="apply", L=(1, 1, 1), mu=0.7, nu=0.1)] * 4
    xs = _inputs(device, spec, seed=13)
    with Server() as srv:
        ref, _ = _serve(srv, spec, xs)
    # footprint of the largest request: in + out + half-spectrum workspace
    # (workspace rounded up to the pool's size class: <= 25% above)
    big = 2 * xs[2].numel() * 8 + int(1.25 * 3 * 96 * 96 * 49 * 16) + 4096
    with Server(hbm_budget_bytes=big, max_batch=8, max_waves=4) as srv:
        out, res = _serve(srv, spec, xs)
        st = srv.stats()
        with pytest.raises(_lib.Bri17Error) as e:
            x = torch.zeros((3, 128, 128, 128), device=device, dtype=torch.float64)
            srv.submit(x, torch.empty_like(x))
        assert e.value.code == _lib.BRI17_ERR_NOMEM
        assert srv.stats()["rejected"] == 1
    print("tiny-budget stats:", st, flush=True)
    assert st["hbm_peak"] <= big
    assert all(r["state"] == "done" for r in res)
    for a, b in zip(ref, out):
        assert torch.equal(a, b)


def test_failures_are_isolated(device):
    import torch
    from bri17_amd import Server, _lib
    spec = _spec(seed=7, n=12)
    xs = _inputs(device, spec, seed=14)
    with Server(max_batch=16) as srv:
        ref, _ = _serve(srv, spec, xs)
        # unsupported grid (rows of 65,537 entries): rejected at submit, nothing queued
        bad = torch.zeros((3, 4, 4, 65537), device=device)
        with pytest.raises(_lib.Bri17Error) as e:
            srv.submit(bad, torch.empty_like(bad))
        assert e.value.code == _lib.BRI17_ERR_UNSUPPORTED
        # invalid material
        with pytest.raises(_lib.Bri17Error):
            srv.submit(xs[0], torch.empty_like(xs[0]), nu=0.5)
        # a burst with one bad request queues nothing
        before = srv.stats()["submitted"]
        with pytest.raises(_lib.Bri17Error) as e:
            srv.submit_many([dict(x=xs[0]), dict(x=bad), dict(x=xs[1])])
        assert e.value.code == _lib.BRI17_ERR_UNSUPPORTED and "request 1" in str(e.value)
        assert srv.stats()["submitted"] == before
        # a NaN input in the middle of a mix only poisons its own result
        poisoned = [x.clone() for x in xs]
        poisoned[5].view(-1)[17] = float("nan")
        torch.cuda.synchronize()
        out, res = _serve(srv, spec, poisoned)
        assert all(r["state"] == "done" for r in res)
        for i in range(len(spec)):
            if i == 5:
                assert torch.isnan(out[i]).any()
            else:
                assert torch.equal(out[i], ref[i]), i
        # a dependency-wait bound of one poll: batched waves may fail, with
        # an error, and later waves are served correctly again
        srv.set_spin_limit(1)
        out, res = _serve(srv, spec, xs)
        states = [r["state"] for r in res]
        print("spin-limit-1 states:", states, flush=True)
        for i, r in enumerate(res):
            assert r["state"] in ("done", "failed"), r
            if r["state"] == "failed":
                assert r["error"] == _lib.BRI17_ERR_STATE and r["wave_size"] >= 1
        srv.set_spin_limit(1 << 24)
        out, res = _serve(srv, spec, xs)
        assert all(r["state"] == "done" for r in res)
        for a, b in zip(out, ref):
            assert torch.equal(a, b)


def test_cancel_queued_request(device):
    import torch
    from bri17_amd import Server, _lib
    x = torch.rand((3, 128, 128, 128), device=device, dtype=torch.float64)
    torch.cuda.synchronize()
    with Server(max_waves=1, max_batch=1) as srv:
        reqs = [srv.submit(x, torch.empty_like(x), "solve") for _ in range(8)]
        reqs[-1].cancel()
        res = [r.wait(timeout=120) for r in reqs]
        st = srv.stats()
    assert [r["state"] for r in res] == ["done"] * 7 + ["cancelled"]
    assert st["cancelled"] == 1 and st["completed"] == 7
    with pytest.raises(_lib.Bri17Error):
        reqs[0].cancel()  # long finished
