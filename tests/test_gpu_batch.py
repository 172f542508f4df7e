"""Batched persistent launch: 64 uneven instances (3D boxes with sides from
{32,48,64,96,128}, 25% 2D grids with sides 2^a 3^b in [64, 512]), per-instance
materials, apply and Green solve mixed in one launch.  Each instance must equal
its solo result bitwise (same runtime-size device code, no atomics on data)
and the fp64 oracle within 1e-5 relative."""
import os

import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

NT = min(16, os.cpu_count() or 1)
SIDES3 = (32, 48, 64, 96, 128)
SIDES2 = (64, 72, 81, 96, 108, 128, 144, 162, 192, 216, 243, 256, 288, 324, 384, 432, 486, 512)


def make_batch_spec(seed=4, n=64, frac2d=0.25):
    rng = np.random.default_rng(seed)
    n2 = int(round(n * frac2d))
    spec = []
    for i in range(n):
        if i < n2:
            shape = tuple(int(rng.choice(SIDES2)) for _ in range(2))
        else:
            shape = tuple(int(rng.choice(SIDES3)) for _ in range(3))
        mu = float(rng.uniform(0.5, 2.0))
        nu = float(rng.uniform(0.0, 0.45))
        mode = "solve" if rng.uniform() < 0.3 else "apply"
        spec.append((shape, mu, nu, mode))
    order = rng.permutation(n)
    return [spec[i] for i in order]


def _inputs(device, spec, seed=9):
    import torch
    g = torch.Generator(device=device).manual_seed(seed)
    xs = []
    for shape, mu, nu, mode in spec:
        x = torch.rand((len(shape),) + shape, generator=g, device=device, dtype=torch.float32) * 2 - 1
        if mode == "solve":
            x -= x.mean(dim=tuple(range(1, len(shape) + 1)), keepdim=True)
        xs.append(x)
    return xs


def test_batch_of_64_uneven_instances(device, oracle):
    import torch
    from bri17_amd import Batch, StiffnessMatrix
    spec = make_batch_spec()
    xs = _inputs(device, spec)
    ops = [StiffnessMatrix(s, None, mu, nu, "float32") for s, mu, nu, _ in spec]
    outs = [torch.empty_like(x) for x in xs]
    b = Batch()
    print("instances:", [s for s, *_ in spec], flush=True)
    b.run([(op, m, x, o) for op, (s, mu, nu, m), x, o in zip(ops, spec, xs, outs)])
    assert b.wait() == 0
    info = b.info()
    print("batch geometry", info)
    assert info["work_items"] > 64
    # solo results through the same runtime-size kernels: bitwise equal
    worst = 0.0
    for i, ((s, mu, nu, m), x, o) in enumerate(zip(spec, xs, outs)):
        # the operator itself where it runs the runtime-size kernels (with its
        # built-in launch shape and radix orders), else forced onto them
        solo = ops[i] if not ops[i].describe()["fixed_size_kernels"] else \
            StiffnessMatrix(s, None, mu, nu, "float32", tuning=dict(force_generic=1))
        r = solo.apply(x) if m == "apply" else solo.solve(x)
        torch.cuda.synchronize()
        assert torch.equal(r, o), (i, s, m)
        if i % 4 == 0:
            x64 = x.double().cpu().numpy()
            ref = (oracle.apply(s, (1.0,) * len(s), mu, nu, x64, nthreads=NT) if m == "apply"
                   else oracle.solve(s, (1.0,) * len(s), mu, nu, x64, nthreads=NT))
            e = rel_err(o.double().cpu().numpy(), ref)
            worst = max(worst, e)
            assert e <= 1e-5, (i, s, m, e)
    print("worst rel err vs fp64 oracle: %.2e" % worst)
    # a second run of the same batch reproduces the first bitwise
    outs2 = [torch.empty_like(x) for x in xs]
    b.run([(op, m, x, o) for op, (s, mu, nu, m), x, o in zip(ops, spec, xs, outs2)])
    assert b.wait() == 0
    assert all(torch.equal(a, c) for a, c in zip(outs, outs2))


def test_batch_in_fp64_matches_oracle(device, oracle):
    import torch
    from bri17_amd import Batch, StiffnessMatrix
    spec = [((12, 16, 10), 1.1, 0.3, "apply"), ((24, 18), 0.7, 0.2, "solve"),
            ((3, 4, 5), 5.6, 0.3, "apply"), ((40, 30, 36), 1.9, 0.41, "solve")]
    xs = [x.double() for x in _inputs(device, spec, seed=3)]
    ops = [StiffnessMatrix(s, None, mu, nu, "float64") for s, mu, nu, _ in spec]
    outs = [torch.empty_like(x) for x in xs]
    b = Batch()
    b.run([(op, m, x, o) for op, (s, mu, nu, m), x, o in zip(ops, spec, xs, outs)])
    assert b.wait() == 0
    for (s, mu, nu, m), x, o in zip(spec, xs, outs):
        x64 = x.cpu().numpy()
        ref = (oracle.apply(s, (1.0,) * len(s), mu, nu, x64) if m == "apply"
               else oracle.solve(s, (1.0,) * len(s), mu, nu, x64))
        assert rel_err(o.cpu().numpy(), ref) <= 1e-12


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_batch_multi_plane_items_and_odd_rows(device, oracle, dtype):
    """Plane items that hold several consecutive planes and odd rows
    transformed in pairs: odd last axes with an even middle axis (pairs
    inside a plane, several planes per item), with an odd middle axis (one
    plane per item, so that no pair straddles two planes), an even last
    axis, a split-3 plan with odd rows and an odd Q, split-2 plans (one
    plane per item, rows strided in memory), and sides with the prime
    radices 11, 13 and (direct-DFT passes) 17, 19, 23.
    Each instance equals its solo runtime-size result bitwise and the fp64
    oracle within tolerance."""
    import torch
    from bri17_amd import Batch, StiffnessMatrix
    # (shape, mu, nu, mode, middle-axis split)
    spec = [((12, 30, 27), 1.3, 0.2, "apply", 1), ((10, 27, 27), 0.8, 0.3, "solve", 1),
            ((8, 16, 45), 1.1, 0.1, "apply", 1), ((9, 24, 32), 1.7, 0.35, "solve", 1),
            ((6, 14, 81), 0.9, 0.25, "apply", 1), ((5, 81, 27), 1.2, 0.15, "apply", 3),
            ((6, 16, 27), 1.4, 0.3, "solve", 2), ((7, 24, 32), 0.6, 0.05, "apply", 2),
            ((11, 26, 13), 1.0, 0.2, "apply", 1), ((13, 22, 11), 0.7, 0.3, "solve", 1),
            ((17, 19, 23), 1.2, 0.25, "apply", 1)]
    g = torch.Generator(device=device).manual_seed(31)
    xs = []
    for shape, _, _, mode, _ in spec:
        x = torch.rand((3,) + shape, generator=g, device=device, dtype=getattr(torch, dtype)) * 2 - 1
        if mode == "solve":
            x -= x.mean(dim=(1, 2, 3), keepdim=True)
        xs.append(x)
    ops = [StiffnessMatrix(s, (1.0, 1.2, 0.9), mu, nu, dtype,
                           tuning=dict(force_generic=1, plane_split=sp))
           for s, mu, nu, _, sp in spec]
    outs = [torch.empty_like(x) for x in xs]
    b = Batch()
    b.run([(op, m, x, o) for op, (s, mu, nu, m, _), x, o in zip(ops, spec, xs, outs)])
    assert b.wait() == 0
    print("batch geometry", b.info(), "splits", [op.describe()["plane"]["split"] for op in ops], flush=True)
    assert [op.describe()["plane"]["split"] for op in ops] == [sp for *_, sp in spec]
    tol = 1e-12 if dtype == "float64" else 1e-5
    for (s, mu, nu, m, _), op, x, o in zip(spec, ops, xs, outs):
        r = op.apply(x) if m == "apply" else op.solve(x)
        torch.cuda.synchronize()
        assert torch.equal(r, o), (s, m)
        fn = oracle.apply if m == "apply" else oracle.solve
        ref = fn(s, (1.0, 1.2, 0.9), mu, nu, x.double().cpu().numpy(), nthreads=NT)
        e = rel_err(o.double().cpu().numpy(), ref)
        print(s, m, "rel err %.2e" % e, flush=True)
        assert e <= tol, (s, m, e)


def test_batch_reruns_reuse_or_rebuild_the_queue(device):
    """A run with the same items as the previous one on the same stream
    reuses the device queue without waiting; a run with other output
    buffers rebuilds it.  Every run's results equal the operators' own
    applies bitwise, including runs enqueued back to back."""
    import torch
    from bri17_amd import Batch, StiffnessMatrix
    spec = make_batch_spec(seed=12, n=12)
    xs = _inputs(device, spec, seed=13)
    ops = [StiffnessMatrix(s, None, mu, nu, "float32") for s, mu, nu, _ in spec]
    ref = []
    for op, (s, mu, nu, m), x in zip(ops, spec, xs):
        solo = op if not op.describe()["fixed_size_kernels"] else \
            StiffnessMatrix(s, None, mu, nu, "float32", tuning=dict(force_generic=1))
        ref.append(solo.apply(x) if m == "apply" else solo.solve(x))
    outs_a = [torch.empty_like(x) for x in xs]
    outs_b = [torch.empty_like(x) for x in xs]
    b = Batch()
    items = lambda outs: [(op, m, x, o) for op, (s, mu, nu, m), x, o in zip(ops, spec, xs, outs)]
    for outs in (outs_a, outs_a, outs_a, outs_b, outs_a, outs_a):
        for o in outs:
            o.zero_()
        b.run(items(outs))
        assert b.wait() == 0
        for r, o in zip(ref, outs):
            assert torch.equal(r, o)
    # back to back: three reruns of one queue in flight at once
    for o in outs_b:
        o.zero_()
    for _ in range(3):
        b.run(items(outs_b))
    assert b.wait() == 0
    for r, o in zip(ref, outs_b):
        assert torch.equal(r, o)


def test_batch_rejects_bad_items(device):
    import torch
    from bri17_amd import Batch, StiffnessMatrix, _lib
    a = StiffnessMatrix((8, 8, 8), None, 1.0, 0.3, "float32")
    c = StiffnessMatrix((8, 8), None, 1.0, 0.3, "float64")
    x = torch.zeros((3, 8, 8, 8), device=device)
    y = torch.zeros_like(x)
    x2 = torch.zeros((2, 8, 8), device=device, dtype=torch.float64)
    b = Batch()
    with pytest.raises(_lib.Bri17Error):
        b.run([(a, "apply", x, y), (a, "solve", y, x)])  # same operator twice
    with pytest.raises(_lib.Bri17Error):
        b.run([(a, "apply", x, y), (c, "apply", x2, x2.clone())])  # mixed dtypes
