"""HIP-graph replays of the operator and of a batch, and the autotuner with
its JSON cache.  Graph replays are checked against the eager result after
every replay, with unrelated eager device work and synchronisations in
between."""
import json
import os

import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(128, 128, 128), (48, 36, 54), (96, 80)])
@pytest.mark.parametrize("mode", ["apply", "solve"])
def test_operator_graph_replays_match_eager(device, shape, mode):
    import torch
    from bri17_amd import StiffnessMatrix
    op = StiffnessMatrix(shape, None, 1.2, 0.3, "float32")
    g = torch.Generator(device=device).manual_seed(3)
    x = torch.randn((len(shape),) + shape, generator=g, device=device)
    eager = op.apply(x) if mode == "apply" else op.solve(x)
    torch.cuda.synchronize()
    out = torch.empty_like(x)
    graph = op.capture(x, out, mode)
    for r in range(4):
        out.fill_(float("nan"))
        _ = (x * 2).sum()  # unrelated eager work
        torch.cuda.synchronize()
        graph.launch()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), "replay %d differs" % r


def test_batch_graph_replays_match_eager(device):
    import torch
    from bri17_amd import Batch, StiffnessMatrix
    spec = [((32, 48, 64), "apply"), ((96, 64), "solve"), ((64, 64, 64), "apply"),
            ((48, 48, 32), "solve"), ((128, 96), "apply"), ((32, 32, 32), "apply")]
    ops = [StiffnessMatrix(s, None, 0.5 + 0.2 * i, 0.1 + 0.05 * i, "float32") for i, (s, _) in enumerate(spec)]
    g = torch.Generator(device=device).manual_seed(8)
    xs = [torch.randn(op.field_shape, generator=g, device=device) for op in ops]
    eager = [torch.empty_like(x) for x in xs]
    b = Batch()
    b.run([(op, m, x, o) for op, (s, m), x, o in zip(ops, spec, xs, eager)])
    assert b.wait() == 0
    outs = [torch.empty_like(x) for x in xs]
    graph = b.capture([(op, m, x, o) for op, (s, m), x, o in zip(ops, spec, xs, outs)])
    for r in range(3):
        for o in outs:
            o.zero_()
        torch.cuda.synchronize()
        graph.launch()
        torch.cuda.synchronize()
        for i, (o, e) in enumerate(zip(outs, eager)):
            assert torch.equal(o, e), "replay %d instance %d differs" % (r, i)


def test_batch_graph_pins_context_and_workspaces(device):
    """A batch graph replays its context's device tables and its members'
    workspaces: while it lives, another run on that context is refused
    (BRI17_ERR_STATE), and so is a strain call that would grow a member's
    workspace; replays stay correct after the refusals, and the context runs
    again once the graph is destroyed."""
    import torch
    from bri17_amd import Batch, StiffnessMatrix, _lib
    spec = [((32, 48, 64), "apply"), ((96, 64), "solve"), ((48, 48, 32), "solve")]
    ops = [StiffnessMatrix(s, None, 0.8 + 0.1 * i, 0.2, "float32") for i, (s, _) in enumerate(spec)]
    g = torch.Generator(device=device).manual_seed(9)
    xs = [torch.randn(op.field_shape, generator=g, device=device) for op in ops]
    eager = [(op.apply(x) if m == "apply" else op.solve(x)) for op, (s, m), x in zip(ops, spec, xs)]
    outs = [torch.empty_like(x) for x in xs]
    items = [(op, m, x, o) for op, (s, m), x, o in zip(ops, spec, xs, outs)]
    b = Batch()
    graph = b.capture(items)
    with pytest.raises(_lib.Bri17Error) as e:
        b.run(items)
    assert e.value.code == _lib.BRI17_ERR_STATE
    with pytest.raises(_lib.Bri17Error) as e:
        ops[0].strain(xs[0])
    assert e.value.code == _lib.BRI17_ERR_STATE
    for r in range(2):
        for o in outs:
            o.fill_(float("nan"))
        torch.cuda.synchronize()
        graph.launch()
        torch.cuda.synchronize()
        for i, (o, ex) in enumerate(zip(outs, eager)):
            assert torch.equal(o, ex), "replay %d instance %d differs" % (r, i)
    graph.close()
    b.run(items)
    assert b.wait() == 0
    for o, ex in zip(outs, eager):
        assert torch.equal(o, ex)


def test_autotune_registers_and_roundtrips_json(device, oracle, tmp_path):
    import torch
    from bri17_amd import StiffnessMatrix, autotune
    autotune.clear()
    shape = (48, 32, 64)
    best, ms, tried = autotune.tune(shape, "float32", reps=3)
    print("autotune", shape, best, "%.4f ms" % ms, tried, "candidates")
    assert ms > 0 and tried >= 10
    path = os.path.join(str(tmp_path), "tuning.json")
    autotune.save(path)
    data = json.load(open(path))
    assert "3:48x32x64:f32" in data
    autotune.clear()
    assert autotune.load(path) == 1
    op = StiffnessMatrix(shape, None, 1.0, 0.3, "float32")
    d = op.describe()
    if best["force_generic"]:
        assert d["fixed_size_kernels"] is False
        if best["pencil_width"]:
            assert d["pencil"]["width"] == best["pencil_width"]
    u = torch.rand((3,) + shape, device=device) * 2 - 1
    got = op.apply(u).double().cpu().numpy()
    ref = oracle.apply(shape, (1.0, 1.0, 1.0), 1.0, 0.3, u.double().cpu().numpy())
    assert rel_err(got, ref) <= 1e-5
    autotune.clear()
