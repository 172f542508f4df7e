"""Strain B u and the eigenstress operator on the device (the field forms of
Hooke::modal_strain_displacement, bri17.hpp:212-236, and of
Hooke::modal_eigenstress_to_opposite_strain, bri17.hpp:308-355).

The strain operator is pinned by the reference's own strain SECTIONs
(tests/test_bri17.cpp:538-605: probed column by column against the assembled
Maxima Be at rtol 1e-15 / atol 1e-14) and by the fp64 oracle at other sizes.
The eigenstress operator is checked against the oracle (itself pinned by an
independent real-space route) and against the per-frequency host function
of the C ABI composed with numpy's FFT."""
import os

import numpy as np
import pytest

from helpers import assemble_strain, assert_equal, rel_err

pytestmark = pytest.mark.gpu

NT = min(16, os.cpu_count() or 1)

CASES = [
    ((3, 4), (3.3, 4.8)),
    ((3, 4, 5), (3.3, 4.8, 6.5)),
    ((64, 64), (1.0, 1.0)),
    ((27, 12, 10), (2.0, 1.0, 0.5)),
    ((48, 36, 54), (1.0, 1.0, 1.0)),
    ((7, 14, 49), (1.0, 2.0, 3.0)),
    ((243, 12), (1.0, 1.0)),
    ((64, 64, 64), (1.0, 1.0, 1.0)),   # fixed-size plan: these modes run the runtime-size kernels
    ((1, 6, 5), (1.0, 1.0, 1.0)),
]


@pytest.mark.parametrize("D", [2, 3])
def test_reference_strain_section_on_device(device, oracle, fixtures, D):
    import torch
    from bri17_amd import StiffnessMatrix
    shape, L = tuple(fixtures["shape%d" % D]), fixtures["L%d" % D]
    op = StiffnessMatrix(shape, L, fixtures["mu"], fixtures["nu"], "float64")
    B_exp = assemble_strain(oracle, shape, L, fixtures["Be%d" % D])
    n = B_exp.shape[1]
    u = torch.zeros((n, D) + shape, dtype=torch.float64, device=device)
    u.view(n, n)[torch.arange(n), torch.arange(n)] = 1.0
    out = torch.empty((n, op.sym) + shape, dtype=torch.float64, device=device)
    for j in range(n):
        op.strain(u[j], out=out[j])
    torch.cuda.synchronize()
    B_act = out.view(n, -1).cpu().numpy().T
    assert_equal(B_exp, B_act, fixtures["rtol"], fixtures["atol"])


@pytest.mark.parametrize("shape,L", CASES)
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_strain_matches_oracle(device, oracle, shape, L, dtype):
    import torch
    from bri17_amd import StiffnessMatrix
    u = oracle.fill_uniform(21, len(shape) * int(np.prod(shape))).reshape((len(shape),) + shape)
    u = u.astype(dtype)
    op = StiffnessMatrix(shape, L, 1.0, 0.3, dtype)
    got = op.strain(torch.from_numpy(u).to(device)).double().cpu().numpy()
    ref = oracle.strain(shape, L, u.astype(np.float64), nthreads=NT)
    err = rel_err(got, ref)
    print(shape, dtype, "strain rel err %.2e" % err)
    assert err <= (1e-12 if dtype == "float64" else 1e-5)


@pytest.mark.parametrize("shape,L", CASES)
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_eigenstress_matches_oracle(device, oracle, shape, L, dtype):
    import torch
    from bri17_amd import StiffnessMatrix
    S = 3 if len(shape) == 2 else 6
    tau = oracle.fill_uniform(22, S * int(np.prod(shape))).reshape((S,) + shape).astype(dtype)
    op = StiffnessMatrix(shape, L, 1.7, 0.23, dtype)
    got = op.eigenstress(torch.from_numpy(tau).to(device)).double().cpu().numpy()
    ref = oracle.eigenstress_strain(shape, L, 1.7, 0.23, tau.astype(np.float64), nthreads=NT)
    err = rel_err(got, ref)
    print(shape, dtype, "eigenstress rel err %.2e" % err)
    assert err <= (1e-12 if dtype == "float64" else 1e-5)


@pytest.mark.parametrize("shape", [(6, 5), (4, 3, 5)])
def test_eigenstress_equals_per_frequency_function(device, shape):
    """The field operator equals the per-k host function of the C ABI
    (Hooke.modal_eigenstress_to_opposite_strain) applied to numpy's DFT of
    tau at every k, followed by the inverse DFT (python/demo.py:24-40)."""
    import torch
    from bri17_amd import StiffnessMatrix, pybri17
    D = len(shape)
    S = 3 if D == 2 else 6
    L = tuple(1.0 + 0.3 * d for d in range(D))
    rng = np.random.default_rng(5)
    tau = rng.standard_normal((S,) + shape)
    grid = getattr(pybri17, "CartesianGrid%df64" % D)(shape, L)
    hooke = getattr(pybri17, "Hooke%df64" % D)(0.8, 0.35, grid)
    th = np.fft.fftn(tau, axes=range(1, D + 1))
    eh = np.zeros_like(th)
    for k in np.ndindex(*shape):
        out = np.zeros(S, dtype=np.complex128)
        hooke.modal_eigenstress_to_opposite_strain(np.array(k, dtype=np.intc),
                                                   np.ascontiguousarray(th[(slice(None),) + k]), out)
        eh[(slice(None),) + k] = out
    ref = np.fft.ifftn(eh, axes=range(1, D + 1))
    assert np.abs(ref.imag).max() <= 1e-12 * np.abs(ref.real).max()
    op = StiffnessMatrix(shape, L, 0.8, 0.35, "float64")
    got = op.eigenstress(torch.from_numpy(tau).to(device)).cpu().numpy()
    assert rel_err(got, ref.real) <= 1e-12


def test_strain_host_pointers_and_graph(device):
    import torch
    from bri17_amd import StiffnessMatrix, _lib
    shape = (24, 20, 18)
    op = StiffnessMatrix(shape, None, 1.1, 0.3, "float32")
    g = torch.Generator(device=device).manual_seed(4)
    u = torch.rand((3,) + shape, generator=g, device=device) * 2 - 1
    eps = op.strain(u)
    host = op.strain(u.cpu().numpy())
    torch.cuda.synchronize()
    assert np.array_equal(host, eps.cpu().numpy())
    out = torch.empty_like(eps)
    graph = op.capture(u, out, "strain")
    for r in range(3):
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        graph.launch()
        torch.cuda.synchronize()
        assert torch.equal(out, eps), "replay %d" % r
    # an operator captured with a DIM-component workspace cannot grow it later
    op2 = StiffnessMatrix(shape, None, 1.1, 0.3, "float32")
    ku = torch.empty_like(u)
    op2.capture(u, ku, "apply")
    with pytest.raises(_lib.Bri17Error) as e:
        op2.strain(u)
    assert e.value.code == _lib.BRI17_ERR_STATE
