"""Parity at the large and odd grid sides of the serving mix, and at material
constants in SI units.

The serving workload draws cubic sides log-uniform in [32, 256] rounded to
2^a 3^b; the sides 144, 162, 192, 216 and 243 (about a quarter of the draws)
run the runtime-size kernels with size-dependent middle-axis splits (243^3:
split 3, 157 KiB plane items).  Each is checked here against the fp64 oracle
(the restatement of tests/test_bri17.cpp:56-107 and of the Green-solve
conventions of bri17.hpp:336-341) at the same tolerances as the other parity
tests: fp32 within 1e-5, fp64 within 1e-12 (relative, 2-norm).

The reference's Hooke<float> solve uses K.llt() (bri17.hpp:340) and works at
any modulus; the device inverts a normalised block, so a steel-like
mu = 8e10 Pa on millimetre and metre cells must give the same relative
accuracy as mu = 1."""
import os

import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

NT = min(16, os.cpu_count() or 1)


def _field(oracle, seed, comps, shape, dist="uniform", zero_mean=False):
    n = comps * int(np.prod(shape))
    x = oracle.fill_uniform(seed, n) if dist == "uniform" else oracle.fill_normal(seed, n)
    x = x.reshape((comps,) + tuple(shape))
    if zero_mean:
        x -= x.mean(axis=tuple(range(1, len(shape) + 1)), keepdims=True)
    return x


def _run(device, op, x, mode):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(x)).to(device)
    out = getattr(op, mode)(t)
    torch.cuda.synchronize()
    return out.double().cpu().numpy()


SERVING_SIDES = [144, 162, 192, 216, 243]


@pytest.mark.parametrize("side", SERVING_SIDES)
@pytest.mark.parametrize("mode", ["apply", "solve"])
def test_serving_sides_fp32(device, oracle, side, mode):
    from bri17_amd import StiffnessMatrix
    shape, L, mu, nu = (side,) * 3, (1.0, 1.0, 1.0), 1.3, 0.27
    x = _field(oracle, 400 + side, 3, shape, "normal", zero_mean=(mode == "solve"))
    x32 = x.astype(np.float32)
    op = StiffnessMatrix(shape, L, mu, nu, "float32")
    print("%d^3 fp32 %s plan %s" % (side, mode, op.describe()["plane"]), flush=True)
    got = _run(device, op, x32, mode)
    fn = oracle.apply if mode == "apply" else oracle.solve
    err = rel_err(got, fn(shape, L, mu, nu, x32.astype(np.float64), nthreads=NT))
    print("  rel err %.3e" % err, flush=True)
    assert err <= 1e-5


@pytest.mark.parametrize("shape", [(162, 162, 162), (216, 256, 162)])
@pytest.mark.parametrize("mode", ["apply", "solve"])
def test_serving_sides_fp64(device, oracle, shape, mode):
    from bri17_amd import StiffnessMatrix
    L, mu, nu = (1.0, 1.2, 0.8), 0.9, 0.31
    x = _field(oracle, 500 + shape[0], 3, shape, "normal", zero_mean=(mode == "solve"))
    op = StiffnessMatrix(shape, L, mu, nu, "float64")
    print("%s fp64 %s plan %s" % (shape, mode, op.describe()["plane"]), flush=True)
    got = _run(device, op, x, mode)
    fn = oracle.apply if mode == "apply" else oracle.solve
    err = rel_err(got, fn(shape, L, mu, nu, x, nthreads=NT))
    print("  rel err %.3e" % err, flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("shape", [(216, 256, 162), (243, 144, 192)])
def test_serving_non_cubic_fp32(device, oracle, shape):
    from bri17_amd import StiffnessMatrix
    L, mu, nu = (1.0, 0.9, 1.1), 1.1, 0.2
    x32 = _field(oracle, 600 + shape[2], 3, shape).astype(np.float32)
    op = StiffnessMatrix(shape, L, mu, nu, "float32")
    got = _run(device, op, x32, "apply")
    err = rel_err(got, oracle.apply(shape, L, mu, nu, x32.astype(np.float64), nthreads=NT))
    print(shape, "fp32 apply rel err %.3e" % err, flush=True)
    assert err <= 1e-5


# (shape, cell lengths): metres and millimetres
SI_CASES = [((64, 64, 64), (1.0, 1.0, 1.0)), ((64, 64, 64), (1e-3, 1e-3, 1e-3)),
            ((48, 36, 54), (2e-3, 1e-3, 3e-3)), ((96, 80), (1e-3, 2e-3))]


@pytest.mark.parametrize("shape,L", SI_CASES)
@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_solve_si_modulus(device, oracle, shape, L, dtype):
    """Green solve at mu = 8e10 (steel, Pa), nu = 0.3."""
    from bri17_amd import StiffnessMatrix
    mu, nu = 8e10, 0.3
    D = len(shape)
    f = _field(oracle, 700, D, shape, "normal", zero_mean=True).astype(dtype)
    got = _run(device, StiffnessMatrix(shape, L, mu, nu, dtype), f, "solve")
    assert np.all(np.isfinite(got))
    err = rel_err(got, oracle.solve(shape, L, mu, nu, f.astype(np.float64), nthreads=NT))
    print(shape, L, dtype, "solve rel err %.3e" % err, flush=True)
    assert err <= (1e-5 if dtype == "float32" else 1e-12)


@pytest.mark.parametrize("shape,L", SI_CASES)
@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_eigenstress_si_modulus(device, oracle, shape, L, dtype):
    from bri17_amd import StiffnessMatrix
    mu, nu = 8e10, 0.3
    S = 3 if len(shape) == 2 else 6
    tau = (1e8 * _field(oracle, 701, S, shape)).astype(dtype)  # stresses of order 100 MPa
    got = _run(device, StiffnessMatrix(shape, L, mu, nu, dtype), tau, "eigenstress")
    assert np.all(np.isfinite(got))
    err = rel_err(got, oracle.eigenstress_strain(shape, L, mu, nu, tau.astype(np.float64), nthreads=NT))
    print(shape, L, dtype, "eigenstress rel err %.3e" % err, flush=True)
    assert err <= (1e-5 if dtype == "float32" else 1e-12)


def test_solve_si_modulus_256_fp32(device, oracle):
    """256^3 fp32 (fixed-size kernels) at mu = 8e10: the largest block of the
    unit cube is about 1e17 and its determinant about 1e51, beyond fp32."""
    from bri17_amd import StiffnessMatrix
    shape, L, mu, nu = (256, 256, 256), (1.0, 1.0, 1.0), 8e10, 0.25
    f32 = _field(oracle, 702, 3, shape, "normal", zero_mean=True).astype(np.float32)
    op = StiffnessMatrix(shape, L, mu, nu, "float32")
    got = _run(device, op, f32, "solve")
    assert np.all(np.isfinite(got))
    f64 = f32.astype(np.float64)
    err = rel_err(got, oracle.solve(shape, L, mu, nu, f64, nthreads=NT))
    print("256^3 fp32 mu=8e10 solve rel err %.3e" % err, flush=True)
    assert err <= 1e-5
    got = _run(device, op, f32, "apply")
    err = rel_err(got, oracle.apply(shape, L, mu, nu, f64, nthreads=NT))
    print("256^3 fp32 mu=8e10 apply rel err %.3e" % err, flush=True)
    assert err <= 1e-5
