"""Parity of the HIP path (libbri17_amd.so through the C ABI) with the CPU
oracle and with the reference's own known-answer tests.

Tolerances: fp64 results within 1e-12 relative (2-norm) of the fp64 oracle;
fp32 results within 1e-5 relative of the fp64 oracle applied to the same
(fp32-rounded) input; the reference's dense-operator test at its own
rtol 1e-15 / atol 1e-14."""
import os

import numpy as np
import pytest

from helpers import assemble_stiffness, assert_equal, rel_err

pytestmark = pytest.mark.gpu

NT = min(16, os.cpu_count() or 1)


def _dev_run(device, op, x, mode="apply"):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(x)).to(device)
    out = op.apply(t) if mode == "apply" else op.solve(t)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _op(shape, L, mu, nu, dtype):
    from bri17_amd import StiffnessMatrix
    return StiffnessMatrix(shape, L, mu, nu, dtype)


@pytest.mark.parametrize("D", [2, 3])
def test_reference_stiffness_section_on_device(device, oracle, fixtures, D):
    """tests/test_bri17.cpp:337-361 / :363-536 re-run on the device apply:
    probe K column by column (:130-150) and compare with the assembly of the
    Maxima Ke (:153-174) at rtol 1e-15, atol 1e-14."""
    import torch
    shape, L = tuple(fixtures["shape%d" % D]), fixtures["L%d" % D]
    op = _op(shape, L, fixtures["mu"], fixtures["nu"], "float64")
    K_exp = assemble_stiffness(oracle, shape, L, fixtures["Ke%d" % D])
    n = K_exp.shape[0]
    u = torch.zeros((n,) + (D,) + shape, dtype=torch.float64, device=device)
    u.view(n, n)[torch.arange(n), torch.arange(n)] = 1.0
    out = torch.empty_like(u)
    for j in range(n):
        op.apply(u[j], out=out[j])
    torch.cuda.synchronize()
    K_act = out.view(n, n).cpu().numpy().T
    assert_equal(K_exp, K_act, fixtures["rtol"], fixtures["atol"])


CASES = [
    ((64, 64), (1.0, 1.0)),            # 2D Q4 64x64 unit square
    ((3, 4), (3.3, 4.8)),              # reference 2D fixture grid
    ((3, 4, 5), (3.3, 4.8, 6.5)),      # reference 3D fixture grid (radix 3 and 5)
    ((6, 8, 10), (1.0, 1.5, 2.0)),
    ((16, 18, 20), (1.0, 1.0, 1.0)),
    ((32, 32, 32), (1.0, 1.0, 1.0)),
    ((27, 12, 10), (2.0, 1.0, 0.5)),
    ((48, 36, 54), (1.0, 1.0, 1.0)),
    ((81, 9, 16), (1.0, 1.0, 1.0)),
    ((7, 14, 49), (1.0, 2.0, 3.0)),
    ((5, 7), (1.0, 1.0)),
    ((243, 12), (1.0, 1.0)),
    ((512, 96), (1.0, 0.2)),
    ((1, 6, 5), (1.0, 1.0, 1.0)),
    ((11, 13, 22), (1.0, 1.2, 0.8)),   # radix 11 and 13 (odd rows of 22 reals)
    ((121, 26), (1.0, 1.0)),
    ((13, 44, 39), (2.0, 1.0, 1.5)),
    ((17, 19, 23), (1.0, 1.1, 0.9)),   # prime factors above 13: direct-DFT passes
    ((97, 34), (1.0, 1.0)),
    ((31, 62, 8), (1.0, 2.0, 0.5)),
    ((64, 64, 64), (1.0, 1.0, 1.0)),    # fixed-size kernels (fp32 and fp64)
]


@pytest.mark.parametrize("shape,dtype", [((64, 64, 64), "float32"), ((128, 128, 128), "float32"),
                                         ((256, 256, 256), "float32"), ((64, 64, 64), "float64"),
                                         ((128, 128, 128), "float64")])
@pytest.mark.parametrize("mode", ["apply", "solve"])
def test_fixed_size_kernels_match_generic(device, shape, dtype, mode):
    """The compile-time-size kernels and the runtime-size kernels implement the
    same pipeline on the same workspace layout: their results agree to
    rounding, on both K u and the Green solve."""
    import torch
    fast = _op(shape, (1.0, 1.3, 0.7), 1.3, 0.27, dtype)
    gen = _op(shape, (1.0, 1.3, 0.7), 1.3, 0.27, dtype)
    gen.close()
    from bri17_amd import StiffnessMatrix
    gen = StiffnessMatrix(shape, (1.0, 1.3, 0.7), 1.3, 0.27, dtype, tuning=dict(force_generic=1))
    assert fast.describe()["fixed_size_kernels"] is True
    assert gen.describe()["fixed_size_kernels"] is False
    g = torch.Generator(device=device).manual_seed(11)
    x = torch.randn((3,) + shape, generator=g, device=device, dtype=getattr(torch, dtype))
    a = fast.apply(x) if mode == "apply" else fast.solve(x)
    b = gen.apply(x) if mode == "apply" else gen.solve(x)
    torch.cuda.synchronize()
    err = float((a - b).norm() / b.norm())
    print(shape, dtype, mode, "fixed vs generic %.2e" % err)
    assert err <= (2e-6 if dtype == "float32" else 1e-13)


def _field(oracle, seed, shape, dist="uniform"):
    D = len(shape)
    n = D * int(np.prod(shape))
    if dist == "uniform":
        return oracle.fill_uniform(seed, n).reshape((D,) + tuple(shape))
    return oracle.fill_normal(seed, n).reshape((D,) + tuple(shape))


@pytest.mark.parametrize("shape,L", CASES)
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_apply_matches_oracle(device, oracle, shape, L, dtype):
    mu, nu = 1.0, 0.3
    u = _field(oracle, 1, shape).astype(dtype)
    print("apply", shape, dtype, flush=True)
    got = _dev_run(device, _op(shape, L, mu, nu, dtype), u)
    ref = oracle.apply(shape, L, mu, nu, u.astype(np.float64), nthreads=NT)
    err = rel_err(got, ref)
    print("  rel err %.3e" % err)
    assert err <= (1e-12 if dtype == "float64" else 1e-5)


@pytest.mark.parametrize("shape,L", CASES)
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_solve_matches_oracle(device, oracle, shape, L, dtype):
    mu, nu = 0.7, 0.25
    f = _field(oracle, 2, shape, "normal")
    f -= f.mean(axis=tuple(range(1, len(shape) + 1)), keepdims=True)
    f = f.astype(dtype)
    print("solve", shape, dtype, flush=True)
    got = _dev_run(device, _op(shape, L, mu, nu, dtype), f, "solve")
    ref = oracle.solve(shape, L, mu, nu, f.astype(np.float64), nthreads=NT)
    err = rel_err(got, ref)
    print("  rel err %.3e" % err)
    assert err <= (1e-12 if dtype == "float64" else 1e-5)


def test_config1_2d_64_against_element_product(device, oracle):
    """2D plane-strain Q4 on the periodic 64x64 unit square, fp64, mu=1,
    nu=0.3: device K u against the reference apply restatement and against
    the assembled element-stiffness product."""
    shape, L, mu, nu = (64, 64), (1.0, 1.0), 1.0, 0.3
    u = _field(oracle, 101, shape)
    got = _dev_run(device, _op(shape, L, mu, nu, "float64"), u)
    assert rel_err(got, oracle.apply(shape, L, mu, nu, u)) <= 1e-12
    Ke = oracle.element_stiffness([1 / 64, 1 / 64], mu, nu)
    assert rel_err(got, oracle.ebe_apply(shape, L, Ke, u)) <= 1e-12


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_config2_3d_128(device, oracle, dtype):
    """3D H8, 128^3 unit cube, mu=1, nu=0.3, u ~ U[-1,1]: single apply against
    the fp64 oracle and the element-by-element product."""
    shape, L, mu, nu = (128, 128, 128), (1.0, 1.0, 1.0), 1.0, 0.3
    u = _field(oracle, 202, shape).astype(dtype)
    got = _dev_run(device, _op(shape, L, mu, nu, dtype), u)
    u64 = u.astype(np.float64)
    ref = oracle.apply(shape, L, mu, nu, u64, nthreads=NT)
    err = rel_err(got, ref)
    print("128^3 %s rel err vs oracle %.3e" % (dtype, err))
    assert err <= (1e-12 if dtype == "float64" else 1e-5)
    if dtype == "float64":
        Ke = oracle.element_stiffness([1 / 128] * 3, mu, nu)
        assert rel_err(got, oracle.ebe_apply(shape, L, Ke, u64, nthreads=NT)) <= 1e-12


def test_config3_3d_256_fp32(device, oracle):
    """3D H8 at 256^3, fp32 with the fp64 oracle, mu=1, nu=0.25, f ~ N(0,1)
    with zero mean: K f and the Green solve K u = f, relative error of each."""
    shape, L, mu, nu = (256, 256, 256), (1.0, 1.0, 1.0), 1.0, 0.25
    f = _field(oracle, 303, shape, "normal")
    f -= f.mean(axis=(1, 2, 3), keepdims=True)
    f32 = f.astype(np.float32)
    f64 = f32.astype(np.float64)
    op = _op(shape, L, mu, nu, "float32")
    ku = _dev_run(device, op, f32)
    e1 = rel_err(ku, oracle.apply(shape, L, mu, nu, f64, nthreads=NT))
    gu = _dev_run(device, op, f32, "solve")
    e2 = rel_err(gu, oracle.solve(shape, L, mu, nu, f64, nthreads=NT))
    print("256^3 fp32 rel err: K.u %.3e  Green %.3e" % (e1, e2))
    assert e1 <= 1e-5 and e2 <= 1e-5


def test_fixed_size_plan_description(device):
    """The description of a fixed-size plan names the kernels it launches: the
    register-block pencil for 256^3 fp32 (N0 = 16^2), the LDS pencil when
    asked for (bit 3 of plane_prefetch) and for the other fixed sizes."""
    from bri17_amd import StiffnessMatrix
    d = StiffnessMatrix((256, 256, 256), None, 1.0, 0.3, "float32").describe()
    assert d["fixed_size_kernels"] and d["fixed_size"]["pencil"] == "register-block"
    assert d["fixed_size"]["pencil_width"] == 16 and d["fixed_size"]["plane_split"] == 2
    d = StiffnessMatrix((256, 256, 256), None, 1.0, 0.3, "float32", tuning=dict(plane_prefetch=9)).describe()
    assert d["fixed_size"]["pencil"] == "lds"
    d = StiffnessMatrix((128, 128, 128), None, 1.0, 0.3, "float32").describe()
    assert d["fixed_size"]["pencil"] == "lds"
    assert "fixed_size" not in StiffnessMatrix((48, 36, 54), None, 1.0, 0.3, "float32").describe()


@pytest.mark.parametrize("shape,dtype", [((256, 256, 256), "float32"), ((128, 128, 128), "float32"),
                                         ((64, 64, 64), "float64")])
def test_fixed_size_launch_variants(device, shape, dtype):
    """Every launch variant of the fixed-size kernels gives the default
    variant's result: the persistent prefetching plane kernels (bits 0 and 1
    of plane_prefetch) and the persistent LDS pencil (bit 2) bitwise, the LDS
    pencil in place of the register-block one (bit 3, 256^3) within 1e-6.  The
    default variant itself is checked against the oracle above."""
    import torch
    from bri17_amd import StiffnessMatrix
    g = torch.Generator(device=device).manual_seed(7)
    u = torch.rand((3,) + shape, generator=g, device=device, dtype=getattr(torch, dtype)) * 2 - 1
    reg = StiffnessMatrix(shape, None, 1.1, 0.3, dtype).describe()["fixed_size"]["pencil"]
    for mode in ("apply", "solve"):
        ref = getattr(StiffnessMatrix(shape, None, 1.1, 0.3, dtype), mode)(u)
        for pf in range(16):
            op = StiffnessMatrix(shape, None, 1.1, 0.3, dtype, tuning=dict(plane_prefetch=pf))
            got = getattr(op, mode)(u)
            torch.cuda.synchronize()
            print(shape, dtype, mode, "plane_prefetch", pf, flush=True)
            if reg == "register-block" and pf & 8:
                assert rel_err(got.cpu().numpy(), ref.cpu().numpy()) <= 1e-6, (mode, pf)
            else:
                assert torch.equal(got, ref), (mode, pf)
            del op


# Single Fourier modes with at least one component N/8 or more: K^(k) is then
# within a factor of about 7 of its largest value, and fp32 rounding of the
# transforms (spread over all frequencies, then multiplied by K^ there) stays
# near 1e-6 of the output.  (For the smoothest modes, e.g. k = (1, 0, 0),
# that factor reaches ~N^2/10 and every fp32 path, fixed-size or runtime-size,
# gives errors of 1e-4 to 1e-3 of the max-norm; broadband fields stay near
# 3e-7.)
MODES_256 = [(0, 0, 0), (40, 0, 0), (128, 0, 0), (0, 128, 0), (0, 0, 128), (40, 128, 0),
             (40, 0, 128), (0, 128, 128), (40, 64, 0), (40, 64, 128), (44, 192, 128),
             (40, 37, 0), (40, 219, 128), (200, 100, 0), (90, 156, 128), (50, 30, 40),
             (130, 250, 127)]


@pytest.mark.parametrize("pf", [3, 11])
def test_fixed_size_single_modes(device, pf):
    """256^3 fp32 on single Fourier modes u = Re(a exp(i k.x)): K u equals
    |h| Re(K^(k) a exp(i k.x)) and the Green solve Re(K^(k)^-1 a exp(i k.x))/|h|
    (0 at k = 0), with K^(k) from the per-frequency function of the C ABI
    (bri17.hpp:247-292).  The modes sit on the k2 = 0 and k2 = N2/2 planes,
    whose columns the plane kernels pack (k1 = 0, N1/4, N1/2, 3N1/4 and
    partner pairs k1, N1 - k1), and elsewhere; pf 3 runs the register-block
    pencil over packed columns, pf 11 the LDS pencil over separate Nyquist
    columns."""
    import torch
    from bri17_amd import StiffnessMatrix, pybri17
    shape, L, mu, nu = (256, 256, 256), (1.0, 1.1, 0.9), 1.2, 0.3
    h = np.array(L) / np.array(shape)
    hv = float(np.prod(h))
    grid = pybri17.CartesianGrid3f64(shape, L)
    hooke = pybri17.Hooke3f64(mu, nu, grid)
    op = StiffnessMatrix(shape, L, mu, nu, "float32", tuning=dict(plane_prefetch=pf))
    rng = np.random.default_rng(11)

    def modal(k):
        K = np.zeros(9, dtype=np.complex128)
        hooke.modal_stiffness_matrix(np.array(k, dtype=np.intc), K)
        return K.reshape(3, 3).real

    ax = [torch.arange(n, device=device, dtype=torch.float64) for n in shape]
    for k in MODES_256:
        a = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        th = 2 * np.pi * (k[0] * ax[0][:, None, None] / shape[0] + k[1] * ax[1][None, :, None] / shape[1]
                          + k[2] * ax[2][None, None, :] / shape[2])
        c, s = torch.cos(th), torch.sin(th)

        def field(v):  # Re(v exp(i theta)) per component
            return torch.stack([float(v[d].real) * c - float(v[d].imag) * s for d in range(3)])

        u = field(a).float()
        K = modal(k)
        # k = 0: K u = 0 and the solve drops the mean; errors are measured
        # against the output of the mode k = (1, 1, 1) of the same amplitude
        Kr = K if any(k) else modal((1, 1, 1))
        for mode in ("apply", "solve"):
            got = getattr(op, mode)(u).double()
            if mode == "apply":
                exp = field(hv * (K @ a))
                ref = np.abs(hv * (Kr @ a)).max()
            else:
                exp = field(np.linalg.solve(K, a) / hv) if any(k) else torch.zeros_like(got)
                ref = np.abs(np.linalg.solve(Kr, a) / hv).max()
            err = float((got - exp).abs().max()) / ref
            print(k, mode, "pf", pf, "max err %.2e" % err, flush=True)
            assert err <= 1e-5, (k, mode)
        del u, c, s, th


@pytest.mark.parametrize("split", [6, 9, 18, 27])
@pytest.mark.parametrize("mode", ["apply", "solve"])
def test_middle_axis_splits_without_fused_butterfly(device, oracle, split, mode):
    """Splits P whose P-point DFT runs as a Stockham transform in the pencil
    kernel agree with the fp64 oracle."""
    import torch
    from bri17_amd import StiffnessMatrix
    shape, L = (10, 54, 12), (1.0, 1.3, 0.7)
    op = StiffnessMatrix(shape, L, 1.4, 0.27, "float64", tuning=dict(plane_split=split, force_generic=1))
    assert op.describe()["plane"]["split"] == split
    g = torch.Generator(device=device).manual_seed(split)
    u = torch.rand((3,) + shape, generator=g, device=device, dtype=torch.float64) * 2 - 1
    got = (op.apply(u) if mode == "apply" else op.solve(u)).cpu().numpy()
    fn = oracle.apply if mode == "apply" else oracle.solve
    ref = fn(shape, L, 1.4, 0.27, u.cpu().numpy())
    assert rel_err(got, ref) <= 1e-12


@pytest.mark.parametrize("n1,split", [(143, 11), (143, 13), (221, 17), (221, 13)])
@pytest.mark.parametrize("mode", ["apply", "solve"])
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_prime_radix_splits(device, oracle, n1, split, mode, dtype):
    """A middle axis of 143 = 11 x 13 or 221 = 13 x 17 split by either
    factor: the plane kernels run the Q-point transform and the pencil the
    split's P-point Stockham transform, with the radix-11 / radix-13
    butterflies and the direct-DFT pass of 17."""
    import torch
    from bri17_amd import StiffnessMatrix
    shape, L = (10, n1, 12), (1.0, 1.3, 0.7)
    op = StiffnessMatrix(shape, L, 1.4, 0.27, dtype, tuning=dict(plane_split=split))
    assert op.describe()["plane"]["split"] == split
    g = torch.Generator(device=device).manual_seed(split)
    u = torch.rand((3,) + shape, generator=g, device=device, dtype=getattr(torch, dtype)) * 2 - 1
    if mode == "solve":
        u -= u.mean(dim=(1, 2, 3), keepdim=True)
    got = (op.apply(u) if mode == "apply" else op.solve(u)).double().cpu().numpy()
    fn = oracle.apply if mode == "apply" else oracle.solve
    ref = fn(shape, L, 1.4, 0.27, u.double().cpu().numpy())
    e = rel_err(got, ref)
    print(shape, split, mode, dtype, "rel err %.2e" % e, flush=True)
    assert e <= (1e-12 if dtype == "float64" else 1e-5)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_plans_exist_for_large_and_odd_grids(device, dtype):
    """Every grid of the benchmark and serving mixes gets a plan whose
    workgroups fit the 160 KiB of LDS of a CU (plan creation only)."""
    from bri17_amd import StiffnessMatrix
    shapes = [(256, 256, 256), (243, 243, 243), (256, 128, 64), (216, 256, 162), (512, 512),
              (486, 324), (343, 49, 125), (7, 5, 3)]
    for s in shapes:
        d = StiffnessMatrix(s, None, 1.0, 0.3, dtype).describe()
        print(s, dtype, d["plane"]["split"], d["plane"]["lds"], d["pencil"]["width"], d["pencil"]["lds"])
        assert d["plane"]["lds"] <= 160 * 1024 and d["pencil"]["lds"] <= 160 * 1024


def test_deterministic_and_linear(device, oracle):
    import torch
    shape = (48, 36, 54)
    op = _op(shape, (1.0, 1.0, 1.0), 1.2, 0.3, "float32")
    u = torch.from_numpy(_field(oracle, 7, shape).astype(np.float32)).to(device)
    v = torch.from_numpy(_field(oracle, 8, shape).astype(np.float32)).to(device)
    a = op.apply(u)
    b = op.apply(u)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    lin = op.apply(u + 2 * v) - a - 2 * op.apply(v)
    assert float(lin.norm() / a.norm()) < 1e-5


def test_round_trip_green_full_size(device, oracle):
    """Size-independent property at 256^3 fp32: K G f = f - mean(f)."""
    import torch
    shape = (256, 256, 256)
    op = _op(shape, (1.0, 1.0, 1.0), 1.0, 0.25, "float32")
    g = torch.Generator(device=device).manual_seed(5)
    f = torch.randn((3,) + shape, device=device, dtype=torch.float32, generator=g)
    f -= f.mean(dim=(1, 2, 3), keepdim=True)
    kgf = op.apply(op.solve(f))
    torch.cuda.synchronize()
    assert float((kgf - f).norm() / f.norm()) < 1e-5
