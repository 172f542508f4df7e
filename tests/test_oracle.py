"""Pins the CPU oracle (oracle/bri17_oracle.c) before it is trusted as the
checker: against the reference's own known-answer tests and fixtures
(tests/test_bri17.cpp SECTIONs "2D/3D stiffness matrix" and "2D/3D
strain-displacement matrix"), a brute-force DFT, numpy.fft, closed forms and
invariants."""
import numpy as np
import pytest

from helpers import assemble_stiffness, assemble_strain, assert_equal, dft_bruteforce, rel_err


@pytest.mark.parametrize("D", [2, 3])
def test_element_stiffness_matches_maxima(oracle, fixtures, D):
    Ke = oracle.element_stiffness(fixtures["spacing%d" % D], fixtures["mu"], fixtures["nu"])
    assert_equal(fixtures["Ke%d" % D], Ke, fixtures["rtol"], fixtures["atol"])


@pytest.mark.parametrize("D", [2, 3])
def test_element_strain_displacement_matches_maxima(oracle, fixtures, D):
    Be = oracle.element_strain_displacement(fixtures["spacing%d" % D])
    assert_equal(fixtures["Be%d" % D], Be, fixtures["rtol"], fixtures["atol"])


@pytest.mark.parametrize("D", [2, 3])
def test_reference_stiffness_section(oracle, fixtures, D):
    """tests/test_bri17.cpp:337-361 (2D) and :363-536 (3D): the FFT-based
    operator probed column by column equals the assembly of the Maxima Ke at
    rtol 1e-15 / atol 1e-14, with |imag| <= 1e-14 (:140)."""
    shape, L = fixtures["shape%d" % D], fixtures["L%d" % D]
    K_exp = assemble_stiffness(oracle, shape, L, fixtures["Ke%d" % D])
    n = K_exp.shape[0]
    K_act = np.zeros_like(K_exp)
    imax = 0.0
    for j in range(n):
        u = np.zeros((D,) + tuple(shape))
        u.flat[j] = 1.0
        Ku, im = oracle.apply(shape, L, fixtures["mu"], fixtures["nu"], u, return_imag=True)
        K_act[:, j] = Ku.ravel()
        imax = max(imax, im)
    assert imax <= fixtures["imag_tol"]
    assert_equal(K_exp, K_act, fixtures["rtol"], fixtures["atol"])


@pytest.mark.parametrize("D", [2, 3])
def test_reference_strain_section(oracle, fixtures, D):
    """tests/test_bri17.cpp:538-605: the FFT-based strain operator equals the
    assembly of the Maxima Be."""
    shape, L = fixtures["shape%d" % D], fixtures["L%d" % D]
    B_exp = assemble_strain(oracle, shape, L, fixtures["Be%d" % D])
    B_act = np.zeros_like(B_exp)
    for j in range(B_exp.shape[1]):
        u = np.zeros((D,) + tuple(shape))
        u.flat[j] = 1.0
        B_act[:, j] = oracle.strain(shape, L, u).ravel()
    assert_equal(B_exp, B_act, fixtures["rtol"], fixtures["atol"])


@pytest.mark.parametrize("shape", [(1,), (2,), (7,), (12,), (16,), (3, 4), (3, 4, 5), (6, 9, 10), (7, 11), (2, 3, 49)])
def test_fft_matches_bruteforce_and_numpy(oracle, shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    X = oracle.fft(x, -1)
    ref = dft_bruteforce(x)
    scale = np.abs(ref).max()
    assert np.abs(X - ref).max() <= 1e-13 * scale
    assert np.abs(X - np.fft.fftn(x)).max() <= 1e-13 * scale
    xb = oracle.fft(X, +1) / x.size
    assert np.abs(xb - x).max() <= 1e-13 * np.abs(x).max()


def test_fft_impulse_gives_twiddles(oracle):
    n = 48
    x = np.zeros(n, complex)
    x[5] = 1.0
    X = oracle.fft(x, -1)
    k = np.arange(n)
    assert np.abs(X - np.exp(-2j * np.pi * ((5 * k) % n) / n)).max() < 4e-15


def _ks(shape):
    return [tuple(k) for k in np.ndindex(*shape)]


@pytest.mark.parametrize("shape,L", [((6, 8), (1.0, 1.3)), ((6, 8, 10), (1.0, 1.2, 0.7))])
def test_modal_stiffness_invariants(oracle, shape, L):
    mu, nu = 1.7, 0.27
    D = len(shape)
    for k in _ks(shape):
        K = oracle.modal_stiffness(shape, L, mu, nu, k)
        assert np.all(K.imag == 0)
        Kr = K.real
        assert np.allclose(Kr, Kr.T, rtol=0, atol=0)
        kn = tuple((-np.array(k)) % np.array(shape))
        assert np.allclose(oracle.modal_stiffness(shape, L, mu, nu, kn).real, Kr, rtol=1e-12, atol=1e-12)
        assert np.allclose(oracle.modal_stiffness(shape, L, 2 * mu, nu, k).real, 2 * Kr, rtol=1e-14, atol=1e-12)
        if any(k):
            assert np.linalg.eigvalsh(Kr).min() > 0
        else:
            assert np.all(Kr == 0)
    if D == 2:
        # special case: 2D off-diagonal = (lambda + mu) Re(B0 conj B1)
        s = mu / (1 - 2 * nu)
        for k in _ks(shape):
            K = oracle.modal_stiffness(shape, L, mu, nu, k).real
            B = oracle.modal_strain_displacement(shape, L, k)
            assert abs(K[0, 1] - s * (B[0] * np.conj(B[1])).real) <= 1e-12 * (1 + abs(K[0, 1]))


def test_modal_strain_displacement_symmetries(oracle):
    shape, L = (6, 8, 10), (1.0, 1.2, 0.7)
    for k in _ks(shape):
        B = oracle.modal_strain_displacement(shape, L, k)
        kp = tuple(np.array(k) + np.array(shape))
        assert np.allclose(oracle.modal_strain_displacement(shape, L, kp), B, atol=1e-12)
        kn = tuple((-np.array(k)) % np.array(shape))
        assert np.allclose(oracle.modal_strain_displacement(shape, L, kn), np.conj(B), atol=1e-12)


@pytest.mark.parametrize("shape,L,mu,nu", [((64, 64), (1.0, 1.0), 1.0, 0.3),
                                           ((6, 8, 10), (1.0, 1.5, 2.0), 1.0, 0.3),
                                           ((9, 4, 5), (0.9, 1.1, 1.3), 2.5, 0.1)])
def test_ebe_product_matches_fft_apply(oracle, shape, L, mu, nu):
    """The FFT apply equals the matrix-free element-by-element product with
    the Gauss-quadrature Ke at the same (h, mu, nu)."""
    D = len(shape)
    h = [l / n for l, n in zip(L, shape)]
    Ke = oracle.element_stiffness(h, mu, nu)
    u = oracle.fill_uniform(11, D * int(np.prod(shape))).reshape((D,) + shape)
    a = oracle.apply(shape, L, mu, nu, u)
    b = oracle.ebe_apply(shape, L, Ke, u)
    assert rel_err(a, b) <= 1e-13


@pytest.mark.parametrize("shape,L", [((12, 10), (1.0, 0.8)), ((6, 8, 10), (1.0, 1.5, 2.0))])
def test_operator_symmetry_positivity_rigid(oracle, shape, L):
    D = len(shape)
    n = D * int(np.prod(shape))
    u = oracle.fill_uniform(3, n).reshape((D,) + shape)
    v = oracle.fill_uniform(4, n).reshape((D,) + shape)
    Ku = oracle.apply(shape, L, 1.3, 0.2, u)
    Kv = oracle.apply(shape, L, 1.3, 0.2, v)
    assert abs(np.vdot(v, Ku) - np.vdot(Kv, u)) <= 1e-12 * abs(np.vdot(v, Ku))
    assert np.vdot(u, Ku) > 0
    c = np.ones((D,) + shape) * np.arange(1, D + 1).reshape((D,) + (1,) * D)
    assert np.abs(oracle.apply(shape, L, 1.3, 0.2, c)).max() < 1e-12


@pytest.mark.parametrize("shape,L", [((12, 10), (1.0, 0.8)), ((6, 8, 10), (1.0, 1.5, 2.0)), ((3, 4, 5), (3.3, 4.8, 6.5))])
def test_green_solve_round_trips(oracle, shape, L):
    D = len(shape)
    n = D * int(np.prod(shape))
    mu, nu = 0.8, 0.35
    f = oracle.fill_normal(5, n).reshape((D,) + shape)
    u = oracle.solve(shape, L, mu, nu, f)
    axes = tuple(range(1, D + 1))
    assert np.abs(u.mean(axis=axes)).max() < 1e-13
    f0 = f - f.mean(axis=axes, keepdims=True)
    assert rel_err(oracle.apply(shape, L, mu, nu, u), f0) <= 1e-12
    w = oracle.fill_uniform(6, n).reshape((D,) + shape)
    w0 = w - w.mean(axis=axes, keepdims=True)
    assert rel_err(oracle.solve(shape, L, mu, nu, oracle.apply(shape, L, mu, nu, w)), w0) <= 1e-12


@pytest.mark.parametrize("shape,L", [((8, 6), (1.0, 0.8)), ((4, 6, 5), (1.0, 1.5, 2.0))])
def test_eigenstress_operator_real_space_route(oracle, shape, L):
    """-eps from the per-k eigenstress operator (bri17.hpp:308-355) equals the
    strain of the displacement solving K u = -|h| B^T tau in real space, built
    from the fixture-pinned element B and the pinned Green solve."""
    D = len(shape)
    S = 3 if D == 2 else 6
    mu, nu = 1.1, 0.3
    N = int(np.prod(shape))
    h = [l / n for l, n in zip(L, shape)]
    hv = float(np.prod(h))
    tau = oracle.fill_normal(9, S * N).reshape((S,) + shape)
    eta = oracle.eigenstress_strain(shape, L, mu, nu, tau)
    Be = oracle.element_strain_displacement(h)
    f = -hv * oracle.ebe_strain_transpose(shape, L, Be, tau)
    u = oracle.solve(shape, L, mu, nu, f)
    eps = oracle.ebe_strain(shape, L, Be, u)
    assert rel_err(eta, -eps) <= 1e-12


def test_generator_is_deterministic(oracle):
    a = oracle.fill_uniform(42, 100000)
    assert np.array_equal(a, oracle.fill_uniform(42, 100000))
    assert a.min() >= -1 and a.max() < 1 and abs(a.mean()) < 0.01
    z = oracle.fill_normal(7, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
