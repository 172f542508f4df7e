"""The ``pybri17`` module name of the reference (PYBIND11_MODULE(pybri17, m),
python/pybri17.cpp:98-107) and its one caller, python/demo.py.

``import pybri17`` resolves to this build's module; the per-frequency loop of
python/demo.py:6-40 runs unchanged on it (on a smaller grid, without the
prints).  On the CPU the loop's output is checked against the oracle's
per-frequency restatement of bri17.hpp:308-355; on the GPU, its inverse DFT
is checked against the device eigenstress operator."""
import numpy as np
import pytest

from helpers import rel_err


def _demo_loop(pybri17, dim, n):
    """python/demo.py:6-40 with shape = dim * (n,) and the loop's prints
    removed; returns (tau, eta_hat, hooke)."""
    sym = (dim * (dim + 1)) // 2
    shape = dim * (n,)
    L = dim * (1.0,)
    patch_ratio = dim * (0.125,)
    grid = getattr(pybri17, "CartesianGrid%df64" % dim)(shape, L)
    mu = 1.0
    nu = 0.3
    hooke = getattr(pybri17, "Hooke%df64" % dim)(mu, nu, grid)
    tau_in = np.zeros((sym,), dtype=grid.dtype)
    tau_in[-1] = 1.0
    tau_out = np.zeros_like(tau_in)
    patch_size = tuple(int(r_ * n_) for r_, n_ in zip(patch_ratio, shape))
    tau = np.empty(shape + (sym,), dtype=np.complex128)
    tau[...] = tau_out
    tau[tuple(slice(n_) for n_ in patch_size) + (...,)] = tau_in
    tau_hat = np.fft.fftn(tau, axes=range(dim))
    eta_hat = np.zeros_like(tau)
    k = np.zeros((dim,), dtype=np.intc)
    for idx in np.ndindex(*shape):
        k[:] = idx
        tau_hat_k = tau_hat[idx]
        eta_hat_k = eta_hat[idx]
        hooke.modal_eigenstress_to_opposite_strain(k, tau_hat_k, eta_hat_k)
    return tau, eta_hat, hooke


def test_import_pybri17_module_attributes():
    import pybri17
    assert pybri17.__version__ == "0.1"
    assert pybri17.__author__ == "S. Brisard"
    for name in ("CartesianGrid2f64", "CartesianGrid3f64", "Hooke2f64", "Hooke3f64",
                 "CartesianGrid2f32", "CartesianGrid3f32", "Hooke2f32", "Hooke3f32"):
        assert hasattr(pybri17, name), name
    grid = pybri17.CartesianGrid2f64((256, 256), (1.0, 1.0))
    assert grid.dtype == np.float64 and grid.dim == 2
    hooke = pybri17.Hooke2f64(1.0, 0.3, grid)
    with pytest.raises(ValueError):  # pybri17.cpp:9-15
        hooke.modal_eigenstress_to_opposite_strain(np.zeros(2, np.intc), np.zeros((3, 2), np.complex128),
                                                   np.zeros(3, np.complex128))


@pytest.mark.parametrize("dim,n", [(2, 32), (3, 8)])
def test_demo_loop_matches_oracle(oracle, dim, n):
    import pybri17
    tau, eta_hat, hooke = _demo_loop(pybri17, dim, n)
    tau_hat = np.fft.fftn(tau, axes=range(dim))
    shape = dim * (n,)
    worst = 0.0
    for idx in np.ndindex(*shape):
        ref = oracle.modal_eigenstress_to_opposite_strain(shape, dim * (1.0,), 1.0, 0.3, idx, tau_hat[idx])
        got = eta_hat[idx]
        scale = max(np.abs(ref).max(), 1e-300)
        worst = max(worst, float(np.abs(got - ref).max() / scale))
    print("demo loop vs oracle, worst per-k relative error %.2e" % worst)
    assert worst <= 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("dim,n", [(2, 64), (3, 16)])
def test_demo_loop_matches_device_eigenstress(device, dim, n):
    """The demo's per-k loop, followed by the inverse DFT, equals the device
    eigenstress operator (StiffnessMatrix.eigenstress) on the same tau within
    1e-12 relative."""
    import torch
    import pybri17
    from bri17_amd import StiffnessMatrix
    tau, eta_hat, hooke = _demo_loop(pybri17, dim, n)
    eta = np.fft.ifftn(eta_hat, axes=range(dim))
    assert np.abs(eta.imag).max() <= 1e-12 * np.abs(eta.real).max()
    # field layout of the operator: component-major (SYM, N...)
    tau_c = np.ascontiguousarray(np.moveaxis(tau.real, -1, 0))
    eta_c = np.moveaxis(eta.real, -1, 0)
    op = StiffnessMatrix(dim * (n,), dim * (1.0,), hooke.mu, hooke.nu, "float64")
    got = op.eigenstress(torch.from_numpy(tau_c).to(device)).cpu().numpy()
    err = rel_err(got, eta_c)
    print("demo loop + inverse DFT vs device eigenstress %dD n=%d: %.2e" % (dim, n, err))
    assert err <= 1e-12
