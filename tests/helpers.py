"""Shared test helpers: dense assembly from element matrices (the
assemble_expected_* functions of tests/test_bri17.cpp:153-174,295-316) and
the reference's elementwise tolerance check (assert_equal, :9-28)."""
import numpy as np


def assemble_stiffness(oracle, shape, L, Ke):
    D = len(shape)
    N = int(np.prod(shape))
    nn = 1 << D
    K = np.zeros((D * N, D * N))
    for cell in range(N):
        cn = oracle.cell_nodes(shape, L, cell)
        idx = [cn[ie % nn] + N * (ie // nn) for ie in range(D * nn)]
        K[np.ix_(idx, idx)] += Ke
    return K


def assemble_strain(oracle, shape, L, Be):
    D = len(shape)
    N = int(np.prod(shape))
    nn = 1 << D
    S = Be.shape[0]
    B = np.zeros((S * N, D * N))
    for cell in range(N):
        cn = oracle.cell_nodes(shape, L, cell)
        cols = [cn[j % nn] + N * (j // nn) for j in range(D * nn)]
        for il in range(S):
            B[il * N + cell, cols] += Be[il]
    return B


def assert_equal(expected, actual, rtol, atol):
    """tests/test_bri17.cpp:9-28: |a - e| <= rtol |e| + atol elementwise."""
    assert expected.shape == actual.shape
    err = np.abs(actual - expected)
    tol = rtol * np.abs(expected) + atol
    bad = np.argwhere(err > tol)
    assert bad.size == 0, "%d entries out of tolerance, first %s: expected %r actual %r" % (
        len(bad), tuple(bad[0]), expected[tuple(bad[0])], actual[tuple(bad[0])])


def rel_err(a, b):
    return float(np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


def dft_bruteforce(x):
    """Textbook unnormalised DFT (sign -1) over all axes, with exact integer
    reduction of the phase (n k mod N) before the exponential."""
    x = np.asarray(x, dtype=np.complex128)
    for ax in range(x.ndim):
        n = x.shape[ax]
        j = np.arange(n)
        W = np.exp(-2j * np.pi * ((np.outer(j, j) % n) / n))
        x = np.moveaxis(np.tensordot(W, np.moveaxis(x, ax, 0), axes=(1, 0)), 0, ax)
    return x
