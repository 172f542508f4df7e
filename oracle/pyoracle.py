"""ctypes front-end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, by ``__graft_entry__.smoke()`` and by the CPU-baseline leg of
``bench.py``, always as the checker / the timed CPU port, never by the product
package ``bri17_amd``.  Functions mirror oracle/bri17_oracle.h, which cites the
reference file:line each one restates.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class _Grid(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("shape", ctypes.c_int * 3), ("L", ctypes.c_double * 3)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        gp = ctypes.POINTER(_Grid)
        L.ora_modal_stiffness.argtypes = [gp, ctypes.c_double, ctypes.c_double, ip, dp]
        L.ora_modal_strain_displacement.argtypes = [gp, ip, dp]
        L.ora_modal_eigenstress_to_opposite_strain.argtypes = [gp, ctypes.c_double, ctypes.c_double, ip, dp, dp]
        L.ora_fft.argtypes = [ctypes.c_int, ip, dp, ctypes.c_int, ctypes.c_int]
        L.ora_apply.argtypes = [gp, ctypes.c_double, ctypes.c_double, dp, dp, dp, ctypes.c_int]
        L.ora_solve.argtypes = [gp, ctypes.c_double, ctypes.c_double, dp, dp, ctypes.c_int]
        L.ora_strain.argtypes = [gp, dp, dp, ctypes.c_int]
        L.ora_eigenstress_strain.argtypes = [gp, ctypes.c_double, ctypes.c_double, dp, dp, ctypes.c_int]
        L.ora_element_stiffness.argtypes = [ctypes.c_int, dp, ctypes.c_double, ctypes.c_double, dp]
        L.ora_element_strain_displacement.argtypes = [ctypes.c_int, dp, dp]
        L.ora_ebe_apply.argtypes = [gp, dp, dp, dp, ctypes.c_int]
        L.ora_ebe_strain.argtypes = [gp, dp, dp, dp, ctypes.c_int]
        L.ora_ebe_strain_transpose.argtypes = [gp, dp, dp, dp, ctypes.c_int]
        L.ora_cell_nodes.argtypes = [gp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
        L.ora_fill_uniform.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_double, ctypes.c_double, dp]
        L.ora_fill_normal.argtypes = [ctypes.c_uint64, ctypes.c_int64, dp]
        L.ora_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _g(shape, L):
    g = _Grid()
    g.dim = len(shape)
    for d in range(len(shape)):
        g.shape[d] = int(shape[d])
        g.L[d] = float(L[d])
    return ctypes.byref(g)


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    a = np.ascontiguousarray(a, dtype=np.intc)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def nthreads_default():
    return lib().ora_num_threads()


def modal_stiffness(shape, L, mu, nu, k):
    D = len(shape)
    out = np.zeros(2 * D * D)
    kk, kp = _ip(k)
    lib().ora_modal_stiffness(_g(shape, L), mu, nu, kp, _dp(out))
    return out.view(np.complex128).reshape(D, D)


def modal_strain_displacement(shape, L, k):
    D = len(shape)
    out = np.zeros(2 * D)
    kk, kp = _ip(k)
    lib().ora_modal_strain_displacement(_g(shape, L), kp, _dp(out))
    return out.view(np.complex128)


def modal_eigenstress_to_opposite_strain(shape, L, mu, nu, k, tau):
    D = len(shape)
    S = 3 if D == 2 else 6
    t = np.ascontiguousarray(tau, dtype=np.complex128).view(np.float64).copy()
    out = np.zeros(2 * S)
    kk, kp = _ip(k)
    lib().ora_modal_eigenstress_to_opposite_strain(_g(shape, L), mu, nu, kp, _dp(t), _dp(out))
    return out.view(np.complex128)


def fft(x, sign=-1, nthreads=1):
    """Unnormalised c2c DFT over all axes of a complex array (FFTW sign)."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    shp, sp = _ip(a.shape)
    lib().ora_fft(a.ndim, sp, _dp(a.view(np.float64)), sign, nthreads)
    return a


def apply(shape, L, mu, nu, u, nthreads=1, return_imag=False):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.empty_like(u)
    im = ctypes.c_double(0.0)
    lib().ora_apply(_g(shape, L), mu, nu, _dp(u), _dp(out), ctypes.pointer(im), nthreads)
    return (out, im.value) if return_imag else out


def solve(shape, L, mu, nu, f, nthreads=1):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.empty_like(f)
    lib().ora_solve(_g(shape, L), mu, nu, _dp(f), _dp(out), nthreads)
    return out


def strain(shape, L, u, nthreads=1):
    D = len(shape)
    S = 3 if D == 2 else 6
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.empty((S,) + tuple(shape))
    lib().ora_strain(_g(shape, L), _dp(u), _dp(out), nthreads)
    return out


def eigenstress_strain(shape, L, mu, nu, tau, nthreads=1):
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    out = np.empty_like(tau)
    lib().ora_eigenstress_strain(_g(shape, L), mu, nu, _dp(tau), _dp(out), nthreads)
    return out


def element_stiffness(h, mu, nu):
    D = len(h)
    nd = D * (1 << D)
    out = np.zeros((nd, nd))
    hh = np.asarray(h, dtype=np.float64)
    lib().ora_element_stiffness(D, _dp(hh), mu, nu, _dp(out))
    return out


def element_strain_displacement(h):
    D = len(h)
    nd = D * (1 << D)
    S = 3 if D == 2 else 6
    out = np.zeros((S, nd))
    hh = np.asarray(h, dtype=np.float64)
    lib().ora_element_strain_displacement(D, _dp(hh), _dp(out))
    return out


def ebe_apply(shape, L, Ke, u, nthreads=1):
    Ke = np.ascontiguousarray(Ke, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.empty_like(u)
    lib().ora_ebe_apply(_g(shape, L), _dp(Ke), _dp(u), _dp(out), nthreads)
    return out


def ebe_strain(shape, L, Be, u, nthreads=1):
    S = Be.shape[0]
    Be = np.ascontiguousarray(Be, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.empty((S,) + tuple(shape))
    lib().ora_ebe_strain(_g(shape, L), _dp(Be), _dp(u), _dp(out), nthreads)
    return out


def ebe_strain_transpose(shape, L, Be, tau, nthreads=1):
    D = len(shape)
    Be = np.ascontiguousarray(Be, dtype=np.float64)
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    out = np.empty((D,) + tuple(shape))
    lib().ora_ebe_strain_transpose(_g(shape, L), _dp(Be), _dp(tau), _dp(out), nthreads)
    return out


def cell_nodes(shape, L, cell):
    out = (ctypes.c_int64 * 8)()
    lib().ora_cell_nodes(_g(shape, L), cell, out)
    return list(out)[: 1 << len(shape)]


def fill_uniform(seed, n, lo=-1.0, hi=1.0):
    out = np.empty(n)
    lib().ora_fill_uniform(seed, n, lo, hi, _dp(out))
    return out


def fill_normal(seed, n):
    out = np.empty(n)
    lib().ora_fill_normal(seed, n, _dp(out))
    return out
