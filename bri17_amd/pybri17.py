"""``pybri17``-compatible Python surface over the C ABI of libbri17_amd.so.

Mirrors the module built from python/pybri17.cpp of sbrisard/bri17 v0.1
(classes CartesianGrid2f64 / CartesianGrid3f64 / Hooke2f64 / Hooke3f64,
:32-107), with the same method names, argument meaning and error behaviour:
array arguments must be one-dimensional and contiguous, otherwise ValueError
("expected one-dimensional, contiguous array", pybri17.cpp:9-15), and output
arrays are filled in place.  Float32 variants (``...f32``) are added.

Output arrays must already have the complex dtype of the class (complex128
for f64, complex64 for f32): pybind11 would silently convert a mismatching
array into a temporary and drop the result, which is reported here as
ValueError instead.
"""
import ctypes

import numpy as np

from . import _lib

# Module attributes of pybri17 (python/pybri17.cpp:100-101), set there from
# the project metadata (metadata/author.txt, metadata/version.txt) that the
# reference's builds pass in (python/setup.py:36-39, python/CMakeLists.txt:4-5):
# they describe the interface, which callers may check.  The implementation
# behind it is bri17_amd (``__implementation__``).
__author__ = "S. Brisard"
__version__ = "0.1"
__implementation__ = "bri17_amd (MI355X / gfx950)"

_TYPEID = {np.float64: "d", np.float32: "f"}


def _check_contiguous(a):
    # pybri17.cpp:9-15
    if not isinstance(a, np.ndarray) or a.ndim != 1 or a.strides[0] != a.itemsize:
        raise ValueError("expected one-dimensional, contiguous array")


def _as_k(k, dim):
    a = np.ascontiguousarray(k, dtype=np.intc)
    if a.ndim != 1 or a.shape[0] < dim:
        raise ValueError("expected one-dimensional, contiguous array")
    return a


def _ostream_float(x):
    # std::ostream default formatting of a floating point value (%g, 6 digits)
    return "%g" % x


class _CartesianGrid:
    dim = None
    dtype = None

    def __init__(self, shape, L):
        shape = tuple(int(n) for n in shape)
        L = tuple(float(x) for x in L)
        if len(shape) != self.dim or len(L) != self.dim:
            raise TypeError("expected %d-tuples for shape and L" % self.dim)
        self._shape = shape
        self._L = tuple(float(self.dtype.type(x)) for x in L)
        # bri17.hpp:54-58 (size is a product of ints)
        self._size = int(np.prod(shape))

    @property
    def shape(self):
        return self._shape

    @property
    def L(self):
        return self._L

    @property
    def size(self):
        return self._size

    def __repr__(self):
        # CartesianGrid::repr, bri17.hpp:61-69
        t = _TYPEID[self.dtype.type]
        return "CartesianGrid<%s,%d>{shape={%s},L={%s}}" % (
            t, self.dim, "".join("%d," % n for n in self._shape),
            "".join(_ostream_float(x) + "," for x in self._L))

    def _shape_c(self):
        return (ctypes.c_int * 3)(*(self._shape + (1,) * (3 - self.dim)))

    def get_cell_nodes(self, cell):
        # bri17.hpp:127-158
        out = (ctypes.c_int * 8)()
        _lib.check(_lib.lib().bri17_grid_cell_nodes(self.dim, self._shape_c(), int(cell), out))
        return list(out)[: 1 << self.dim]

    def get_node_at(self, *idx):
        # bri17.hpp:78-93
        if len(idx) != self.dim:
            raise TypeError("get_node_at expects %d indices" % self.dim)
        i = (ctypes.c_int * 3)(*(tuple(int(v) for v in idx) + (0,) * (3 - self.dim)))
        node = ctypes.c_int(0)
        _lib.check(_lib.lib().bri17_grid_node_at(self.dim, self._shape_c(), i, ctypes.byref(node)))
        return node.value


class _Hooke:
    grid_class = None

    def __init__(self, mu, nu, grid):
        if not isinstance(grid, self.grid_class):
            raise TypeError("expected a %s" % self.grid_class.__name__)
        dt = grid.dtype.type
        self._mu = float(dt(mu))
        self._nu = float(dt(nu))
        self._grid = grid

    mu = property(lambda self: self._mu)
    nu = property(lambda self: self._nu)
    grid = property(lambda self: self._grid)

    def __repr__(self):
        # Hooke::repr, bri17.hpp:197-202 (the reference leaves the brace open
        # and appends a newline; kept as is)
        t = _TYPEID[self._grid.dtype.type]
        return "Hooke<%s,%d>{mu=%s,nu=%s,grid=%r\n" % (
            t, self._grid.dim, _ostream_float(self._mu), _ostream_float(self._nu), self._grid)

    # --- helpers
    def _f64(self):
        return self._grid.dtype.type is np.float64

    def _L_c(self):
        g = self._grid
        ct = ctypes.c_double if self._f64() else ctypes.c_float
        return (ct * 3)(*(g.L + (1.0,) * (3 - g.dim)))

    def _check_out(self, a, n):
        _check_contiguous(a)
        want = np.complex128 if self._f64() else np.complex64
        if a.dtype != want:
            raise ValueError("expected an array of dtype %s" % np.dtype(want).name)
        if a.shape[0] < n:
            raise ValueError("output array too small: need %d entries" % n)

    def _ptr(self, a):
        ct = ctypes.c_double if self._f64() else ctypes.c_float
        return a.ctypes.data_as(ctypes.POINTER(ct))

    def modal_strain_displacement(self, k, B):
        """B^(k): DIM complex entries written to B (bri17.hpp:212-236)."""
        d = self._grid.dim
        _check_contiguous(np.asarray(k)) if isinstance(k, np.ndarray) else None
        kk = _as_k(k, d)
        self._check_out(B, d)
        L = _lib.lib()
        fn = L.bri17_modal_strain_displacement_f64 if self._f64() else L.bri17_modal_strain_displacement_f32
        _lib.check(fn(d, self._grid._shape_c(), self._L_c(),
                      kk.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), self._ptr(B)))

    def modal_stiffness_matrix(self, k, K):
        """K^(k): DIM*DIM complex entries, row-major, written to K
        (bri17.hpp:247-292; bound as "modal_stiffness_matrix", pybri17.cpp:82)."""
        d = self._grid.dim
        _check_contiguous(np.asarray(k)) if isinstance(k, np.ndarray) else None
        kk = _as_k(k, d)
        self._check_out(K, d * d)
        L = _lib.lib()
        if self._f64():
            rc = L.bri17_modal_stiffness_f64(d, self._grid._shape_c(), self._L_c(), self._mu, self._nu,
                                             kk.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), self._ptr(K))
        else:
            rc = L.bri17_modal_stiffness_f32(d, self._grid._shape_c(), self._L_c(), self._mu, self._nu,
                                             kk.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), self._ptr(K))
        _lib.check(rc)

    # the C++ name, for callers written against the header
    modal_stiffness = modal_stiffness_matrix

    def modal_eigenstress_to_opposite_strain(self, k, tau, eta):
        """-eps^(k) induced by the eigenstress tau^(k), Voigt-Mandel
        (bri17.hpp:308-355).  As in pybri17.cpp:91-92, only tau and eta are
        checked for layout."""
        d = self._grid.dim
        sym = 3 if d == 2 else 6
        _check_contiguous(tau)
        self._check_out(eta, sym)
        kk = _as_k(k, d)
        want = np.complex128 if self._f64() else np.complex64
        t = np.ascontiguousarray(tau, dtype=want)
        if t.shape[0] < sym:
            raise ValueError("tau too small: need %d entries" % sym)
        L = _lib.lib()
        fn = (L.bri17_modal_eigenstress_to_opposite_strain_f64 if self._f64()
              else L.bri17_modal_eigenstress_to_opposite_strain_f32)
        _lib.check(fn(d, self._grid._shape_c(), self._L_c(), self._mu, self._nu,
                      kk.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), self._ptr(t), self._ptr(eta)))


def _make(dim, dt):
    suffix = "%d%s" % (dim, "f64" if dt is np.float64 else "f32")
    G = type("CartesianGrid" + suffix, (_CartesianGrid,), {"dim": dim, "dtype": np.dtype(dt)})
    H = type("Hooke" + suffix, (_Hooke,), {"grid_class": G})
    return G, H


CartesianGrid2f64, Hooke2f64 = _make(2, np.float64)
CartesianGrid3f64, Hooke3f64 = _make(3, np.float64)
CartesianGrid2f32, Hooke2f32 = _make(2, np.float32)
CartesianGrid3f32, Hooke3f32 = _make(3, np.float32)

__all__ = ["CartesianGrid2f64", "CartesianGrid3f64", "Hooke2f64", "Hooke3f64",
           "CartesianGrid2f32", "CartesianGrid3f32", "Hooke2f32", "Hooke3f32"]
