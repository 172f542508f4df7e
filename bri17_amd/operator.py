"""StiffnessMatrix: the matrix-free periodic stiffness operator on the GPU.

Python face of ``bri17_operator`` (include/bri17_amd.h).  ``apply`` is the
``K u`` of StiffnessMatrixFactory::compute_Ku (tests/test_bri17.cpp:56-107);
``solve`` is the Green solve ``K u = f`` with the zero-mean convention of
sphinx/theory.rst:208-212; ``strain`` is ``B u``
(tests/test_bri17.cpp:194-238); ``eigenstress`` maps an eigenstress field to
the opposite of the strain it induces (bri17.hpp:308-355).  Displacement and
force fields have shape (DIM, N0, ..., N_{DIM-1}); strain and stress fields
(SYM, N0, ...), SYM = 3 (2D) or 6 (3D), Voigt-Mandel order.

Device tensors (torch, on a ROCm device) run asynchronously on the current
stream; numpy arrays go through the synchronous host-pointer entry points.
"""
import ctypes

import numpy as np

from . import _lib

_DTYPES = {"float32": _lib.BRI17_F32, "float64": _lib.BRI17_F64}
_MODES = {"apply": 0, "solve": 1, "strain": 2, "eigenstress": 3}
_ENTRY = {0: "bri17_operator_apply", 1: "bri17_operator_solve", 2: "bri17_operator_strain",
          3: "bri17_operator_eigenstress"}


class StiffnessMatrix:
    def __init__(self, shape, L=None, mu=1.0, nu=0.3, dtype="float32", tuning=None):
        shape = tuple(int(n) for n in shape)
        if len(shape) not in (2, 3):
            raise ValueError("shape must have 2 or 3 entries")
        if L is None:
            L = (1.0,) * len(shape)
        L = tuple(float(x) for x in L)
        dtype = np.dtype(dtype).name
        if dtype not in _DTYPES:
            raise ValueError("dtype must be float32 or float64")
        self.shape = shape
        self.L = L
        self.dim = len(shape)
        self.dtype = dtype
        self._mu = float(mu)
        self._nu = float(nu)
        h = ctypes.c_void_p()
        sh = (ctypes.c_int * 3)(*(shape + (1,) * (3 - self.dim)))
        LL = (ctypes.c_double * 3)(*(L + (1.0,) * (3 - self.dim)))
        if tuning is None:
            _lib.check(_lib.lib().bri17_operator_create(self.dim, sh, LL, self._mu, self._nu,
                                                        _DTYPES[dtype], ctypes.byref(h)))
        else:
            t = tuning if isinstance(tuning, _lib.Tuning) else _lib.Tuning.make(**tuning)
            _lib.check(_lib.lib().bri17_operator_create_tuned(self.dim, sh, LL, self._mu, self._nu,
                                                              _DTYPES[dtype], ctypes.byref(t),
                                                              ctypes.byref(h)))
        self._h = h
        n = ctypes.c_int64()
        ws = ctypes.c_size_t()
        _lib.check(_lib.lib().bri17_operator_info(h, ctypes.byref(n), ctypes.byref(ws)))
        self.num_dofs = n.value
        self.workspace_bytes = ws.value

    # ------------------------------------------------------------------ state
    @property
    def mu(self):
        return self._mu

    @property
    def nu(self):
        return self._nu

    def set_material(self, mu, nu):
        _lib.check(_lib.lib().bri17_operator_set_material(self._h, float(mu), float(nu)))
        self._mu, self._nu = float(mu), float(nu)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().bri17_operator_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def describe(self):
        """Plan geometry as a dict (kernel launch shapes, radix orders)."""
        import json
        buf = ctypes.create_string_buffer(4096)
        _lib.check(_lib.lib().bri17_operator_describe(self._h, buf, len(buf)))
        return json.loads(buf.value.decode())

    @property
    def handle(self):
        return self._h

    @property
    def field_shape(self):
        return (self.dim,) + self.shape

    # --------------------------------------------------------------- compute
    @property
    def sym(self):
        """Number of Voigt-Mandel components of a symmetric tensor field."""
        return 3 if self.dim == 2 else 6

    def _shapes(self, mode):
        cin = self.sym if mode == 3 else self.dim
        cout = self.sym if mode >= 2 else self.dim
        return (cin,) + self.shape, (cout,) + self.shape

    def _run(self, x, out, mode, stream):
        L = _lib.lib()
        ishape, oshape = self._shapes(mode)
        name = _ENTRY[mode]
        if isinstance(x, np.ndarray):
            want = np.dtype(self.dtype)
            if x.dtype != want or x.shape != ishape:
                raise ValueError("expected a %s array of shape %s" % (want.name, ishape))
            x = np.ascontiguousarray(x)
            if out is None:
                out = np.empty(oshape, dtype=want)
            elif out.dtype != want or out.shape != oshape or not out.flags.c_contiguous:
                raise ValueError("out must be a contiguous %s array of shape %s" % (want.name, oshape))
            _lib.check(getattr(L, name + "_host")(self._h, x.ctypes.data, out.ctypes.data))
            return out
        import torch  # device path
        if not isinstance(x, torch.Tensor) or not x.is_cuda:
            raise TypeError("expected a numpy array or a device torch.Tensor")
        want = getattr(torch, self.dtype)
        if x.dtype != want or tuple(x.shape) != ishape or not x.is_contiguous():
            raise ValueError("expected a contiguous %s tensor of shape %s" % (self.dtype, ishape))
        if out is None:
            out = torch.empty(oshape, dtype=want, device=x.device)
        elif (out.dtype != want or tuple(out.shape) != oshape or not out.is_contiguous()
              or out.device != x.device):
            raise ValueError("bad out tensor")
        if stream is None:
            stream = torch.cuda.current_stream(x.device).cuda_stream
        with torch.cuda.device(x.device):
            _lib.check(getattr(L, name)(self._h, ctypes.c_void_p(x.data_ptr()),
                                        ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream)))
        return out

    def capture(self, x, out, mode="apply"):
        """HIP graph of one call on fixed device tensors (runs once eagerly,
        then captures); replay with ``Graph.launch``."""
        import torch
        m = _MODES[mode] if isinstance(mode, str) else int(mode)
        want = getattr(torch, self.dtype)
        for t, shp in zip((x, out), self._shapes(m)):
            if (not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != want
                    or tuple(t.shape) != shp or not t.is_contiguous()):
                raise ValueError("expected contiguous %s device tensors of shape %s" % (self.dtype, shp))
        h = ctypes.c_void_p()
        with torch.cuda.device(x.device):
            _lib.check(_lib.lib().bri17_graph_create(self._h, m, ctypes.c_void_p(x.data_ptr()),
                                                     ctypes.c_void_p(out.data_ptr()), ctypes.byref(h)))
        return Graph(h, keep=(self, x, out))

    def apply(self, u, out=None, stream=None):
        """K u (tests/test_bri17.cpp:56-107)."""
        return self._run(u, out, 0, stream)

    def solve(self, f, out=None, stream=None):
        """u with K u = f - mean(f) and mean(u) = 0 (sphinx/theory.rst:208-212)."""
        return self._run(f, out, 1, stream)

    def strain(self, u, out=None, stream=None):
        """eps = B u, (SYM, N...) Voigt-Mandel cell field
        (tests/test_bri17.cpp:194-238)."""
        return self._run(u, out, 2, stream)

    def eigenstress(self, tau, out=None, stream=None):
        """Opposite of the strain induced by the eigenstress field tau,
        (SYM, N...) -> (SYM, N...) (bri17.hpp:308-355 at every frequency)."""
        return self._run(tau, out, 3, stream)

    __matmul__ = apply


class Graph:
    """An instantiated HIP graph (bri17_graph_*); keeps its operator and
    buffers alive."""

    def __init__(self, handle, keep=()):
        self._h = handle
        self._keep = keep

    def launch(self, stream=None):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.lib().bri17_graph_launch(self._h, ctypes.c_void_p(stream)))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().bri17_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
