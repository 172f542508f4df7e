"""StiffnessMatrix: the matrix-free periodic stiffness operator on the GPU.

Python face of ``bri17_operator`` (include/bri17_amd.h).  ``apply`` is the
``K u`` of StiffnessMatrixFactory::compute_Ku (tests/test_bri17.cpp:56-107);
``solve`` is the Green solve ``K u = f`` with the zero-mean convention of
sphinx/theory.rst:208-212.  Fields have shape (DIM, N0, ..., N_{DIM-1}).

Device tensors (torch, on a ROCm device) run asynchronously on the current
stream; numpy arrays go through the synchronous host-pointer entry points.
"""
import ctypes

import numpy as np

from . import _lib

_DTYPES = {"float32": _lib.BRI17_F32, "float64": _lib.BRI17_F64}


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
    def _run(self, x, out, mode, stream):
        L = _lib.lib()
        if isinstance(x, np.ndarray):
            want = np.dtype(self.dtype)
            if x.dtype != want or x.shape != self.field_shape:
                raise ValueError("expected a %s array of shape %s" % (want.name, self.field_shape))
            x = np.ascontiguousarray(x)
            if out is None:
                out = np.empty_like(x)
            elif out.dtype != want or out.shape != self.field_shape or not out.flags.c_contiguous:
                raise ValueError("out must be a contiguous %s array of shape %s" % (want.name, self.field_shape))
            fn = L.bri17_operator_apply_host if mode == 0 else L.bri17_operator_solve_host
            _lib.check(fn(self._h, x.ctypes.data, out.ctypes.data))
            return out
        import torch  # device path
        if not isinstance(x, torch.Tensor) or not x.is_cuda:
            raise TypeError("expected a numpy array or a device torch.Tensor")
        want = getattr(torch, self.dtype)
        if x.dtype != want or tuple(x.shape) != self.field_shape or not x.is_contiguous():
            raise ValueError("expected a contiguous %s tensor of shape %s" % (self.dtype, self.field_shape))
        if out is None:
            out = torch.empty_like(x)
        elif (out.dtype != want or tuple(out.shape) != self.field_shape or not out.is_contiguous()
              or out.device != x.device):
            raise ValueError("bad out tensor")
        if stream is None:
            stream = torch.cuda.current_stream(x.device).cuda_stream
        fn = L.bri17_operator_apply if mode == 0 else L.bri17_operator_solve
        with torch.cuda.device(x.device):
            _lib.check(fn(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                          ctypes.c_void_p(stream)))
        return out

    def capture(self, x, out, mode="apply"):
        """HIP graph of one apply / solve on fixed device tensors (runs once
        eagerly, then captures); replay with ``Graph.launch``."""
        import torch
        want = getattr(torch, self.dtype)
        for t in (x, out):
            if (not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != want
                    or tuple(t.shape) != self.field_shape or not t.is_contiguous()):
                raise ValueError("expected contiguous %s device tensors of shape %s" % (self.dtype, self.field_shape))
        h = ctypes.c_void_p()
        with torch.cuda.device(x.device):
            _lib.check(_lib.lib().bri17_graph_create(self._h, 0 if mode == "apply" else 1,
                                                     ctypes.c_void_p(x.data_ptr()),
                                                     ctypes.c_void_p(out.data_ptr()), ctypes.byref(h)))
        return Graph(h, keep=(self, x, out))

    def apply(self, u, out=None, stream=None):
        """K u (tests/test_bri17.cpp:56-107)."""
        return self._run(u, out, 0, stream)

    def solve(self, f, out=None, stream=None):
        """u with K u = f - mean(f) and mean(u) = 0 (sphinx/theory.rst:208-212)."""
        return self._run(f, out, 1, stream)

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
