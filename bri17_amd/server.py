"""Server: continuous batching of operator applies and Green solves.

Python face of ``bri17_server_*`` (include/bri17_amd.h).  Requests of mixed
grids, dtypes, materials and modes are queued from any thread; the library's
worker thread admits them in arrival order against an HBM budget and runs
them in waves on several streams while earlier waves are still on the
device.  Heavy requests (``heavy_nodes`` nodes or more, default 2**22) run
at most ``max_heavy_waves`` at a time (default 1, -1 for no limit); the
requests behind a heavy one that has to wait for that start first.  No
reference counterpart: the reference applies one operator per call on the
CPU (tests/test_bri17.cpp:56-107).

Device tensors are read on the server's own streams, so their producer must
have finished when ``submit`` is called (``sync=True`` synchronises torch's
current stream first).  numpy arrays are staged through device buffers of
the server's pool.
"""
import ctypes
import time

import numpy as np

from . import _lib

_DT = {"float32": _lib.BRI17_F32, "float64": _lib.BRI17_F64}
_MODES = {"apply": 0, "solve": 1}


class ServedRequest:
    """Handle of one submitted request; keeps its buffers alive until it ends."""

    def __init__(self, server, rid, inp, out):
        self._server = server
        self.id = rid
        self.inp = inp
        self.out = out
        self.result = None

    def _finish(self, res):
        self.result = {"state": _lib.REQ_STATES[res.state], "error": res.error,
                       "queued_s": res.queued_s, "latency_s": res.latency_s,
                       "wave": res.wave, "wave_size": res.wave_size}
        self.inp = None
        self._server._live.pop(self.id, None)
        return self.result

    def poll(self):
        """Result dict once the request has ended, else None."""
        if self.result is not None:
            return self.result
        res = _lib.RequestResult()
        _lib.check(_lib.lib().bri17_server_poll(self._server._h, self.id, ctypes.byref(res)))
        return self._finish(res) if res.state >= 2 else None

    def wait(self, timeout=None):
        if self.result is not None:
            return self.result
        res = _lib.RequestResult()
        _lib.check(_lib.lib().bri17_server_wait(self._server._h, self.id,
                                                -1.0 if timeout is None else float(timeout),
                                                ctypes.byref(res)))
        return self._finish(res)

    def cancel(self):
        _lib.check(_lib.lib().bri17_server_cancel(self._server._h, self.id))


class Server:
    def __init__(self, hbm_budget_bytes=0, max_batch=0, max_waves=0, solo_nodes=0, spin_limit=0,
                 heavy_nodes=0, max_heavy_waves=0):
        cfg = _lib.ServerConfig(int(hbm_budget_bytes), int(max_batch), int(max_waves), int(solo_nodes),
                                int(spin_limit), int(heavy_nodes), int(max_heavy_waves))
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().bri17_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._live = {}  # id -> request: buffers stay referenced until the request ends

    @staticmethod
    def _request(x, out, mode, mu, nu, L, sync):
        """bri17_request of one field; returns (struct, out)."""
        if isinstance(x, np.ndarray):
            memory = _lib.BRI17_MEM_HOST
            if not x.flags.c_contiguous:
                raise ValueError("x must be C-contiguous")
            if out is None:
                out = np.empty_like(x)
            dtype, shape = x.dtype.name, x.shape
            ok = isinstance(out, np.ndarray) and out.shape == x.shape and out.dtype == x.dtype \
                and out.flags.c_contiguous
            pin, pout = x.ctypes.data, out.ctypes.data
        else:
            import torch
            if not isinstance(x, torch.Tensor) or not x.is_cuda or not x.is_contiguous():
                raise TypeError("x must be a numpy array or a contiguous device torch.Tensor")
            memory = _lib.BRI17_MEM_DEVICE
            if out is None:
                out = torch.empty_like(x)
            dtype, shape = str(x.dtype).replace("torch.", ""), tuple(x.shape)
            ok = (isinstance(out, torch.Tensor) and out.is_cuda and out.is_contiguous()
                  and tuple(out.shape) == shape and out.dtype == x.dtype and out.device == x.device)
            pin, pout = x.data_ptr(), out.data_ptr()
            if sync:
                torch.cuda.current_stream(x.device).synchronize()
        if not ok:
            raise ValueError("out must match x in kind, shape, dtype and contiguity")
        if dtype not in _DT:
            raise ValueError("dtype must be float32 or float64")
        dim = len(shape) - 1
        if dim not in (2, 3) or shape[0] != dim:
            raise ValueError("x must have shape (DIM, N0, ..., N_{DIM-1}) with DIM 2 or 3")
        L = tuple(float(v) for v in (L if L is not None else (1.0,) * dim))
        r = _lib.Request()
        r.dim = dim
        for d in range(dim):
            r.shape[d] = int(shape[1 + d])
            r.L[d] = L[d]
        r.mu, r.nu = float(mu), float(nu)
        r.mode = _MODES[mode] if isinstance(mode, str) else int(mode)
        r.dtype = _DT[dtype]
        r.memory = memory
        r.inp, r.out = pin, pout
        return r, out

    def submit(self, x, out=None, mode="apply", mu=1.0, nu=0.3, L=None, sync=False):
        """Queue K x (mode "apply") or the Green solve of x (mode "solve").
        x: (DIM, N0, ..., N_{DIM-1}) float32/float64, a device torch.Tensor or
        a numpy array; out: same kind (allocated if None)."""
        r, out = self._request(x, out, mode, mu, nu, L, sync)
        rid = ctypes.c_int64()
        _lib.check(_lib.lib().bri17_server_submit(self._h, ctypes.byref(r), ctypes.byref(rid)))
        req = ServedRequest(self, rid.value, x, out)
        self._live[req.id] = req
        return req

    def submit_many(self, items, sync=False):
        """Queue several requests at once (all or nothing).  items: dicts
        with keys x, and optionally out, mode, mu, nu, L."""
        built = [self._request(it["x"], it.get("out"), it.get("mode", "apply"), it.get("mu", 1.0),
                               it.get("nu", 0.3), it.get("L"), sync and i == 0)
                 for i, it in enumerate(items)]
        arr = (_lib.Request * len(built))(*[r for r, _ in built])
        ids = (ctypes.c_int64 * len(built))()
        _lib.check(_lib.lib().bri17_server_submit_many(self._h, arr, len(built), ids))
        reqs = []
        for i, (it, (_, out)) in enumerate(zip(items, built)):
            req = ServedRequest(self, ids[i], it["x"], out)
            self._live[req.id] = req
            reqs.append(req)
        return reqs

    def set_spin_limit(self, limit):
        _lib.check(_lib.lib().bri17_server_set_spin_limit(self._h, int(limit)))

    def drain(self):
        _lib.check(_lib.lib().bri17_server_drain(self._h))

    def trim(self):
        _lib.check(_lib.lib().bri17_server_trim(self._h))

    def stats(self):
        st = _lib.ServerStats()
        _lib.check(_lib.lib().bri17_server_get_stats(self._h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in _lib.ServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().bri17_server_destroy(self._h)  # cancels queued, waits for running
            self._h = None
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_arrivals(rate_hz, n, seed=0):
    """Arrival offsets (s) of a Poisson process of the given rate."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_hz, size=n))


def sleep_until(t_abs):
    """Wait until time.perf_counter() reaches t_abs (sleep, then spin)."""
    while True:
        dt = t_abs - time.perf_counter()
        if dt <= 0:
            return
        if dt > 2e-3:
            time.sleep(dt - 1e-3)
