"""Batches of operators applied or solved in one persistent kernel launch
(bri17_batch_* of include/bri17_amd.h)."""
import ctypes

from . import _lib

MODES = {"apply": 0, "solve": 1}


class Batch:
    """Reusable launch context.  ``run`` takes (operator, mode, x, out)
    tuples with device tensors; every operator at most once per batch."""

    def __init__(self):
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().bri17_batch_create(ctypes.byref(h)))
        self._h = h

    def _items(self, items):
        import torch
        arr = (_lib.BatchItem * len(items))()
        dev = None
        for i, (op, mode, x, out) in enumerate(items):
            want = getattr(torch, op.dtype)
            for t in (x, out):
                if (not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != want
                        or tuple(t.shape) != op.field_shape or not t.is_contiguous()):
                    raise ValueError("item %d: expected contiguous %s device tensors of shape %s"
                                     % (i, op.dtype, op.field_shape))
            dev = x.device
            arr[i].op = op.handle.value
            arr[i].mode = MODES[mode] if isinstance(mode, str) else int(mode)
            arr[i].inp = x.data_ptr()
            arr[i].out = out.data_ptr()
        return arr, dev

    def run(self, items, stream=None):
        import torch
        arr, dev = self._items(items)
        if stream is None:
            stream = torch.cuda.current_stream(dev).cuda_stream
        with torch.cuda.device(dev):
            _lib.check(_lib.lib().bri17_batch_run(self._h, ctypes.byref(arr), len(items),
                                                  ctypes.c_void_p(stream)))

    def capture(self, items):
        """HIP graph replaying this batch (runs it once eagerly first)."""
        import torch
        from .operator import Graph
        arr, dev = self._items(items)
        h = ctypes.c_void_p()
        with torch.cuda.device(dev):
            _lib.check(_lib.lib().bri17_batch_capture(self._h, ctypes.byref(arr), len(items), ctypes.byref(h)))
        return Graph(h, keep=(self, items))

    def wait(self):
        """Block until the last run finished; returns the dependency-timeout flag."""
        t = ctypes.c_int(0)
        _lib.check(_lib.lib().bri17_batch_wait(self._h, ctypes.byref(t)))
        return t.value

    def set_spin_limit(self, limit):
        _lib.check(_lib.lib().bri17_batch_set_spin_limit(self._h, int(limit)))

    def info(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        d = ctypes.c_size_t()
        _lib.check(_lib.lib().bri17_batch_info(self._h, ctypes.byref(a), ctypes.byref(b),
                                               ctypes.byref(c), ctypes.byref(d)))
        return {"work_items": a.value, "workgroups": b.value, "threads": c.value, "lds_bytes": d.value}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().bri17_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
