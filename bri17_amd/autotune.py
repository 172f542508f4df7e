"""Launch-shape autotuner with a JSON cache.

``tune(shape, dtype)`` times the candidate launch shapes of a grid on the
device (bri17_autotune: fixed-size kernels and their persistent variants;
runtime-size kernels over the middle-axis split, pencil tile width and radix
orders) and registers the fastest with the library, so operators created
afterwards for that grid use it.  ``save`` / ``load`` keep the choices in a
JSON file keyed "DIM:N0xN1[xN2]:f32|f64".
"""
import ctypes
import json

from . import _lib

_DT = {"float32": _lib.BRI17_F32, "float64": _lib.BRI17_F64}
_FIELDS = ("plane_split", "pencil_width", "lds_budget_kb", "xcd_remap", "force_generic", "plane_prefetch")

_registered = {}


def _key(shape, dtype):
    return "%d:%s:%s" % (len(shape), "x".join(str(int(n)) for n in shape),
                         "f32" if dtype == "float32" else "f64")


def tuning_to_dict(t):
    d = {f: getattr(t, f) for f in _FIELDS}
    d["radices"] = [[r for r in t.radices[i] if r] for i in range(3)]
    return d


def tuning_from_dict(d):
    t = _lib.Tuning.make(plane_split=d.get("plane_split", 0), pencil_width=d.get("pencil_width", 0),
                         lds_budget_kb=d.get("lds_budget_kb", 0), xcd_remap=d.get("xcd_remap", -1),
                         radices=tuple(d.get("radices", [None, None, None])),
                         force_generic=d.get("force_generic", 0),
                         plane_prefetch=d.get("plane_prefetch", -1))
    return t


def register(shape, dtype, tuning):
    t = tuning if isinstance(tuning, _lib.Tuning) else tuning_from_dict(tuning)
    sh = (ctypes.c_int * 3)(*(tuple(shape) + (1,) * (3 - len(shape))))
    _lib.check(_lib.lib().bri17_tuning_cache_set(len(shape), sh, _DT[dtype], ctypes.byref(t)))
    _registered[_key(shape, dtype)] = tuning_to_dict(t)


def clear():
    _lib.check(_lib.lib().bri17_tuning_cache_clear())
    _registered.clear()


def tune(shape, dtype="float32", L=None, reps=10, register_result=True):
    """Returns (tuning dict, ms per apply, candidates tried)."""
    shape = tuple(int(n) for n in shape)
    L = tuple(L) if L is not None else (1.0,) * len(shape)
    sh = (ctypes.c_int * 3)(*(shape + (1,) * (3 - len(shape))))
    LL = (ctypes.c_double * 3)(*(L + (1.0,) * (3 - len(shape))))
    best = _lib.Tuning()
    ms = ctypes.c_double()
    tried = ctypes.c_int()
    _lib.check(_lib.lib().bri17_autotune(len(shape), sh, LL, _DT[dtype], int(reps), ctypes.byref(best),
                                         ctypes.byref(ms), ctypes.byref(tried)))
    if register_result:
        register(shape, dtype, best)
    return tuning_to_dict(best), ms.value, tried.value


def save(path):
    with open(path, "w") as f:
        json.dump(_registered, f, indent=1, sort_keys=True)


def load(path):
    with open(path) as f:
        data = json.load(f)
    for key, d in data.items():
        dim, dims, dt = key.split(":")
        shape = tuple(int(n) for n in dims.split("x"))
        assert len(shape) == int(dim)
        register(shape, "float32" if dt == "f32" else "float64", d)
    return len(data)
