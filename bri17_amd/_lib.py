"""ctypes binding of libbri17_amd.so (the C ABI declared in include/bri17_amd.h).

The library is the only compute path: if it is missing this module raises
instead of falling back to anything else.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbri17_amd.so")

BRI17_OK = 0
BRI17_ERR_INVALID = 1
BRI17_ERR_UNSUPPORTED = 2
BRI17_ERR_HIP = 3
BRI17_ERR_NOMEM = 4
BRI17_ERR_STATE = 5
BRI17_F32 = 0
BRI17_F64 = 1

# name -> (restype, argtypes); mirrors include/bri17_amd.h
_c = ctypes
_ip = _c.POINTER(_c.c_int)
_dp = _c.POINTER(_c.c_double)
_fp = _c.POINTER(_c.c_float)
_vp = _c.c_void_p
SIGNATURES = {
    "bri17_last_error": (_c.c_char_p, []),
    "bri17_version": (_c.c_char_p, []),
    "bri17_grid_node_at": (_c.c_int, [_c.c_int, _ip, _ip, _ip]),
    "bri17_grid_cell_nodes": (_c.c_int, [_c.c_int, _ip, _c.c_int, _ip]),
    "bri17_modal_strain_displacement_f64": (_c.c_int, [_c.c_int, _ip, _dp, _ip, _dp]),
    "bri17_modal_strain_displacement_f32": (_c.c_int, [_c.c_int, _ip, _fp, _ip, _fp]),
    "bri17_modal_stiffness_f64": (_c.c_int, [_c.c_int, _ip, _dp, _c.c_double, _c.c_double, _ip, _dp]),
    "bri17_modal_stiffness_f32": (_c.c_int, [_c.c_int, _ip, _fp, _c.c_float, _c.c_float, _ip, _fp]),
    "bri17_modal_eigenstress_to_opposite_strain_f64": (
        _c.c_int, [_c.c_int, _ip, _dp, _c.c_double, _c.c_double, _ip, _dp, _dp]),
    "bri17_modal_eigenstress_to_opposite_strain_f32": (
        _c.c_int, [_c.c_int, _ip, _fp, _c.c_float, _c.c_float, _ip, _fp, _fp]),
    "bri17_operator_create": (
        _c.c_int, [_c.c_int, _ip, _dp, _c.c_double, _c.c_double, _c.c_int, _c.POINTER(_vp)]),
    "bri17_operator_destroy": (_c.c_int, [_vp]),
    "bri17_operator_set_material": (_c.c_int, [_vp, _c.c_double, _c.c_double]),
    "bri17_operator_info": (_c.c_int, [_vp, _c.POINTER(_c.c_int64), _c.POINTER(_c.c_size_t)]),
    "bri17_operator_apply": (_c.c_int, [_vp, _vp, _vp, _vp]),
    "bri17_operator_solve": (_c.c_int, [_vp, _vp, _vp, _vp]),
    "bri17_operator_apply_host": (_c.c_int, [_vp, _vp, _vp]),
    "bri17_operator_solve_host": (_c.c_int, [_vp, _vp, _vp]),
    "bri17_operator_strain": (_c.c_int, [_vp, _vp, _vp, _vp]),
    "bri17_operator_eigenstress": (_c.c_int, [_vp, _vp, _vp, _vp]),
    "bri17_operator_strain_host": (_c.c_int, [_vp, _vp, _vp]),
    "bri17_operator_eigenstress_host": (_c.c_int, [_vp, _vp, _vp]),
    "bri17_operator_run_noalloc": (_c.c_int, [_vp, _c.c_int, _vp, _vp, _vp]),
    "bri17_operator_create_tuned": (
        _c.c_int, [_c.c_int, _ip, _dp, _c.c_double, _c.c_double, _c.c_int, _vp, _c.POINTER(_vp)]),
    "bri17_operator_describe": (_c.c_int, [_vp, _c.c_char_p, _c.c_size_t]),
    "bri17_batch_create": (_c.c_int, [_c.POINTER(_vp)]),
    "bri17_batch_destroy": (_c.c_int, [_vp]),
    "bri17_batch_run": (_c.c_int, [_vp, _vp, _c.c_int, _vp]),
    "bri17_batch_wait": (_c.c_int, [_vp, _ip]),
    "bri17_batch_set_spin_limit": (_c.c_int, [_vp, _c.c_int]),
    "bri17_batch_info": (_c.c_int, [_vp, _ip, _ip, _ip, _c.POINTER(_c.c_size_t)]),
    "bri17_graph_create": (_c.c_int, [_vp, _c.c_int, _vp, _vp, _c.POINTER(_vp)]),
    "bri17_batch_capture": (_c.c_int, [_vp, _vp, _c.c_int, _c.POINTER(_vp)]),
    "bri17_graph_launch": (_c.c_int, [_vp, _vp]),
    "bri17_graph_destroy": (_c.c_int, [_vp]),
    "bri17_autotune": (_c.c_int, [_c.c_int, _ip, _dp, _c.c_int, _c.c_int, _vp, _dp, _ip]),
    "bri17_tuning_cache_set": (_c.c_int, [_c.c_int, _ip, _c.c_int, _vp]),
    "bri17_tuning_cache_clear": (_c.c_int, []),
    "bri17_operator_profile": (_c.c_int, [_vp, _c.c_int]),
    "bri17_operator_profile_read": (_c.c_int, [_vp, _ip, _dp]),
    "bri17_server_create": (_c.c_int, [_vp, _c.POINTER(_vp)]),
    "bri17_server_destroy": (_c.c_int, [_vp]),
    "bri17_server_submit": (_c.c_int, [_vp, _vp, _c.POINTER(_c.c_int64)]),
    "bri17_server_submit_many": (_c.c_int, [_vp, _vp, _c.c_int, _c.POINTER(_c.c_int64)]),
    "bri17_server_poll": (_c.c_int, [_vp, _c.c_int64, _vp]),
    "bri17_server_wait": (_c.c_int, [_vp, _c.c_int64, _c.c_double, _vp]),
    "bri17_server_cancel": (_c.c_int, [_vp, _c.c_int64]),
    "bri17_server_drain": (_c.c_int, [_vp]),
    "bri17_server_get_stats": (_c.c_int, [_vp, _vp]),
    "bri17_server_trim": (_c.c_int, [_vp]),
    "bri17_server_set_spin_limit": (_c.c_int, [_vp, _c.c_int]),
}

BRI17_MEM_DEVICE = 0
BRI17_MEM_HOST = 1
REQ_STATES = {0: "queued", 1: "running", 2: "done", 3: "failed", 4: "cancelled"}



class BatchItem(ctypes.Structure):
    """bri17_batch_item of include/bri17_amd.h."""
    _fields_ = [("op", ctypes.c_void_p), ("mode", ctypes.c_int), ("inp", ctypes.c_void_p),
                ("out", ctypes.c_void_p)]


class Tuning(ctypes.Structure):
    """bri17_tuning of include/bri17_amd.h."""
    _fields_ = [("plane_split", ctypes.c_int), ("pencil_width", ctypes.c_int),
                ("lds_budget_kb", ctypes.c_int), ("xcd_remap", ctypes.c_int),
                ("radices", (ctypes.c_int * 8) * 3), ("force_generic", ctypes.c_int),
                ("plane_prefetch", ctypes.c_int)]

    @classmethod
    def make(cls, plane_split=0, pencil_width=0, lds_budget_kb=0, xcd_remap=-1, radices=(None, None, None),
             force_generic=0, plane_prefetch=-1):
        t = cls()
        t.force_generic = force_generic
        t.plane_prefetch = plane_prefetch
        t.plane_split, t.pencil_width, t.lds_budget_kb, t.xcd_remap = plane_split, pencil_width, lds_budget_kb, xcd_remap
        for r, rs in enumerate(radices):
            for i, v in enumerate(rs or ()):
                t.radices[r][i] = v
        return t


class ServerConfig(ctypes.Structure):
    """bri17_server_config of include/bri17_amd.h."""
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("max_batch", ctypes.c_int),
                ("max_waves", ctypes.c_int), ("solo_nodes", ctypes.c_int64),
                ("spin_limit", ctypes.c_int), ("heavy_nodes", ctypes.c_int64),
                ("max_heavy_waves", ctypes.c_int)]


class Request(ctypes.Structure):
    """bri17_request of include/bri17_amd.h."""
    _fields_ = [("dim", ctypes.c_int), ("shape", ctypes.c_int * 3), ("L", ctypes.c_double * 3),
                ("mu", ctypes.c_double), ("nu", ctypes.c_double), ("mode", ctypes.c_int),
                ("dtype", ctypes.c_int), ("memory", ctypes.c_int), ("inp", ctypes.c_void_p),
                ("out", ctypes.c_void_p)]


class RequestResult(ctypes.Structure):
    """bri17_request_result of include/bri17_amd.h."""
    _fields_ = [("state", ctypes.c_int), ("error", ctypes.c_int), ("queued_s", ctypes.c_double),
                ("latency_s", ctypes.c_double), ("wave", ctypes.c_int64), ("wave_size", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    """bri17_server_stats of include/bri17_amd.h."""
    _fields_ = [(n, ctypes.c_int64) for n in ("submitted", "completed", "failed", "rejected", "cancelled",
                                              "waves", "batched_waves", "solo_waves", "queued", "running")] + \
               [(n, ctypes.c_size_t) for n in ("hbm_budget", "hbm_admitted", "hbm_cached", "hbm_peak")] + \
               [("plans_resident", ctypes.c_int64), ("peak_waves_in_flight", ctypes.c_int64)]


_lib = None


class Bri17Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("bri17_amd error %d: %s" % (code, msg))
        self.code = code


def lib():
    """Load libbri17_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libbri17_amd.so is not built (%s); run `python -c \"import __graft_entry__ as g; "
                "g.build()\"` from the repository root" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != BRI17_OK:
        msg = lib().bri17_last_error()
        raise Bri17Error(rc, msg.decode() if msg else "")
    return rc
