"""CPU-side checks of the C ABI and the pybri17 mirror (no GPU work):
the library loads, exports every symbol the headers declare, and its host
per-frequency functions match the oracle; argument validation mirrors
python/pybri17.cpp."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = []
    inc = os.path.join(ROOT, "include")
    for fn in sorted(os.listdir(inc)):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(bri17_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from bri17_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_version_and_error_string():
    from bri17_amd import _lib
    assert _lib.lib().bri17_version().decode().startswith("0.1")
    sh = (ctypes.c_int * 3)(4, 4, 0)
    rc = _lib.lib().bri17_grid_cell_nodes(3, sh, 0, (ctypes.c_int * 8)())
    assert rc == _lib.BRI17_ERR_INVALID
    assert b"invalid" in _lib.lib().bri17_last_error()


def test_operator_create_rejects_bad_grids():
    from bri17_amd import _lib
    h = ctypes.c_void_p()
    L = (ctypes.c_double * 3)(1, 1, 1)
    # a row of 65,537 (prime) entries: beyond the LDS and the threads of a workgroup
    for dim, shape, code in [(4, (4, 4, 4), 1), (3, (4, 4, 65537), 2), (3, (0, 4, 4), 1)]:
        sh = (ctypes.c_int * 3)(*shape)
        rc = _lib.lib().bri17_operator_create(dim, sh, L, 1.0, 0.3, 0, ctypes.byref(h))
        assert rc == code, (dim, shape, rc)


@pytest.mark.parametrize("cls,dim", [("CartesianGrid2f64", 2), ("CartesianGrid3f64", 3)])
def test_grid_mirrors_reference(oracle, cls, dim):
    from bri17_amd import pybri17
    shape = (3, 4) if dim == 2 else (3, 4, 5)
    L = tuple(float(n) for n in shape)
    g = getattr(pybri17, cls)(shape, L)
    assert g.dim == dim and g.shape == shape and g.size == int(np.prod(shape))
    assert g.dtype == np.float64
    for cell in range(g.size):
        assert g.get_cell_nodes(cell) == oracle.cell_nodes(shape, L, cell)
    idx = tuple(n - 1 for n in shape)
    assert g.get_node_at(*idx) == g.size - 1
    assert repr(g).startswith("CartesianGrid<d,%d>{shape={" % dim)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_per_frequency_functions_match_oracle(oracle, dim, prec):
    from bri17_amd import pybri17
    shape = (6, 8) if dim == 2 else (6, 8, 10)
    L = (1.0, 1.3) if dim == 2 else (1.0, 1.3, 0.7)
    mu, nu = 1.7, 0.27
    g = getattr(pybri17, "CartesianGrid%d%s" % (dim, prec))(shape, L)
    h = getattr(pybri17, "Hooke%d%s" % (dim, prec))(mu, nu, g)
    ct = np.complex128 if prec == "f64" else np.complex64
    tol = 1e-12 if prec == "f64" else 3e-6
    S = 3 if dim == 2 else 6
    rng = np.random.default_rng(0)
    for k in np.ndindex(*shape):
        k = np.array(k, dtype=np.intc)
        K = np.zeros(dim * dim, ct)
        h.modal_stiffness_matrix(k, K)
        Kr = oracle.modal_stiffness(shape, g.L, h.mu, h.nu, k)
        assert np.abs(K.reshape(dim, dim) - Kr).max() <= tol * (1 + np.abs(Kr).max())
        B = np.zeros(dim, ct)
        h.modal_strain_displacement(k, B)
        Br = oracle.modal_strain_displacement(shape, g.L, k)
        assert np.abs(B - Br).max() <= tol * (1 + np.abs(Br).max())
        tau = (rng.standard_normal(S) + 1j * rng.standard_normal(S)).astype(ct)
        eta = np.zeros(S, ct)
        h.modal_eigenstress_to_opposite_strain(k, tau, eta)
        er = oracle.modal_eigenstress_to_opposite_strain(shape, g.L, h.mu, h.nu, k, tau.astype(np.complex128))
        assert np.abs(eta - er).max() <= 10 * tol * (1 + np.abs(er).max())


def test_pybri17_layout_errors_are_value_errors():
    """python/pybri17.cpp:9-15: arrays must be 1-D and contiguous."""
    from bri17_amd import pybri17
    g = pybri17.CartesianGrid2f64((4, 4), (1.0, 1.0))
    h = pybri17.Hooke2f64(1.0, 0.3, g)
    k = np.array([1, 2], dtype=np.intc)
    with pytest.raises(ValueError):
        h.modal_stiffness_matrix(k, np.zeros((2, 2), complex))
    with pytest.raises(ValueError):
        h.modal_stiffness_matrix(k, np.zeros(8, complex)[::2])
    with pytest.raises(ValueError):
        h.modal_strain_displacement(k, np.zeros(2, np.float64))
    with pytest.raises(ValueError):
        h.modal_eigenstress_to_opposite_strain(k, np.zeros((3, 1), complex), np.zeros(3, complex))
    assert repr(h) == "Hooke<d,2>{mu=1,nu=0.3,grid=CartesianGrid<d,2>{shape={4,4,},L={1,1,}}\n"


def test_ctypes_structs_match_the_c_layout(tmp_path):
    """sizeof / offsetof of every struct of the C ABI, compiled with the host
    C compiler, equal those of the ctypes mirrors."""
    import shutil
    import subprocess
    from bri17_amd import _lib
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    mirrors = {"bri17_batch_item": _lib.BatchItem, "bri17_tuning": _lib.Tuning,
               "bri17_server_config": _lib.ServerConfig, "bri17_request": _lib.Request,
               "bri17_request_result": _lib.RequestResult, "bri17_server_stats": _lib.ServerStats}
    renames = {"inp": "in"}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "bri17_amd.h"', "int main(void) {"]
    for cname, cls in mirrors.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, renames.get(f, f)))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                         text=True).stdout.split("\n") if l)
    for cname, cls in mirrors.items():
        assert int(out[cname]) == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(out["%s.%s" % (cname, f)]) == getattr(cls, f).offset, (cname, f)


def test_server_entry_points_validate_arguments():
    from bri17_amd import _lib
    L = _lib.lib()
    rid = ctypes.c_int64()
    assert L.bri17_server_submit(None, ctypes.byref(_lib.Request()), ctypes.byref(rid)) == _lib.BRI17_ERR_INVALID
    assert L.bri17_server_poll(None, 1, None) == _lib.BRI17_ERR_INVALID
    assert L.bri17_server_get_stats(None, None) == _lib.BRI17_ERR_INVALID
    assert L.bri17_server_destroy(None) == _lib.BRI17_OK
    import torch
    if not torch.cuda.is_available():
        h = ctypes.c_void_p()
        assert L.bri17_server_create(None, ctypes.byref(h)) == _lib.BRI17_ERR_HIP
