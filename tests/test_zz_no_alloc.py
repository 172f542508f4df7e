"""bri17_operator_run_noalloc: the operator call that allocates nothing (its
workspace is the library's static device scratch array).  No reference
counterpart; checked against the allocating entry points of the same
operator (bri17_operator_apply / _solve / _strain / _eigenstress), which run
the same kernels: bitwise equal, on two calls in a row with the second on a
smaller grid, for every mode and both dtypes.  Sizes above the scratch are
refused before anything is launched."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu

ENTRY = {0: "bri17_operator_apply", 1: "bri17_operator_solve", 2: "bri17_operator_strain",
         3: "bri17_operator_eigenstress"}


def _call(fn, op, x, out):
    import torch
    from bri17_amd import _lib
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(fn(op.handle, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()), st))


def _noalloc(op, mode, x, out):
    import torch
    from bri17_amd import _lib
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return _lib.lib().bri17_operator_run_noalloc(op.handle, mode, ctypes.c_void_p(x.data_ptr()),
                                                  ctypes.c_void_p(out.data_ptr()), st)


# (larger grid, smaller grid) per dtype: fixed-size and runtime-size kernels
PAIRS = {"float32": [((128, 128, 128), (48, 36, 54)), ((96, 80), (36, 27))],
         "float64": [((64, 64, 64), (30, 24, 20)), ((128, 96), (40, 18))]}


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_noalloc_matches_allocating_entry_point(device, dtype, mode):
    import torch
    from bri17_amd import StiffnessMatrix, _lib
    L = _lib.lib()
    tdt = getattr(torch, dtype)
    g = torch.Generator(device=device).manual_seed(31 + mode)
    for big, small in PAIRS[dtype]:
        for shape in (big, small):  # two calls in a row, the second smaller
            op = StiffnessMatrix(shape, None, 1.3, 0.27, dtype)
            D = len(shape)
            sym = 3 if D == 2 else 6
            cin = sym if mode == 3 else D
            cout = sym if mode >= 2 else D
            x = torch.randn((cin,) + shape, generator=g, device=device, dtype=tdt)
            ref = torch.empty((cout,) + shape, device=device, dtype=tdt)
            got = torch.full((cout,) + shape, float("nan"), device=device, dtype=tdt)
            print("noalloc", dtype, "mode", mode, shape, flush=True)
            _call(getattr(L, ENTRY[mode]), op, x, ref)
            _lib.check(_noalloc(op, mode, x, got))
            torch.cuda.synchronize()
            assert torch.equal(got, ref), (shape, dtype, mode)
            del op


def test_noalloc_refuses_workspace_above_scratch(device):
    """256^3 fp32 K u needs a 201 MB workspace, above the 64 MiB scratch:
    refused with BRI17_ERR_NOMEM before any launch; the output is untouched."""
    import torch
    from bri17_amd import StiffnessMatrix, _lib
    op = StiffnessMatrix((256, 256, 256), None, 1.0, 0.3, "float32")
    x = torch.zeros((3, 256, 256, 256), device=device)
    out = torch.full_like(x, 7.0)
    rc = _noalloc(op, 0, x, out)
    assert rc == _lib.BRI17_ERR_NOMEM
    assert b"static scratch" in _lib.lib().bri17_last_error()
    torch.cuda.synchronize()
    assert bool((out == 7.0).all())
    assert _noalloc(op, 4, x, out) == _lib.BRI17_ERR_INVALID
    assert _noalloc(op, 0, x, x) == _lib.BRI17_ERR_INVALID
