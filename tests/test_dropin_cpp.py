"""The C++ drop-in header (include/bri17/bri17.hpp) used the way the
reference's own C++ test uses bri17.hpp (tests/test_bri17.cpp:318-536)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "test_dropin")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _run(mode):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/test_dropin is not built (run __graft_entry__.build())")
    return subprocess.run([EXE, mode, GOLDEN], capture_output=True, text=True, timeout=120)


def test_header_grid_and_per_frequency_functions(oracle):
    r = _run("host")
    assert r.returncode == 0, r.stdout + r.stderr
    shape, L = (3, 4, 5), (3.3, 4.8, 6.5)
    lines = r.stdout.splitlines()
    assert lines[0] == "repr CartesianGrid<d,3>{shape={3,4,5,},L={3.3,4.8,6.5,}}"
    assert lines[1] == "node %d %d" % (2 * 4 + 3, (2 * 4 + 3) * 5 + 4)
    n = 0
    for ln in lines:
        f = ln.split()
        if f[0] == "cell":
            assert [int(x) for x in f[2:]] == oracle.cell_nodes(shape, L, int(f[1]))
        elif f[0] in ("K", "B"):
            k = [int(x) for x in f[1:4]]
            v = np.array([float(x) for x in f[4:]]).view(np.complex128)
            ref = (oracle.modal_stiffness(shape, L, 5.6, 0.3, k).ravel() if f[0] == "K"
                   else oracle.modal_strain_displacement(shape, L, k))
            assert np.abs(v - ref).max() <= 1e-13 * (1 + np.abs(ref).max())
            n += 1
        elif f[0] == "Kf":
            ref = oracle.modal_stiffness(shape, L, 5.6, 0.3, [1, 2, 3]).real
            assert abs(float(f[1]) - ref[0, 0]) <= 1e-6 * ref[0, 0]
    assert n == 18


@pytest.mark.gpu
def test_header_reference_assembly_sections_on_device(device):
    """2D and 3D stiffness and strain SECTIONs of the reference test through
    bri17::StiffnessMatrix<double, DIM> on the GPU, at rtol 1e-15 / atol 1e-14."""
    r = _run("device")
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
