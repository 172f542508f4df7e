"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/cpp/test_host_sanitize: the C ABI's host translation units (per-
frequency functions, grid helpers, argument validation, planning, batches,
graphs, the request server) rebuilt with ASan + UBSan on the host pass only
(device code unchanged); tests/cpp/test_oracle_sanitize: the CPU oracle
rebuilt with the same sanitizers.  Built by `make -C tests/cpp sanitize`
(part of __graft_entry__.build()).  A run passes when the program's own
checks pass and no sanitizer reports an error (ASan aborts on the first
error; UBSan is built with -fno-sanitize-recover)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")


def _run(exe, args=(), leaks=True, timeout=300):
    path = os.path.join(CPP, exe)
    if not os.path.exists(path):
        pytest.fail("%s is not built (run `make -C tests/cpp sanitize`)" % path)
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=%d:abort_on_error=0:halt_on_error=1" % (1 if leaks else 0)
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    r = subprocess.run([path] + list(args), capture_output=True, text=True, timeout=timeout, env=env)
    out = r.stdout + r.stderr
    print(out[-3000:])
    assert r.returncode == 0, out[-3000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out
    assert "OK (0 failures)" in r.stdout
    return out


def test_oracle_under_sanitizers():
    _run("test_oracle_sanitize")


def test_host_abi_under_sanitizers():
    """Host paths only: no device work is submitted (without a GPU the
    creation calls return their error codes)."""
    _run("test_host_sanitize")


@pytest.mark.gpu
def test_host_abi_under_sanitizers_with_device_work(device):
    """The same binary driving operators, a batch with its graph, and the
    request server fed by four submitting threads on the GPU; leak checking
    is off (the HIP runtime keeps process-lifetime allocations)."""
    _run("test_host_sanitize", ["device"], leaks=False)
