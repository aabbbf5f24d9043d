"""The C-ABI host layer under AddressSanitizer.

``seismic_cpml_amd/_build/cpml_host_check_asan`` is the library's host code
compiled with ``-fsanitize=address`` (device code unchanged) and linked with
``tests/native/cpml_host_check.cpp``, which calls every entry point: argument
validation, NULL handles, error messages, the host set-up helpers and file
writers (CPU), and every set-up / run / read-back path of the solver, the
batch and the viscoelastic and ADE-RK4 steps on small grids (GPU).  ASan
aborts the program on the first out-of-bounds access, use-after-free or
double free in host code; the program itself fails on any unexpected status.
"""
import os
import subprocess
from pathlib import Path

import pytest

EXE = Path(__file__).resolve().parents[1] / "seismic_cpml_amd" / "_build" / "cpml_host_check_asan"


def _run(mode, tmp_path, leaks):
    assert EXE.is_file(), f"{EXE} not built (python -m seismic_cpml_amd.build)"
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = f"detect_leaks={int(leaks)}:abort_on_error=0:halt_on_error=1"
    r = subprocess.run([str(EXE), mode, str(tmp_path)], capture_output=True, text=True,
                       timeout=600, env=env)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out, out[-6000:]
    assert r.returncode == 0 and "HOST CHECK OK" in r.stdout, out[-4000:]
    return r.stdout


def test_host_layer_under_asan_cpu(tmp_path):
    # leak checks where no GPU driver is present (the HIP runtime keeps its
    # own allocations until exit on a GPU machine)
    out = _run("cpu", tmp_path, leaks=not Path("/dev/kfd").exists())
    assert (tmp_path / "x.pnm").is_file() and (tmp_path / "x.su").is_file()
    assert "0 failed" in out


@pytest.mark.gpu
def test_host_layer_under_asan_gpu(tmp_path):
    # the HIP runtime keeps allocations until exit: leak reports are its own
    out = _run("gpu", tmp_path, leaks=False)
    assert "gpu paths" in out and "0 failed" in out
