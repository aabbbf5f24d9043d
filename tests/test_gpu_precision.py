"""FP32 on the GPU against the reference's double-precision build.

Bar (the single-precision requirement, written here): seismograms within 1e-4
relative L2 of the reference's float64 build.  Final wavefields are held to
the same 1e-4 where the wave is still in the grid (the reduced layered cases,
600 steps); after the wave has left the grid (the shipped programs, 2000 and
4000 steps, and the reduced 3D case) the residual field is rounding noise
and the reference's own float32 build differs from its float64 build by up
to 4.1e-3 (2D4 sigma_xx) -- there the GPU's float32 fields must be no
further from the float64 reference than 1.5 times what the reference's own
float32 build is.
"""
import numpy as np
import pytest

from case_inputs import fixture
from golden.cases import CASES
from gpu_runner import rel_l2, run_gpu

pytestmark = pytest.mark.gpu

BAR = 1e-4


def _fields(case):
    if CASES[case].get("NZ"):
        return ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")
    return ("vx", "vy", "sxx", "syy", "sxy")


@pytest.mark.parametrize("case", list(CASES))
def test_f32_gpu_against_f64_reference_build(case):
    f64, f32 = fixture(case, "f64"), fixture(case, "f32")
    g = run_gpu(case, "f32", energy=False)
    for k in ("sisvx", "sisvy"):
        assert rel_l2(g[k], f64[k]) <= BAR, k
    step = 3 if CASES[case].get("NZ") else 4
    sl = (slice(None, None, step),) * (3 if CASES[case].get("NZ") else 2)
    for k in _fields(case):
        err = rel_l2(g[k][sl], f64[f"sub_{k}"])
        own = rel_l2(f32[f"sub_{k}"], f64[f"sub_{k}"])
        assert err <= max(BAR, 1.5 * own), (k, err, own)
