"""PNM snapshots of the shipped 2D programs (2D2:716-731, 861-999) and of the
reduced 3D program (3D:1231-1234, the cut plane k = NZ/2): the oracle's
colouring (oracle/color_image.py) of the oracle's own fields against the
images the compiled reference wrote (tests/golden/images_<prog>_<prec>.npz:
2D2 at steps 5, 100, 200 and 2D4 at steps 5, 200; display_3d_<prec>.npz: 3D
at steps 5, 100, 200, 300; f64 and f32 runs).  Bar: every pixel equal.

The 3D timestamp files (3D:1219-1229) are checked here for what the oracle
fixes: the step and time lines, and the velocity-norm and energy lines
against the oracle's fields and energy."""
import sys
from pathlib import Path

import numpy as np
import pytest

from case_inputs import inputs, run_oracle
from golden.cases import CASES

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
from color_image import color_image  # noqa: E402

GOLDEN = Path(__file__).resolve().parent / "golden"


SNAPSHOTS = [("2d2", 5), ("2d2", 100), ("2d2", 200), ("2d4", 5), ("2d4", 200)]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("prog,it", SNAPSHOTS)
def test_oracle_images_match_reference(prog, it, prec):
    g = np.load(GOLDEN / f"images_{prog}_{prec}.npz")
    case = f"{prog}_shipped"
    d = inputs(case, prec, it)
    o = run_oracle(case, prec, d=d, energy=False, nthreads=4)
    for comp, key in (("Vx", "vx"), ("Vy", "vy")):
        img = color_image(o[key], d["ISOURCE"], d["JSOURCE"], d["ix_rec"], d["iy_rec"],
                          CASES[case]["NPML"], ghost=1 if prog == "2d4" else 0)
        np.testing.assert_array_equal(img, g[f"{comp}_{it}"], err_msg=f"{comp} step {it}")


def test_golden_images_are_not_trivial():
    g = np.load(GOLDEN / "images_2d2_f64.npz")
    for it in (100, 200):
        px = g[f"Vx_{it}"].reshape(-1, 3)
        red = (px[:, 0] > 0) & (px[:, 1] == 0) & (px[:, 2] == 0)
        blue = (px[:, 2] > 0) & (px[:, 0] == 0) & (px[:, 1] == 0)
        assert red.sum() > 100 and blue.sum() > 100


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("it", [5, 100, 200, 300])
def test_oracle_3d_images_match_reference(it, prec):
    g = np.load(GOLDEN / f"display_3d_{prec}.npz")
    d = inputs("3d_small", prec, it)
    o = run_oracle("3d_small", prec, d=d, energy=False, nthreads=4)
    k = o["vx"].shape[0] // 2 - 1  # global plane NZ/2
    for comp, key in (("Vx", "vx"), ("Vy", "vy")):
        img = color_image(o[key][k], d["ISOURCE"], d["JSOURCE"], d["ix_rec"], d["iy_rec"],
                          CASES["3d_small"]["NPML"])
        np.testing.assert_array_equal(img, g[f"{comp}_{it}"], err_msg=f"{comp} step {it}")


def test_oracle_3d_final_plane_images():
    """The step-300 images from the reference's own final cut plane."""
    for prec in ("f64", "f32"):
        g = np.load(GOLDEN / f"display_3d_{prec}.npz")
        d = inputs("3d_small", prec, 300)
        for comp, key in (("Vx", "vx_plane"), ("Vy", "vy_plane")):
            img = color_image(g[key], d["ISOURCE"], d["JSOURCE"], g["ix_rec"], g["iy_rec"],
                              CASES["3d_small"]["NPML"])
            np.testing.assert_array_equal(img, g[f"{comp}_300"])


def timestamp_fields(text):
    """(label, value text) for each line of a 3D timestamp file."""
    out = []
    for line in text.rstrip("\n").split("\n"):
        if "hh:mm:ss" in line:
            out.append((line, None))
            continue
        label, _, rest = line.rpartition("=") if "=" in line else line.partition(":")
        out.append((label, rest.split()[0] if "Time:" in line or "=" in line
                    else line.split("#")[1].strip()))
    return out


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_oracle_3d_timestamps(prec):
    """Step, time, velocity norm and total energy lines of the reference's
    timestamp files against the oracle: norm = max |v| over the grid, energy
    = the oracle's total energy at that step."""
    g = np.load(GOLDEN / f"display_3d_{prec}.npz")
    rtol = 1e-12 if prec == "f64" else 1e-5
    dt = np.float32(1.6e-3) if prec == "f32" else 1.6e-3
    for it in (5, 100, 200, 300):
        f = timestamp_fields(str(g[f"timestamp_{it}"]))
        assert len(f) == 7
        assert int(f[0][1]) == it
        assert np.float32(f[1][1]) == np.float32((it - 1) * dt)
        d = inputs("3d_small", prec, it)
        o = run_oracle("3d_small", prec, d=d, energy=True, nthreads=4)
        vn = np.sqrt(o["vx"].astype(np.float64) ** 2 + o["vy"].astype(np.float64) ** 2
                     + o["vz"].astype(np.float64) ** 2).max()
        assert abs(float(f[2][1]) / vn - 1) < rtol, (it, f[2], vn)
        assert abs(float(f[3][1]) / o["etotal"][it - 1] - 1) < 10 * rtol, (it, f[3])
        assert float(f[4][1]) >= 0 and float(f[6][1]) >= 0
