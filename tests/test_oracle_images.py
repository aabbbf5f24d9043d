"""PNM snapshots of the shipped 2D programs (2D2:716-731, 861-999): the
oracle's colouring (oracle/color_image.py) of the oracle's own fields against
the images the compiled reference wrote (tests/golden/images_<prog>_<prec>.npz:
2D2 at steps 5, 100, 200 and 2D4 at steps 5, 200, of the f64 and f32 runs).
Bar: every pixel equal."""
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
