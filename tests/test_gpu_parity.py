"""Parity of the HIP path with the oracle and with the compiled reference.

Bars (the accuracy requirements of this port, written here):
  * GPU vs oracle on the same inputs: bit-exact for every field, memory
    variable and seismogram sample (2D and 3D, float64 and float32); the
    kernels keep the reference's association and are built without FMA
    contraction.  Energy is summed in a different order (double, tree) than the
    reference's sequential sum: rel-L2 1e-12 (f64), 1e-4 (f32, where the
    oracle and the reference accumulate in float).
  * GPU vs reference fixtures: 2D seismograms bit-exact in both precisions;
    3D (whose (lambda+2mu) coefficient is formed as DELTAT*lambda +
    2*DELTAT*mu instead of DELTAT*rho*cp^2) within 1e-12 rel-L2 in float64
    and 1e-5 in float32.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

from case_inputs import fixture, inputs, run_oracle
from golden.cases import ALL_CASES, CASES
from gpu_runner import rel_l2, run_gpu

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
from color_image import color_image  # noqa: E402

pytestmark = pytest.mark.gpu

ALL = [(c, p) for c in CASES for p in ("f64", "f32")]


def _fields(case):
    if ALL_CASES[case].get("NZ"):
        return ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")
    return ("vx", "vy", "sxx", "syy", "sxy")


@pytest.mark.parametrize("case,prec", ALL)
def test_gpu_matches_oracle_bit_exact(case, prec):
    nstep = min(CASES[case]["NSTEP"], 1200)
    d = inputs(case, prec, nstep)
    g = run_gpu(case, prec, d=d, memory=True)
    o = run_oracle(case, prec, d=d, nthreads=8, derived_lp2m=True)
    for k in _fields(case):
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    for k in [k for k in g if k.startswith("m_")]:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    np.testing.assert_array_equal(g["sisvx"], o["sisvx"])
    np.testing.assert_array_equal(g["sisvy"], o["sisvy"])
    tol = 1e-12 if prec == "f64" else 1e-4
    if ALL_CASES[case].get("NZ"):
        assert rel_l2(g["ekin"] + g["epot"], o["etotal"]) < tol
    else:
        assert rel_l2(g["ekin"], o["ekin"]) < tol
        assert rel_l2(g["epot"], o["epot"]) < tol


@pytest.mark.parametrize("case,prec", ALL)
def test_gpu_matches_reference_fixture(case, prec):
    f = fixture(case, prec)
    g = run_gpu(case, prec)
    is3d = bool(CASES[case].get("NZ"))
    if not is3d:
        np.testing.assert_array_equal(g["sisvx"], f["sisvx"])
        np.testing.assert_array_equal(g["sisvy"], f["sisvy"])
        for k in _fields(case):
            np.testing.assert_array_equal(g[k][::4, ::4], f[f"sub_{k}"], err_msg=k)
        etol = 1e-12 if prec == "f64" else 1e-4
        assert rel_l2(g["ekin"], f["ekin"]) < etol
        assert rel_l2(g["epot"], f["epot"]) < etol
    else:
        tol = 1e-12 if prec == "f64" else 1e-5
        assert rel_l2(g["sisvx"], f["sisvx"]) < tol
        assert rel_l2(g["sisvy"], f["sisvy"]) < tol
        for k in _fields(case):
            assert rel_l2(g[k][::3, ::3, ::3], f[f"sub_{k}"]) < tol, k
        assert rel_l2(g["ekin"] + g["epot"], f["etotal"]) < (1e-12 if prec == "f64" else 1e-4)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_eighth_order_matches_oracle(prec):
    """8th-order Holberg operator (ADE8:263-266) in the leapfrog C-PML loop:
    no reference trajectory exists; pinned to the oracle bit-exactly."""
    import oracle_lib as orc

    from gpu_runner import make_solver

    case = "2d2_layered"
    d = inputs(case, prec, 400)
    c = dict(CASES[case])
    o = orc.run2d(c["NX"], c["NY"], 8, d["DELTAX"], d["DELTAY"], d["DELTAT"], d["nstep"],
                  d["lam"], d["mu"], d["rho"], d["prof_x"], d["prof_y"], d["ISOURCE"],
                  d["JSOURCE"], d["fx"], d["fy"], d["ix_rec"], d["iy_rec"], prec=prec,
                  nthreads=8)
    CASES[case]["order"] = 8
    try:
        with make_solver(case, prec, d, energy=False) as s:
            s.run()
            for n, k in (("vx", "vx"), ("vy", "vy"), ("sigmaxx", "sxx"), ("sigmaxy", "sxy")):
                np.testing.assert_array_equal(s.field(n), o[k], err_msg=n)
            sx, sy = s.seismograms()
            np.testing.assert_array_equal(sx, o["sisvx"])
    finally:
        CASES[case]["order"] = c["order"]


def test_graph_eager_and_timed_paths_identical():
    case, prec = "2d4_layered", "f32"
    d = inputs(case, prec, 300)
    a = run_gpu(case, prec, d=d, graph_steps=16)
    b = run_gpu(case, prec, d=d, graph_steps=0)
    c = run_gpu(case, prec, d=d, timing=True)
    for k in ("vx", "vy", "sxx", "sisvx"):
        np.testing.assert_array_equal(a[k], b[k])
        np.testing.assert_array_equal(a[k], c[k])
    (n0, t0), (n1, t1) = c["ktime"]
    assert n0 == 300 and n1 == 300 and t0 > 0 and t1 > 0


def test_run_in_chunks_equals_one_run():
    from gpu_runner import make_solver

    case, prec = "3d_small", "f64"
    d = inputs(case, prec, 120)
    with make_solver(case, prec, d) as s:
        s.run(1, 120)
        whole = s.field("vz")
    with make_solver(case, prec, d) as s:
        for a, b in ((1, 7), (8, 60), (61, 120)):
            s.run(a, b)
        parts = s.field("vz")
    np.testing.assert_array_equal(whole, parts)


@pytest.mark.parametrize("kchunk,tile", [(1, 4), (7, 8), (64, 16), (5, 16)])
def test_3d_tiling_does_not_change_results(kchunk, tile):
    case, prec = "3d_small", "f32"
    d = inputs(case, prec, 80)
    ref = run_gpu(case, prec, d=d, energy=False)
    g = run_gpu(case, prec, d=d, energy=False, kchunk=kchunk, tile=tile)
    for k in _fields(case):
        np.testing.assert_array_equal(g[k], ref[k])


def test_linearity_and_zero_source():
    """Doubling the force doubles every field exactly (x2 is exact in binary
    floating point); a zero force leaves every field at zero."""
    case, prec = "2d2_layered", "f64"
    d = inputs(case, prec, 300)
    a = run_gpu(case, prec, d=d, energy=False)
    d2 = dict(d)
    d2["fx"], d2["fy"] = d["fx"] * 2, d["fy"] * 2
    b = run_gpu(case, prec, d=d2, energy=False)
    for k in ("vx", "vy", "sxx", "syy", "sxy"):
        np.testing.assert_array_equal(b[k], 2 * a[k])
    d0 = dict(d)
    d0["fx"], d0["fy"] = d["fx"] * 0, d["fy"] * 0
    z = run_gpu(case, prec, d=d0, energy=False)
    for k in ("vx", "vy", "sxx", "syy", "sxy"):
        assert not np.any(z[k])


def test_slab_widths_follow_profiles():
    case, prec = "2d4_shipped", "f64"
    d = inputs(case, prec, 2)
    g = run_gpu(case, prec, d=d, energy=False)
    ax = (d["prof_x"][0] != 0) | (d["prof_x"][3] != 0)
    ay = (d["prof_y"][0] != 0) | (d["prof_y"][3] != 0)
    nx, ny = len(ax), len(ay)
    wl, wr, hl, hr = g["slabs"][:4]
    assert ax[:wl].any() and not ax[wl:nx - wr].any()
    assert ay[:hl].any() and not ay[hl:ny - hr].any()


def test_cfl_violation_is_reported():
    from seismic_cpml_amd import CpmlError
    from seismic_cpml_amd.solver import CPML_ERR_CFL

    from gpu_runner import make_solver

    case, prec = "2d2_shipped", "f64"
    d = inputs(case, prec, 10)
    d["DELTAT"] = d["DELTAT"] * 3  # Courant 0.93 -> 2.8
    with pytest.raises(CpmlError) as e:
        make_solver(case, prec, d)
    assert e.value.code == CPML_ERR_CFL
    assert "time step is too large" in e.value.message


def test_blowup_is_reported_at_step_5():
    """2D2:716-726 checks max |v| at it == 5 and every IT_DISPLAY steps; a
    threshold below the actual norm forces the error path from the host."""
    from seismic_cpml_amd import CpmlError
    from seismic_cpml_amd.solver import CPML_ERR_BLOWUP

    from gpu_runner import make_solver

    case, prec = "2d2_shipped", "f64"
    d = inputs(case, prec, 50)
    with make_solver(case, prec, d) as s:
        s.set_stability(100, 1e-30)
        with pytest.raises(CpmlError) as e:
            s.run()
        assert e.value.code == CPML_ERR_BLOWUP
        assert "step 5" in e.value.message


def test_determinism_repeat_runs():
    case, prec = "3d_small", "f32"
    d = inputs(case, prec, 100)
    a = run_gpu(case, prec, d=d)
    b = run_gpu(case, prec, d=d)
    for k in ("vx", "vz", "syz", "ekin", "epot"):
        np.testing.assert_array_equal(a[k], b[k])


@pytest.mark.parametrize("tile", [4, 8, 16, 32, 64])
@pytest.mark.parametrize("case", ["2d4_layered", "2d2_layered"])
def test_2d_tiling_does_not_change_results(case, tile):
    """Rows per tile (one or several row groups per block) leave every field
    bit-identical to the oracle."""
    prec = "f32"
    d = inputs(case, prec, 200)
    g = run_gpu(case, prec, d=d, energy=False, tile=tile)
    o = run_oracle(case, prec, d=d, nthreads=8, energy=False)
    for k in _fields(case):
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    np.testing.assert_array_equal(g["sisvx"], o["sisvx"])


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_3d_interior_and_shell_tiles_match_oracle(prec):
    """A grid where some tiles touch no absorbing slab: those run in the launch
    without C-PML code, the rest in the shell launch.  Bit-exact against the
    oracle for every field, memory variable and seismogram sample."""
    case = "3d_wide"
    d = inputs(case, prec, 200)
    g = run_gpu(case, prec, d=d, memory=True)
    o = run_oracle(case, prec, d=d, nthreads=8, derived_lp2m=True)
    for k in _fields(case):
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    mem = [k for k in g if k.startswith("m_")]
    assert len(mem) == 18
    for k in mem:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    for k in ("sisvx", "sisvy"):
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    assert np.abs(g["sisvx"]).max() > 0
    tol = 1e-12 if prec == "f64" else 1e-4
    assert rel_l2(g["ekin"] + g["epot"], o["etotal"]) < tol


@pytest.mark.parametrize("kchunk,tile", [(1, 4), (3, 8), (16, 16), (8, 16), (40, 4)])
def test_3d_interior_box_tilings_agree(kchunk, tile):
    case, prec = "3d_wide", "f32"
    d = inputs(case, prec, 120)
    ref = run_gpu(case, prec, d=d, energy=False)
    g = run_gpu(case, prec, d=d, energy=False, kchunk=kchunk, tile=tile)
    for k in _fields(case) + ("sisvx", "sisvy"):
        np.testing.assert_array_equal(g[k], ref[k], err_msg=k)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("prog,steps", [("2d2", (5, 100, 200)), ("2d4", (5, 200)),
                                        ("3d", (5, 100, 200, 300))])
def test_gpu_images_match_reference(prog, steps, prec):
    """PNM snapshots computed on the device (cpml_gpu_color_image) against the
    reference's image files (tests/golden/images_<prog>_<prec>.npz, and for 3D
    the cut-plane images of the reduced program, display_3d_<prec>.npz): every
    pixel equal; the 4th-order program images its ghost-padded arrays.  The
    3D fields match the reference to rounding, not bit for bit, so there the
    device image must also equal the oracle's colouring of the device's own
    cut plane, and in f32 at most 0.5% of the pixels may differ from the
    reference's (a value within rounding of a colour-level boundary)."""
    from gpu_runner import make_solver
    name = f"display_3d_{prec}.npz" if prog == "3d" else f"images_{prog}_{prec}.npz"
    g = np.load(Path(__file__).resolve().parent / "golden" / name)
    case = "3d_small" if prog == "3d" else f"{prog}_shipped"
    d = inputs(case, prec, max(steps))
    with make_solver(case, prec, d, energy=False) as s:
        done = 0
        for it in steps:
            s.run(done + 1, it)
            done = it
            for comp, fn in (("Vx", 1), ("Vy", 2)):
                img = s.color_image(fn)
                if prog == "3d":
                    plane = s.field(comp.lower())[s.nz // 2 - 1]
                    own = color_image(plane, d["ISOURCE"], d["JSOURCE"], d["ix_rec"],
                                      d["iy_rec"], ALL_CASES[case]["NPML"])
                    np.testing.assert_array_equal(img, own, err_msg=f"{comp} step {it}")
                if prog == "3d" and prec == "f32":
                    bad = (img != g[f"{comp}_{it}"]).any(-1).sum()
                    assert bad <= 0.005 * img.shape[0] * img.shape[1], (comp, it, bad)
                    continue
                np.testing.assert_array_equal(img, g[f"{comp}_{it}"],
                                              err_msg=f"{comp} step {it}")
