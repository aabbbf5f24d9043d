"""Parity on heterogeneous media: every kernel family against the oracle on a
medium that varies along every axis, so that an offset error on a material
array (a neighbour's lambda, mu or rho, a plane or row stride) cannot hide.

* 3D second order (seismic_CPML_3D_isotropic_MPI_OpenMP.f90:836-1052), the
  benchmark's medium at test size (96 x 80 x 72, 12 layers along z, +-5%
  smoothed perturbation, Courant 0.5), with K_MAX_PML = 1 and with K > 1 in
  the slabs (the d/K branch of every memory-variable update), through the
  two-phase kernels and the fused step.  Bar: bit-exact on the 9 fields, the
  18 memory variables and every seismogram sample; energy 1e-12 (f64) / 5e-4
  (f32) rel-L2 (summed in another order).
* 2D orders 2, 4 and 8 on the layered 2D cases with a smoothed perturbation
  (the shipped layered fixtures vary along y only).
* The 2D viscoelastic steps (orders 2 and 4) and the ADE-PML RK4 step on
  their reduced cases with a perturbed medium.
"""
import numpy as np
import pytest

import layered_media as lm
import oracle_lib as orc
from case_inputs import inputs, run_oracle, visco2d_inputs
from gpu_runner import rel_l2, run_gpu

pytestmark = pytest.mark.gpu

F3 = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")


def _check3d(g, o, prec):
    for k in F3:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    mem = [k for k in g if k.startswith("m_")]
    assert len(mem) == 18
    for k in mem:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    for k in ("sisvx", "sisvy"):
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    assert np.abs(g["sisvx"]).max() > 0
    # the oracle (like the reference) sums the 3D energy sequentially in the
    # working precision; over this 76 x 60 x 52 window a float sum carries
    # relative rounding of a few 1e-4, the GPU's double tree sum does not
    tol = 1e-12 if prec == "f64" else 5e-4
    assert rel_l2(g["ekin"] + g["epot"], o["etotal"]) < tol


@pytest.mark.parametrize("kmax", [1.0, 3.0])
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("fused", [False, True])
def test_3d_layered_matches_oracle_bit_exact(prec, kmax, fused):
    d = lm.layered3d_inputs(prec, 200, kmax=kmax)
    if kmax != 1.0:
        assert (d["prof_x"][2] != 1).any() and (d["prof_z"][5] != 1).any()
    g = run_gpu("3d_layered", prec, d=d, memory=True, fused=fused)
    o = run_oracle("3d_layered", prec, d=d, nthreads=8, derived_lp2m=True)
    _check3d(g, o, prec)


@pytest.mark.parametrize("kchunk,tile", [(5, 8), (24, 16)])
def test_3d_layered_tilings_agree(kchunk, tile):
    prec = "f32"
    d = lm.layered3d_inputs(prec, 90, kmax=3.0)
    a = run_gpu("3d_layered", prec, d=d, energy=False)
    for fused in (False, True):
        b = run_gpu("3d_layered", prec, d=d, energy=False, kchunk=kchunk, tile=tile,
                    fused=fused)
        for k in F3 + ("sisvx", "sisvy"):
            np.testing.assert_array_equal(b[k], a[k], err_msg=f"{k} fused={fused}")


@pytest.mark.parametrize("order", [2, 4, 8])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_2d_perturbed_layers_match_oracle(order, prec):
    from golden.cases import CASES

    from gpu_runner import make_solver

    case = "2d4_layered" if order == 4 else "2d2_layered"
    d = inputs(case, prec, 300)
    for q, k in enumerate(("lam", "mu", "rho")):
        d[k] = lm.perturbed(d[k], 100 + q)
    c = CASES[case]
    o = orc.run2d(c["NX"], c["NY"], order, d["DELTAX"], d["DELTAY"], d["DELTAT"], d["nstep"],
                  d["lam"], d["mu"], d["rho"], d["prof_x"], d["prof_y"], d["ISOURCE"],
                  d["JSOURCE"], d["fx"], d["fy"], d["ix_rec"], d["iy_rec"], ewin=d["ewin"],
                  prec=prec, nthreads=8)
    saved = c["order"]
    c["order"] = order
    try:
        with make_solver(case, prec, d) as s:
            s.run()
            for n, k in (("vx", "vx"), ("vy", "vy"), ("sigmaxx", "sxx"), ("sigmayy", "syy"),
                         ("sigmaxy", "sxy")):
                np.testing.assert_array_equal(s.field(n), o[k], err_msg=n)
            sx, sy = s.seismograms()[:2]
            np.testing.assert_array_equal(sx, o["sisvx"])
            np.testing.assert_array_equal(sy, o["sisvy"])
            ek, ep = s.energy()
    finally:
        c["order"] = saved
    tol = 1e-12 if prec == "f64" else 1e-4
    assert rel_l2(ek, o["ekin"]) < tol and rel_l2(ep, o["epot"]) < tol


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_visco2d_perturbed_medium_matches_oracle(prec, order):
    from test_gpu_visco2d import KEYS
    from test_gpu_visco2d import run_gpu as run_visco
    from test_gpu_visco2d import run_oracle as run_visco_oracle

    d, _ = visco2d_inputs(prec, order)
    for q, k in enumerate(("lam", "mu", "rho")):
        d[k] = lm.perturbed(d[k], 200 + q)
    g = run_visco(d, prec, nstep=400)
    o = run_visco_oracle(d, prec, nstep=400)
    for k in KEYS:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_ade8_perturbed_medium_matches_oracle(prec):
    import test_gpu_ade8 as A

    f, d = A.inputs(prec)
    f = dict(f)
    for q, k in enumerate(("lam", "mu", "rho")):
        f[k] = lm.perturbed(f[k], 300 + q)
    g = A.run_gpu(f, d, prec)
    o = A.run_oracle(f, d, prec)
    sl = slice(5, 5 + A.N)
    for k in ("vx", "vy", "sxx", "syy", "sxy"):
        np.testing.assert_array_equal(g[k], o[k][sl, sl], err_msg=k)
    np.testing.assert_array_equal(g["sisvx"], o["sisvx"])
    np.testing.assert_array_equal(g["sisvy"], o["sisvy"])
