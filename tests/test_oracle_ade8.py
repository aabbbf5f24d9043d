"""ADE-PML, RK4, 8th order (seismic_ADEPML_2D_elastic_RK4_eighth_order.f90):
the oracle's restatement (runade8) against the compiled reference's own
outputs (tests/golden/ade8_small_f64.npz: 81 x 81, 300 steps, 81 receivers,
the program's stage-dependent ADE coefficients, medium and receiver indices).
Bar: bit-exact on every seismogram sample, final field and memory array.

This is the checker for the ADE8 path; the GPU path of this program is not
built (the GPU 8th-order operator is the leapfrog C-PML one)."""
import numpy as np

import oracle_lib as orc
from case_inputs import GOLDEN
from golden.make_golden import field_digest

KEYS = ("vx", "vy", "sxx", "syy", "sxy", "m_dvx_dx", "m_dvx_dy", "m_dvy_dx", "m_dvy_dy",
        "m_dsxx_dx", "m_dsyy_dy", "m_dsxy_dx", "m_dsxy_dy")


def run(f, nstep=300, prec="f64"):
    prof = {k: f[k] for k in ("a_x", "b_x", "K_x", "ah_x", "bh_x", "Kh_x", "a_y", "b_y", "K_y",
                              "ah_y", "bh_y", "Kh_y")}
    # ADE8:165-229: DELTAX = DELTAY = 10, DELTAT = 3e-3, f0 = 10, t0 = 1/f0,
    # factor 1e4, angle 90 degrees, source at ((NX-1)/2, (NY-1)/2), NPML 10
    return orc.runade8(81, 81, 10.0, 10.0, 3e-3, nstep, 10, f["lam"], f["mu"], f["rho"], prof,
                       40, 40, 10.0, 1.0 / 10.0, 1e4, 90.0, f["ix_rec"], f["iy_rec"], prec=prec,
                       nthreads=4)


def test_ade8_oracle_bit_exact():
    f = np.load(GOLDEN / "ade8_small_f64.npz")
    o = run(f)
    np.testing.assert_array_equal(o["sisvx"], f["sisvx"])
    np.testing.assert_array_equal(o["sisvy"], f["sisvy"])
    assert np.abs(f["sisvy"]).max() > 0
    for k in KEYS:
        np.testing.assert_array_equal(o[k][::4, ::4], f[f"sub_{k}"], err_msg=k)
        assert field_digest(o[k]) == str(f[f"sha_{k}"]), k

