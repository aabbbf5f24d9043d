"""2D viscoelastic C-PML (seismic_CPML_2D_velocity_and_stress_second_order_
viscoelastic.f90 and the fourth-order program, attenuation on): the oracle
(run2dve) against the compiled reference's own outputs
(tests/golden/2dve_small_f64.npz, 2dve4_small_f64.npz: 161 x 161, 700
steps, Qp 65 / Qs 55 with 3 standard linear solids fitted by the reference's
SolvOpt).  Bar: bit-exact on every seismogram sample (vx, vy, pressure),
final field, C-PML memory and attenuation memory variable."""
import numpy as np
import pytest

from case_inputs import visco2d_inputs
from golden.make_golden import field_digest
import oracle_lib as orc

KEYS = ("vx", "vy", "sxx", "syy", "sxy", "m_dvx_dx", "m_dvx_dy", "m_dvy_dx", "m_dvy_dy",
        "m_dsxx_dx", "m_dsyy_dy", "m_dsxy_dx", "m_dsxy_dy", "r_e1", "r_e11", "r_e13")


def run(d, nstep=None, prec="f64"):
    return orc.run2dve(d["NX"], d["NY"], d["DELTAX"], d["DELTAY"], d["DELTAT"],
                       nstep or d["nstep"], d["lam"], d["mu"], d["rho"], d["prof_x"],
                       d["prof_y"], d["taus"], d["ISOURCE"], d["JSOURCE"], d["fx"], d["fy"],
                       d["ix_rec"], d["iy_rec"], prec=prec, nthreads=4, order=d["order"])


@pytest.mark.parametrize("order", [2, 4])
def test_visco2d_oracle_bit_exact(order):
    d, f = visco2d_inputs("f64", order)
    o = run(d)
    for k in ("sisvx", "sisvy", "sispressure"):
        np.testing.assert_array_equal(o[k], f[k], err_msg=k)
    assert np.abs(f["sisvy"]).max() > 0
    for k in KEYS:
        np.testing.assert_array_equal(o[k][..., ::4, ::4], f[f"sub_{k}"], err_msg=k)
        assert field_digest(o[k]) == str(f[f"sha_{k}"]), k


def test_visco2d_attenuation_matters_and_vanishes():
    """The attenuation terms change the result (tau_epsilon = tau_sigma, i.e.
    no attenuation, differs from the fitted SLS by more than 1% in the
    seismograms), and with tau_epsilon = tau_sigma the memory variables stay
    exactly zero."""
    d, f = visco2d_inputs("f64")
    d["taus"] = (d["taus"][1], d["taus"][1], d["taus"][3], d["taus"][3])
    o = run(d)
    for k in ("r_e1", "r_e11", "r_e13"):
        assert not np.any(o[k]), k
    rel = np.linalg.norm(o["sisvy"] - f["sisvy"]) / np.linalg.norm(f["sisvy"])
    assert rel > 1e-2, rel
