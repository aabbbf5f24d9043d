"""2D viscoelastic C-PML on the GPU (cpml_gpu_set_attenuation; kernels
k_stress2dve / k_velocity2dve / k_pressure2d, orders 2 and 4: the second- and
fourth-order viscoelastic programs) against the oracle (run2dve), which
equals the compiled reference bit for bit (tests/test_oracle_visco2d.py).

Bars: float64 -- every field, C-PML memory variable, attenuation memory
variable and seismogram sample (vx, vy, pressure) bit-exact against the oracle
and the seismograms bit-exact against the reference fixture; float32 --
bit-exact against the float32 oracle on the same inputs.  Graph replay and
eager stepping agree bit for bit.
"""
import numpy as np
import pytest

import oracle_lib as orc
from case_inputs import visco2d_inputs
from gpu_runner import F2D, MEM2D, ORACLE_NAME, oracle_mem_name

pytestmark = pytest.mark.gpu


def run_gpu(d, prec, nstep=None, graph_steps=16):
    from seismic_cpml_amd import CpmlGpu
    nstep = nstep or d["nstep"]
    with CpmlGpu(2, d["order"], 8 if prec == "f64" else 4, d["NX"], d["NY"], 1, d["DELTAX"],
                 d["DELTAY"], 1.0, d["DELTAT"], d["NPML"], nstep) as s:
        s.set_medium(d["lam"], d["mu"], d["rho"])
        s.set_pml(d["prof_x"], d["prof_y"])
        s.set_attenuation(*d["taus"])
        s.set_source(d["ISOURCE"], d["JSOURCE"], 1, d["fx"][:nstep], d["fy"][:nstep])
        s.set_receivers(d["ix_rec"], d["iy_rec"])
        s.set_graph_steps(graph_steps)
        s.run()
        g = {ORACLE_NAME[n]: s.field(n) for n in F2D}
        g.update({oracle_mem_name(m): s.memory(m) for m in MEM2D})
        for k, name in enumerate(("r_e1", "r_e11", "r_e13")):
            g[name] = s.attenuation_memory(k)
        g["sisvx"], g["sisvy"] = s.seismograms()[:2]
        g["sispressure"] = s.pressure()
    return g


def run_oracle(d, prec, nstep=None):
    nstep = nstep or d["nstep"]
    return orc.run2dve(d["NX"], d["NY"], d["DELTAX"], d["DELTAY"], d["DELTAT"], nstep,
                       d["lam"], d["mu"], d["rho"], d["prof_x"], d["prof_y"], d["taus"],
                       d["ISOURCE"], d["JSOURCE"], d["fx"][:nstep], d["fy"][:nstep],
                       d["ix_rec"], d["iy_rec"], prec=prec, nthreads=8, order=d["order"])


KEYS = ("vx", "vy", "sxx", "syy", "sxy", "m_dvx_dx", "m_dvx_dy", "m_dvy_dx", "m_dvy_dy",
        "m_dsxx_dx", "m_dsyy_dy", "m_dsxy_dx", "m_dsxy_dy", "r_e1", "r_e11", "r_e13",
        "sisvx", "sisvy", "sispressure")


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_visco2d_matches_oracle_bit_exact(prec, order):
    d, f = visco2d_inputs(prec, order)
    g = run_gpu(d, prec)
    o = run_oracle(d, prec)
    for k in KEYS:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    assert np.abs(g["sispressure"]).max() > 0 and np.abs(g["r_e13"]).max() > 0
    if prec == "f64":
        for k in ("sisvx", "sisvy", "sispressure"):
            np.testing.assert_array_equal(g[k], f[k], err_msg=k)


def test_visco2d_graph_equals_eager():
    d, _ = visco2d_inputs("f32")
    a = run_gpu(d, "f32", nstep=150, graph_steps=16)
    b = run_gpu(d, "f32", nstep=150, graph_steps=0)
    for k in KEYS:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_attenuation_and_order4_error_paths():
    """Refusals of the new entry points, reported before any launch: the
    attenuation on a 3D handle, non-positive or missing tau values, more than
    8 SLS, pressure without attenuation; the fused step and the learned
    boundary on a 3D 4th-order handle.  Each leaves the handle usable."""
    from seismic_cpml_amd import CpmlError, CpmlGpu, solver
    t = np.array([0.02, 0.004, 0.0009])
    with CpmlGpu(3, 4, 4, 32, 32, 32, 5.0, 5.0, 5.0, 1e-4, 4, 10) as s3:
        with pytest.raises(CpmlError) as e:
            s3.set_attenuation(t, t, t, t)
        assert e.value.code == solver.CPML_ERR_ARG
        with pytest.raises(CpmlError) as e:
            s3.set_fused(True)
        assert e.value.code == solver.CPML_ERR_ARG
    with CpmlGpu(2, 2, 8, 32, 32, 1, 5.0, 5.0, 1.0, 1e-4, 4, 10) as s2:
        with pytest.raises(CpmlError) as e:
            s2.pressure()
        assert e.value.code == solver.CPML_ERR_STATE
        with pytest.raises(CpmlError) as e:
            s2.set_attenuation(t, -t, t, t)
        assert e.value.code == solver.CPML_ERR_ARG
        nine = np.full(9, 0.01)
        with pytest.raises(CpmlError) as e:
            s2.set_attenuation(nine, nine, nine, nine)
        assert e.value.code == solver.CPML_ERR_ARG
        s2.set_attenuation(t, t * 0.9, t, t * 0.9)  # a valid call still succeeds
        s2.set_attenuation([], [], [], [])          # and back to elastic
    with CpmlGpu(2, 8, 8, 32, 32, 1, 5.0, 5.0, 1.0, 1e-4, 4, 10) as s8:
        with pytest.raises(CpmlError) as e:
            s8.set_attenuation(t, t, t, t)
        assert e.value.code == solver.CPML_ERR_ARG
