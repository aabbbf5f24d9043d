"""ADE-PML RK4 8th order on the GPU (cpml_gpu_set_ade_rk4: four k_ade_stage
launches per step, X ping-ponging between the handle's two state buffers,
and the 4-node vx record) against the oracle (runade8), which equals the
compiled reference bit for bit
(tests/test_oracle_ade8.py), on the reduced ADE8 case (81 x 81, 300 steps, 81
receivers).  Bars: float64 -- every final field and seismogram sample
bit-exact against the oracle, and the seismograms bit-exact against the
reference fixture; float32 -- bit-exact against the float32 oracle."""
import numpy as np
import pytest

import oracle_lib as orc
from case_inputs import GOLDEN

pytestmark = pytest.mark.gpu

N, NSTEP, DX, DT, F0, FACTOR, ANGLE = 81, 300, 10.0, 3e-3, 10.0, 1e4, 90.0
TABLES = ("a_x", "b_x", "ah_x", "bh_x", "a_y", "b_y", "ah_y", "bh_y")
NAMES = ("vx", "vy", "sigmaxx", "sigmayy", "sigmaxy")


def inputs(prec):
    f = np.load(GOLDEN / "ade8_small_f64.npz")
    dt = np.float64 if prec == "f64" else np.float32
    R = dt
    t0 = R(1.0) / R(F0)
    rk41 = np.array([0.0, 0.5, 0.5, 1.0], dt)
    times = ((np.arange(NSTEP, dtype=dt)[:, None] + rk41[None, :]) * R(DT)).astype(dt)
    fx4, fy4 = orc.stf_at(times.ravel(), R(F0), t0, R(FACTOR), R(ANGLE), 1, prec)
    return f, dict(dtype=dt, t0=t0, fx4=fx4.reshape(NSTEP, 4), fy4=fy4.reshape(NSTEP, 4))


# the program's arrays run from -4 to N+4: Fortran index 1 is element 5
def run_gpu(f, d, prec, graph_steps=16, timing=False, segments=None, reload_at=None):
    from seismic_cpml_amd import CpmlGpu
    dt = d["dtype"]
    sl = slice(5, 5 + N)
    with CpmlGpu(2, 8, 8 if prec == "f64" else 4, N, N, 1, DX, DX, 1.0, DT, 10, NSTEP) as s:
        s.set_medium(*(f[k][sl, sl].astype(dt) for k in ("lam", "mu", "rho")))
        prof = [np.stack([f[f"a_{ax}"][0, sl], f[f"b_{ax}"][0, sl], f[f"K_{ax}"][sl],
                          f[f"ah_{ax}"][0, sl], f[f"bh_{ax}"][0, sl], f[f"Kh_{ax}"][sl]]).astype(dt)
                for ax in ("x", "y")]
        s.set_pml(*prof)
        s.set_ade_rk4({k: f[k][:, sl].astype(dt) for k in TABLES}, d["fx4"], d["fy4"])
        s.set_source(40, 40, 1, np.zeros(NSTEP, dt), np.zeros(NSTEP, dt))
        s.set_receivers(f["ix_rec"], f["iy_rec"])
        s.set_graph_steps(graph_steps)
        if timing:
            s.set_kernel_timing(True)
        for b, e in segments or [(1, NSTEP)]:
            s.run(b, e)
            if reload_at == e:  # read the state back and write it again (checkpoint)
                for n in NAMES:
                    s.set_field(n, s.field(n))
        g = {k: s.field(n) for k, n in zip(("vx", "vy", "sxx", "syy", "sxy"), NAMES)}
        g["sisvx"], g["sisvy"] = s.seismograms()[:2]
    return g


def run_oracle(f, d, prec):
    dt = d["dtype"]
    prof = {k: f[k].astype(dt) for k in ("a_x", "b_x", "K_x", "ah_x", "bh_x", "Kh_x", "a_y",
                                          "b_y", "K_y", "ah_y", "bh_y", "Kh_y")}
    return orc.runade8(N, N, dt(DX), dt(DX), dt(DT), NSTEP, 10, f["lam"].astype(dt),
                       f["mu"].astype(dt), f["rho"].astype(dt), prof, 40, 40, dt(F0), d["t0"],
                       dt(FACTOR), dt(ANGLE), f["ix_rec"], f["iy_rec"], prec=prec, nthreads=8)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_ade8_matches_oracle_bit_exact(prec):
    f, d = inputs(prec)
    g = run_gpu(f, d, prec)
    o = run_oracle(f, d, prec)
    sl = slice(5, 5 + N)
    for k in ("vx", "vy", "sxx", "syy", "sxy"):
        np.testing.assert_array_equal(g[k], o[k][sl, sl], err_msg=k)
    np.testing.assert_array_equal(g["sisvx"], o["sisvx"])
    np.testing.assert_array_equal(g["sisvy"], o["sisvy"])
    assert np.abs(g["sisvy"]).max() > 0
    if prec == "f64":
        np.testing.assert_array_equal(g["sisvx"], f["sisvx"])
        np.testing.assert_array_equal(g["sisvy"], f["sisvy"])


def test_ade8_graph_equals_eager():
    f, d = inputs("f32")
    a = run_gpu(f, d, "f32", graph_steps=16)
    b = run_gpu(f, d, "f32", graph_steps=0)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_ade8_timed_run_is_the_same_step():
    """With per-kernel timing on (events around every step) the handle still
    takes the RK4 step: results equal the untimed run bit for bit."""
    f, d = inputs("f32")
    a = run_gpu(f, d, "f32")
    b = run_gpu(f, d, "f32", timing=True)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


@pytest.mark.parametrize("graph_steps", [0, 5])
def test_ade8_odd_segments_and_checkpoint(graph_steps):
    """X alternates between two buffers every step: runs split into segments
    of odd lengths (graphs of an even number of steps, eager remainders),
    with the state read back and written again after an odd step, equal the
    single run bit for bit."""
    f, d = inputs("f32")
    a = run_gpu(f, d, "f32", graph_steps=16)
    b = run_gpu(f, d, "f32", graph_steps=graph_steps,
                segments=[(1, 7), (8, 150), (151, 151), (152, NSTEP)], reload_at=7)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
