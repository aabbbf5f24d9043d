"""3D 4th order on the GPU (the elastic part of
seismic_CPML_3D_viscoelastic_MPI.f90, k_stress3d4 / k_velocity3d4) against the
oracle's single-domain restatement (oracle run3d4 with zcut = 0).

The oracle itself equals the compiled reference bit for bit when it
reproduces the reference's 2-rank slab-interface reads (zcut = NZ/2,
tests/test_oracle_golden.py); the GPU computes the single domain, which the
same oracle code gives with zcut = 0.

Bars: every field, memory variable and seismogram sample bit-exact (float64
and float32; the kernels keep the reference's association and are built
without FMA contraction); energy (double tree sum on the GPU, sequential in
the working precision in the oracle) within 1e-12 (f64) / 1e-4 (f32) rel-L2.
"""
import numpy as np
import pytest

import oracle_lib as orc
from case_inputs import inputs, run_oracle
from gpu_runner import rel_l2, run_gpu

pytestmark = pytest.mark.gpu

CASE = "3d4_small"
FIELDS = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_3d4_matches_oracle_bit_exact(prec):
    d = inputs(CASE, prec)
    g = run_gpu(CASE, prec, d=d, memory=True)
    o = run_oracle(CASE, prec, d=d, nthreads=8)
    for k in FIELDS:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    mem = [k for k in g if k.startswith("m_")]
    assert len(mem) == 18
    for k in mem:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    np.testing.assert_array_equal(g["sisvx"], o["sisvx"])
    np.testing.assert_array_equal(g["sisvy"], o["sisvy"])
    assert np.abs(g["sisvy"]).max() > 0
    tol = 1e-12 if prec == "f64" else 1e-4
    assert rel_l2(g["ekin"] + g["epot"], o["etotal"]) < tol


@pytest.mark.parametrize("kchunk", [1, 7, 40])
def test_3d4_z_chunks_do_not_change_results(kchunk):
    d = inputs(CASE, "f32", 150)
    ref = run_gpu(CASE, "f32", d=d, energy=False)
    g = run_gpu(CASE, "f32", d=d, energy=False, kchunk=kchunk, graph_steps=0)
    for k in FIELDS + ("sisvx", "sisvy"):
        np.testing.assert_array_equal(g[k], ref[k], err_msg=k)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_3d4_ragged_grid_matches_oracle(prec):
    """70 x 37 x 23 (a partial second tile along x, ragged rows and z chunk),
    K > 1 in the absorbing layers, a heterogeneous density, source off the
    grid centre."""
    from seismic_cpml_amd import CpmlGpu
    from gpu_runner import F3D, MEM3D, ORACLE_NAME, oracle_mem_name
    d = inputs(CASE, prec, 120)
    nx, ny, nz, npml, nstep = 70, 37, 23, 6, 120
    dt = d["dtype"]
    prof = {}
    for ax, n in (("x", nx), ("y", ny), ("z", nz)):
        p = orc.profile(n, d["DELTAX"], d["DELTAT"], npml, d["cp"], d["f0"], prec,
                        clamp_alpha=ax == "x")
        p[2] = np.where(p[0] != 0, dt(2.5), dt(1))
        p[5] = np.where(p[3] != 0, dt(1.75), dt(1))
        prof[ax] = p.astype(dt)
    shp = (nz, ny, nx)
    rho = (d["density"] * (1 + 0.1 * np.random.default_rng(7).random(shp))).astype(dt)
    mu = np.full(shp, d["mu"].flat[0], dt)
    lam = np.full(shp, d["lam"].flat[0], dt)
    isrc, jsrc, ksrc = 41, 17, 11
    ixr, iyr, izr = (np.array(a, np.int32) for a in ([20, 50], [12, 30], [11, 11]))
    with CpmlGpu(3, 4, 8 if prec == "f64" else 4, nx, ny, nz, d["DELTAX"], d["DELTAY"],
                 d["DELTAZ"], d["DELTAT"], npml, nstep) as s:
        s.set_medium(lam, mu, rho)
        s.set_pml(prof["x"], prof["y"], prof["z"])
        s.set_source(isrc, jsrc, ksrc, d["fx"], d["fy"])
        s.set_receivers(ixr, iyr, izr)
        s.run()
        g = {ORACLE_NAME[n]: s.field(n) for n in F3D}
        g.update({oracle_mem_name(m): s.memory(m) for m in MEM3D})
        sis = s.seismograms()
    o = orc.run3d4(nx, ny, nz, d["DELTAX"], d["DELTAY"], d["DELTAZ"], d["DELTAT"], nstep, lam,
                   mu, rho, prof["x"], prof["y"], prof["z"], isrc, jsrc, ksrc, d["fx"], d["fy"],
                   ixr, iyr, izr, prec=prec, nthreads=8)
    for k in g:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    np.testing.assert_array_equal(sis[0], o["sisvx"])
    np.testing.assert_array_equal(sis[1], o["sisvy"])
    assert np.abs(o["sisvy"]).max() > 0
