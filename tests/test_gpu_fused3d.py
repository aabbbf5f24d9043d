"""The fused 3D step (one kernel per time step, ping-pong state buffers)
against the oracle: bit-exact for every field, memory variable and
seismogram sample, in both precisions, for every tiling and both the
graph-replayed and eager launch paths."""
import numpy as np
import pytest

from case_inputs import inputs, run_oracle
from gpu_runner import make_solver, rel_l2, run_gpu

pytestmark = pytest.mark.gpu

FIELDS = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_fused_matches_oracle_bit_exact(prec):
    case = "3d_small"
    d = inputs(case, prec, 300)
    g = run_gpu(case, prec, d=d, memory=True, fused=True)
    o = run_oracle(case, prec, d=d, nthreads=8, derived_lp2m=True)
    for k in FIELDS:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    for k in [k for k in g if k.startswith("m_")]:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    np.testing.assert_array_equal(g["sisvx"], o["sisvx"])
    np.testing.assert_array_equal(g["sisvy"], o["sisvy"])
    tol = 1e-12 if prec == "f64" else 1e-4
    assert rel_l2(g["ekin"] + g["epot"], o["etotal"]) < tol


@pytest.mark.parametrize("kchunk,tile,graph", [(8, 8, 16), (3, 8, 0), (64, 16, 5), (5, 16, 16)])
def test_fused_tiling_and_launch_paths(kchunk, tile, graph):
    case, prec = "3d_small", "f32"
    d = inputs(case, prec, 90)
    ref = run_gpu(case, prec, d=d, energy=False)
    g = run_gpu(case, prec, d=d, energy=False, kchunk=kchunk, tile=tile, fused=True,
                graph_steps=graph)
    for k in FIELDS + ("sisvx", "sisvy"):
        np.testing.assert_array_equal(g[k], ref[k], err_msg=k)


def test_fused_chunked_runs_and_switch_back():
    """Runs in uneven chunks (odd step counts flip the state buffer), then the
    two-kernel path continues from the fused state."""
    case, prec = "3d_small", "f64"
    d = inputs(case, prec, 120)
    ref = run_gpu(case, prec, d=d, energy=False)
    with make_solver(case, prec, d, energy=False, fused=True) as s:
        s.run(1, 7)
        s.run(8, 51)
        s.set_fused(False)
        s.run(52, 80)
        s.set_fused(True)
        s.run(81, 120)
        for n, k in (("vx", "vx"), ("sigmayz", "syz"), ("vz", "vz")):
            np.testing.assert_array_equal(s.field(n), ref[k])


def test_fused_odd_grid_partial_tiles():
    """A grid that is not a multiple of the 64 x 8 tile (partial tiles in x
    and y, z chunks that do not divide NZ) against the two-kernel path."""
    case, prec = "3d_small", "f64"
    d = inputs(case, prec, 60)
    a = run_gpu(case, prec, d=d, energy=False, kchunk=7)
    b = run_gpu(case, prec, d=d, energy=False, kchunk=7, fused=True, tile=16)
    for k in FIELDS:
        np.testing.assert_array_equal(a[k], b[k])


def test_fused_wide_grid_matches_two_phase():
    """160 x 48 x 40: the two-phase path splits into interior-box and shell
    launches here; the fused step covers the grid in one launch."""
    case, prec = "3d_wide", "f32"
    d = inputs(case, prec, 80)
    a = run_gpu(case, prec, d=d, energy=False)
    b = run_gpu(case, prec, d=d, energy=False, fused=True)
    for k in FIELDS + ("sisvx", "sisvy"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
