"""Time-blocked 3D leapfrog (cpml_gpu_set_time_block on a 3D second-order
handle: NT steps per launch, intermediate levels in registers and LDS, state
ping-pong) against the oracle and the single-step kernels.

Bar: bit-exact on the 9 fields, the 18 memory variables and every seismogram
sample, float32 and float64, 1 and 2 steps per launch, on the benchmark's
medium at test size (layered along z, +-5% smoothed perturbation) with
K_MAX_PML 1 and 3, with tile widths and chunk depths that cut the grid
unevenly (tiles of every slab class, including corners), step counts that
leave single steps over, graph replay and eager launches, the stability
check splitting the run, and time blocking switched off part-way.
"""
import numpy as np
import pytest

import layered_media as lm
from case_inputs import inputs, run_oracle
from gpu_runner import F3D, MEM3D, ORACLE_NAME, make_solver, oracle_mem_name, run_gpu

pytestmark = pytest.mark.gpu

F3 = tuple(ORACLE_NAME[n] for n in F3D)


def _run(case, prec, d, nt, ow=0, kc=0, graph=16, chunks=None, stability=None, memory=True):
    with make_solver(case, prec, d, energy=False, graph_steps=graph) as s:
        s.set_time_block(nt, ow, kc)
        if stability:
            s.set_stability(*stability)
        if chunks:
            for a, b in chunks:
                s.run(a, b)
        else:
            s.run()
        g = {ORACLE_NAME[n]: s.field(n) for n in F3D}
        if memory:
            g.update({oracle_mem_name(m): s.memory(m) for m in MEM3D})
        g["sisvx"], g["sisvy"], g["sisvz"] = s.seismograms()
    return g


def _check(g, o, keys=None):
    keys = keys or [k for k in g if k in o]
    for k in keys:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    assert np.abs(g["sisvx"]).max() > 0


@pytest.mark.parametrize("kmax", [1.0, 3.0])
@pytest.mark.parametrize("nt", [1, 2])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_tb3d_layered_matches_oracle_bit_exact(prec, nt, kmax):
    d = lm.layered3d_inputs(prec, 121, kmax=kmax)  # odd count: a single step left over
    o = run_oracle("3d_layered", prec, d=d, nthreads=8, energy=False, derived_lp2m=True)
    g = _run("3d_layered", prec, d, nt)
    _check(g, o, list(F3) + [oracle_mem_name(m) for m in MEM3D] + ["sisvx", "sisvy"])


@pytest.mark.parametrize("ow,kc,graph", [(20, 5, 0), (60, 16, 16), (8, 200, 6), (33, 9, 16)])
def test_tb3d_tiles_chunks_and_launch_paths(ow, kc, graph):
    prec = "f32"
    d = lm.layered3d_inputs(prec, 80, kmax=3.0)
    ref = run_gpu("3d_layered", prec, d=d, energy=False, memory=True)
    g = _run("3d_layered", prec, d, 2, ow=ow, kc=kc, graph=graph)
    _check(g, ref, list(F3) + [oracle_mem_name(m) for m in MEM3D] + ["sisvx", "sisvy", "sisvz"])


@pytest.mark.parametrize("case", ["3d_small", "3d_wide"])
def test_tb3d_reference_cases_match_two_phase(case):
    prec = "f64"
    d = inputs(case, prec, 90)
    ref = run_gpu(case, prec, d=d, energy=False, memory=True)
    g = _run(case, prec, d, 2)
    _check(g, ref, list(F3) + ["sisvx", "sisvy", "sisvz"])


def test_tb3d_uneven_runs_stability_and_switch_off():
    prec = "f32"
    d = lm.layered3d_inputs(prec, 100, kmax=1.0)
    o = run_oracle("3d_layered", prec, d=d, nthreads=8, energy=False, derived_lp2m=True)
    with make_solver("3d_layered", prec, d, energy=False) as s:
        s.set_stability(9, 1e25)
        s.set_time_block(2)
        s.run(1, 4)
        s.run(5, 40)
        s.set_time_block(0)
        s.run(41, 63)
        s.set_time_block(2, 24, 11)
        s.run(64, 100)
        for n in F3D:
            np.testing.assert_array_equal(s.field(n), o[ORACLE_NAME[n]], err_msg=n)
        sx, sy, _ = s.seismograms()
        np.testing.assert_array_equal(sx, o["sisvx"])
        np.testing.assert_array_equal(sy, o["sisvy"])
