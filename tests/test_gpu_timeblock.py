"""Time-blocked 2D leapfrog (cpml_gpu_set_time_block: NT steps per launch,
every intermediate level in registers, state ping-pong) against the oracle.

Bar: bit-exact on every field, memory variable and seismogram sample (the
kernel keeps the single-step arithmetic statement by statement), for orders
2, 4 and 8, float32 and float64, every step count the library builds, on
heterogeneous media (the layered cases with a smoothed perturbation), with
strip widths and chunk heights that cut the grid unevenly, step counts that
leave single steps over, graph replay and eager launches, the stability
check splitting the run, and time blocking switched off part-way.
"""
import numpy as np
import pytest

import layered_media as lm
import oracle_lib as orc
from case_inputs import inputs
from golden.cases import CASES
from gpu_runner import MEM2D, make_solver, oracle_mem_name

pytestmark = pytest.mark.gpu

FIELDS = (("vx", "vx"), ("vy", "vy"), ("sigmaxx", "sxx"), ("sigmayy", "syy"), ("sigmaxy", "sxy"))
MAX_NT = {("f32", 2): 4, ("f32", 4): 3, ("f32", 8): 2, ("f64", 2): 2, ("f64", 4): 2,
          ("f64", 8): 2}


def _case(order):
    return "2d4_layered" if order == 4 else "2d2_layered"


def _inputs(order, prec, nstep):
    d = inputs(_case(order), prec, nstep)
    for q, k in enumerate(("lam", "mu", "rho")):
        d[k] = lm.perturbed(d[k], 40 + q)
    return d


def _oracle(order, prec, d):
    c = CASES[_case(order)]
    return orc.run2d(c["NX"], c["NY"], order, d["DELTAX"], d["DELTAY"], d["DELTAT"], d["nstep"],
                     d["lam"], d["mu"], d["rho"], d["prof_x"], d["prof_y"], d["ISOURCE"],
                     d["JSOURCE"], d["fx"], d["fy"], d["ix_rec"], d["iy_rec"], prec=prec,
                     nthreads=8)


def _run(order, prec, d, nt, ow=0, jc=0, graph=16, chunks=None, stability=None):
    c = CASES[_case(order)]
    saved = c["order"]
    c["order"] = order
    try:
        with make_solver(_case(order), prec, d, energy=False, graph_steps=graph) as s:
            s.set_time_block(nt, ow, jc)
            if stability:
                s.set_stability(*stability)
            if chunks:
                for a, b in chunks:
                    s.run(a, b)
            else:
                s.run()
            g = {k: s.field(n) for n, k in FIELDS}
            g.update({oracle_mem_name(m): s.memory(m) for m in MEM2D})
            g["sisvx"], g["sisvy"] = s.seismograms()[:2]
    finally:
        c["order"] = saved
    return g


def _check(g, o):
    for k in [k for _, k in FIELDS] + [oracle_mem_name(m) for m in MEM2D] + ["sisvx", "sisvy"]:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    assert np.abs(g["sisvx"]).max() > 0


CASES_NT = [(o, p, nt) for o in (2, 4, 8) for p in ("f32", "f64")
            for nt in range(1, MAX_NT[(p, o)] + 1)]


@pytest.mark.parametrize("order,prec,nt", CASES_NT)
def test_time_block_matches_oracle_bit_exact(order, prec, nt):
    d = _inputs(order, prec, 301)  # 301 steps: single steps left over for nt > 1
    o = _oracle(order, prec, d)
    g = _run(order, prec, d, nt)
    _check(g, o)


@pytest.mark.parametrize("ow,jc,graph", [(8, 7, 0), (17, 33, 16), (0, 1000, 5), (24, 16, 16)])
@pytest.mark.parametrize("order", [2, 4])
def test_time_block_strips_chunks_and_launch_paths(order, ow, jc, graph):
    prec = "f32"
    nt = 2
    d = _inputs(order, prec, 200)
    o = _oracle(order, prec, d)
    g = _run(order, prec, d, nt, ow=ow, jc=jc, graph=graph)
    _check(g, o)


def test_time_block_uneven_runs_and_stability_checks():
    """run() in uneven pieces and a stability check every 7 steps (and at
    step 5): every block ends at a check, the rest goes by single steps."""
    order, prec = 4, "f32"
    d = _inputs(order, prec, 180)
    o = _oracle(order, prec, d)
    g = _run(order, prec, d, 3, chunks=((1, 10), (11, 11), (12, 95), (96, 180)),
             stability=(7, 1e25))
    _check(g, o)


def test_time_block_switched_off_midway():
    order, prec = 2, "f64"
    d = _inputs(order, prec, 160)
    o = _oracle(order, prec, d)
    c = CASES[_case(order)]
    with make_solver(_case(order), prec, d, energy=False) as s:
        s.set_time_block(2)
        s.run(1, 51)
        s.set_time_block(0)
        s.run(52, 100)
        s.set_time_block(2, 16, 40)
        s.run(101, 160)
        for n, k in FIELDS:
            np.testing.assert_array_equal(s.field(n), o[k], err_msg=n)
        np.testing.assert_array_equal(s.seismograms()[0], o["sisvx"])
    assert c["order"] == 2


def test_time_block_shipped_grid_and_timed_path():
    """2D2 as shipped (101 x 641, float64): a tall grid cut into many chunks;
    the per-kernel timing path launches the same blocks."""
    case, prec = "2d2_shipped", "f64"
    d = inputs(case, prec, 400)
    c = CASES[case]
    o = orc.run2d(c["NX"], c["NY"], 2, d["DELTAX"], d["DELTAY"], d["DELTAT"], d["nstep"],
                  d["lam"], d["mu"], d["rho"], d["prof_x"], d["prof_y"], d["ISOURCE"],
                  d["JSOURCE"], d["fx"], d["fy"], d["ix_rec"], d["iy_rec"], prec=prec,
                  nthreads=8)
    for timing in (False, True):
        with make_solver(case, prec, d, energy=False) as s:
            s.set_time_block(2, 0, 64)
            if timing:
                s.set_kernel_timing(True)
            s.run()
            for n, k in FIELDS:
                np.testing.assert_array_equal(s.field(n), o[k], err_msg=n)
            np.testing.assert_array_equal(s.seismograms()[1], o["sisvy"])
            if timing:
                n0, t0 = s.kernel_time(0)
                assert n0 == 200 and t0 > 0
