"""Batched shots with time blocking (cpml_batch_set_time_block: NT steps of
every shot per launch over a flat (shot, strip, chunk) item list, ping-pong
pools), per-shot energy and the per-shot stability check, against the oracle.

Bars: every shot's fields and seismograms bit-exact against the oracle's
single run of that shot, for orders 2, 4 and 8, float32 and float64, one
wavefront per item and persistent grids, several chunk heights; per-shot
energy 1e-12 (f64) / 1e-4 (f32) rel-L2 against the oracle's energy of the
same shot (summed in another order), also with blocks whose ends fall on the
energy cadence; the stability check names the shot that blew up.
"""
import numpy as np
import pytest

import oracle_lib as orc
from gpu_runner import rel_l2
from test_gpu_batch import NSTEP, check_against_oracle, make_batch, shots_for

pytestmark = pytest.mark.gpu

MAX_NT = {("f32", 2): 4, ("f32", 4): 3, ("f32", 8): 2, ("f64", 2): 2, ("f64", 4): 2,
          ("f64", 8): 2}


@pytest.mark.parametrize("order", [2, 4, 8])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_batch_time_block_matches_oracle(order, prec):
    for nt in sorted({1, MAX_NT[(prec, order)]}):
        with make_batch(order, prec, shots_for(order, prec)) as b:
            b.set_time_block(nt)
            b.run()
            check_against_oracle(b, order, prec)


@pytest.mark.parametrize("bpc,jc,ow,graph", [(2, 64, 0, 16), (0, 17, 24, 0), (4, 200, 8, 5)])
def test_batch_time_block_shapes(bpc, jc, ow, graph):
    order, prec = 4, "f32"
    with make_batch(order, prec, shots_for(order, prec)) as b:
        b.set_tiling(4, bpc, graph)
        b.set_time_block(3, ow, jc)
        b.run(1, 100)
        b.set_time_block(0)  # back to single steps part-way
        b.run(101, 161)
        b.set_time_block(2, 0, 64)
        b.run(162, NSTEP)
        check_against_oracle(b, order, prec)


def _oracle_energy(order, prec, s):
    npml = s["NPOINTS_PML"]
    nx, ny = s["NX"], s["NY"]
    if order == 4:  # 2D4:696
        ew = (npml, nx - npml + 1, npml, ny - npml + 1)
    else:  # 2D2:695
        ew = (npml + 1, nx - npml, npml + 1, ny - npml)
    o = orc.run2d(nx, ny, order, s["DELTAX"], s["DELTAY"], s["DELTAT"], NSTEP, s["lam"], s["mu"],
                  s["rho"], s["prof_x"], s["prof_y"], s["ISOURCE"], s["JSOURCE"], s["fx"],
                  s["fy"], s["ix_rec"], s["iy_rec"], ewin=ew, prec=prec, nthreads=8)
    return o["ekin"], o["epot"]


@pytest.mark.parametrize("order,prec,nt,cadence", [(2, "f64", 0, 1), (4, "f32", 2, 2),
                                                   (2, "f64", 2, 4), (8, "f32", 2, 1)])
def test_batch_energy_per_shot(order, prec, nt, cadence):
    shots = shots_for(order, prec)
    tol = 1e-12 if prec == "f64" else 1e-4
    with make_batch(order, prec, shots) as b:
        b.set_energy(cadence)
        if nt:
            b.set_time_block(nt)
        b.run()
        check_against_oracle(b, order, prec)
        for q, s in enumerate(shots):
            ek, ep = b.energy(q)
            ok, op = _oracle_energy(order, prec, s)
            idx = np.arange(cadence - 1, NSTEP, cadence)
            assert rel_l2(ek[idx], ok[idx]) < tol, (q, "kinetic")
            assert rel_l2(ep[idx], op[idx]) < tol, (q, "potential")
            off = np.setdiff1d(np.arange(NSTEP), idx)
            assert not np.any(ek[off]) and not np.any(ep[off])


@pytest.mark.parametrize("nt", [0, 2])
def test_batch_stability_names_the_shot(nt):
    """A shot without a source stays at rest (max |v| = 0); the other one
    passes a threshold of 1e-30 at step 5, and the error names it."""
    from seismic_cpml_amd import CpmlError
    from seismic_cpml_amd.solver import CPML_ERR_BLOWUP

    order, prec = 2, "f32"
    s = shots_for(order, prec)
    shots = [s[1], s[0]]
    with make_batch(order, prec, shots) as b:
        b.set_source(0, 0, 0, shots[0]["fx"], shots[0]["fy"])
        b.set_stability(100, 1e-30)
        if nt:
            b.set_time_block(nt)
        with pytest.raises(CpmlError) as e:
            b.run()
        assert e.value.code == CPML_ERR_BLOWUP
        assert "shot 1" in e.value.message and "step 5" in e.value.message
        b.set_stability(100, 1e25)
        b.run()  # below the threshold: no error


def test_batch_autotune_explores_time_blocking(tmp_path):
    order, prec = 4, "f32"
    with make_batch(order, prec, shots_for(order, prec)) as b:
        b.autotune(trial_steps=24, cache_path=tmp_path / "t.txt")
        ty, bpc, tb, jc = b.launch_shape()
        assert 0 <= tb <= 3 and jc > 0
        b.run()
        check_against_oracle(b, order, prec)
    line = (tmp_path / "t.txt").read_text().split()
    assert "tb" in line
