"""Batched multi-shot 2D path (cpml_batch_*) against the oracle.

Each batch mixes shots of different grid sizes (one not a multiple of the
tile in either direction), spacings, time steps, media, sources and receiver
sets.  Every shot of a batch must equal the oracle's single run of that shot,
bit for bit (fields and seismograms), for every order and precision and for
every launch shape (tile rows, persistent grid, graph length), including runs
split into chunks and runs after the autotuner.
"""
import numpy as np
import pytest

import oracle_lib as orc
from case_inputs import inputs
from golden.cases import CASES, values
from seismic_cpml_amd import CpmlBatch, CpmlError, solver

pytestmark = pytest.mark.gpu

NSTEP = 300
FIELD_KEYS = (("vx", "vx"), ("vy", "vy"), ("sigmaxx", "sxx"), ("sigmayy", "syy"),
              ("sigmaxy", "sxy"))


def _shot_from_case(case, prec, nstep):
    d = inputs(case, prec, nstep)
    c = CASES[case]
    return dict(NX=c["NX"], NY=c["NY"], DELTAX=float(d["DELTAX"]), DELTAY=float(d["DELTAY"]),
                DELTAT=float(d["DELTAT"]), NPOINTS_PML=c["NPML"], lam=d["lam"], mu=d["mu"],
                rho=d["rho"], prof_x=d["prof_x"], prof_y=d["prof_y"], ISOURCE=d["ISOURCE"],
                JSOURCE=d["JSOURCE"], fx=d["fx"], fy=d["fy"], ix_rec=d["ix_rec"],
                iy_rec=d["iy_rec"])


def shots_for(order, prec, nstep=NSTEP):
    """Four shots: a layered model, the same with the source moved, the
    shipped homogeneous model (other size and time step), and a 70 x 50 crop
    of the layered model with its own profiles and receivers."""
    lay = "2d4_layered" if order == 4 else "2d2_layered"
    shp = "2d4_shipped" if order == 4 else "2d2_shipped"
    s0 = _shot_from_case(lay, prec, nstep)
    s1 = dict(s0, ISOURCE=30, JSOURCE=70)
    s2 = _shot_from_case(shp, prec, nstep)
    if order == 8:
        # the shipped DELTAT is beyond the 8th-order operator's stability
        # limit (Courant 0.93 times Holberg's sum |c_k| = 1.36)
        s2["DELTAT"] = s2["DELTAT"] / 2
    v = values(lay, prec)
    nx, ny, npml = 70, 50, 8
    s3 = dict(s0, NX=nx, NY=ny, NPOINTS_PML=npml,
              lam=np.ascontiguousarray(s0["lam"][:ny, :nx]),
              mu=np.ascontiguousarray(s0["mu"][:ny, :nx]),
              rho=np.ascontiguousarray(s0["rho"][:ny, :nx]),
              prof_x=orc.profile(nx, v["DELTAX"], v["DELTAT"], npml, v["cp"], v["f0"], prec,
                                 clamp_alpha=True),
              prof_y=orc.profile(ny, v["DELTAY"], v["DELTAT"], npml, v["cp"], v["f0"], prec,
                                 clamp_alpha=False),
              ISOURCE=35, JSOURCE=25, ix_rec=np.array([10, 40, 70], np.int32),
              iy_rec=np.array([10, 30, 50], np.int32))
    return [s0, s1, s2, s3]


_ORACLE = {}


def oracle_results(order, prec):
    key = (order, prec)
    if key not in _ORACLE:
        out = []
        for s in shots_for(order, prec):
            out.append(orc.run2d(s["NX"], s["NY"], order, s["DELTAX"], s["DELTAY"], s["DELTAT"],
                                 NSTEP, s["lam"], s["mu"], s["rho"], s["prof_x"], s["prof_y"],
                                 s["ISOURCE"], s["JSOURCE"], s["fx"], s["fy"], s["ix_rec"],
                                 s["iy_rec"], prec=prec, nthreads=8))
        _ORACLE[key] = out
    return _ORACLE[key]


def make_batch(order, prec, shots, nstep=NSTEP):
    b = CpmlBatch(order, 8 if prec == "f64" else 4, shots, nstep)
    for q, s in enumerate(shots):
        b.set_medium(q, s["lam"], s["mu"], s["rho"])
        b.set_pml(q, s["prof_x"], s["prof_y"])
        b.set_source(q, s["ISOURCE"], s["JSOURCE"], s["fx"], s["fy"])
        b.set_receivers(q, s["ix_rec"], s["iy_rec"])
    return b


def check_against_oracle(b, order, prec):
    for q, o in enumerate(oracle_results(order, prec)):
        for name, k in FIELD_KEYS:
            np.testing.assert_array_equal(b.field(q, name), o[k], err_msg=f"shot {q} {name}")
        sx, sy = b.seismograms(q)
        np.testing.assert_array_equal(sx, o["sisvx"], err_msg=f"shot {q} sisvx")
        np.testing.assert_array_equal(sy, o["sisvy"], err_msg=f"shot {q} sisvy")


@pytest.mark.parametrize("order", [2, 4, 8])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_batch_matches_oracle_bit_exact(order, prec):
    with make_batch(order, prec, shots_for(order, prec)) as b:
        b.run()
        check_against_oracle(b, order, prec)


@pytest.mark.parametrize("tile_y,bpc,graph", [(8, 0, 16), (16, 0, 0), (4, 1, 7), (8, 4, 5),
                                              (16, 8, 32), (32, 0, 16), (64, 2, 3)])
def test_batch_launch_shapes_identical(tile_y, bpc, graph):
    order, prec = 4, "f32"
    with make_batch(order, prec, shots_for(order, prec)) as b:
        b.set_tiling(tile_y, bpc, graph)
        b.run()
        check_against_oracle(b, order, prec)


def test_batch_run_in_chunks():
    order, prec = 2, "f64"
    with make_batch(order, prec, shots_for(order, prec)) as b:
        b.run(1, 137)
        b.set_tiling(8, 2, 16)  # the launch shape may change between chunks
        b.run(138, NSTEP)
        check_against_oracle(b, order, prec)


def test_batch_autotune_resets_state_and_caches(tmp_path):
    order, prec = 4, "f32"
    cache = tmp_path / "tune.txt"
    shots = shots_for(order, prec)
    with make_batch(order, prec, shots) as b:
        ty, bpc, ms = b.autotune(trial_steps=12, cache_path=cache)
        assert ty in (4, 8, 16, 32, 64) and 0 <= bpc <= 64 and ms > 0
        b.run()
        check_against_oracle(b, order, prec)
    assert cache.read_text().strip()
    with make_batch(order, prec, shots) as b:
        ty2, bpc2, ms2 = b.autotune(trial_steps=12, cache_path=cache)
        assert (ty2, bpc2, ms2) == (ty, bpc, -1.0)
        b.run()
        check_against_oracle(b, order, prec)


def test_batch_shot_without_source_stays_at_rest():
    order, prec = 2, "f32"
    shots = shots_for(order, prec)[:2]
    with make_batch(order, prec, shots) as b:
        b.set_source(1, 0, 0, shots[1]["fx"], shots[1]["fy"])
        b.run()
        for name, _ in FIELD_KEYS:
            assert not b.field(1, name).any()
        o = oracle_results(order, prec)[0]
        np.testing.assert_array_equal(b.field(0, "vx"), o["vx"])


def test_batch_errors():
    order, prec = 2, "f64"
    shots = shots_for(order, prec)[:2]
    with CpmlBatch(order, 8, shots, NSTEP) as b:
        s = shots[0]
        with pytest.raises(CpmlError) as e:
            b.run()
        assert e.value.code == solver.CPML_ERR_STATE
        with pytest.raises(CpmlError) as e:
            b.set_source(0, s["NX"], 3, s["fx"], s["fy"])
        assert e.value.code == solver.CPML_ERR_ARG
        with pytest.raises(CpmlError) as e:
            b.set_receivers(1, [1, s["NX"] + 1], [1, 1])
        assert e.value.code == solver.CPML_ERR_ARG
        with pytest.raises(CpmlError) as e:
            b.set_medium(0, s["lam"] * 100, s["mu"] * 100, s["rho"])
        assert e.value.code == solver.CPML_ERR_CFL
        assert "unstable" in e.value.message
        with pytest.raises(CpmlError) as e:
            b.set_tiling(5, 0, 16)
        assert e.value.code == solver.CPML_ERR_ARG
        with pytest.raises(CpmlError):
            b.field(2, "vx")
