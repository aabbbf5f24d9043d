"""Build the inputs of a golden case (profiles, source, receivers, medium) with
the oracle's L2 routines, and run the oracle on it.

Used by the parity tests: the same inputs go to the oracle and to the HIP
path, and the oracle's outputs are checked against the reference fixtures.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE / "golden"))

import oracle_lib as orc  # noqa: E402
from cases import ALL_CASES, values  # noqa: E402

GOLDEN = HERE / "golden"


def fixture(case: str, prec: str):
    return np.load(GOLDEN / f"{case}_{prec}.npz")


def inputs(case: str, prec: str, nstep: int | None = None) -> dict:
    """All inputs of ``case`` in precision ``prec``, built the reference's way."""
    v = values(case, prec)
    c = ALL_CASES[case]
    dt = np.float64 if prec == "f64" else np.float32
    nx, ny, npml = c["NX"], c["NY"], c["NPML"]
    nz = c.get("NZ")
    nstep = c["NSTEP"] if nstep is None else nstep
    d = dict(v)
    d["nstep"] = nstep
    d["dtype"] = dt
    # profiles: 2D2:283-465 (x clamps alpha, y does not; 2D4 y origin quirk)
    d["prof_x"] = orc.profile(nx, v["DELTAX"], v["DELTAT"], npml, v["cp"], v["f0"], prec,
                              clamp_alpha=True)
    d["prof_y"] = orc.profile(ny, v["DELTAY"], v["DELTAT"], npml, v["cp"], v["f0"], prec,
                              origin_max_uses_n=c.get("yorigin_n", False), clamp_alpha=False)
    if nz is not None:
        d["prof_z"] = orc.profile(nz, v["DELTAZ"], v["DELTAT"], npml, v["cp"], v["f0"], prec,
                                  clamp_alpha=False)
    if nz is not None and c["order"] == 4:
        return _inputs_3d4(case, prec, d, c)
    kind = 2 if c["src"] == "ricker" else 1
    d["fx"], d["fy"] = orc.stf(nstep, v["DELTAT"], v["f0"], v["t0"], v["factor"], v["ANGLE"],
                               kind, prec)
    d["ix_rec"], d["iy_rec"] = orc.receivers(nx, ny, v["DELTAX"], v["DELTAY"], c["NREC"],
                                             v["xdeb"], v["ydeb"], v["xfin"], v["yfin"], prec)
    if nz is None:
        if c.get("model"):
            f = fixture(case, prec)
            d["rho"], d["mu"], d["lam"] = f["rho"], f["mu"], f["lam"]
        else:
            d["rho"] = np.full((ny, nx), v["density"], dt)
            d["mu"] = np.full((ny, nx), v["mu"], dt)
            d["lam"] = np.full((ny, nx), v["lambda"], dt)
        if c["order"] == 4:  # 2D4:696-704
            d["ewin"] = (npml, nx - npml + 1, npml, ny - npml + 1)
        else:  # 2D2:695-703
            d["ewin"] = (npml + 1, nx - npml, npml + 1, ny - npml)
    else:
        shp = (nz, ny, nx)
        d["rho"] = np.full(shp, v["density"], dt)
        d["mu"] = np.full(shp, v["mu"], dt)
        d["lam"] = np.full(shp, v["lambda"], dt)
        d["lp2m"] = np.full(shp, v["lambdaplustwomu"], dt)
        d["ksrc"] = nz // 2  # 3D:1074-1081 (cut plane, global k = NZ/2)
        d["iz_rec"] = np.full(c["NREC"], nz // 2, np.int32)
        d["ewin"] = (npml + 1, nx - npml, npml + 1, ny - npml, npml + 1, nz - npml)
    return d


def _inputs_3d4(case: str, prec: str, d: dict, c: dict) -> dict:
    """3D 4th order: the reference's own profiles and receiver indices (float64
    fixture, rounded for float32), source at (ISOURCE, JSOURCE, NZ/2) and
    receivers on that plane (3D-visco:1350-1352, 1374-1377), energy window
    NPOINTS_PML..N-NPOINTS_PML+1 on every axis (3D-visco:1383-1388)."""
    f = fixture(case, "f64")
    dt = d["dtype"]
    nx, ny, nz, npml = c["NX"], c["NY"], c["NZ"], c["NPML"]
    for k in ("prof_x", "prof_y", "prof_z"):
        d[k] = f[k].astype(dt)
    d["fx"], d["fy"] = orc.stf(d["nstep"], d["DELTAT"], d["f0"], d["t0"], d["factor"],
                               d["ANGLE"], 1, prec)
    d["ix_rec"], d["iy_rec"] = f["ix_rec"].astype(np.int32), f["iy_rec"].astype(np.int32)
    shp = (nz, ny, nx)
    d["rho"] = np.full(shp, d["density"], dt)
    d["mu"] = np.full(shp, d["mu"], dt)
    d["lam"] = np.full(shp, d["lambda"], dt)
    d["ksrc"] = nz // 2
    d["iz_rec"] = np.full(c["NREC"], nz // 2, np.int32)
    d["ewin"] = (npml, nx - npml + 1, npml, ny - npml + 1, npml, nz - npml + 1)
    return d


def run_oracle(case: str, prec: str, nstep: int | None = None, d: dict | None = None,
               nthreads: int = 0, energy=True, derived_lp2m: bool = False,
               zcut: int = 0) -> dict:
    """derived_lp2m: 3D only -- use DELTAT*lambda + 2*(DELTAT*mu) for the
    (lambda+2mu) coefficient, as the HIP kernels do, instead of the reference's
    DELTAT*(rho*cp*cp) (3D:145, 298).  zcut: 3D 4th order only -- reproduce
    the reference's slab-interface reads after plane zcut (0: single domain,
    what the HIP kernels compute)."""
    d = inputs(case, prec, nstep) if d is None else d
    c = ALL_CASES[case]
    if c.get("NZ") is not None and c["order"] == 4:
        return orc.run3d4(c["NX"], c["NY"], c["NZ"], d["DELTAX"], d["DELTAY"], d["DELTAZ"],
                          d["DELTAT"], d["nstep"], d["lam"], d["mu"], d["rho"], d["prof_x"],
                          d["prof_y"], d["prof_z"], d["ISOURCE"], d["JSOURCE"], d["ksrc"],
                          d["fx"], d["fy"], d["ix_rec"], d["iy_rec"], d["iz_rec"],
                          ewin=d["ewin"] if energy else None, energy_slabs=c["nproc"],
                          prec=prec, zcut=zcut, nthreads=nthreads)
    if c.get("NZ") is None:
        return orc.run2d(c["NX"], c["NY"], c["order"], d["DELTAX"], d["DELTAY"], d["DELTAT"],
                         d["nstep"], d["lam"], d["mu"], d["rho"], d["prof_x"], d["prof_y"],
                         d["ISOURCE"], d["JSOURCE"], d["fx"], d["fy"], d["ix_rec"], d["iy_rec"],
                         ewin=d["ewin"] if energy else None, prec=prec, nthreads=nthreads)
    return orc.run3d(c["NX"], c["NY"], c["NZ"], d["DELTAX"], d["DELTAY"], d["DELTAZ"],
                     d["DELTAT"], d["nstep"], d["lam"], d["mu"], d["rho"],
                     None if derived_lp2m else d["lp2m"],
                     d["prof_x"], d["prof_y"], d["prof_z"], d["ISOURCE"], d["JSOURCE"],
                     d["ksrc"], d["fx"], d["fy"], d["ix_rec"], d["iy_rec"], d["iz_rec"],
                     ewin=d["ewin"] if energy else None, energy_slabs=c.get("nproc", 1),
                     prec=prec, nthreads=nthreads)


def visco2d_inputs(prec: str, order: int = 2):
    """Inputs of the reduced 2D viscoelastic case (2D-visco2:149-315 with the
    oracle/build_ref.py ref2dve_f64 overrides; order 4: the same case of the
    4th-order program, ref2dve4_f64) and its fixture: profiles, tau
    values, medium and receivers are the reference's own (float64 fixture,
    rounded for float32); force table = Ricker (2D-visco2:895) divided by
    DELTAX*DELTAY (2D-visco2:899), angle 0 (force along y only)."""
    f = np.load(GOLDEN / ("2dve_small_f64.npz" if order == 2 else "2dve4_small_f64.npz"))
    dt = np.float64 if prec == "f64" else np.float32
    R = dt
    d = dict(NX=161, NY=161, nstep=700, NPML=10, dtype=dt, order=order)
    d["DELTAX"] = d["DELTAY"] = R(1.5)
    d["DELTAT"] = R(2.2e-4) if prec == "f64" else np.float32(2.2e-4)
    f0 = R(35.0)
    t0 = R(1.20) / f0
    fx, fy = orc.stf(d["nstep"], d["DELTAT"], f0, t0, R(1.0), R(0.0), 2, prec)
    area = d["DELTAX"] * d["DELTAY"]
    d["fx"], d["fy"] = (fx / area).astype(dt), (fy / area).astype(dt)
    d["ISOURCE"] = d["JSOURCE"] = int(120.0 / 1.5 + 1)
    for k in ("prof_x", "prof_y", "lam", "mu", "rho"):
        d[k] = f[k].astype(dt)
    d["taus"] = tuple(f[k].astype(dt) for k in ("tau_eps_nu1", "tau_sig_nu1", "tau_eps_nu2",
                                                 "tau_sig_nu2"))
    d["ix_rec"], d["iy_rec"] = f["ix_rec"].astype(np.int32), f["iy_rec"].astype(np.int32)
    return d, f
