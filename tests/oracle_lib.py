"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: tests, the smoke check and the benchmark's CPU
baseline use it; the product path never does.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
LIB_PATH = ROOT / "oracle" / "liboracle.so"

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.is_file():
            subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
        _lib = C.CDLL(str(LIB_PATH))
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _real(prec):
    return (np.float64, C.c_double, "d") if prec == "f64" else (np.float32, C.c_float, "f")


def profile(n, delta, deltat, npml, cp, f0, prec="f64", use_min=True, use_max=True,
            origin_max_uses_n=False, clamp_alpha=True):
    """Returns (6, n): a, b, K, a_half, b_half, K_half."""
    dt, cr, s = _real(prec)
    out = np.zeros((6, n), dt)
    f = getattr(lib(), f"oracle_profile_{s}")
    f(C.c_int(n), cr(delta), cr(deltat), C.c_int(npml), cr(cp), cr(f0), C.c_int(use_min),
      C.c_int(use_max), C.c_int(origin_max_uses_n), C.c_int(clamp_alpha),
      *[_ptr(out[q]) for q in range(6)])
    return out


def stf(nstep, deltat, f0, t0, factor, angle, kind, prec="f64"):
    """kind: 0 Gaussian, 1 first derivative of Gaussian, 2 Ricker. Returns fx, fy."""
    dt, cr, s = _real(prec)
    fx = np.zeros(nstep, dt)
    fy = np.zeros(nstep, dt)
    getattr(lib(), f"oracle_stf_{s}")(C.c_int(nstep), cr(deltat), cr(f0), cr(t0), cr(factor),
                                      cr(angle), C.c_int(kind), _ptr(fx), _ptr(fy))
    return fx, fy


def receivers(nx, ny, dx, dy, nrec, xdeb, ydeb, xfin, yfin, prec="f64"):
    dt, cr, s = _real(prec)
    ix = np.zeros(nrec, np.int32)
    iy = np.zeros(nrec, np.int32)
    getattr(lib(), f"oracle_receivers_{s}")(C.c_int(nx), C.c_int(ny), cr(dx), cr(dy),
                                            C.c_int(nrec), cr(xdeb), cr(ydeb), cr(xfin),
                                            cr(yfin), _ptr(ix), _ptr(iy))
    return ix, iy


STATE2D = ("vx", "vy", "sxx", "syy", "sxy", "m_dvx_dx", "m_dvx_dy", "m_dvy_dx", "m_dvy_dy",
           "m_dsxx_dx", "m_dsyy_dy", "m_dsxy_dx", "m_dsxy_dy")
STATE3D = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz",
           "m_dvx_dx", "m_dvx_dy", "m_dvx_dz", "m_dvy_dx", "m_dvy_dy", "m_dvy_dz",
           "m_dvz_dx", "m_dvz_dy", "m_dvz_dz", "m_dsxx_dx", "m_dsyy_dy", "m_dszz_dz",
           "m_dsxy_dx", "m_dsxy_dy", "m_dsxz_dx", "m_dsxz_dz", "m_dsyz_dy", "m_dsyz_dz")


def run2d(nx, ny, order, dx, dy, deltat, nstep, lam, mu, rho, prof_x, prof_y, isrc, jsrc, fx, fy,
          ixr, iyr, ewin=None, prec="f64", state=None, it_begin=1, it_end=None, nthreads=0):
    """Run the 2D loop; arrays (ny, nx).  Returns dict with fields, memory,
    seismograms (nrec, nstep) and energies."""
    dt, cr, s = _real(prec)
    it_end = nstep if it_end is None else it_end
    nrec = len(ixr)
    if state is None:
        state = np.zeros((13, ny, nx), dt)
    state = np.ascontiguousarray(state, dt)
    sisvx = np.zeros((max(nrec, 1), nstep), dt)
    sisvy = np.zeros((max(nrec, 1), nstep), dt)
    ekin = np.zeros(nstep, dt)
    epot = np.zeros(nstep, dt)
    ew = None if ewin is None else np.asarray(ewin, np.int32)
    args = [np.ascontiguousarray(a, dt) for a in (lam, mu, rho, prof_x, prof_y, fx, fy)]
    ixr = np.ascontiguousarray(ixr, np.int32)
    iyr = np.ascontiguousarray(iyr, np.int32)
    rc = getattr(lib(), f"oracle_run2d_{s}")(
        C.c_int(nx), C.c_int(ny), C.c_int(order), cr(dx), cr(dy), cr(deltat),
        C.c_int(it_begin), C.c_int(it_end), _ptr(args[0]), _ptr(args[1]), _ptr(args[2]),
        _ptr(args[3]), _ptr(args[4]), C.c_int(isrc), C.c_int(jsrc), _ptr(args[5]),
        _ptr(args[6]), C.c_int(nrec), _ptr(ixr), _ptr(iyr), C.c_int(nstep), _ptr(sisvx),
        _ptr(sisvy), _ptr(ew), _ptr(ekin), _ptr(epot), _ptr(state), C.c_int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle_run2d returned {rc}")
    out = {k: state[q] for q, k in enumerate(STATE2D)}
    out.update(state=state, sisvx=sisvx[:nrec], sisvy=sisvy[:nrec], ekin=ekin, epot=epot)
    return out


def run3d(nx, ny, nz, dx, dy, dz, deltat, nstep, lam, mu, rho, lp2m, prof_x, prof_y, prof_z,
          isrc, jsrc, ksrc, fx, fy, ixr, iyr, izr, ewin=None, energy_slabs=1, prec="f64",
          state=None, it_begin=1, it_end=None, nthreads=0, want_vz=True):
    """Run the 3D loop; arrays (nz, ny, nx)."""
    dt, cr, s = _real(prec)
    it_end = nstep if it_end is None else it_end
    nrec = len(ixr)
    if state is None:
        state = np.zeros((27, nz, ny, nx), dt)
    state = np.ascontiguousarray(state, dt)
    sis = [np.zeros((max(nrec, 1), nstep), dt) for _ in range(3)]
    etot = np.zeros(nstep, dt)
    ew = None if ewin is None else np.asarray(ewin, np.int32)
    lam, mu, rho = (np.ascontiguousarray(a, dt) for a in (lam, mu, rho))
    lp = None if lp2m is None else np.ascontiguousarray(lp2m, dt)
    px, py, pz, fx, fy = (np.ascontiguousarray(a, dt) for a in (prof_x, prof_y, prof_z, fx, fy))
    ixr, iyr, izr = (np.ascontiguousarray(a, np.int32) for a in (ixr, iyr, izr))
    rc = getattr(lib(), f"oracle_run3d_{s}")(
        C.c_int(nx), C.c_int(ny), C.c_int(nz), cr(dx), cr(dy), cr(dz), cr(deltat),
        C.c_int(it_begin), C.c_int(it_end), _ptr(lam), _ptr(mu), _ptr(rho), _ptr(lp),
        _ptr(px), _ptr(py), _ptr(pz), C.c_int(isrc), C.c_int(jsrc), C.c_int(ksrc), _ptr(fx),
        _ptr(fy), C.c_int(nrec), _ptr(ixr), _ptr(iyr), _ptr(izr), C.c_int(nstep),
        _ptr(sis[0]), _ptr(sis[1]), _ptr(sis[2]) if want_vz else None, _ptr(ew),
        C.c_int(energy_slabs), _ptr(etot), _ptr(state), C.c_int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle_run3d returned {rc}")
    out = {k: state[q] for q, k in enumerate(STATE3D)}
    out.update(state=state, sisvx=sis[0][:nrec], sisvy=sis[1][:nrec], sisvz=sis[2][:nrec],
               etotal=etot)
    return out


def run3d4(nx, ny, nz, dx, dy, dz, deltat, nstep, lam, mu, rho, prof_x, prof_y, prof_z,
           isrc, jsrc, ksrc, fx, fy, ixr, iyr, izr, ewin=None, energy_slabs=1, prec="f64",
           zcut=0, state=None, it_begin=1, it_end=None, nthreads=0):
    """Run the 4th-order 3D loop (elastic part of the 3D viscoelastic program);
    arrays (nz, ny, nx).  zcut > 0 reproduces the reference's 2-rank halo
    reads at the slab interface after plane zcut."""
    dt, cr, s = _real(prec)
    it_end = nstep if it_end is None else it_end
    nrec = len(ixr)
    if state is None:
        state = np.zeros((27, nz, ny, nx), dt)
    state = np.ascontiguousarray(state, dt)
    sis = [np.zeros((max(nrec, 1), nstep), dt) for _ in range(2)]
    etot = np.zeros(nstep, dt)
    ew = None if ewin is None else np.asarray(ewin, np.int32)
    lam, mu, rho = (np.ascontiguousarray(a, dt) for a in (lam, mu, rho))
    px, py, pz, fx, fy = (np.ascontiguousarray(a, dt) for a in (prof_x, prof_y, prof_z, fx, fy))
    ixr, iyr, izr = (np.ascontiguousarray(a, np.int32) for a in (ixr, iyr, izr))
    rc = getattr(lib(), f"oracle_run3d4_{s}")(
        C.c_int(nx), C.c_int(ny), C.c_int(nz), cr(dx), cr(dy), cr(dz), cr(deltat),
        C.c_int(it_begin), C.c_int(it_end), _ptr(lam), _ptr(mu), _ptr(rho),
        _ptr(px), _ptr(py), _ptr(pz), C.c_int(isrc), C.c_int(jsrc), C.c_int(ksrc), _ptr(fx),
        _ptr(fy), C.c_int(nrec), _ptr(ixr), _ptr(iyr), _ptr(izr), C.c_int(nstep),
        _ptr(sis[0]), _ptr(sis[1]), _ptr(ew), C.c_int(energy_slabs), _ptr(etot),
        C.c_int(zcut), _ptr(state), C.c_int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle_run3d4 returned {rc}")
    out = {k: state[q] for q, k in enumerate(STATE3D)}
    out.update(state=state, sisvx=sis[0][:nrec], sisvy=sis[1][:nrec], etotal=etot)
    return out


STATE2DVE = STATE2D_BASE = ("vx", "vy", "sxx", "syy", "sxy", "m_dvx_dx", "m_dvx_dy", "m_dvy_dx",
                            "m_dvy_dy", "m_dsxx_dx", "m_dsyy_dy", "m_dsxy_dx", "m_dsxy_dy")


def run2dve(nx, ny, dx, dy, deltat, nstep, lam, mu, rho, prof_x, prof_y, taus, isrc, jsrc,
            fx, fy, ixr, iyr, prec="f64", state=None, it_begin=1, it_end=None, nthreads=0,
            order=2):
    """2D viscoelastic C-PML loop (2nd order); arrays (ny, nx).  taus: tau_epsilon
    and tau_sigma of nu1 and nu2, each (nsls,).  Returns fields, C-PML memories,
    r_e1 / r_e11 / r_e13 (nsls, ny, nx) and sisvx, sisvy, sispressure."""
    dt, cr, s = _real(prec)
    it_end = nstep if it_end is None else it_end
    nrec = len(ixr)
    te1, ts1, te2, ts2 = (np.ascontiguousarray(a, dt) for a in taus)
    nsls = len(te1)
    if state is None:
        state = np.zeros((13 + 3 * nsls, ny, nx), dt)
    state = np.ascontiguousarray(state, dt)
    sis = [np.zeros((max(nrec, 1), nstep), dt) for _ in range(3)]
    lam, mu, rho = (np.ascontiguousarray(a, dt) for a in (lam, mu, rho))
    px, py, fx, fy = (np.ascontiguousarray(a, dt) for a in (prof_x, prof_y, fx, fy))
    ixr, iyr = (np.ascontiguousarray(a, np.int32) for a in (ixr, iyr))
    rc = getattr(lib(), f"oracle_run2dve_{s}")(
        C.c_int(nx), C.c_int(ny), C.c_int(order), cr(dx), cr(dy), cr(deltat),
        C.c_int(it_begin), C.c_int(it_end), _ptr(lam), _ptr(mu), _ptr(rho), _ptr(px), _ptr(py),
        C.c_int(nsls),
        _ptr(te1), _ptr(ts1), _ptr(te2), _ptr(ts2), C.c_int(isrc), C.c_int(jsrc), _ptr(fx),
        _ptr(fy), C.c_int(nrec), _ptr(ixr), _ptr(iyr), C.c_int(nstep), _ptr(sis[0]),
        _ptr(sis[1]), _ptr(sis[2]), _ptr(state), C.c_int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle_run2dve returned {rc}")
    out = {k: state[q] for q, k in enumerate(STATE2DVE)}
    out["r_e1"] = state[13:13 + nsls]
    out["r_e11"] = state[13 + nsls:13 + 2 * nsls]
    out["r_e13"] = state[13 + 2 * nsls:13 + 3 * nsls]
    out.update(state=state, sisvx=sis[0][:nrec], sisvy=sis[1][:nrec],
               sispressure=sis[2][:nrec])
    return out


def runade8(nx, ny, dx, dy, deltat, nstep, npml, lam, mu, rho, prof, isrc, jsrc, f0, t0,
            factor, angle, ixr, iyr, prec="f64", state=None, nthreads=0):
    """ADE-PML RK4 8th-order loop (seismic_ADEPML_2D_elastic_RK4_eighth_order.f90).
    Arrays keep the program's 4-cell ghost border: (ny+9, nx+9).  prof: dict
    with a_x, b_x, ah_x, bh_x (4, nx+9), K_x, Kh_x (nx+9) and the same for y."""
    dt, cr, s = _real(prec)
    nrec = len(ixr)
    shp = (ny + 9, nx + 9)
    if state is None:
        state = np.zeros((13,) + shp, dt)
    state = np.ascontiguousarray(state, dt)
    sis = [np.zeros((max(nrec, 1), nstep), dt) for _ in range(2)]
    lam, mu, rho = (np.ascontiguousarray(a, dt) for a in (lam, mu, rho))
    P = {k: np.ascontiguousarray(v, dt) for k, v in prof.items()}
    ixr, iyr = (np.ascontiguousarray(a, np.int32) for a in (ixr, iyr))
    rc = getattr(lib(), f"oracle_runade8_{s}")(
        C.c_int(nx), C.c_int(ny), cr(dx), cr(dy), cr(deltat), C.c_int(1), C.c_int(nstep),
        C.c_int(npml), _ptr(lam), _ptr(mu), _ptr(rho),
        *[_ptr(P[k]) for k in ("a_x", "b_x", "K_x", "ah_x", "bh_x", "Kh_x",
                               "a_y", "b_y", "K_y", "ah_y", "bh_y", "Kh_y")],
        C.c_int(isrc), C.c_int(jsrc), cr(f0), cr(t0), cr(factor), cr(angle), C.c_int(nrec),
        _ptr(ixr), _ptr(iyr), C.c_int(nstep), _ptr(sis[0]), _ptr(sis[1]), _ptr(state),
        C.c_int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle_runade8 returned {rc}")
    out = {k: state[q] for q, k in enumerate(STATE2DVE)}
    out.update(state=state, sisvx=sis[0][:nrec], sisvy=sis[1][:nrec])
    return out


def stf_at(t, f0, t0, factor, angle, kind, prec="f64"):
    """Source time function at the given times (e.g. the RK4 stage times)."""
    dt, cr, s = _real(prec)
    t = np.ascontiguousarray(t, dt)
    fx = np.zeros(t.size, dt)
    fy = np.zeros(t.size, dt)
    getattr(lib(), f"oracle_stf_at_{s}")(C.c_int(t.size), _ptr(t), cr(f0), cr(t0), cr(factor),
                                         cr(angle), C.c_int(kind), _ptr(fx), _ptr(fy))
    return fx, fy
