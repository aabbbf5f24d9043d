"""Run a golden case through the HIP path (C ABI via seismic_cpml_amd)."""
from __future__ import annotations

import numpy as np

from seismic_cpml_amd import CpmlGpu

from case_inputs import inputs
from golden.cases import ALL_CASES

F2D = ("vx", "vy", "sigmaxx", "sigmayy", "sigmaxy")
F3D = ("vx", "vy", "vz", "sigmaxx", "sigmayy", "sigmazz", "sigmaxy", "sigmaxz", "sigmayz")
# oracle state names for the same fields
ORACLE_NAME = {"vx": "vx", "vy": "vy", "vz": "vz", "sigmaxx": "sxx", "sigmayy": "syy",
               "sigmazz": "szz", "sigmaxy": "sxy", "sigmaxz": "sxz", "sigmayz": "syz"}
MEM2D = ("dvx_dx", "dvx_dy", "dvy_dx", "dvy_dy", "dsigmaxx_dx", "dsigmayy_dy", "dsigmaxy_dx",
         "dsigmaxy_dy")
MEM3D = ("dvx_dx", "dvx_dy", "dvx_dz", "dvy_dx", "dvy_dy", "dvy_dz", "dvz_dx", "dvz_dy",
         "dvz_dz", "dsigmaxx_dx", "dsigmayy_dy", "dsigmazz_dz", "dsigmaxy_dx", "dsigmaxy_dy",
         "dsigmaxz_dx", "dsigmaxz_dz", "dsigmayz_dy", "dsigmayz_dz")


def oracle_mem_name(name: str) -> str:
    return "m_" + name.replace("sigma", "s")


def make_solver(case: str, prec: str, d: dict, energy=True, graph_steps=16, kchunk=None,
                tile=None, fused=False):
    c = ALL_CASES[case]
    p = 8 if prec == "f64" else 4
    ndim = 3 if c.get("NZ") else 2
    s = CpmlGpu(ndim, c["order"], p, c["NX"], c["NY"], c.get("NZ", 1), d["DELTAX"], d["DELTAY"],
                d["DELTAZ"], d["DELTAT"], c["NPML"], d["nstep"])
    s.set_medium(d["lam"], d["mu"], d["rho"])
    s.set_pml(d["prof_x"], d["prof_y"], d.get("prof_z"))
    s.set_source(d["ISOURCE"], d["JSOURCE"], d.get("ksrc", 1), d["fx"], d["fy"])
    s.set_receivers(d["ix_rec"], d["iy_rec"], d.get("iz_rec"))
    if energy:
        s.set_energy(1, window=d["ewin"])
    s.set_graph_steps(graph_steps)
    if kchunk:
        s.set_kchunk(kchunk)
    if tile:
        s.set_tile(tile)
    if fused:
        s.set_fused(True)
    return s


def run_gpu(case: str, prec: str, nstep: int | None = None, d: dict | None = None,
            energy=True, graph_steps=16, kchunk=None, memory=False, timing=False,
            tile=None, fused=False) -> dict:
    d = inputs(case, prec, nstep) if d is None else d
    with make_solver(case, prec, d, energy, graph_steps, kchunk, tile, fused) as s:
        if timing:
            s.set_kernel_timing(True)
        s.run()
        out = {}
        names = F3D if s.ndim == 3 else F2D
        for n in names:
            out[ORACLE_NAME[n]] = s.field(n)
        sis = s.seismograms()
        out["sisvx"], out["sisvy"] = sis[0], sis[1]
        if s.ndim == 3:
            out["sisvz"] = sis[2]
        if energy:
            out["ekin"], out["epot"] = s.energy()
        if memory:
            for m in (MEM3D if s.ndim == 3 else MEM2D):
                out[oracle_mem_name(m)] = s.memory(m)
        out["slabs"] = s.slab_widths()
        if timing:
            out["ktime"] = [s.kernel_time(0), s.kernel_time(1)]
    return out


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))
