"""Build and run the compiled reference programs (test infrastructure only).

This recipe compiles the reference Fortran programs *from their sources under
/root/reference* with ``amdflang -O3`` into ``oracle/_ref/`` and runs them to
produce golden vectors.  Only binaries are written under ``oracle/_ref/``; the
textually patched sources live in a temporary directory that is deleted after
compilation, so no reference source is copied into the repository.

Patches applied to each variant (all textual, on a temporary copy):

* ``parameter`` values (grid size, steps, source, receivers ...), because the
  reference programs are configured by compile-time constants only;
* single precision: ``double precision`` -> ``real``, ``1.d0`` -> ``1.e0``,
  ``dble(``/``sngl(`` -> ``real(``, ``MPI_DOUBLE_PRECISION`` -> ``MPI_REAL``
  (the reference's own instruction, 2D2:131-133);
* optional: medium read from ``model.bin`` instead of the homogeneous loop
  (2D2:468-474), Ricker source (the commented line 2D2:654), image output
  disabled (2D2:728-731) except in the variant that pins the images;
* a binary dump written after the time loop (profiles, seismograms, energy,
  final wavefields and memory variables) so results can be compared at full
  precision -- the reference itself only writes single-precision ASCII.  The
  ADE8 variant has no dump: it runs only for the SU files it writes.

The 3D program is run with 2 MPI ranks (MPICH from /opt/conda), the smallest
decomposition it accepts (3D:381-394).
"""
from __future__ import annotations

import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

REF_DIR = Path(os.environ.get("CPML_REFERENCE_DIR", "/root/reference"))
HERE = Path(__file__).resolve().parent
OUT_DIR = HERE / "_ref"
MPI_BIN = Path("/opt/conda/bin")
MPI_INC = Path("/opt/conda/include")
MPI_LIB = Path("/opt/conda/lib")

SRC_2D2 = "seismic_CPML_2D_isotropic_second_order.f90"
SRC_ADE8 = "seismic_ADEPML_2D_elastic_RK4_eighth_order.f90"
SRC_2D4 = "seismic_CPML_2D_isotropic_fourth_order.f90"
SRC_3D = "seismic_CPML_3D_isotropic_MPI_OpenMP.f90"
SRC_3DV = "seismic_CPML_3D_viscoelastic_MPI.f90"
SRC_2DVE = "seismic_CPML_2D_velocity_and_stress_second_order_viscoelastic.f90"
SRC_2DVE4 = "seismic_CPML_2D_velocity_and_stress_fourth_order_viscoelastic.f90"

# Reduced heterogeneous 2D case: 128 x 96, 5 m, seeded layered medium read
# from model.bin, Ricker source at the centre, 8 receivers along the top
# interior row (snapped by the reference's own nearest-node search).
_LAYERED_2D = {
    "NX": "128", "NY": "96", "DELTAX": "5.d0", "NPOINTS_PML": "10", "NSTEP": "600",
    "cp": "5500.d0", "DELTAT": "3.2d-4", "f0": "10.d0", "ISOURCE": "64", "JSOURCE": "48",
    "NREC": "8", "xdeb": "50.d0", "ydeb": "425.d0", "xfin": "585.d0", "yfin": "425.d0",
}
# Reduced 3D case (homogeneous, as the reference): 40 x 44 x 36, 8-point PML,
# 300 steps; receivers moved inside the grid.
_SMALL_3D = {
    "NX": "40", "NY": "44", "NZ": "36", "NPROC": "2", "NPOINTS_PML": "8", "NSTEP": "300",
    "xdeb": "xsource - 100.d0", "ydeb": "300.d0", "xfin": "xsource", "yfin": "100.d0",
}

VARIANTS: dict[str, dict] = {}
for _p in ("f64", "f32"):
    VARIANTS[f"ref2d2_shipped_{_p}"] = dict(src=SRC_2D2, prec=_p, params={}, ndim=2, ghost=0)
    VARIANTS[f"ref2d4_shipped_{_p}"] = dict(src=SRC_2D4, prec=_p, params={}, ndim=2, ghost=1)
    VARIANTS[f"ref2d2_layered_{_p}"] = dict(src=SRC_2D2, prec=_p, params=_LAYERED_2D, ndim=2,
                                            ghost=0, model=True, ricker=True)
    VARIANTS[f"ref2d4_layered_{_p}"] = dict(src=SRC_2D4, prec=_p, params=_LAYERED_2D, ndim=2,
                                            ghost=1, model=True, ricker=True)
    VARIANTS[f"ref3d_small_{_p}"] = dict(src=SRC_3D, prec=_p, params=_SMALL_3D, ndim=3,
                                         ghost=0, nproc=2)
# The 3D program timed as the CPU baseline of the 768^3 benchmark: its own
# homogeneous model on a 768 x 768 x 64-plane sample of the benchmark grid,
# NPML 20, two MPI ranks (the program requires an even count, 3D:388) of
# OpenMP threads (its PARALLEL DO loops, 3D:825), energy every step as the
# program computes it, no display steps; built for 10 and 310 steps, so that
# the difference of the two wall times is the time of 300 steps without the
# set-up (the program's arrays are static: more planes overflow the 2 GiB
# reach of its data section).
_BENCH_3D = {"NX": "768", "NY": "768", "NZ": "64", "NPROC": "2", "NPOINTS_PML": "20",
             "IT_DISPLAY": "100000"}
for _n in (10, 310):
    VARIANTS[f"ref3d_bench{_n}_f32"] = dict(src=SRC_3D, prec="f32",
                                           params=dict(_BENCH_3D, NSTEP=str(_n)), ndim=3,
                                           ghost=0, nproc=2, dump=False, openmp=True)
# 2D2 and 2D4 as shipped with their PNM snapshots left on (2D2:716-731,
# 861-999; 2D4:729-731), 200 steps: images of Vx and Vy at steps 5, 100, 200.
for _p in ("f64", "f32"):
    VARIANTS[f"ref2d2_images_{_p}"] = dict(src=SRC_2D2, prec=_p, params={"NSTEP": "200"},
                                           ndim=2, ghost=0, images=True)
    VARIANTS[f"ref2d4_images_{_p}"] = dict(src=SRC_2D4, prec=_p, params={"NSTEP": "200"},
                                           ndim=2, ghost=1, images=True)
    VARIANTS[f"ref3d_images_{_p}"] = dict(src=SRC_3D, prec=_p, params=_SMALL_3D, ndim=3,
                                          ghost=0, images=True)
# The 4th-order 3D program (3D-visco) in its elastic limit: after the
# attenuation fit, tau_epsilon := tau_sigma, so phi = 0 and Mu_nu = 1 exactly,
# the memory variables e* stay zero and the update is the elastic one
# (3D-visco:466-477, 977-1300).  Reduced grid, 2 ranks, receivers moved near
# the source (the shipped ones lie 500 m and 2260 m away, 3D-visco:824-829).
_SMALL_3DV = {
    "NX": "40", "NY": "60", "NZ": "40", "NPROC": "2", "NPOINTS_PML": "8", "NSTEP": "400",
    "IT_DISPLAY": "100",
}
_RECEIVERS_3DV = (("xrec(1)=xsource+500.d0", "xrec(1)=xsource+20.d0"),
                  ("yrec(1)=ysource+500.d0", "yrec(1)=ysource+40.d0"),
                  ("yrec(2)=ysource+2260.d0", "yrec(2)=ysource+120.d0"),
                  ("xrec(3)=xsource+500.d0", "xrec(3)=xsource+20.d0"),
                  ("yrec(3)=ysource+2260.d0", "yrec(3)=ysource+120.d0"))
VARIANTS["ref3dv4_elastic_f64"] = dict(src=SRC_3DV, prec="f64", params=_SMALL_3DV, ndim=3,
                                       ghost=0, nproc=2, visco=True)
# 2D viscoelastic C-PML, 2nd order (2D-visco2), as shipped (Qp 65, Qs 55,
# N_SLS 3, attenuation on) on a reduced 161 x 161 grid, 700 steps, source at
# the centre, 3 receivers; float64 only (its SolvOpt include stays double).
# ref2dve4_f64: the same case of the 4th-order program (2D-visco4).
VARIANTS["ref2dve_f64"] = dict(src=SRC_2DVE, prec="f64", ndim=2, ghost=0, visco2d=True,
                               params={"NX": "161", "NY": "161", "NSTEP": "700",
                                       "xsource": "120.d0", "ysource": "120.d0", "NREC": "3",
                                       "xdeb": "60.d0", "ydeb": "190.d0", "xfin": "190.d0",
                                       "yfin": "60.d0"})
VARIANTS["ref2dve4_f64"] = dict(VARIANTS["ref2dve_f64"], src=SRC_2DVE4, ghost=1)
# ADE8 (RK4, ADE-PML, 8th order) on 81 x 81 for 300 steps with a binary dump
# after its time loop (and its SU files, for which it needs its folders).
VARIANTS["refade8_f64"] = dict(src=SRC_ADE8, prec="f64", ndim=2, ghost=0, ade8=True,
                               params={"NX": "81", "NY": "81", "NSTEP": "300",
                                       "IT_DISPLAY": "100"})
# ADE8 on a 61 x 61 grid for 60 steps, snapshots every 30: run only for the SU
# files it writes (seismograms ADE8:987-1003, snapshots ADE8:935-953); no dump.
VARIANTS["refade8_su_f64"] = dict(src=SRC_ADE8, prec="f64", ndim=2, ghost=0, dump=False,
                                  params={"NX": "61", "NY": "61", "NSTEP": "60",
                                          "IT_DISPLAY": "30"})


def _set_params(text: str, params: dict) -> str:
    lines = text.split("\n")
    for name, value in params.items():
        pat = re.compile(r"((?:::|,)\s*)(" + re.escape(name) + r")(\s*=\s*)([^,!\n]+)")
        hits = 0
        for n, line in enumerate(lines):
            if "parameter" not in line.lower() or line.lstrip().startswith("!"):
                continue
            new, k = pat.subn(lambda m: m.group(1) + m.group(2) + m.group(3) + value + " ",
                              line, count=1)
            if k:
                lines[n] = new
                hits += 1
                break
        if hits != 1:
            raise RuntimeError(f"parameter {name} not found")
    return "\n".join(lines)


def _to_single(text: str) -> str:
    out = []
    for line in text.split("\n"):
        code, bang, comment = line.partition("!") if "'" not in line else (line, "", "")
        code = re.sub(r"(?i)double\s+precision", "real", code)
        code = re.sub(r"(?<![\w.])(\d+\.?\d*|\.\d+)[dD]([+-]?\d+)", r"\1e\2", code)
        code = re.sub(r"\bdble\(", "real(", code)
        code = re.sub(r"\bsngl\(", "real(", code)
        code = code.replace("MPI_DOUBLE_PRECISION", "MPI_REAL")
        out.append(code + bang + comment)
    return "\n".join(out)


_MODEL_LOOP = re.compile(
    r"  do j = 1,NY\n    do i = 1,NX\n        rho\(i,j\) = density\n"
    r"        mu\(i,j\) = density\*cs\*cs\n"
    r"        lambda\(i,j\) = density\*\(cp\*cp - 2\.d0\*cs\*cs\)\n    enddo\n  enddo\n")
_MODEL_READ = ("  open(unit=31,file='model.bin',access='stream',form='unformatted',status='old')\n"
               "  read(31) rho(1:NX,1:NY),mu(1:NX,1:NY),lambda(1:NX,1:NY)\n"
               "  close(31)\n")
_FDGAUSS = "  source_term = - factor * 2.d0*a*(t-t0)*exp(-a*(t-t0)**2)\n"
_RICKER_C = "! source_term = factor * (1.d0 - 2.d0*a*(t-t0)**2)*exp(-a*(t-t0)**2)\n"


def _dump_2d(src: str) -> str:
    s = "sigma_" if src == SRC_2D2 else "sigma"
    m = "memory_dsigma_" if src == SRC_2D2 else "memory_dsigma"
    return (
        "  open(unit=77,file='ref_dump.bin',access='stream',form='unformatted',status='replace')\n"
        "  write(77) a_x,b_x,K_x,a_x_half,b_x_half,K_x_half\n"
        "  write(77) a_y,b_y,K_y,a_y_half,b_y_half,K_y_half\n"
        "  write(77) sisvx,sisvy\n"
        "  write(77) total_energy_kinetic,total_energy_potential\n"
        f"  write(77) vx,vy,{s}xx,{s}yy,{s}xy\n"
        "  write(77) memory_dvx_dx,memory_dvx_dy,memory_dvy_dx,memory_dvy_dy\n"
        f"  write(77) {m}xx_dx,{m}yy_dy,{m}xy_dx,{m}xy_dy\n"
        "  write(77) ix_rec,iy_rec\n"
        "  close(77)\n")


_DUMP_3D = (
    "  write(ref_dumpname,\"('ref_dump_',i3.3,'.bin')\") rank\n"
    "  open(unit=77,file=ref_dumpname,access='stream',form='unformatted',status='replace')\n"
    "  write(77) a_x,b_x,K_x,a_x_half,b_x_half,K_x_half\n"
    "  write(77) a_y,b_y,K_y,a_y_half,b_y_half,K_y_half\n"
    "  write(77) a_z,b_z,K_z,a_z_half,b_z_half,K_z_half\n"
    "  write(77) vx(:,:,1:NZ_LOCAL),vy(:,:,1:NZ_LOCAL),vz(:,:,1:NZ_LOCAL)\n"
    "  write(77) sigmaxx(:,:,1:NZ_LOCAL),sigmayy(:,:,1:NZ_LOCAL),sigmazz(:,:,1:NZ_LOCAL)\n"
    "  write(77) sigmaxy(:,:,1:NZ_LOCAL),sigmaxz(:,:,1:NZ_LOCAL),sigmayz(:,:,1:NZ_LOCAL)\n"
    "  write(77) memory_dvx_dx,memory_dvx_dy,memory_dvx_dz,memory_dvy_dx,memory_dvy_dy\n"
    "  write(77) memory_dvy_dz,memory_dvz_dx,memory_dvz_dy,memory_dvz_dz\n"
    "  write(77) memory_dsigmaxx_dx,memory_dsigmayy_dy,memory_dsigmazz_dz,memory_dsigmaxy_dx\n"
    "  write(77) memory_dsigmaxy_dy,memory_dsigmaxz_dx,memory_dsigmaxz_dz\n"
    "  write(77) memory_dsigmayz_dy,memory_dsigmayz_dz\n"
    "  if (rank == rank_cut_plane) then\n"
    "    write(77) sisvx,sisvy,total_energy\n"
    "    write(77) ix_rec,iy_rec\n"
    "  endif\n"
    "  close(77)\n")


_DUMP_3DV = (
    "  write(ref_dumpname,\"('ref_dump_',i3.3,'.bin')\") rank\n"
    "  open(unit=77,file=ref_dumpname,access='stream',form='unformatted',status='replace')\n"
    "  write(77) a_x,b_x,K_x,a_x_half,b_x_half,K_x_half\n"
    "  write(77) a_y,b_y,K_y,a_y_half,b_y_half,K_y_half\n"
    "  write(77) a_z,b_z,K_z,a_z_half,b_z_half,K_z_half\n"
    "  write(77) vx(1:NX,1:NY,1:NZ_LOCAL),vy(1:NX,1:NY,1:NZ_LOCAL),vz(1:NX,1:NY,1:NZ_LOCAL)\n"
    "  write(77) sigmaxx(1:NX,1:NY,1:NZ_LOCAL),sigmayy(1:NX,1:NY,1:NZ_LOCAL)\n"
    "  write(77) sigmazz(1:NX,1:NY,1:NZ_LOCAL),sigmaxy(1:NX,1:NY,1:NZ_LOCAL)\n"
    "  write(77) sigmaxz(1:NX,1:NY,1:NZ_LOCAL),sigmayz(1:NX,1:NY,1:NZ_LOCAL)\n"
    + "".join(f"  write(77) {m}(1:NX,1:NY,1:NZ_LOCAL)\n" for m in (
        "memory_dvx_dx", "memory_dvx_dy", "memory_dvx_dz", "memory_dvy_dx", "memory_dvy_dy",
        "memory_dvy_dz", "memory_dvz_dx", "memory_dvz_dy", "memory_dvz_dz",
        "memory_dsigmaxx_dx", "memory_dsigmayy_dy", "memory_dsigmazz_dz",
        "memory_dsigmaxy_dx", "memory_dsigmaxy_dy", "memory_dsigmaxz_dx",
        "memory_dsigmaxz_dz", "memory_dsigmayz_dy", "memory_dsigmayz_dz")) +
    "  if (rank == rank_cut_plane) then\n"
    "    write(77) sisvx,sisvy,total_energy\n"
    "    write(77) ix_rec,iy_rec\n"
    "  endif\n"
    "  close(77)\n")
_ELASTIC_LIMIT = ("  tau_epsilon_nu1 = tau_sigma_nu1\n"
                  "  tau_epsilon_nu2 = tau_sigma_nu2\n")


def _patch_visco(text: str) -> str:
    """Elastic limit, receivers, mpif.h instead of the mpi module (whose .mod
    is not built for amdflang), images off."""
    lines = text.split("\n")
    n = next(i for i, l in enumerate(lines) if "tau_epsilon_nu2,tau_sigma_nu2)" in l)
    lines.insert(n + 1, _ELASTIC_LIMIT.rstrip("\n"))
    text = "\n".join(lines)
    for old, new in _RECEIVERS_3DV:
        if text.count(old) != 1:
            raise RuntimeError(f"receiver line {old!r} not found")
        text = text.replace(old, new)
    if text.count("  use mpi\n") != 1:
        raise RuntimeError("use mpi not found")
    text = text.replace("  use mpi\n", "")
    text = text.replace("  implicit none\n", "  implicit none\n  include 'mpif.h'\n"
                        "  character(len=64) :: ref_dumpname\n", 1)
    text = text.replace("call create_color_image(", "if (.false.) call create_color_image(")
    lines = text.split("\n")
    mark = next(n for n, l in enumerate(lines) if "end of time loop" in l)
    mark = next(n for n in range(mark + 1, len(lines)) if lines[n].strip() == "enddo")
    lines.insert(mark + 1, _DUMP_3DV.rstrip("\n"))
    return "\n".join(lines)


_DUMP_2DVE = (
    "  open(unit=77,file='ref_dump.bin',access='stream',form='unformatted',status='replace')\n"
    "  write(77) a_x,b_x,K_x,a_x_half,b_x_half,K_x_half\n"
    "  write(77) a_y,b_y,K_y,a_y_half,b_y_half,K_y_half\n"
    "  write(77) sisvx,sisvy,sispressure\n"
    "  write(77) vx,vy,sigma_xx,sigma_yy,sigma_xy\n"
    "  write(77) memory_dvx_dx,memory_dvx_dy,memory_dvy_dx,memory_dvy_dy\n"
    "  write(77) memory_dsigma_xx_dx,memory_dsigma_yy_dy,memory_dsigma_xy_dx,memory_dsigma_xy_dy\n"
    "  write(77) memory_variable_R_e1_dot,memory_variable_R_e11_dot,memory_variable_R_e13_dot\n"
    "  write(77) tau_epsilon_nu1,tau_sigma_nu1,tau_epsilon_nu2,tau_sigma_nu2\n"
    "  write(77) lambda_unrelaxed,mu_unrelaxed,rho\n"
    "  write(77) ix_rec,iy_rec\n"
    "  close(77)\n")


_DUMP_ADE8 = (
    "  open(unit=77,file='ref_dump.bin',access='stream',form='unformatted',status='replace')\n"
    "  write(77) a_x,b_x,K_x,a_x_half,b_x_half,K_x_half\n"
    "  write(77) a_y,b_y,K_y,a_y_half,b_y_half,K_y_half\n"
    "  write(77) sisvx,sisvy\n"
    "  write(77) vx,vy,sigmaxx,sigmayy,sigmaxy\n"
    "  write(77) memory_dvx_dx,memory_dvx_dy,memory_dvy_dx,memory_dvy_dy\n"
    "  write(77) memory_dsigmaxx_dx,memory_dsigmayy_dy,memory_dsigmaxy_dx,memory_dsigmaxy_dy\n"
    "  write(77) lambda,mu,rho\n"
    "  write(77) ix_rec,iy_rec\n"
    "  close(77)\n")


def _patch_ade8(text: str) -> str:
    """Binary dump after the time loop (ADE8:983)."""
    lines = text.split("\n")
    mark = next(n for n, l in enumerate(lines) if "end of time loop" in l)
    lines.insert(mark + 1, _DUMP_ADE8.rstrip("\n"))
    return "\n".join(lines)


def _patch_visco2d(text: str) -> str:
    """Images and the file-removal shell call off, binary dump after the loop."""
    text = text.replace("call create_color_image(", "if (.false.) call create_color_image(")
    old = "  call system('rm -f Vx_file*.dat Vy_file*.dat image*.pnm image*.gif')"
    if text.count(old) != 1:
        raise RuntimeError("system call line not found")
    text = text.replace(old, "")
    lines = text.split("\n")
    mark = next(n for n, l in enumerate(lines) if "end of time loop" in l)
    lines.insert(mark + 1, _DUMP_2DVE.rstrip("\n"))
    return "\n".join(lines)


def patch(text: str, v: dict) -> str:
    """Return the patched program text for variant ``v``."""
    text = _set_params(text, v["params"])
    if v.get("visco"):
        return _patch_visco(text)
    if v.get("visco2d"):
        return _patch_visco2d(text)
    if v.get("ade8"):
        return _patch_ade8(text)
    if v.get("model"):
        text, k = _MODEL_LOOP.subn(_MODEL_READ, text)
        if k != 1:
            raise RuntimeError("material loop not found")
    if v.get("ricker"):
        if text.count(_FDGAUSS) != 1 or text.count(_RICKER_C) != 1:
            raise RuntimeError("source lines not found")
        text = text.replace(_FDGAUSS, "! " + _FDGAUSS.lstrip())
        text = text.replace(_RICKER_C, _RICKER_C[2:].rjust(len(_RICKER_C)))
    if not v.get("images"):
        text = text.replace("call create_color_image(", "if (.false.) call create_color_image(")
    if not v.get("dump", True):
        return _to_single(text) if v["prec"] == "f32" else text
    lines = text.split("\n")
    mark = next(n for n, l in enumerate(lines) if "end of time loop" in l)
    if "enddo" not in lines[mark]:
        mark = next(n for n in range(mark + 1, len(lines)) if lines[n].strip() == "enddo")
    dump = _dump_2d(v["src"]) if v["ndim"] == 2 else _DUMP_3D
    lines.insert(mark + 1, dump.rstrip("\n"))
    if v["ndim"] == 3:
        inc = next(n for n, l in enumerate(lines) if "include 'mpif.h'" in l)
        lines.insert(inc + 1, "  character(len=64) :: ref_dumpname")
    text = "\n".join(lines)
    if v["prec"] == "f32":
        text = _to_single(text)
    return text


def available() -> bool:
    return (REF_DIR / SRC_2D2).is_file() and shutil.which("amdflang") is not None


def build(names=None, verbose=False) -> list[str]:
    """Compile the requested variants into oracle/_ref/.  Returns the names built."""
    if not available():
        raise RuntimeError(f"reference sources or amdflang not available ({REF_DIR})")
    OUT_DIR.mkdir(exist_ok=True)
    built = []
    with tempfile.TemporaryDirectory(prefix="cpml_ref_src_") as tmp:
        for name in names or VARIANTS:
            v = VARIANTS[name]
            src = (REF_DIR / v["src"]).read_text()
            f90 = Path(tmp) / f"{name}.f90"
            f90.write_text(patch(src, v))
            exe = OUT_DIR / name
            cmd = ["amdflang", "-O3", "-o", str(exe), str(f90)]
            if v.get("visco2d"):  # 2D-visco2:1337 includes the SolvOpt file
                cmd = ["amdflang", "-O3", f"-I{REF_DIR}", "-o", str(exe), str(f90)]
            if v["ndim"] == 3:
                # -I REF_DIR: 3D-visco:1578 includes the reference's SolvOpt file
                cmd = ["amdflang", "-O3", f"-I{MPI_INC}", f"-I{REF_DIR}", "-o", str(exe),
                       str(f90), f"-L{MPI_LIB}", "-lmpifort", "-lmpi", f"-Wl,-rpath,{MPI_LIB}"]
            if v.get("openmp"):
                cmd.insert(1, "-fopenmp")
            r = subprocess.run(cmd, cwd=tmp, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"compile {name} failed:\n{r.stdout}\n{r.stderr}")
            meta = {k: v[k] for k in ("src", "prec", "ndim", "ghost")}
            meta["params"] = v["params"]
            (OUT_DIR / f"{name}.json").write_text(json.dumps(meta, indent=1))
            built.append(name)
            if verbose:
                print("built", exe)
    return built


def _shipped_ints(v: dict) -> dict:
    """Integer parameters of a variant after patching (NX, NY, NZ, NSTEP, NREC, NPML)."""
    text = (REF_DIR / v["src"]).read_text()
    vals = {}
    for name in ("NX", "NY", "NZ", "NSTEP", "NREC", "NPOINTS_PML", "NPROC"):
        if name in v["params"]:
            vals[name] = int(eval(v["params"][name]))  # plain integers only
            continue
        m = re.search(r"integer, parameter :: " + name + r" = ([^!\n]+)", text)
        if m:
            vals[name] = int(eval(m.group(1).replace("/", "//")))
    return vals


def run(name: str, workdir: Path, model: tuple | None = None, timeout: int = 600) -> dict:
    """Run a built variant in ``workdir`` and return its dump as numpy arrays."""
    v = VARIANTS[name]
    exe = OUT_DIR / name
    if not exe.is_file():
        raise RuntimeError(f"{exe} not built")
    workdir = Path(workdir)
    workdir.mkdir(parents=True, exist_ok=True)
    if v.get("visco2d"):
        return _run_visco2d(v, exe, workdir, timeout)
    if v.get("ade8"):
        return _run_ade8(v, exe, workdir, timeout)
    if not v.get("dump", True):  # ADE8 writes into these folders (ADE8:942, 994)
        for sub in ("seismograms", "snapshots"):
            (workdir / sub).mkdir(exist_ok=True)
        with open(workdir / "stdout.txt", "w") as log:
            r = subprocess.run([str(exe)], cwd=workdir, stdout=log, stderr=subprocess.STDOUT,
                               timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"{name} exited with {r.returncode}; see {workdir}/stdout.txt")
        return {}
    dt = np.float64 if v["prec"] == "f64" else np.float32
    ints = _shipped_ints(v)
    nx, ny, nstep, nrec = ints["NX"], ints["NY"], ints["NSTEP"], ints["NREC"]
    if model is not None:
        with open(workdir / "model.bin", "wb") as f:
            for arr in model:  # rho, mu, lambda, each (ny, nx) C-order == Fortran (nx, ny)
                f.write(np.ascontiguousarray(arr, dtype=dt).tobytes())
    if v["ndim"] == 2:
        cmd = [str(exe)]
    else:
        cmd = [str(MPI_BIN / "mpirun"), "-np", str(v.get("nproc", 2)), str(exe)]
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = f"{MPI_LIB}:" + env.get("LD_LIBRARY_PATH", "")
    with open(workdir / "stdout.txt", "w") as log:
        r = subprocess.run(cmd, cwd=workdir, stdout=log, stderr=subprocess.STDOUT,
                           env=env, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"{name} exited with {r.returncode}; see {workdir}/stdout.txt")
    out: dict = {"params": ints}
    if v["ndim"] == 2:
        g = v["ghost"]
        raw = (workdir / "ref_dump.bin").read_bytes()
        off = 0

        def take(count, typ=dt):
            nonlocal off
            a = np.frombuffer(raw, dtype=typ, count=count, offset=off)
            off += a.nbytes
            return a.copy()

        px = take(6 * nx).reshape(6, nx)
        py = take(6 * ny).reshape(6, ny)
        out["prof_x"], out["prof_y"] = px, py
        out["sisvx"] = take(nstep * nrec).reshape(nrec, nstep)
        out["sisvy"] = take(nstep * nrec).reshape(nrec, nstep)
        out["ekin"] = take(nstep)
        out["epot"] = take(nstep)
        shp = (ny + 2 * g, nx + 2 * g)
        cnt = shp[0] * shp[1]
        sl = (slice(g, g + ny), slice(g, g + nx))
        for key in ("vx", "vy", "sxx", "syy", "sxy"):
            out[key] = take(cnt).reshape(shp)[sl].copy()
        for key in ("m_dvx_dx", "m_dvx_dy", "m_dvy_dx", "m_dvy_dy",
                    "m_dsxx_dx", "m_dsyy_dy", "m_dsxy_dx", "m_dsxy_dy"):
            out[key] = take(cnt).reshape(shp)[sl].copy()
        out["ix_rec"] = take(nrec, np.int32)
        out["iy_rec"] = take(nrec, np.int32)
        assert off == len(raw), (off, len(raw))
    else:
        nz, nproc = ints["NZ"], v.get("nproc", 2)
        nzl = nz // nproc
        fields = {k: [] for k in ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")}
        mems = ("m_dvx_dx", "m_dvx_dy", "m_dvx_dz", "m_dvy_dx", "m_dvy_dy", "m_dvy_dz",
                "m_dvz_dx", "m_dvz_dy", "m_dvz_dz", "m_dsxx_dx", "m_dsyy_dy", "m_dszz_dz",
                "m_dsxy_dx", "m_dsxy_dy", "m_dsxz_dx", "m_dsxz_dz", "m_dsyz_dy", "m_dsyz_dz")
        for k in mems:
            fields[k] = []
        for rank in range(nproc):
            raw = (workdir / f"ref_dump_{rank:03d}.bin").read_bytes()
            off = 0

            def take(count, typ=dt):
                nonlocal off
                a = np.frombuffer(raw, dtype=typ, count=count, offset=off)
                off += a.nbytes
                return a.copy()

            px = take(6 * nx).reshape(6, nx)
            py = take(6 * ny).reshape(6, ny)
            pz = take(6 * nz).reshape(6, nz)
            if rank == 0:
                out["prof_x"], out["prof_y"], out["prof_z"] = px, py, pz
            for key in list(fields):
                fields[key].append(take(nx * ny * nzl).reshape(nzl, ny, nx))
            if rank == nproc // 2 - 1:
                out["sisvx"] = take(nstep * nrec).reshape(nrec, nstep)
                out["sisvy"] = take(nstep * nrec).reshape(nrec, nstep)
                out["etotal"] = take(nstep)
                out["ix_rec"] = take(nrec, np.int32)
                out["iy_rec"] = take(nrec, np.int32)
            assert off == len(raw), (rank, off, len(raw))
        for key, parts in fields.items():
            out[key] = np.concatenate(parts, axis=0)
    return out


def _run_ade8(v: dict, exe: Path, workdir: Path, timeout: int) -> dict:
    """Run the ADE8 variant.  Arrays keep the program's 4-cell ghost border:
    fields, memories, media (ny+9, nx+9); profiles a, b (ny|nx+9, 4 stages)
    as [stage][index], K (n+9)."""
    nx, ny, nstep = (int(v["params"][k]) for k in ("NX", "NY", "NSTEP"))
    nrec = nx
    for sub in ("seismograms", "snapshots"):
        (workdir / sub).mkdir(exist_ok=True)
    with open(workdir / "stdout.txt", "w") as log:
        r = subprocess.run([str(exe)], cwd=workdir, stdout=log, stderr=subprocess.STDOUT,
                           timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"exited with {r.returncode}; see {workdir}/stdout.txt")
    raw = (workdir / "ref_dump.bin").read_bytes()
    off = 0

    def take(count, typ=np.float64):
        nonlocal off
        a = np.frombuffer(raw, dtype=typ, count=count, offset=off)
        off += a.nbytes
        return a.copy()

    out = {}
    for ax, n in (("x", nx), ("y", ny)):
        g = n + 9
        out[f"a_{ax}"] = take(4 * g).reshape(g, 4).T.copy()
        out[f"b_{ax}"] = take(4 * g).reshape(g, 4).T.copy()
        out[f"K_{ax}"] = take(g)
        out[f"ah_{ax}"] = take(4 * g).reshape(g, 4).T.copy()
        out[f"bh_{ax}"] = take(4 * g).reshape(g, 4).T.copy()
        out[f"Kh_{ax}"] = take(g)
    out["sisvx"] = take(nstep * nrec).reshape(nrec, nstep)
    out["sisvy"] = take(nstep * nrec).reshape(nrec, nstep)
    shp = (ny + 9, nx + 9)
    for k in ("vx", "vy", "sxx", "syy", "sxy", "m_dvx_dx", "m_dvx_dy", "m_dvy_dx", "m_dvy_dy",
              "m_dsxx_dx", "m_dsyy_dy", "m_dsxy_dx", "m_dsxy_dy", "lam", "mu", "rho"):
        out[k] = take(shp[0] * shp[1]).reshape(shp)
    out["ix_rec"] = take(nrec, np.int32)
    out["iy_rec"] = take(nrec, np.int32)
    assert off == len(raw), (off, len(raw))
    return out


def _run_visco2d(v: dict, exe: Path, workdir: Path, timeout: int) -> dict:
    """Run the 2D viscoelastic variant; its dump as (ny, nx) arrays, the
    attenuation memory variables as (N_SLS, ny, nx)."""
    nx, ny = int(v["params"]["NX"]), int(v["params"]["NY"])
    nstep, nrec, nsls = int(v["params"]["NSTEP"]), int(v["params"]["NREC"]), 3
    with open(workdir / "stdout.txt", "w") as log:
        r = subprocess.run([str(exe)], cwd=workdir, stdout=log, stderr=subprocess.STDOUT,
                           timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"exited with {r.returncode}; see {workdir}/stdout.txt")
    raw = (workdir / "ref_dump.bin").read_bytes()
    off = 0

    def take(count, typ=np.float64):
        nonlocal off
        a = np.frombuffer(raw, dtype=typ, count=count, offset=off)
        off += a.nbytes
        return a.copy()

    g = v["ghost"]  # the 4th-order program's fields and media are 0:NX+1, 0:NY+1

    def grid():
        a = take((nx + 2 * g) * (ny + 2 * g)).reshape(ny + 2 * g, nx + 2 * g)
        return a[g:g + ny, g:g + nx].copy()

    out = {"prof_x": take(6 * nx).reshape(6, nx), "prof_y": take(6 * ny).reshape(6, ny)}
    for k in ("sisvx", "sisvy", "sispressure"):
        out[k] = take(nstep * nrec).reshape(nrec, nstep)
    for k in ("vx", "vy", "sxx", "syy", "sxy"):
        out[k] = grid()
    for k in ("m_dvx_dx", "m_dvx_dy", "m_dvy_dx", "m_dvy_dy",
              "m_dsxx_dx", "m_dsyy_dy", "m_dsxy_dx", "m_dsxy_dy"):
        out[k] = take(nx * ny).reshape(ny, nx)
    for k in ("r_e1", "r_e11", "r_e13"):
        out[k] = take(nx * ny * nsls).reshape(nsls, ny, nx)
    for k in ("tau_eps_nu1", "tau_sig_nu1", "tau_eps_nu2", "tau_sig_nu2"):
        out[k] = take(nsls)
    for k in ("lam", "mu", "rho"):
        out[k] = grid()
    out["ix_rec"] = take(nrec, np.int32)
    out["iy_rec"] = take(nrec, np.int32)
    assert off == len(raw), (off, len(raw))
    return out


if __name__ == "__main__":
    names = sys.argv[1:] or None
    for n in build(names, verbose=True):
        pass
