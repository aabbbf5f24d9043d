"""Generate the golden fixtures in tests/golden/ by running the compiled
reference programs (built by oracle/build_ref.py from /root/reference).

    python tests/golden/make_golden.py            # all cases, both precisions
    python tests/golden/make_golden.py images     # only the 2D PNM snapshots
    python tests/golden/make_golden.py display3d  # only the 3D display-step files
    python tests/golden/make_golden.py su         # only the ADE8 SU files
    python tests/golden/make_golden.py visco2d    # only the 2D viscoelastic case
    python tests/golden/make_golden.py ade8       # only the ADE-PML RK4 case

Each fixture ``<case>_<prec>.npz`` holds the reference's own outputs at full
precision: damping profiles, receiver indices, seismograms, energy curve(s),
the SHA-256 digest of each final wavefield with a regular subsample of it
(every 4th point per axis in 2D, every 3rd in 3D), and for each C-PML memory array the masks of x/y/z
lines on which it is non-zero (the arrays themselves are not stored).  For the
layered cases the medium (rho, mu, lambda) fed to the program is stored too.
``3d4_small_f64.npz`` is the 4th-order 3D case: the viscoelastic 3D program
in its elastic limit (float64 only).
``images_<2d2|2d4>_<prec>.npz`` hold the PNM snapshots (header and pixels)
that the shipped 2D2 and 2D4 programs write at steps 5, 100 and 200;
``display_3d_<prec>.npz`` the timestamp files and cut-plane PNM snapshots of
the reduced 3D program at its display steps; ``su_ade8.npz`` the SU
seismogram and snapshot files of a reduced run of the 8th-order program;
``2dve_small_f64.npz`` / ``2dve4_small_f64.npz`` the reduced 2D
viscoelastic (attenuation) case of the 2nd- / 4th-order program;
``ade8_small_f64.npz`` the reduced ADE-PML RK4 8th-order case.
This script needs /root/reference and amdflang; the fixtures it writes do not.
"""
from __future__ import annotations

import hashlib
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(HERE))

import build_ref  # noqa: E402
from cases import ALL_CASES, CASES, CASES_3D4, VARIANT_OF, layered_model  # noqa: E402

FIELDS_2D = ("vx", "vy", "sxx", "syy", "sxy")
MEM_2D = ("m_dvx_dx", "m_dvx_dy", "m_dvy_dx", "m_dvy_dy",
          "m_dsxx_dx", "m_dsyy_dy", "m_dsxy_dx", "m_dsxy_dy")
FIELDS_3D = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")


def field_digest(a) -> str:
    """SHA-256 of the array's bytes (C order, little endian) with -0.0 folded to +0.0."""
    a = np.ascontiguousarray(a + a.dtype.type(0))
    return hashlib.sha256(a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes()).hexdigest()


def make(case: str, prec: str, workroot: Path) -> Path:
    c = ALL_CASES[case]
    variant = f"{VARIANT_OF[case]}_{prec}"
    model = None
    if c.get("model"):
        model = layered_model(c["seed"], c["NX"], c["NY"], prec)
    out = build_ref.run(variant, workroot / variant, model=model)
    save = {}
    for k in ("prof_x", "prof_y", "prof_z", "sisvx", "sisvy", "ekin", "epot", "etotal",
              "ix_rec", "iy_rec"):
        if k in out:
            save[k] = out[k]
    if model is not None:
        save["rho"], save["mu"], save["lam"] = model
    fields = FIELDS_2D if c.get("NZ") is None else FIELDS_3D
    sub = (slice(None, None, 4),) * 2 if c.get("NZ") is None else (slice(None, None, 3),) * 3
    for k in fields:
        save[f"sha_{k}"] = np.array(field_digest(out[k]))
        save[f"sub_{k}"] = out[k][sub]
    mems = [k for k in out if k.startswith("m_")]
    for k in mems:
        a = out[k]
        nz = a != 0
        if a.ndim == 2:  # (ny, nx)
            save[f"nzmask_{k}_x"] = nz.any(axis=0)
            save[f"nzmask_{k}_y"] = nz.any(axis=1)
        else:  # (nz, ny, nx)
            save[f"nzmask_{k}_x"] = nz.any(axis=(0, 1))
            save[f"nzmask_{k}_y"] = nz.any(axis=(0, 2))
            save[f"nzmask_{k}_z"] = nz.any(axis=(1, 2))
    path = HERE / f"{case}_{prec}.npz"
    np.savez_compressed(path, **save)
    return path


IMAGE_STEPS = (5, 100, 200)


def read_pnm(path: Path):
    """Header lines (verbatim) and pixels (rows top to bottom, RGB uint8) of a
    P3 image."""
    lines = path.read_text().split("\n")
    head = "\n".join(lines[:3]) + "\n"
    nx, ny = (int(t) for t in lines[1].split())
    px = np.array([[int(t) for t in l.split()] for l in lines[3:3 + nx * ny]], np.uint8)
    return head, px.reshape(ny, nx, 3)


def make_images(workroot: Path, prec: str = "f64", prog: str = "2d2") -> Path:
    """images_<prog>_<prec>.npz: the PNM snapshots of the shipped 2D2 or 2D4
    program (200 steps) -- header text and pixels of Vx and Vy at steps 5,
    100, 200."""
    wd = workroot / f"ref{prog}_images_{prec}"
    build_ref.run(f"ref{prog}_images_{prec}", wd)
    save = {}
    for it in IMAGE_STEPS:
        if not (wd / f"image{it:06d}_Vx.pnm").is_file():
            continue  # 2D4 displays every 200 steps (2D4:187)
        for comp in ("Vx", "Vy"):
            head, px = read_pnm(wd / f"image{it:06d}_{comp}.pnm")
            save[f"{comp}_{it}"] = px
            save["header"] = np.array(head)
    path = HERE / f"images_{prog}_{prec}.npz"
    np.savez_compressed(path, **save)
    return path


DISPLAY_STEPS_3D = (5, 100, 200, 300)


def make_display3d(workroot: Path, prec: str = "f64") -> Path:
    """display_3d_<prec>.npz: what the reduced 3D program (the 3d_small
    configuration, 2 ranks) writes at its display steps 5, 100, 200, 300
    (3D:1183-1234) -- the text of each timestamp file, the PNM snapshots of
    Vx and Vy on the cut plane, and the final Vx and Vy of that plane
    (global k = NZ/2), which is what the step-300 images show."""
    wd = workroot / f"ref3d_images_{prec}"
    out = build_ref.run(f"ref3d_images_{prec}", wd)
    save = {}
    for it in DISPLAY_STEPS_3D:
        save[f"timestamp_{it}"] = np.array((wd / f"timestamp{it:06d}").read_text())
        for comp in ("Vx", "Vy"):
            head, px = read_pnm(wd / f"image{it:06d}_{comp}.pnm")
            save[f"{comp}_{it}"] = px
            save["header"] = np.array(head)
    k = out["vx"].shape[0] // 2 - 1
    save["vx_plane"] = out["vx"][k]
    save["vy_plane"] = out["vy"][k]
    save["ix_rec"], save["iy_rec"] = out["ix_rec"], out["iy_rec"]
    path = HERE / f"display_3d_{prec}.npz"
    np.savez_compressed(path, **save)
    return path


def make_visco2d(workroot: Path, order: int = 2) -> Path:
    """2dve_small_f64.npz (2dve4_small_f64.npz): the 2D viscoelastic program
    of 2nd (4th) order, attenuation on, on 161 x 161 for 700 steps: its
    profiles, tau values of the fitted SLS, medium, receivers, seismograms
    (vx, vy, pressure), and the digest and every-4th-point subsample of each
    final field, C-PML memory and attenuation memory variable."""
    tag = "2dve" if order == 2 else "2dve4"
    out = build_ref.run(f"ref{tag}_f64", workroot / f"ref{tag}_f64")
    save = {k: out[k] for k in ("prof_x", "prof_y", "sisvx", "sisvy", "sispressure",
                                "tau_eps_nu1", "tau_sig_nu1", "tau_eps_nu2", "tau_sig_nu2",
                                "lam", "mu", "rho", "ix_rec", "iy_rec")}
    for k in ("vx", "vy", "sxx", "syy", "sxy", "m_dvx_dx", "m_dvx_dy", "m_dvy_dx", "m_dvy_dy",
              "m_dsxx_dx", "m_dsyy_dy", "m_dsxy_dx", "m_dsxy_dy", "r_e1", "r_e11", "r_e13"):
        save[f"sha_{k}"] = np.array(field_digest(out[k]))
        save[f"sub_{k}"] = out[k][..., ::4, ::4]
    path = HERE / f"{tag}_small_f64.npz"
    np.savez_compressed(path, **save)
    return path


def make_ade8(workroot: Path) -> Path:
    """ade8_small_f64.npz: the ADE-PML RK4 8th-order program on 81 x 81 for
    300 steps: its stage-dependent profiles, medium and receivers (with the
    4-cell ghost border), the seismograms of its 81 receivers, and the digest
    and every-4th-point subsample of each final field and memory array."""
    out = build_ref.run("refade8_f64", workroot / "refade8_f64")
    keys = ("a_x", "b_x", "K_x", "ah_x", "bh_x", "Kh_x", "a_y", "b_y", "K_y", "ah_y", "bh_y",
            "Kh_y", "sisvx", "sisvy", "lam", "mu", "rho", "ix_rec", "iy_rec")
    save = {k: out[k] for k in keys}
    for k in ("vx", "vy", "sxx", "syy", "sxy", "m_dvx_dx", "m_dvx_dy", "m_dvy_dx", "m_dvy_dy",
              "m_dsxx_dx", "m_dsyy_dy", "m_dsxy_dx", "m_dsxy_dy"):
        save[f"sha_{k}"] = np.array(field_digest(out[k]))
        save[f"sub_{k}"] = out[k][::4, ::4]
    path = HERE / "ade8_small_f64.npz"
    np.savez_compressed(path, **save)
    return path


SU_FILES = ("seismograms/seisVx.su", "seismograms/seisVy.su", "snapshots/00030snapVx.su",
            "snapshots/00060snapVy.su")


def make_su(workroot: Path) -> Path:
    """su_ade8.npz: the SU files (raw bytes) that the 8th-order ADE-PML program
    writes on a 61 x 61 grid in 60 steps -- its two seismogram files
    (ADE8:987-1003) and a snapshot of each component (ADE8:935-953)."""
    wd = workroot / "refade8_su_f64"
    build_ref.run("refade8_su_f64", wd)
    save = {Path(f).name.replace(".su", ""): np.frombuffer((wd / f).read_bytes(), np.uint8)
            for f in SU_FILES}
    path = HERE / "su_ade8.npz"
    np.savez_compressed(path, **save)
    return path


def main(argv):
    if argv == ["ade8"]:
        build_ref.build(["refade8_f64"])
        with tempfile.TemporaryDirectory(prefix="cpml_golden_") as tmp:
            path = make_ade8(Path(tmp))
            print(f"{path.name}: {path.stat().st_size / 1024:.0f} KiB")
        return
    if argv == ["visco2d"]:
        build_ref.build(["ref2dve_f64", "ref2dve4_f64"])
        with tempfile.TemporaryDirectory(prefix="cpml_golden_") as tmp:
            for order in (2, 4):
                path = make_visco2d(Path(tmp), order)
                print(f"{path.name}: {path.stat().st_size / 1024:.0f} KiB")
        return
    if argv == ["su"]:
        build_ref.build(["refade8_su_f64"])
        with tempfile.TemporaryDirectory(prefix="cpml_golden_") as tmp:
            path = make_su(Path(tmp))
            print(f"{path.name}: {path.stat().st_size / 1024:.0f} KiB")
        return
    imgs = [(g, p) for g in ("2d2", "2d4") for p in ("f64", "f32")]
    if argv == ["images"]:
        build_ref.build([f"ref{g}_images_{p}" for g, p in imgs])
        with tempfile.TemporaryDirectory(prefix="cpml_golden_") as tmp:
            for g, p in imgs:
                path = make_images(Path(tmp), p, g)
                print(f"{path.name}: {path.stat().st_size / 1024:.0f} KiB")
        return
    if argv == ["display3d"]:
        build_ref.build([f"ref3d_images_{p}" for p in ("f64", "f32")])
        with tempfile.TemporaryDirectory(prefix="cpml_golden_") as tmp:
            for p in ("f64", "f32"):
                path = make_display3d(Path(tmp), p)
                print(f"{path.name}: {path.stat().st_size / 1024:.0f} KiB")
        return
    names = argv or list(CASES) + list(CASES_3D4)
    precs = {n: ALL_CASES[n].get("precs", ("f64", "f32")) for n in names}
    build_ref.build([f"{VARIANT_OF[n]}_{p}" for n in names for p in precs[n]])
    with tempfile.TemporaryDirectory(prefix="cpml_golden_") as tmp:
        for n in names:
            for p in precs[n]:
                path = make(n, p, Path(tmp))
                print(f"{path.name}: {path.stat().st_size / 1024:.0f} KiB")
        if not argv:
            build_ref.build([f"ref{g}_images_{p}" for g, p in imgs])
            for g, p in imgs:
                path = make_images(Path(tmp), p, g)
                print(f"{path.name}: {path.stat().st_size / 1024:.0f} KiB")
            build_ref.build([f"ref3d_images_{p}" for p in ("f64", "f32")])
            for p in ("f64", "f32"):
                path = make_display3d(Path(tmp), p)
                print(f"{path.name}: {path.stat().st_size / 1024:.0f} KiB")
            build_ref.build(["refade8_su_f64"])
            path = make_su(Path(tmp))
            print(f"{path.name}: {path.stat().st_size / 1024:.0f} KiB")
            build_ref.build(["ref2dve_f64", "ref2dve4_f64"])
            for order in (2, 4):
                path = make_visco2d(Path(tmp), order)
                print(f"{path.name}: {path.stat().st_size / 1024:.0f} KiB")
            build_ref.build(["refade8_f64"])
            path = make_ade8(Path(tmp))
            print(f"{path.name}: {path.stat().st_size / 1024:.0f} KiB")


if __name__ == "__main__":
    main(sys.argv[1:])
