"""The Fortran driver programs (cpml_gpu module over the C ABI) against the
compiled reference's outputs.

The 2D drivers run the shipped configurations of 2D2 and 2D4; their
seismograms must equal the reference's bit-for-bit in both precisions and
their ASCII Vx_file_001.dat must match the reference's single-precision
values, and their PNM snapshots the reference's image files byte for byte.
The reduced 3D build matches the 3D golden case within the 3D tolerances
(1e-12 float64, 1e-5 float32 rel-L2), and its timestamp files and cut-plane
snapshots the reference's.  The 2D viscoelastic driver's seismograms (vx, vy,
pressure) and final fields equal the reference's bit for bit.  The 3D
4th-order driver's profiles and receivers equal the reference's, and its
seismograms the oracle's single-domain run.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from case_inputs import fixture
from gpu_runner import rel_l2

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parents[1] / "seismic_cpml_amd" / "fortran" / "bin"


def run_driver(name, tmp_path):
    exe = BIN / name
    assert exe.is_file(), f"{exe} not built (python -m seismic_cpml_amd.build)"
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "End of the simulation" in r.stdout
    return r.stdout


def read_dump(path, dt, nstep, nrec, nplane):
    raw = path.read_bytes()
    es = np.dtype(dt).itemsize
    off = 0

    def take(n, t):
        nonlocal off
        a = np.frombuffer(raw, dtype=t, count=n, offset=off)
        off += n * np.dtype(t).itemsize
        return a

    out = {"sisvx": take(nstep * nrec, dt).reshape(nrec, nstep),
           "sisvy": take(nstep * nrec, dt).reshape(nrec, nstep),
           "ekin": take(nstep, np.float64), "epot": take(nstep, np.float64),
           "vx": take(nplane, dt)}
    assert es in (4, 8)
    return out


def check_images(tmp_path, order, prec):
    """The driver's PNM snapshots at the steps the reference's images cover
    (tests/golden/images_2d[24]_<prec>.npz): the same file bytes."""
    g = np.load(Path(__file__).resolve().parent / "golden" / f"images_2d{order}_{prec}.npz")
    steps = sorted({int(k.split("_")[1]) for k in g.files if k.startswith("Vx_")})
    for it in steps:
        for comp in ("Vx", "Vy"):
            text = (tmp_path / f"image{it:06d}_{comp}.pnm").read_text()
            px = g[f"{comp}_{it}"].reshape(-1, 3)
            want = str(g["header"]) + "".join(f"{r:3d} {gg:3d} {b:3d}\n" for r, gg, b in px)
            assert text == want, f"image {it} {comp}"


@pytest.mark.parametrize("order,prec", [(2, "f64"), (2, "f32"), (4, "f64"), (4, "f32")])
def test_2d_driver_matches_reference(order, prec, tmp_path):
    name = f"seismic_cpml_2d_{'second' if order == 2 else 'fourth'}_{prec}"
    case = f"2d{order}_shipped"
    out = run_driver(name, tmp_path)
    assert "Max norm velocity vector" in out
    check_images(tmp_path, order, prec)
    f = fixture(case, prec)
    dt = np.float64 if prec == "f64" else np.float32
    nstep = 2000 if order == 2 else 4000
    d = read_dump(tmp_path / "gpu_dump.bin", dt, nstep, 2, 101 * 641)
    np.testing.assert_array_equal(d["sisvx"], f["sisvx"])
    np.testing.assert_array_equal(d["sisvy"], f["sisvy"])
    etol = 1e-12 if prec == "f64" else 1e-4
    assert rel_l2(d["ekin"], f["ekin"]) < etol
    assert rel_l2(d["epot"], f["epot"]) < etol
    np.testing.assert_array_equal(d["vx"].reshape(641, 101)[::4, ::4], f["sub_vx"])
    # the ASCII seismogram in the reference's format
    txt = np.loadtxt(tmp_path / "Vx_file_001.dat")
    assert txt.shape == (nstep, 2)
    np.testing.assert_array_equal(txt[:, 1].astype(np.float32),
                                  f["sisvx"][0].astype(np.float32))
    assert (tmp_path / "energy.dat").is_file()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_3d_driver_matches_reference(prec, tmp_path):
    run_driver(f"seismic_cpml_3d_small_{prec}", tmp_path)
    f = fixture("3d_small", prec)
    dt = np.float64 if prec == "f64" else np.float32
    d = read_dump(tmp_path / "gpu_dump.bin", dt, 300, 2, 40 * 44)
    tol = 1e-12 if prec == "f64" else 1e-5
    assert rel_l2(d["sisvx"], f["sisvx"]) < tol
    assert rel_l2(d["sisvy"], f["sisvy"]) < tol
    assert rel_l2(d["ekin"] + d["epot"], f["etotal"]) < (1e-12 if prec == "f64" else 1e-4)
    check_display_3d(tmp_path, prec)


def check_display_3d(tmp_path, prec):
    """The timestamp files and cut-plane PNM snapshots at the display steps
    against the reference's (tests/golden/display_3d_<prec>.npz).  Timestamp
    lines: the labels and the step and time text equal, the velocity norm and
    energy within the 3D tolerances (the fields match to rounding), the
    elapsed-time lines in the reference's form.  Images: every pixel equal in
    f64, at most 0.5% of the pixels in f32 (a value within rounding of a
    colour-level boundary)."""
    import re
    g = np.load(Path(__file__).resolve().parent / "golden" / f"display_3d_{prec}.npz")
    rtol = 1e-10 if prec == "f64" else 1e-5
    for it in (5, 100, 200, 300):
        ours = (tmp_path / f"timestamp{it:06d}").read_text().rstrip("\n").split("\n")
        ref = str(g[f"timestamp_{it}"]).rstrip("\n").split("\n")
        assert len(ours) == len(ref) == 7
        assert ours[:2] == ref[:2], (ours[:2], ref[:2])
        for n in (2, 3):
            la, _, va = ours[n].rpartition("=")
            lb, _, vb = ref[n].rpartition("=")
            assert la == lb
            assert abs(float(va) / float(vb) - 1) < (rtol if n == 2 else 10 * rtol), (ours[n], ref[n])
        for n in (4, 6):
            assert ours[n].rpartition("=")[0] == ref[n].rpartition("=")[0]
            float(ours[n].rpartition("=")[2])
        assert re.fullmatch(r" Elapsed time in hh:mm:ss = [ \d]{4} h \d\d m \d\d s", ours[5])
        for comp in ("Vx", "Vy"):
            text = (tmp_path / f"image{it:06d}_{comp}.pnm").read_text()
            assert text.startswith(str(g["header"]))
            px = np.array(text.split()[4:], np.int64).reshape(g[f"{comp}_{it}"].shape)
            bad = (px != g[f"{comp}_{it}"]).any(-1).sum()
            assert bad <= (0 if prec == "f64" else 0.005 * px.shape[0] * px.shape[1]), (it, comp, bad)


def test_visco2d_driver_matches_reference(tmp_path):
    """The viscoelastic driver (cpml_gpu_set_attenuation / cpml_gpu_get_pressure
    from Fortran) on the reduced 161 x 161, 700-step case: vx, vy and pressure
    seismograms and the final vx, vy equal the reference's bit for bit
    (tests/golden/2dve_small_f64.npz)."""
    from golden.make_golden import field_digest
    run_driver("seismic_cpml_2d_visco_small", tmp_path)
    g = np.load(Path(__file__).resolve().parent / "golden" / "2dve_small_f64.npz")
    raw = np.fromfile(tmp_path / "gpu_dump.bin", np.float64)
    nstep, nrec, n = 700, 3, 161
    off = 0
    for k in ("sisvx", "sisvy", "sispressure"):
        a = raw[off:off + nstep * nrec].reshape(nrec, nstep)
        off += nstep * nrec
        np.testing.assert_array_equal(a, g[k], err_msg=k)
    for k in ("vx", "vy"):
        a = raw[off:off + n * n].reshape(n, n)
        off += n * n
        assert field_digest(a) == str(g[f"sha_{k}"]), k
    assert off == raw.size


def test_3d4_driver_matches_oracle_and_reference_setup(tmp_path):
    """The 4th-order 3D driver on the reduced case (40 x 60 x 40, 400 steps):
    its profiles (K_MAX_PML 7, Rcoef 1e-4) and receiver indices equal the
    reference's (tests/golden/3d4_small_f64.npz), and its seismograms and
    final cut plane equal the oracle's single-domain run bit for bit."""
    from case_inputs import run_oracle
    run_driver("seismic_cpml_3d4_small", tmp_path)
    g = fixture("3d4_small", "f64")
    raw = (tmp_path / "gpu_dump.bin").read_bytes()
    off = 0

    def take(n, t=np.float64):
        nonlocal off
        a = np.frombuffer(raw, t, count=n, offset=off)
        off += a.nbytes
        return a

    for key, n in (("prof_x", 40), ("prof_y", 60), ("prof_z", 40)):
        np.testing.assert_array_equal(take(6 * n).reshape(6, n), g[key], err_msg=key)
    np.testing.assert_array_equal(take(3, np.int32), g["ix_rec"])
    np.testing.assert_array_equal(take(3, np.int32), g["iy_rec"])
    sisvx = take(400 * 3).reshape(3, 400)
    sisvy = take(400 * 3).reshape(3, 400)
    plane = take(40 * 60).reshape(60, 40)
    assert off == len(raw)
    o = run_oracle("3d4_small", "f64", nthreads=8, energy=False)
    np.testing.assert_array_equal(sisvx, o["sisvx"])
    np.testing.assert_array_equal(sisvy, o["sisvy"])
    np.testing.assert_array_equal(plane, o["vx"][40 // 2 - 1])


@pytest.mark.parametrize("order,prec", [(2, "f64"), (4, "f32")])
def test_2d_driver_resumes_from_checkpoint_exactly(order, prec, tmp_path):
    """The RESUME_AT build saves the fields and C-PML memory variables after a
    display step (cpml_gpu_get_field / cpml_gpu_get_memory), closes the
    handle, opens a new one and restores them (set_field / set_memory): every
    output file equals the uninterrupted run's byte for byte."""
    base = f"seismic_cpml_2d_{'second' if order == 2 else 'fourth'}_{prec}"
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    run_driver(base, tmp_path / "a")
    out = run_driver(base + "_resume", tmp_path / "b")
    assert "resumed from the checkpoint" in out
    names = sorted(p.name for p in (tmp_path / "a").iterdir())
    assert "gpu_dump.bin" in names and "energy.dat" in names
    assert names == sorted(p.name for p in (tmp_path / "b").iterdir())
    for n in names:
        assert (tmp_path / "a" / n).read_bytes() == (tmp_path / "b" / n).read_bytes(), n



def test_binding_checks_kinds_shapes_and_ghost_layout(tmp_path):
    """The generic wrappers refuse real(4) arrays on a double handle and grid
    arrays of the wrong shape (CPML_ERR_ARG with a message), and the
    reference's explicit-lower-bound arrays (ADE8 -4:NX+4, 2D4 0:NX+1) go
    through set/get of fields and memory variables in place, interior only,
    ghost cells untouched (cpml_gpu_binding_check.F90)."""
    exe = BIN / "cpml_gpu_binding_check"
    assert exe.is_file(), f"{exe} not built (python -m seismic_cpml_amd.build)"
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "BINDING OK" in r.stdout, r.stdout[-3000:] + r.stderr[-2000:]
    assert "real(4)" in r.stdout and "(41,30)" in r.stdout
