"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, validates arguments on the host before touching HIP, and its
host-side set-up helpers reproduce the reference (through the oracle, which
is pinned bit-exactly to the compiled reference)."""
import re
from pathlib import Path

import numpy as np
import pytest

import seismic_cpml_amd as sc
from seismic_cpml_amd import solver

import oracle_lib as orc
from case_inputs import fixture
from golden.cases import CASES, values

ROOT = Path(__file__).resolve().parents[1]


def header_symbols():
    text = (ROOT / "include" / "cpml_gpu.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(cpml_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = sc.library()
    syms = header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(solver.SIGNATURES), set(syms) ^ set(solver.SIGNATURES)


def _init(*args):
    import ctypes as C

    h = C.c_void_p()
    rc = sc.library().cpml_gpu_init(C.byref(h), *args)
    return rc, h


@pytest.mark.parametrize("args,code", [
    ((4, 2, 8, 10, 10, 1, 1.0, 1.0, 1.0, 1e-3, 2, 10), solver.CPML_ERR_ARG),   # ndim
    ((2, 6, 8, 10, 10, 1, 1.0, 1.0, 1.0, 1e-3, 2, 10), solver.CPML_ERR_ARG),   # order
    ((3, 8, 8, 10, 10, 10, 1.0, 1.0, 1.0, 1e-3, 2, 10), solver.CPML_ERR_ARG),  # 3D order 8
    ((2, 2, 2, 10, 10, 1, 1.0, 1.0, 1.0, 1e-3, 2, 10), solver.CPML_ERR_ARG),   # prec
    ((2, 2, 8, 1, 10, 1, 1.0, 1.0, 1.0, 1e-3, 2, 10), solver.CPML_ERR_STENCIL),
    ((2, 2, 8, 10, 10, 1, -1.0, 1.0, 1.0, 1e-3, 2, 10), solver.CPML_ERR_ARG),
    ((2, 2, 8, 10, 10, 1, 1.0, 1.0, 1.0, 0.0, 2, 10), solver.CPML_ERR_ARG),
    ((2, 2, 8, 10, 10, 1, 1.0, 1.0, 1.0, 1e-3, 2, 0), solver.CPML_ERR_ARG),    # NSTEP
    ((2, 2, 8, 10, 10, 1, 1.0, 1.0, 1.0, 1e-3, 6, 10), solver.CPML_ERR_STENCIL),  # PML too wide
])
def test_init_validates_on_host(args, code):
    rc, h = _init(*args)
    assert rc == code
    assert not h.value
    msg = sc.solver._err(None, rc).message
    assert msg


def test_init_without_gpu_reports_nodevice():
    import torch

    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(sc.CpmlError) as e:
        sc.CpmlGpu(2, 2, 8, 32, 32, 1, 10.0, 10.0, 10.0, 1e-3, 4, 10)
    assert e.value.code == solver.CPML_ERR_NODEVICE


def test_courant_number_matches_reference_formula():
    # 2D2:513 with the shipped parameters: 3300 * 2e-3 * sqrt(2)/10
    c = sc.courant_number(2, 3300.0, 2e-3, 10.0, 10.0)
    assert c == 3300.0 * 2e-3 * np.sqrt(1.0 / 10.0 ** 2 + 1.0 / 10.0 ** 2)
    c3 = sc.courant_number(3, 3300.0, 1.6e-3, 10.0, 10.0, 10.0)
    assert abs(c3 - 3300 * 1.6e-3 * np.sqrt(3) / 10) < 1e-15


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_host_profile_helper_matches_reference(case, prec):
    """The library's host profile builder against the reference's own profiles."""
    v = values(case, prec)
    c = CASES[case]
    f = fixture(case, prec)
    p = 8 if prec == "f64" else 4
    px = sc.pml_profile(c["NX"], v["DELTAX"], v["DELTAT"], c["NPML"], v["cp"], v["f0"], p,
                        clamp_alpha=True)
    py = sc.pml_profile(c["NY"], v["DELTAY"], v["DELTAT"], c["NPML"], v["cp"], v["f0"], p,
                        origin_max_uses_n=c.get("yorigin_n", False), clamp_alpha=False)
    np.testing.assert_array_equal(px, f["prof_x"])
    np.testing.assert_array_equal(py, f["prof_y"])
    if "prof_z" in f.files:
        pz = sc.pml_profile(c["NZ"], v["DELTAZ"], v["DELTAT"], c["NPML"], v["cp"], v["f0"], p,
                            clamp_alpha=False)
        np.testing.assert_array_equal(pz, f["prof_z"])


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_host_stf_and_receivers_match_oracle(case, prec):
    v = values(case, prec)
    c = CASES[case]
    p = 8 if prec == "f64" else 4
    kind = 2 if c["src"] == "ricker" else 1
    fx, fy = sc.source_time_function(c["NSTEP"], v["DELTAT"], v["f0"], v["t0"], v["factor"],
                                     v["ANGLE"], kind, p)
    ox, oy = orc.stf(c["NSTEP"], v["DELTAT"], v["f0"], v["t0"], v["factor"], v["ANGLE"], kind,
                     prec)
    np.testing.assert_array_equal(fx, ox)
    np.testing.assert_array_equal(fy, oy)
    ix, iy = sc.nearest_receivers(c["NX"], c["NY"], v["DELTAX"], v["DELTAY"], c["NREC"],
                                  v["xdeb"], v["ydeb"], v["xfin"], v["yfin"], p)
    f = fixture(case, prec)
    np.testing.assert_array_equal(ix, f["ix_rec"])
    np.testing.assert_array_equal(iy, f["iy_rec"])


def _batch_init(order, prec, nx, ny, dx, dt, npml, nstep):
    import ctypes as C

    n = len(nx)
    arr = lambda v, t: np.ascontiguousarray(v, t)  # noqa: E731
    NX, NY, NP = arr(nx, np.int32), arr(ny, np.int32), arr(npml, np.int32)
    DX, DT = arr(dx, np.float64), arr(dt, np.float64)
    h = C.c_void_p()
    rc = sc.library().cpml_batch_init(C.byref(h), order, prec, n, solver._ptr(NX),
                                      solver._ptr(NY), solver._ptr(DX), solver._ptr(DX),
                                      solver._ptr(DT), solver._ptr(NP), nstep)
    return rc, h


@pytest.mark.parametrize("args,code", [
    ((6, 8, [32], [32], [1.0], [1e-3], [4], 10), solver.CPML_ERR_ARG),        # order
    ((2, 2, [32], [32], [1.0], [1e-3], [4], 10), solver.CPML_ERR_ARG),        # prec
    ((2, 8, [], [], [], [], [], 10), solver.CPML_ERR_ARG),                    # no shots
    ((2, 8, [32, 1], [32, 9], [1.0, 1.0], [1e-3] * 2, [4, 0], 10), solver.CPML_ERR_STENCIL),
    ((2, 8, [32, 32], [32, 32], [1.0, -1.0], [1e-3] * 2, [4, 4], 10), solver.CPML_ERR_ARG),
    ((4, 4, [32, 32], [32, 32], [1.0, 1.0], [1e-3, 0.0], [4, 4], 10), solver.CPML_ERR_ARG),
    ((2, 8, [32, 32], [32, 20], [1.0, 1.0], [1e-3] * 2, [4, 11], 10), solver.CPML_ERR_STENCIL),
    ((2, 8, [32], [32], [1.0], [1e-3], [4], 0), solver.CPML_ERR_ARG),         # NSTEP
])
def test_batch_init_validates_on_host(args, code):
    rc, h = _batch_init(*args)
    assert rc == code
    assert not h.value
    assert solver._err(None, rc, "cpml_batch_last_error").message


def test_batch_without_gpu_reports_nodevice():
    import torch

    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(sc.CpmlError) as e:
        sc.CpmlBatch(2, 8, [dict(NX=32, NY=32, DELTAX=10.0, DELTAY=10.0, DELTAT=1e-3,
                                 NPOINTS_PML=4)], 10)
    assert e.value.code == solver.CPML_ERR_NODEVICE


def test_pnm_writer_matches_reference_file_format(tmp_path):
    """cpml_pnm_write (host only) writes the reference's P3 text: the header
    and pixel lines of tests/golden/images_2d2_f64.npz byte for byte."""
    from pathlib import Path
    from seismic_cpml_amd import write_pnm
    g = np.load(Path(__file__).resolve().parent / "golden" / "images_2d2_f64.npz")
    px = g["Vx_100"]
    write_pnm(tmp_path / "a.pnm", px)
    want = str(g["header"]) + "".join(f"{r:3d} {gg:3d} {b:3d}\n" for r, gg, b in px.reshape(-1, 3))
    assert (tmp_path / "a.pnm").read_text() == want


def _su_traces(raw, ns):
    """(header words, traces) of an SU file with ns samples per trace."""
    rec = np.frombuffer(raw, np.uint8).reshape(-1, 240 + 4 * ns)
    return rec[:, :240].copy().view(np.int16), rec[:, 240:].copy().view(np.float32)


@pytest.mark.parametrize("name,ns,dt_value", [
    ("seisVx", 60, 3.e-3 * 1e6), ("seisVy", 60, 3.e-3 * 1e6),
    ("00030snapVx", 61, 10. * 1e3), ("00060snapVy", 61, 10. * 1e3)])
def test_su_writer_matches_reference_files(name, ns, dt_value, tmp_path):
    """cpml_su_write (host only) against the SU files the 8th-order program
    wrote (tests/golden/su_ade8.npz, a 61 x 61 run of 60 steps): its traces,
    written again with the reference's header values (ns; DELTAT*1e6 for
    seismograms, DELTAY*1e3 for snapshots), give the same bytes, from float32
    traces and from float64 ones (which round to the same floats)."""
    from seismic_cpml_amd import write_su
    raw = np.load(Path(__file__).resolve().parent / "golden" / "su_ade8.npz")[name].tobytes()
    head, tr = _su_traces(raw, ns)
    assert tr.shape[0] == 61 and (head[:, 57] == ns).all()
    assert np.count_nonzero(head) == 2 * 61  # only words 58 and 59 are set
    assert np.abs(tr).max() > 0
    for a in (tr, tr.astype(np.float64)):
        write_su(tmp_path / "a.su", a, dt_value)
        assert (tmp_path / "a.su").read_bytes() == raw


def test_su_writer_rounds_and_refuses(tmp_path):
    """float64 samples are rounded to nearest float32 (Fortran real()); the dt
    word is truncated toward zero (integer(2) assignment); an ns or dt word
    outside 0..32767 is refused rather than wrapped."""
    from seismic_cpml_amd import write_su
    x = np.array([[1 + 2.0 ** -24 + 2.0 ** -40, -1e-30, 3.0]])  # just above a float tie
    write_su(tmp_path / "b.su", x, 2999.9)
    head, tr = _su_traces((tmp_path / "b.su").read_bytes(), 3)
    assert head[0, 58] == 2999 and head[0, 57] == 3
    np.testing.assert_array_equal(tr[0], x[0].astype(np.float32))
    assert tr[0, 0] == np.float32(1 + 2.0 ** -23)
    for bad in ((np.zeros((1, 32768)), 1.0), (x, 32768.0), (x, -1.0)):
        with pytest.raises(solver.CpmlError):
            write_su(tmp_path / "c.su", *bad)


def test_noalloc_norm_refuses_bad_arguments_before_any_gpu_call():
    """cpml_velocity_norm_noalloc checks its arguments on the host: a grid
    above its 2^36-point limit and an empty axis are refused with
    CPML_ERR_ARG and a message, before any HIP call (none happens here)."""
    import ctypes as C

    from seismic_cpml_amd.solver import library

    lib = library()
    p = C.c_void_p(256)
    assert lib.cpml_velocity_norm_noalloc(4, 3, 1 << 13, 1 << 13, 1 << 12, p, p, p, p,
                                          None) == 1
    buf = C.create_string_buffer(256)
    lib.cpml_gpu_last_error(None, buf, 256)
    assert b"2^36" in buf.value
    assert lib.cpml_velocity_norm_noalloc(4, 3, 0, 4, 4, p, p, p, p, None) == 1
    assert lib.cpml_velocity_norm_noalloc(8, 2, 4, 4, 1, p, None, None, p, None) == 1
    assert lib.cpml_velocity_norm_noalloc(2, 2, 4, 4, 1, p, p, None, p, None) == 1
