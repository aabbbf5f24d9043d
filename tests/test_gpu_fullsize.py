"""The benchmark configuration itself (3D 768^3 FP32, NPML 20, layered +-5%
smoothed medium, Ricker 8 Hz, 1024 receivers) checked through properties
that hold at any size, since the oracle cannot run it in seconds:

* the fused step and the time-blocked step (2 steps per launch) equal the
  two-phase kernels bit for bit on every field and seismogram sample after
  20 steps (three independent kernel designs, with the 32-bit per-chunk
  buffer offsets at 1.8 GB per field); fields are compared through 128-bit
  digests of their bytes;
* linearity: doubling the force doubles every field exactly (x2 is exact in
  binary floating point and every update is linear in the fields) except
  where a product or sum inside an update falls below the smallest normal
  float (1.2e-38) in one run and not in the other: there the two roundings
  differ by a subnormal step (1.4e-45), and such absolute errors, carried
  by the later steps, stay far below 1e-30 (seen: a handful of 4.5e8 values
  off by 1.4e-45, one of them next to a normal value).  Values above 1e-30
  must double bit for bit; the rest must agree to 2 x 1.2e-38;
* a zero force leaves every field exactly zero;
* with the source moved next to a corner (4 cells inside the inner edges of
  the x, y and z slabs, force x1e6 so that the front is far above underflow
  when it reaches them), the time-blocked step equals the two-phase kernels
  on every field, every one of the 18 memory variables and every seismogram
  sample: the slab classes of the time-blocked launch at full size.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("vx", "vy", "vz", "sigmaxx", "sigmayy", "sigmazz", "sigmaxy", "sigmaxz", "sigmayz")
NSTEP = 20


def _digest(a):
    import xxhash

    return xxhash.xxh3_128_hexdigest(np.ascontiguousarray(a).view(np.uint8))


@pytest.fixture(scope="module")
def workload():
    import torch

    from seismic_cpml_amd import workloads

    w = workloads.layered_3d(n=768, nstep=NSTEP)
    yield w
    w.medium_dev = None
    torch.cuda.empty_cache()


def _solver(w, scale=1.0, fused=False, time_block=0, src=None):
    s = w.make_solver()
    if scale != 1.0 or src is not None:
        fx, fy = w.force
        s.set_source(*(src or w.src), fx * np.float32(scale), fy * np.float32(scale))
    if fused:
        s.set_kchunk(48)
        s.set_tile(16)
        s.set_fused(True)
    if time_block:
        s.set_time_block(time_block)
    s.run()
    return s


@pytest.fixture(scope="module")
def two_phase(workload):
    s = _solver(workload)
    digests = {name: _digest(s.field(name)) for name in FIELDS}
    yield s, digests, s.seismograms()
    s.close()


@pytest.mark.parametrize("path", ["fused", "time_blocked"])
def test_fullsize_paths_equal_two_phase(workload, two_phase, path):
    a, digests, seis = two_phase
    assert a.velocity_norm() > 0
    b = _solver(workload, fused=path == "fused", time_block=2 if path == "time_blocked" else 0)
    try:
        for name in FIELDS:
            assert _digest(b.field(name)) == digests[name], f"{path} {name}"
        for x, y in zip(seis, b.seismograms()):
            np.testing.assert_array_equal(x, y)
    finally:
        b.close()


def test_fullsize_linearity(workload, two_phase):
    a = two_phase[0]
    c = _solver(workload, scale=2.0)
    tiny = np.finfo(np.float32).tiny
    try:
        for name in FIELDS:
            fa2, fc = 2 * a.field(name), c.field(name)
            if np.array_equal(fc, fa2):
                continue
            big = np.abs(fa2) >= 1e-30
            np.testing.assert_array_equal(fc[big], fa2[big], err_msg=f"x2 {name}")
            assert np.abs(fc[~big] - fa2[~big]).max(initial=0.0) <= 2 * tiny, name
    finally:
        c.close()


def test_fullsize_zero_force_leaves_zero_fields(workload):
    z = _solver(workload, scale=0.0)
    try:
        assert z.velocity_norm() == 0.0
        for name in FIELDS:
            assert not np.any(z.field(name)), name
    finally:
        z.close()


def test_fullsize_corner_source_time_blocked_equals_two_phase(workload):
    from gpu_runner import MEM3D

    src = (workload.npml + 4,) * 3
    out = []
    for tb in (0, 2):
        s = _solver(workload, scale=1e6, time_block=tb, src=src)
        try:
            d = {n: _digest(s.field(n)) for n in FIELDS}
            touched = 0
            for m in MEM3D:  # full-grid arrays of 1.8 GB: one at a time
                a = s.memory(m)
                d[m] = _digest(a)
                touched += int(np.any(a))
                del a
            out.append((d, s.seismograms(), touched))
        finally:
            s.close()
    (da, sa, ta), (db, sb, tb_) = out
    assert ta == tb_ and ta >= 9  # the wave reached the slabs' memory variables
    for k in da:
        assert da[k] == db[k], k
    for x, y in zip(sa, sb):
        np.testing.assert_array_equal(x, y)
