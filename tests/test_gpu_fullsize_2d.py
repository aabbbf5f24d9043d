"""The 2D benchmark configurations at their full sizes, checked through
properties that hold at any size (the oracle cannot run them in seconds):

* 2D 4096^2 4th order FP32 (NPML 20, 12 seeded layers, Ricker 10 Hz, 256
  receivers on the top row), the benchmark's 4000 steps: the time-blocked
  step (2 steps per launch, 52-column strips, 40-row chunks: the benchmark's
  path) equals the single-step kernels bit for bit on every field and
  seismogram sample; fields are compared through 128-bit digests of their
  bytes.  The benchmark's centred source does not reach the absorbing slabs
  in 4000 steps, so a second case moves the source next to a corner (40
  cells inside both slabs) and runs 1500 steps, through the C-PML of two
  faces and their corner.  A zero force leaves every field exactly zero.
* The batched survey (128 shots, seeded grids of 512..2048 cells per axis,
  each with its own 12-layer model, source position and Ricker f0, 4th
  order FP32, 3000 steps): after the autotuner has picked the launch shape,
  every shot's seismograms and final fields equal that shot run alone with
  the plain solver, bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("vx", "vy", "sigmaxx", "sigmayy", "sigmaxy")


def _digest(a):
    import xxhash

    return xxhash.xxh3_128_hexdigest(np.ascontiguousarray(a).view(np.uint8))


@pytest.fixture(scope="module")
def layered():
    from seismic_cpml_amd import workloads

    return workloads.layered_2d(n=4096, nstep=4000)


def _run(w, nstep, time_block, src=None, scale=1.0):
    s = w.make_solver()
    if src is not None or scale != 1.0:
        fx, fy = w.force
        i, j = src if src is not None else w.src[:2]
        s.set_source(i, j, 1, fx * np.float32(scale), fy * np.float32(scale))
    if time_block:
        s.set_time_block(time_block)
    s.run(1, nstep)
    out = {name: _digest(s.field(name)) for name in FIELDS}
    out["seis"] = s.seismograms()
    out["vmax"] = s.velocity_norm()
    s.close()
    return out


@pytest.mark.parametrize("case", ["benchmark", "corner"])
def test_fullsize_2d_time_blocked_equals_single_step(layered, case):
    src, nstep = (None, 4000) if case == "benchmark" else ((layered.npml + 40,) * 2, 1500)
    a = _run(layered, nstep, 0, src)
    b = _run(layered, nstep, 2, src)
    assert a["vmax"] > 0
    for name in FIELDS:
        assert a[name] == b[name], name
    for x, y in zip(a["seis"], b["seis"]):
        np.testing.assert_array_equal(x, y)


def test_fullsize_2d_zero_force_leaves_zero_fields(layered):
    s = layered.make_solver()
    fx, fy = layered.force
    s.set_source(*layered.src[:2], 1, np.zeros_like(fx), np.zeros_like(fy))
    s.set_time_block(2)
    s.run(1, 200)
    for name in FIELDS:
        assert not np.any(s.field(name)), name
    s.close()


def test_fullsize_batch_matches_single_runs():
    import torch

    from seismic_cpml_amd import workloads

    w = workloads.batch_2d(nshots=128, nstep=3000)
    b = w.make_batch()
    b.autotune(trial_steps=20)
    b.run(1, w.nstep)
    torch.cuda.synchronize()
    moved = 0
    for q in range(len(w.shots)):
        h = w.make_solver(q)
        h.run()
        sx, sy = h.seismograms()
        bx, by = b.seismograms(q)
        np.testing.assert_array_equal(bx, sx, err_msg=f"shot {q} vx")
        np.testing.assert_array_equal(by, sy, err_msg=f"shot {q} vy")
        for name in FIELDS:
            assert _digest(b.field(q, name)) == _digest(h.field(name)), (q, name)
        moved += int(h.velocity_norm() > 0)
        h.close()
    b.close()
    assert moved == len(w.shots)
