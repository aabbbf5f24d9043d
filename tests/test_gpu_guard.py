"""NaN-poisoned guard bands: every kernel family runs once with the guard
cells of its device arrays (the row padding outside the grid and its ghost
points: cpml_gpu_guard_cells) filled with NaN, and once without.

Bars: the poisoned run's fields, memory variables and seismograms equal the
clean run's bit for bit (no guard value reaches a result: a NaN would spread
through every later step), and every guard cell still holds its NaN after the
run (no store lands outside the grid).  The clean run is itself bit-exact
against the oracle in the parity tests of each family.
"""
import numpy as np
import pytest

import layered_media as lm
from case_inputs import inputs, visco2d_inputs
from seismic_cpml_amd import CpmlGpu

pytestmark = pytest.mark.gpu


@pytest.fixture
def poisoned(monkeypatch):
    """Patch CpmlGpu.run: poison the guards before each run, check them after."""
    seen = []
    run = CpmlGpu.run

    def guarded(self, *a, **k):
        n = self.poison_guards()
        assert n > 0
        run(self, *a, **k)
        changed = self.guards_changed()
        seen.append((n, changed))
        assert changed == 0, f"{changed} of {n} guard cells overwritten"

    monkeypatch.setattr(CpmlGpu, "run", guarded)
    return seen


def _same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert np.isfinite(b[k]).all(), k
            np.testing.assert_array_equal(b[k], a[k], err_msg=k)


def _twice(fn, monkeypatch_fixture_request):
    clean = fn()
    seen = monkeypatch_fixture_request.getfixturevalue("poisoned")
    dirty = fn()
    assert seen, "run() was not called"
    _same(clean, dirty)


@pytest.mark.parametrize("order,prec,nt", [(2, "f64", 0), (4, "f32", 0), (8, "f32", 0),
                                           (2, "f32", 4), (4, "f64", 2), (8, "f32", 2)])
def test_guards_2d(order, prec, nt, request):
    from test_gpu_timeblock import _inputs, _run

    d = _inputs(order, prec, 160)
    _twice(lambda: _run(order, prec, d, nt), request)


@pytest.mark.parametrize("path", ["two_phase", "fused", "tb1", "tb2"])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_guards_3d(path, prec, request):
    from test_gpu_timeblock3d import _run

    from gpu_runner import run_gpu

    d = lm.layered3d_inputs(prec, 40, kmax=3.0)
    if path.startswith("tb"):
        fn = lambda: _run("3d_layered", prec, d, int(path[2:]))  # noqa: E731
    else:
        fn = lambda: run_gpu("3d_layered", prec, d=d, memory=True,  # noqa: E731
                             fused=path == "fused")
    _twice(fn, request)


def test_guards_3d4(request):
    from gpu_runner import run_gpu

    d = inputs("3d4_small", "f32", 120)
    _twice(lambda: run_gpu("3d4_small", "f32", d=d, memory=True), request)


@pytest.mark.parametrize("order", [2, 4])
def test_guards_visco2d(order, request):
    from test_gpu_visco2d import run_gpu as run_visco

    d, _ = visco2d_inputs("f32", order)
    _twice(lambda: run_visco(d, "f32", nstep=200), request)


def test_guards_ade8(request):
    from test_gpu_ade8 import inputs as ade_inputs
    from test_gpu_ade8 import run_gpu as run_ade

    f, d = ade_inputs("f32")
    _twice(lambda: run_ade(f, d, "f32"), request)
