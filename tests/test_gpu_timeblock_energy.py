"""Time blocking with the energy curve on.

2D: the time-blocked kernel sums the energy terms of every level it computes
(each wavefront over its owned points in the window, the items' sums then
added in a fixed order), so any cadence is kept, the reference's every-step
curve included.  Bars: fields and seismograms bit-exact against the
single-step run; the energy curve within 1e-12 relative of the single-step
curve (the same terms per point, in the field precision, summed in double in
another order), at exactly the same steps.
3D: the energy of each block's last step comes from the state in memory, so
blocks are used when the cadence is a multiple of the steps per launch.
Bar: energy, fields and seismograms bit for bit against the single-step run.
"""
import numpy as np
import pytest

import layered_media as lm
from golden.cases import CASES
from gpu_runner import F3D, make_solver
from test_gpu_timeblock import _case, _inputs

pytestmark = pytest.mark.gpu


def _run2d(order, d, nt, cadence, prec="f32"):
    c = CASES[_case(order)]
    saved = c["order"]
    c["order"] = order
    try:
        with make_solver(_case(order), prec, d, energy=False, graph_steps=16) as s:
            s.set_energy(cadence, window=d["ewin"])
            s.set_time_block(nt)
            s.run()
            kin, pot = s.energy()
            return kin, pot, s.field("vx"), s.field("sigmaxy"), s.seismograms()[0]
    finally:
        c["order"] = saved


@pytest.mark.parametrize("order,prec", [(2, "f32"), (4, "f32"), (8, "f32"), (4, "f64")])
@pytest.mark.parametrize("nt,cadence", [(1, 1), (2, 1), (2, 3), (3, 1)])
def test_energy_with_time_blocks_2d(order, prec, nt, cadence):
    if nt > {2: 4, 4: 3, 8: 2}[order] or (prec == "f64" and nt > 2):
        pytest.skip("steps per launch beyond what this order and precision build")
    d = _inputs(order, prec, 160)
    a = _run2d(order, d, 0, cadence, prec)
    b = _run2d(order, d, nt, cadence, prec)
    for x, y in zip(a[2:], b[2:]):  # fields and seismograms
        np.testing.assert_array_equal(x, y)
    for x, y in zip(a[:2], b[:2]):  # kinetic, potential
        assert np.array_equal(x != 0, y != 0)
        np.testing.assert_allclose(y, x, rtol=1e-12, atol=0)
    kin = a[0]
    assert np.count_nonzero(kin) == np.count_nonzero(kin[cadence - 1::cadence]) > 0


def _run3d(d, nt, cadence):
    with make_solver("3d_layered", "f32", d, energy=False, graph_steps=16) as s:
        s.set_energy(cadence)
        s.set_time_block(nt)
        s.run()
        kin, pot = s.energy()
        return [kin, pot] + [s.field(n) for n in F3D] + list(s.seismograms())


def test_energy_with_time_blocks_3d():
    d = lm.layered3d_inputs("f32", 80, kmax=3.0)
    a = _run3d(d, 0, 2)
    b = _run3d(d, 2, 2)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert np.count_nonzero(a[0]) > 0
