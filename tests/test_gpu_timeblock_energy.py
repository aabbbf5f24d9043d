"""Time blocking with the energy curve on (cpml_gpu_set_energy with a cadence
that is a multiple of the steps per launch: the energy of each block's last
step is computed from the state the block leaves in memory).

Bars: with energy every 2 steps, the time-blocked run (1 and 2 steps per
launch) gives the single-step run's energy curve, fields and seismograms bit
for bit, in 2D (orders 2, 4, 8) and 3D; with energy every step, 2 steps per
launch are not used and the run equals the single-step run.
"""
import numpy as np
import pytest

import layered_media as lm
from golden.cases import CASES
from gpu_runner import F3D, make_solver
from test_gpu_timeblock import _case, _inputs

pytestmark = pytest.mark.gpu


def _run2d(order, d, nt, cadence):
    c = CASES[_case(order)]
    saved = c["order"]
    c["order"] = order
    try:
        with make_solver(_case(order), "f32", d, energy=False, graph_steps=16) as s:
            s.set_energy(cadence, window=d["ewin"])
            s.set_time_block(nt)
            s.run()
            kin, pot = s.energy()
            return kin, pot, s.field("vx"), s.field("sigmaxy"), s.seismograms()[0]
    finally:
        c["order"] = saved


@pytest.mark.parametrize("order", [2, 4, 8])
@pytest.mark.parametrize("nt,cadence", [(1, 2), (2, 2), (2, 4), (2, 1)])
def test_energy_with_time_blocks_2d(order, nt, cadence):
    d = _inputs(order, "f32", 160)
    a = _run2d(order, d, 0, cadence)
    b = _run2d(order, d, nt, cadence)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
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
