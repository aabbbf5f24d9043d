"""The oracle against the compiled reference's own outputs (tests/golden).

Bar: bit-exact.  The oracle restates 2D2/2D4/3D statement by statement, and
the fixtures hold what the reference programs compute, so every profile,
receiver index, seismogram sample, energy value and final field must agree
exactly (fields through their SHA-256 digests and subsamples).
"""
import numpy as np
import pytest

from case_inputs import fixture, inputs, run_oracle
from golden.cases import CASES
from golden.make_golden import FIELDS_2D, FIELDS_3D, field_digest

ALL = [(c, p) for c in CASES for p in ("f64", "f32")]


@pytest.mark.parametrize("case,prec", ALL)
def test_profiles_and_receivers(case, prec):
    f = fixture(case, prec)
    d = inputs(case, prec, nstep=1)
    for k in ("prof_x", "prof_y", "prof_z"):
        if k in f.files:
            np.testing.assert_array_equal(d[k], f[k], err_msg=k)
    np.testing.assert_array_equal(d["ix_rec"], f["ix_rec"])
    np.testing.assert_array_equal(d["iy_rec"], f["iy_rec"])


@pytest.mark.parametrize("case,prec", ALL)
def test_time_loop_bit_exact(case, prec):
    f = fixture(case, prec)
    o = run_oracle(case, prec, nthreads=4)
    for k in ("sisvx", "sisvy", "ekin", "epot", "etotal"):
        if k in f.files:
            np.testing.assert_array_equal(o[k], f[k], err_msg=k)
    fields = FIELDS_2D if CASES[case].get("NZ") is None else FIELDS_3D
    step = 4 if CASES[case].get("NZ") is None else 3
    for k in fields:
        sub = o[k][(slice(None, None, step),) * o[k].ndim]
        np.testing.assert_array_equal(sub, f[f"sub_{k}"], err_msg=k)
        assert field_digest(o[k]) == str(f[f"sha_{k}"]), k


@pytest.mark.parametrize("case,prec", [a for a in ALL if a[1] == "f64"])
def test_memory_variables_zero_outside_slabs(case, prec):
    """Where a (and a_half) vanish the memory variable must stay exactly 0
    (2D2:221 'could declare these arrays in PML only'); the reference's own
    non-zero masks must lie inside {a != 0} of the matching profile."""
    f = fixture(case, prec)
    axes = {"x": f["prof_x"], "y": f["prof_y"]}
    if "prof_z" in f.files:
        axes["z"] = f["prof_z"]
    for key in f.files:
        if not key.startswith("nzmask_"):
            continue
        name, axis = key[len("nzmask_m_"):-2], key[-1]
        deriv_axis = name[-1]
        if axis != deriv_axis:
            continue
        prof = axes[axis]
        allowed = (prof[0] != 0) | (prof[3] != 0)
        assert not np.any(f[key] & ~allowed), key


def test_3d4_time_loop_bit_exact():
    """4th-order 3D (elastic limit of the 3D viscoelastic program, 2 ranks):
    with the reference's slab-interface reads reproduced (zcut = NZ/2) the
    oracle equals every seismogram sample, energy value and final field."""
    f = fixture("3d4_small", "f64")
    o = run_oracle("3d4_small", "f64", nthreads=4, zcut=40 // 2)
    for k in ("sisvx", "sisvy", "etotal"):
        np.testing.assert_array_equal(o[k], f[k], err_msg=k)
    for k in FIELDS_3D:
        np.testing.assert_array_equal(o[k][::3, ::3, ::3], f[f"sub_{k}"], err_msg=k)
        assert field_digest(o[k]) == str(f[f"sha_{k}"]), k


def test_3d4_reference_halo_reads_matter():
    """The single-domain stencil (zcut = 0, what the GPU computes) differs from
    the 2-rank reference by a few percent after 400 steps, and the difference
    starts at the interface (the source plane NZ/2 is the last plane of the
    lower rank): equal after step 1 (only the source has acted), after step 2
    different only on global planes zcut..zcut+3."""
    f = fixture("3d4_small", "f64")
    o = run_oracle("3d4_small", "f64", nthreads=4)
    rel = np.linalg.norm(o["sisvy"] - f["sisvy"]) / np.linalg.norm(f["sisvy"])
    assert 1e-3 < rel < 0.2, rel
    for n, planes in ((1, []), (2, [19, 20, 21, 22])):
        a = run_oracle("3d4_small", "f64", nstep=n, nthreads=4)
        b = run_oracle("3d4_small", "f64", nstep=n, nthreads=4, zcut=20)
        diff = set()
        for k in FIELDS_3D:
            diff |= set(np.nonzero((a[k] != b[k]).any(axis=(1, 2)))[0].tolist())
        assert sorted(diff) == planes, (n, sorted(diff))


@pytest.mark.parametrize("case", list(CASES))
def test_reference_f32_build_against_its_f64_build(case):
    """The bar for single precision is 1e-4 relative L2 against the
    reference's double-precision build.  The reference's own float32 build
    (the textual variant its authors prescribe, 2D2:131-133) meets it on the
    seismograms of every case; its final wavefields meet it where the wave is
    still in the grid (the layered cases) but not after the wave has left
    (shipped 2D2/2D4: up to 4.1e-3 on sigma_xx; reduced 3D: 1.5e-3 on vz),
    where the field left is rounding noise of the absorbed wave."""
    a, b = fixture(case, "f32"), fixture(case, "f64")

    def rel(x, y):
        x, y = np.asarray(x, np.float64), np.asarray(y, np.float64)
        return np.linalg.norm(x - y) / np.linalg.norm(y)

    for k in ("sisvx", "sisvy"):
        assert rel(a[k], b[k]) <= 1e-4, k
    fields = [k for k in a.keys() if k.startswith("sub_")]
    worst = max(rel(a[k], b[k]) for k in fields)
    if "layered" in case:
        assert worst <= 1e-4
    else:
        assert worst > 1e-4
