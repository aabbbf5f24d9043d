"""cpml_velocity_norm_noalloc (the allocation-free variant of
cpml_gpu_velocity_norm, partial maxima in the library's fixed device array)
against cpml_gpu_velocity_norm through the C ABI: equal bit for bit (a
maximum of the same squared norms and one square root), on two calls in a
row, the second on a smaller grid; the scratch array keeps the first call's
partial maxima, so a stale entry would show."""
import ctypes as C

import pytest

from case_inputs import inputs
from gpu_runner import make_solver

pytestmark = pytest.mark.gpu


def _noalloc_norm(s, stream):
    import torch

    from seismic_cpml_amd.solver import _check, library

    names = ("vx", "vy", "vz") if s.ndim == 3 else ("vx", "vy")
    dev = [torch.from_numpy(s.field(n)).cuda() for n in names]
    out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ptr = [d.data_ptr() for d in dev] + ([0] if s.ndim == 2 else [])
    _check(library().cpml_velocity_norm_noalloc(
        s.prec, s.ndim, s.nx, s.ny, s.nz, C.c_void_p(ptr[0]), C.c_void_p(ptr[1]),
        C.c_void_p(ptr[2]) if ptr[2] else None, C.c_void_p(out.data_ptr()),
        C.c_void_p(stream.cuda_stream)))
    stream.synchronize()
    return float(out.item())


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_norm_equals_velocity_norm(prec):
    import torch

    st = torch.cuda.Stream()
    # a large grid first (many partial blocks), then a smaller one
    for case, nstep in (("3d_wide", 120), ("2d2_layered", 150)):
        d = inputs(case, prec, nstep)
        with make_solver(case, prec, d, energy=False) as s:
            s.run()
            ref = s.velocity_norm()
            got = _noalloc_norm(s, st)
        assert ref > 0
        assert got == ref, (case, got, ref)
