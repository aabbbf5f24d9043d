"""seismic_cpml_amd -- the seismic C-PML time loop on AMD Instinct MI355X.

Native core: ``libcpml_gpu.so`` (hand-written HIP kernels for gfx950 behind
the C ABI in include/cpml_gpu.h), bound from Fortran by the ``cpml_gpu``
module (fortran/cpml_gpu.f90) and from Python by :mod:`.solver`.
"""
# PyTorch (used here only for device-memory plumbing) ships its own HIP
# runtime; it must be loaded before libcpml_gpu initialises the system one,
# or torch later finds no device.  Import it first when it is installed.
try:  # noqa: SIM105
    import torch  # noqa: F401
except ImportError:
    pass

from .solver import (  # noqa: E402,F401
    CpmlAbcNet,
    CpmlBatch,
    CpmlError,
    CpmlGpu,
    courant_number,
    library,
    no_absorption,
    nearest_receivers,
    pml_profile,
    source_time_function,
    write_pnm,
    write_su,
)

__all__ = ["CpmlAbcNet", "CpmlBatch", "CpmlError", "CpmlGpu", "courant_number", "library",
           "nearest_receivers", "no_absorption", "pml_profile", "source_time_function",
           "write_pnm", "write_su"]
