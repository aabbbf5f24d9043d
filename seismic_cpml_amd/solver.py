"""Python mirror of the ``cpml_gpu`` interface over the C ABI (ctypes).

The reference programs configure the time loop with Fortran ``parameter``s
and run it inline; ``CpmlGpu`` takes the same names as arguments and runs the
loop on the GPU through libcpml_gpu.so.  Errors raise :class:`CpmlError`
carrying the library's status code and message (the reference's ``stop``
strings for the CFL and blow-up checks).

Arrays are numpy arrays in C order with the *last* axis x, i.e. ``(ny, nx)``
or ``(nz, ny, nx)`` -- the same memory as the reference's Fortran
``(NX, NY[, NZ])`` arrays.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libcpml_gpu.so"

CPML_OK = 0
CPML_ERR_ARG = 1
CPML_ERR_CFL = 2
CPML_ERR_STENCIL = 3
CPML_ERR_OOM = 4
CPML_ERR_HIP = 5
CPML_ERR_BLOWUP = 6
CPML_ERR_STATE = 7
CPML_ERR_NODEVICE = 8

FIELDS = ("vx", "vy", "vz", "sigmaxx", "sigmayy", "sigmazz", "sigmaxy", "sigmaxz", "sigmayz")
MEMORY = ("dvx_dx", "dvx_dy", "dvx_dz", "dvy_dx", "dvy_dy", "dvy_dz", "dvz_dx", "dvz_dy",
          "dvz_dz", "dsigmaxx_dx", "dsigmayy_dy", "dsigmazz_dz", "dsigmaxy_dx", "dsigmaxy_dy",
          "dsigmaxz_dx", "dsigmaxz_dz", "dsigmayz_dy", "dsigmayz_dz")

# (name, restype, argtypes) of every exported symbol declared in include/cpml_gpu.h
_VP, _I, _D = C.c_void_p, C.c_int, C.c_double
_IP, _DP = C.POINTER(C.c_int), C.POINTER(C.c_double)
SIGNATURES = {
    "cpml_gpu_init": [C.POINTER(_VP), _I, _I, _I, _I, _I, _I, _D, _D, _D, _D, _I, _I],
    "cpml_gpu_finalize": [_VP],
    "cpml_gpu_set_medium": [_VP, _VP, _VP, _VP],
    "cpml_gpu_set_medium_device": [_VP, _VP, _VP, _VP],
    "cpml_gpu_set_pml": [_VP] + [_VP] * 18,
    "cpml_gpu_set_source": [_VP, _I, _I, _I, _VP, _VP],
    "cpml_gpu_set_receivers": [_VP, _I, _VP, _VP, _VP],
    "cpml_gpu_set_energy": [_VP, _I, _VP, _I],
    "cpml_gpu_set_stability": [_VP, _I, _D],
    "cpml_gpu_run": [_VP, _I, _I],
    "cpml_gpu_get_seismograms": [_VP, _VP, _VP, _VP],
    "cpml_gpu_get_energy": [_VP, _VP, _VP],
    "cpml_gpu_get_field": [_VP, _I, _VP],
    "cpml_gpu_set_field": [_VP, _I, _VP],
    "cpml_gpu_get_memory": [_VP, _I, _VP],
    "cpml_gpu_set_memory": [_VP, _I, _VP],
    "cpml_gpu_velocity_norm": [_VP, _DP],
    "cpml_velocity_norm_noalloc": [_I, _I, _I, _I, _I, _VP, _VP, _VP, _VP, _VP],
    "cpml_gpu_color_image": [_VP, C.c_int, C.c_void_p],
    "cpml_pnm_write": [C.c_char_p, C.c_int, C.c_int, C.c_void_p],
    "cpml_su_write": [C.c_char_p, _I, _I, C.c_double, C.c_void_p, _I],
    "cpml_gpu_set_attenuation": [_VP, _I, _VP, _VP, _VP, _VP],
    "cpml_gpu_get_attenuation_memory": [_VP, _I, _VP],
    "cpml_gpu_get_pressure": [_VP, _VP],
    "cpml_gpu_set_ade_rk4": [_VP] + [_VP] * 10,
    "cpml_gpu_set_graph_steps": [_VP, _I],
    "cpml_gpu_set_kchunk": [_VP, _I],
    "cpml_gpu_set_tile": [_VP, _I],
    "cpml_gpu_set_fused": [_VP, _I],
    "cpml_gpu_set_time_block": [_VP, _I, _I, _I],
    "cpml_gpu_autotune": [_VP, _I, _IP, _IP, _DP],
    "cpml_gpu_set_kernel_timing": [_VP, _I],
    "cpml_gpu_kernel_time": [_VP, _I, _IP, _DP],
    "cpml_gpu_elapsed_ms": [_VP, _DP],
    "cpml_gpu_device_bytes": [_VP, _DP],
    "cpml_gpu_slab_widths": [_VP, _IP],
    "cpml_gpu_last_error": [_VP, C.c_char_p, _I],
    "cpml_gpu_check_kind": [_VP, _I],
    "cpml_batch_check_kind": [_VP, _I],
    "cpml_gpu_grid": [_VP, _IP, _IP, _IP],
    "cpml_gpu_set_host_layout": [_VP, _I, _I],
    "cpml_gpu_guard_cells": [_VP, _I, C.POINTER(C.c_longlong)],
    "cpml_gpu_check_grid_array": [_VP, _I, _I, C.POINTER(C.c_int64)],
    "cpml_batch_grid": [_VP, _I, _IP, _IP],
    "cpml_batch_init": [C.POINTER(_VP), _I, _I, _I, _VP, _VP, _VP, _VP, _VP, _VP, _I],
    "cpml_batch_finalize": [_VP],
    "cpml_batch_set_medium": [_VP, _I, _VP, _VP, _VP],
    "cpml_batch_set_pml": [_VP, _I] + [_VP] * 12,
    "cpml_batch_set_source": [_VP, _I, _I, _I, _VP, _VP],
    "cpml_batch_set_receivers": [_VP, _I, _I, _VP, _VP],
    "cpml_batch_set_tiling": [_VP, _I, _I, _I],
    "cpml_batch_autotune": [_VP, _I, C.c_char_p, _IP, _IP, _DP],
    "cpml_batch_set_time_block": [_VP, _I, _I, _I],
    "cpml_batch_launch_shape": [_VP, _IP, _IP, _IP, _IP],
    "cpml_batch_set_energy": [_VP, _I],
    "cpml_batch_get_energy": [_VP, _I, _VP, _VP],
    "cpml_batch_set_stability": [_VP, _I, _D],
    "cpml_batch_run": [_VP, _I, _I],
    "cpml_batch_get_seismograms": [_VP, _I, _VP, _VP],
    "cpml_batch_get_field": [_VP, _I, _I, _VP],
    "cpml_batch_elapsed_ms": [_VP, _DP],
    "cpml_batch_last_error": [_VP, C.c_char_p, _I],
    "cpml_abc_net_create": [C.POINTER(_VP), _VP, _VP],
    "cpml_abc_net_destroy": [_VP],
    "cpml_abc_net_get_weights": [_VP, _VP],
    "cpml_abc_net_set_weights": [_VP, _VP],
    "cpml_gpu_set_boundary": [_VP, _I, _VP],
    "cpml_abc_train_init": [_VP, _D, _D, _D, _D, _I],
    "cpml_abc_train_step": [_VP, _VP, _VP, _I, _I, _I, _I, _VP],
    "cpml_abc_train_noise": [_VP, _D],
    "cpml_abc_grad": [_VP, _VP, _VP, _I, _I, _I, _VP, _VP],
    "cpml_abc_correction": [_VP, _VP, _I, _VP],
    "cpml_abc_last_error": [_VP, C.c_char_p, _I],
    "cpml_pml_profile": [_I, _I, _D, _D, _I, _D, _D, _I, _I, _I, _I] + [_VP] * 6,
    "cpml_source_time_function": [_I, _I, _D, _D, _D, _D, _D, _I, _VP, _VP],
    "cpml_nearest_receivers": [_I, _I, _I, _D, _D, _I, _D, _D, _D, _D, _VP, _VP],
    "cpml_courant_number": [_I, _D, _D, _D, _D, _D, _DP],
}

_lib = None


class CpmlError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


def library():
    """Load libcpml_gpu.so; fails loudly when the HIP extension is not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.is_file():
            raise RuntimeError(f"{LIB_PATH} not found: build it with "
                               "`python -m seismic_cpml_amd.build` (hipcc, gfx950)")
        lib = C.CDLL(str(LIB_PATH))
        for name, args in SIGNATURES.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = C.c_int
        _lib = lib
    return _lib


def write_pnm(path, rgb):
    """Write rgb (rows, columns, 3) uint8, top row first, as the reference's
    PNM P3 text file (2D2:900-903, 989)."""
    rgb = np.ascontiguousarray(rgb, np.uint8)
    ny, nx = rgb.shape[:2]
    rc = library().cpml_pnm_write(str(path).encode(), nx, ny, _ptr(rgb))
    if rc:
        raise CpmlError(rc, f"cannot write {path}")


def write_su(path, traces, dt_value):
    """Write traces (ntr, ns) float32/float64 as the 8th-order program's SU
    file (ADE8:987-1003, 935-953): header word 58 = ns, word 59 =
    int(dt_value) -- DELTAT*1e6 for seismograms, DELTAY*1e3 for snapshots."""
    a = np.asarray(traces)
    a = np.ascontiguousarray(a, np.float64 if a.dtype == np.float64 else np.float32)
    if a.ndim != 2:
        raise ValueError("traces must be (ntr, ns)")
    rc = library().cpml_su_write(str(path).encode(), a.shape[0], a.shape[1], float(dt_value),
                                 _ptr(a), a.itemsize)
    if rc:
        raise CpmlError(rc, f"cannot write {path} (ns and the dt word must be 0..32767)")


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _err(h, code, getter="cpml_gpu_last_error"):
    buf = C.create_string_buffer(1024)
    getattr(library(), getter)(h, buf, 1024)
    return CpmlError(code, buf.value.decode(errors="replace"))


def _check(code, h=None, getter="cpml_gpu_last_error"):
    if code != CPML_OK:
        raise _err(h, code, getter)


def _dtype(prec):
    return np.float32 if prec == 4 else np.float64


# ---------------------------------------------------------------- helpers --

def pml_profile(n, delta, deltat, npoints_pml, cp, f0, prec=8, use_min=True, use_max=True,
                origin_max_uses_n=False, clamp_alpha=True):
    """Damping profile along one axis (reference 2D2:283-465).  Returns a
    (6, n) array: a, b, K, a_half, b_half, K_half."""
    out = np.zeros((6, n), _dtype(prec))
    _check(library().cpml_pml_profile(prec, n, float(delta), float(deltat), npoints_pml,
                                      float(cp), float(f0), int(use_min), int(use_max),
                                      int(origin_max_uses_n), int(clamp_alpha),
                                      *[_ptr(out[q]) for q in range(6)]))
    return out


def source_time_function(nstep, deltat, f0, t0, factor, angle_force, kind=1, prec=8):
    """force_x, force_y per step (reference 2D2:644-657); kind 0 Gaussian,
    1 first derivative of a Gaussian, 2 Ricker."""
    fx = np.zeros(nstep, _dtype(prec))
    fy = np.zeros(nstep, _dtype(prec))
    _check(library().cpml_source_time_function(prec, nstep, float(deltat), float(f0),
                                               float(t0), float(factor), float(angle_force),
                                               kind, _ptr(fx), _ptr(fy)))
    return fx, fy


def nearest_receivers(nx, ny, deltax, deltay, nrec, xdeb, ydeb, xfin, yfin, prec=8):
    """1-based grid indices of receivers on a line (reference 2D2:483-509)."""
    ix = np.zeros(max(nrec, 1), np.int32)
    iy = np.zeros(max(nrec, 1), np.int32)
    _check(library().cpml_nearest_receivers(prec, nx, ny, float(deltax), float(deltay), nrec,
                                            float(xdeb), float(ydeb), float(xfin), float(yfin),
                                            _ptr(ix), _ptr(iy)))
    return ix[:nrec], iy[:nrec]


def courant_number(ndim, cp, deltat, deltax, deltay, deltaz=1.0):
    out = C.c_double()
    _check(library().cpml_courant_number(ndim, float(cp), float(deltat), float(deltax),
                                         float(deltay), float(deltaz), C.byref(out)))
    return out.value


# ----------------------------------------------------------------- solver --

class CpmlGpu:
    """One C-PML time-loop instance on the current GPU (cpml_gpu_init ...)."""

    def __init__(self, ndim, order, prec, NX, NY, NZ, DELTAX, DELTAY, DELTAZ, DELTAT,
                 NPOINTS_PML, NSTEP):
        self.ndim, self.order, self.prec = ndim, order, prec
        self.nx, self.ny, self.nz = NX, NY, (NZ if ndim == 3 else 1)
        self.nstep = NSTEP
        self.dtype = _dtype(prec)
        self.nrec = 0
        self.nsls = 0
        self._h = C.c_void_p()
        _check(library().cpml_gpu_init(C.byref(self._h), ndim, order, prec, NX, NY,
                                       NZ if ndim == 3 else 1, float(DELTAX), float(DELTAY),
                                       float(DELTAZ), float(DELTAT), NPOINTS_PML, NSTEP))

    # context management
    def close(self):
        if self._h:
            library().cpml_gpu_finalize(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.ny, self.nx) if self.ndim == 2 else (self.nz, self.ny, self.nx)

    def _grid(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        if a.shape != self.shape:
            raise ValueError(f"array shape {a.shape} != grid {self.shape}")
        return a

    def _vec(self, a, n):
        a = np.ascontiguousarray(a, self.dtype)
        if a.shape != (n,):
            raise ValueError(f"profile length {a.shape} != {n}")
        return a

    def _call(self, fn, *args):
        _check(getattr(library(), fn)(self._h, *args), self._h)

    def set_medium(self, lam, mu, rho):
        lam, mu, rho = self._grid(lam), self._grid(mu), self._grid(rho)
        self._call("cpml_gpu_set_medium", _ptr(lam), _ptr(mu), _ptr(rho))

    def set_medium_device(self, d_lam, d_mu, d_rho):
        """Device pointers (ints) to interior-only arrays, e.g. tensor.data_ptr()."""
        self._call("cpml_gpu_set_medium_device", C.c_void_p(d_lam), C.c_void_p(d_mu),
                   C.c_void_p(d_rho))

    def set_pml(self, prof_x, prof_y, prof_z=None):
        """Each profile is (6, n): a, b, K, a_half, b_half, K_half."""
        px = [self._vec(prof_x[q], self.nx) for q in range(6)]
        py = [self._vec(prof_y[q], self.ny) for q in range(6)]
        pz = ([self._vec(prof_z[q], self.nz) for q in range(6)] if self.ndim == 3
              else [None] * 6)
        self._call("cpml_gpu_set_pml", *[_ptr(a) for a in px + py + pz])

    def set_source(self, ISOURCE, JSOURCE, KSOURCE, force_x, force_y):
        fx, fy = self._vec(force_x, self.nstep), self._vec(force_y, self.nstep)
        self._call("cpml_gpu_set_source", ISOURCE, JSOURCE, KSOURCE, _ptr(fx), _ptr(fy))

    def set_receivers(self, ix_rec, iy_rec, iz_rec=None):
        ix = np.ascontiguousarray(ix_rec, np.int32)
        iy = np.ascontiguousarray(iy_rec, np.int32)
        iz = None if iz_rec is None else np.ascontiguousarray(iz_rec, np.int32)
        self.nrec = len(ix)
        self._call("cpml_gpu_set_receivers", len(ix), _ptr(ix), _ptr(iy), _ptr(iz))

    def set_energy(self, cadence=1, window=None, corrected3d=False):
        w = None if window is None else np.ascontiguousarray(
            list(window) + [1] * (6 - len(window)), np.int32)
        self._call("cpml_gpu_set_energy", cadence, _ptr(w), int(corrected3d))

    def set_stability(self, it_display=100, threshold=1e25):
        self._call("cpml_gpu_set_stability", it_display, float(threshold))

    def set_graph_steps(self, n):
        self._call("cpml_gpu_set_graph_steps", n)

    def set_kchunk(self, n):
        self._call("cpml_gpu_set_kchunk", n)

    def set_tile(self, rows):
        self._call("cpml_gpu_set_tile", rows)

    def set_time_block(self, steps, strip_width=0, chunk_rows=0):
        self._call("cpml_gpu_set_time_block", steps, strip_width, chunk_rows)

    def autotune(self, trial_steps=20):
        """Time single steps against every time-blocked shape on this grid and
        keep the fastest (cpml_gpu_autotune); the state is left as it was.
        Returns (steps_per_launch, chunk, ms of trial_steps steps)."""
        nt, jc, ms = C.c_int(), C.c_int(), C.c_double()
        self._call("cpml_gpu_autotune", trial_steps, C.byref(nt), C.byref(jc), C.byref(ms))
        return nt.value, jc.value, ms.value

    def set_fused(self, on=True):
        self._call("cpml_gpu_set_fused", int(on))

    def set_boundary(self, net=None):
        """Attach a learned strip corrector (:class:`CpmlAbcNet`), or detach
        with None.  The handle must have been given profiles without
        absorption (see :func:`no_absorption`)."""
        if net is None:
            self._call("cpml_gpu_set_boundary", 0, None)
        else:
            self._call("cpml_gpu_set_boundary", 1, net._h)
        self._abc = net

    def set_kernel_timing(self, on=True):
        self._call("cpml_gpu_set_kernel_timing", int(on))

    def run(self, it_begin=1, it_end=None):
        self._call("cpml_gpu_run", it_begin, self.nstep if it_end is None else it_end)

    def seismograms(self):
        """(sisvx, sisvy[, sisvz]) each (NREC, NSTEP) (memory of Fortran (NSTEP, NREC))."""
        n = max(self.nrec, 1)
        out = [np.zeros((n, self.nstep), self.dtype) for _ in range(self.ndim)]
        z = _ptr(out[2]) if self.ndim == 3 else None
        self._call("cpml_gpu_get_seismograms", _ptr(out[0]), _ptr(out[1]), z)
        return tuple(o[: self.nrec] for o in out)

    def energy(self):
        kin = np.zeros(self.nstep)
        pot = np.zeros(self.nstep)
        self._call("cpml_gpu_get_energy", _ptr(kin), _ptr(pot))
        return kin, pot

    def field(self, name):
        out = np.zeros(self.shape, self.dtype)
        self._call("cpml_gpu_get_field", FIELDS.index(name), _ptr(out))
        return out

    def set_field(self, name, a):
        a = self._grid(a)
        self._call("cpml_gpu_set_field", FIELDS.index(name), _ptr(a))

    def set_memory(self, name, a):
        a = self._grid(a)
        self._call("cpml_gpu_set_memory", MEMORY.index(name), _ptr(a))

    def memory(self, name):
        out = np.zeros(self.shape, self.dtype)
        self._call("cpml_gpu_get_memory", MEMORY.index(name), _ptr(out))
        return out

    def set_attenuation(self, tau_epsilon_nu1, tau_sigma_nu1, tau_epsilon_nu2, tau_sigma_nu2):
        """2D second order: the viscoelastic step of
        seismic_CPML_2D_velocity_and_stress_second_order_viscoelastic.f90
        with these relaxation times (one per SLS); empty arrays: elastic."""
        t = [np.ascontiguousarray(a, np.float64) for a in (tau_epsilon_nu1, tau_sigma_nu1,
                                                             tau_epsilon_nu2, tau_sigma_nu2)]
        self.nsls = len(t[0])
        self._call("cpml_gpu_set_attenuation", self.nsls, *[_ptr(a) for a in t])

    def set_ade_rk4(self, tables, force_x4, force_y4):
        """2D order 8: the ADE-PML RK4 step of
        seismic_ADEPML_2D_elastic_RK4_eighth_order.f90.  tables: dict of
        a_x, b_x, ah_x, bh_x (4, NX) and a_y, b_y, ah_y, bh_y (4, NY), stage
        first; force_x4, force_y4: (NSTEP, 4) force at each stage time."""
        def fort(a, n):  # Fortran a(4, n): stage fastest = C (n, 4)
            a = np.asarray(a)
            assert a.shape == (4, n)
            return np.ascontiguousarray(a.T, self.dtype)
        t = [fort(tables[k], self.nx) for k in ("a_x", "b_x", "ah_x", "bh_x")]
        t += [fort(tables[k], self.ny) for k in ("a_y", "b_y", "ah_y", "bh_y")]
        f = [np.ascontiguousarray(np.asarray(a).reshape(self.nstep, 4), self.dtype)
             for a in (force_x4, force_y4)]
        self._call("cpml_gpu_set_ade_rk4", *[_ptr(a) for a in t + f])

    def attenuation_memory(self, which):
        """R_e1 (0), R_e11 (1) or R_e13 (2) as (N_SLS, ny, nx)."""
        out = np.zeros((self.nsls, self.ny, self.nx), self.dtype)
        self._call("cpml_gpu_get_attenuation_memory", int(which), _ptr(out))
        return out

    def pressure(self):
        """Pressure seismograms (nrec, nstep) of the viscoelastic step."""
        out = np.zeros((max(self.nrec, 1), self.nstep), self.dtype)
        self._call("cpml_gpu_get_pressure", _ptr(out))
        return out[:self.nrec]

    def velocity_norm(self):
        v = C.c_double()
        self._call("cpml_gpu_velocity_norm", C.byref(v))
        return v.value

    def poison_guards(self):
        """Fill the guard cells (row padding outside the grid and its ghost
        points) of the device arrays with NaN; returns their number."""
        n = C.c_longlong()
        self._call("cpml_gpu_guard_cells", 1, C.byref(n))
        return n.value

    def guards_changed(self):
        """Number of guard cells that no longer hold a NaN."""
        n = C.c_longlong()
        self._call("cpml_gpu_guard_cells", 0, C.byref(n))
        return n.value

    def color_image(self, field_number):
        """PNM snapshot of Vx (1) or Vy (2), computed on the device
        (2D2:861-999): uint8 (rows, columns, 3), top row first; order 4 gives
        the ghost-padded (NY+2, NX+2) image of 2D4:729-731, a 3D grid the
        image of its plane k = NZ/2 (3D:1231-1234)."""
        g = 1 if self.order == 4 and self.ndim == 2 else 0
        out = np.zeros((self.ny + 2 * g, self.nx + 2 * g, 3), np.uint8)
        self._call("cpml_gpu_color_image", int(field_number), _ptr(out))
        return out

    def kernel_time(self, kernel):
        n, ms = C.c_int(), C.c_double()
        self._call("cpml_gpu_kernel_time", kernel, C.byref(n), C.byref(ms))
        return n.value, ms.value

    def elapsed_ms(self):
        v = C.c_double()
        self._call("cpml_gpu_elapsed_ms", C.byref(v))
        return v.value

    def device_bytes(self):
        v = C.c_double()
        self._call("cpml_gpu_device_bytes", C.byref(v))
        return v.value

    def slab_widths(self):
        w = (C.c_int * 6)()
        self._call("cpml_gpu_slab_widths", w)
        return list(w)


def no_absorption(n, prec=4):
    """(6, n) profile without damping: a = b = 0, K = 1 (the grid is then
    updated by the plain staggered-grid step)."""
    p = np.zeros((6, n), _dtype(prec))
    p[2] = 1
    p[5] = 1
    return p


# ------------------------------------------------------- learned boundary --

ABC_WEIGHTS = 6 * 3 * 27 * 16 * 16


class CpmlAbcNet:
    """Per-face strip corrector of the learned thin absorbing boundary
    (cpml_abc_*).  weights: float32 (6, 3, 27, 16, 16) [face][layer][tap]
    [in][out]; scales: 9 positive field scales (vx vy vz sxx syy szz sxy sxz
    syz)."""

    def __init__(self, scales, weights=None):
        sc = np.ascontiguousarray(scales, np.float32)
        if sc.shape != (9,):
            raise ValueError("scales must have 9 values")
        w = None if weights is None else self._w(weights)
        self._h = C.c_void_p()
        self.scales = sc
        # C-PML cells of the strip the corrector was trained over (0: none);
        # recorded with the weights by the training driver
        self.strip_cpml = 0
        _check(library().cpml_abc_net_create(C.byref(self._h), _ptr(w), _ptr(sc)), None,
               "cpml_abc_last_error")

    @staticmethod
    def _w(weights):
        w = np.ascontiguousarray(weights, np.float32)
        if w.size != ABC_WEIGHTS:
            raise ValueError(f"weights must have {ABC_WEIGHTS} values")
        return w

    def _call(self, fn, *args):
        _check(getattr(library(), fn)(self._h, *args), self._h, "cpml_abc_last_error")

    def close(self):
        if self._h:
            library().cpml_abc_net_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def weights(self):
        w = np.zeros((6, 3, 27, 16, 16), np.float32)
        self._call("cpml_abc_net_get_weights", _ptr(w))
        return w

    @weights.setter
    def weights(self, w):
        w = self._w(w)
        self._call("cpml_abc_net_set_weights", _ptr(w))

    def train_init(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, batch=8):
        self._call("cpml_abc_train_init", float(lr), float(beta1), float(beta2), float(eps),
                   int(batch))

    def train_noise(self, rel_noise):
        self._call("cpml_abc_train_noise", float(rel_noise))

    def train_step(self, student, teacher, offset, it):
        loss = np.zeros(6)
        self._call("cpml_abc_train_step", student._h, teacher._h, *[int(o) for o in offset],
                   int(it), _ptr(loss))
        return loss

    def grad(self, student, teacher, offset):
        loss = np.zeros(6)
        g = np.zeros((6, 3, 27, 16, 16), np.float32)
        self._call("cpml_abc_grad", student._h, teacher._h, *[int(o) for o in offset],
                   _ptr(loss), _ptr(g))
        return loss, g

    def correction(self, h, use_mfma=True):
        """Corrector output for h's current state: list of 6 arrays
        (9, 4, nb, na) in scaled units."""
        n = (h.nx, h.ny, h.nz)
        shapes = []
        for f in range(6):
            ax = f // 2
            ta, tb = (1, 2) if ax == 0 else ((0, 2) if ax == 1 else (0, 1))
            shapes.append((9, 4, n[tb], n[ta]))
        out = np.zeros(sum(int(np.prod(s)) for s in shapes), np.float32)
        self._call("cpml_abc_correction", h._h, int(bool(use_mfma)), _ptr(out))
        res, o = [], 0
        for s in shapes:
            k = int(np.prod(s))
            res.append(out[o:o + k].reshape(s))
            o += k
        return res


# ------------------------------------------------------------------ batch --

class CpmlBatch:
    """Many independent 2D shots advanced together (cpml_batch_init ...).

    ``shots`` is a list of dicts with keys NX, NY, DELTAX, DELTAY, DELTAT,
    NPOINTS_PML; every shot is then configured with the same calls as
    :class:`CpmlGpu` plus a leading shot index.  A batched shot's results
    equal a :class:`CpmlGpu` run of that shot alone, bit for bit."""

    BATCH_FIELDS = ("vx", "vy", "sigmaxx", "sigmayy", "sigmaxy")

    def __init__(self, order, prec, shots, NSTEP):
        self.order, self.prec, self.nstep = order, prec, NSTEP
        self.dtype = _dtype(prec)
        self.shots = [dict(s) for s in shots]
        self.nrec = [0] * len(shots)
        n = len(shots)
        col = lambda k, t: np.ascontiguousarray([s[k] for s in shots], t)  # noqa: E731
        nx, ny = col("NX", np.int32), col("NY", np.int32)
        dx, dy, dt = col("DELTAX", np.float64), col("DELTAY", np.float64), col("DELTAT",
                                                                                np.float64)
        npml = col("NPOINTS_PML", np.int32)
        self._h = C.c_void_p()
        _check(library().cpml_batch_init(C.byref(self._h), order, prec, n, _ptr(nx), _ptr(ny),
                                         _ptr(dx), _ptr(dy), _ptr(dt), _ptr(npml), NSTEP),
               None, "cpml_batch_last_error")

    def close(self):
        if self._h:
            library().cpml_batch_finalize(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.shots)

    def _call(self, fn, *args):
        _check(getattr(library(), fn)(self._h, *args), self._h, "cpml_batch_last_error")

    def _shape(self, q):
        if not 0 <= q < len(self.shots):
            raise CpmlError(CPML_ERR_ARG, f"shot index {q} out of range")
        return (self.shots[q]["NY"], self.shots[q]["NX"])

    def _arr(self, a, shape):
        a = np.ascontiguousarray(a, self.dtype)
        if a.shape != shape:
            raise ValueError(f"array shape {a.shape} != {shape}")
        return a

    def set_medium(self, q, lam, mu, rho):
        sh = self._shape(q)
        lam, mu, rho = (self._arr(x, sh) for x in (lam, mu, rho))
        self._call("cpml_batch_set_medium", q, _ptr(lam), _ptr(mu), _ptr(rho))

    def set_pml(self, q, prof_x, prof_y):
        ny, nx = self._shape(q)
        px = [self._arr(prof_x[k], (nx,)) for k in range(6)]
        py = [self._arr(prof_y[k], (ny,)) for k in range(6)]
        self._call("cpml_batch_set_pml", q, *[_ptr(a) for a in px + py])

    def set_source(self, q, ISOURCE, JSOURCE, force_x, force_y):
        fx = self._arr(force_x, (self.nstep,))
        fy = self._arr(force_y, (self.nstep,))
        self._call("cpml_batch_set_source", q, ISOURCE, JSOURCE, _ptr(fx), _ptr(fy))

    def set_receivers(self, q, ix_rec, iy_rec):
        ix = np.ascontiguousarray(ix_rec, np.int32)
        iy = np.ascontiguousarray(iy_rec, np.int32)
        self.nrec[q] = len(ix)
        self._call("cpml_batch_set_receivers", q, len(ix), _ptr(ix), _ptr(iy))

    def set_tiling(self, tile_y=4, blocks_per_cu=0, graph_steps=16):
        self._call("cpml_batch_set_tiling", tile_y, blocks_per_cu, graph_steps)

    def autotune(self, trial_steps=20, cache_path=None):
        """Returns (tile_y, blocks_per_cu, best_ms); best_ms is -1 on a cache hit."""
        ty, bpc, ms = C.c_int(), C.c_int(), C.c_double()
        path = None if cache_path is None else str(cache_path).encode()
        self._call("cpml_batch_autotune", trial_steps, path, C.byref(ty), C.byref(bpc),
                   C.byref(ms))
        return ty.value, bpc.value, ms.value

    def set_time_block(self, steps, strip_width=0, chunk_rows=0):
        self._call("cpml_batch_set_time_block", steps, strip_width, chunk_rows)

    def launch_shape(self):
        """(tile_y, blocks_per_cu, time_block, chunk_rows) in use."""
        v = [C.c_int() for _ in range(4)]
        self._call("cpml_batch_launch_shape", *[C.byref(x) for x in v])
        return tuple(x.value for x in v)

    def set_energy(self, cadence=1):
        self._call("cpml_batch_set_energy", cadence)

    def energy(self, q):
        kin = np.zeros(self.nstep)
        pot = np.zeros(self.nstep)
        self._call("cpml_batch_get_energy", q, _ptr(kin), _ptr(pot))
        return kin, pot

    def set_stability(self, it_display=100, threshold=1e25):
        self._call("cpml_batch_set_stability", it_display, threshold)

    def run(self, it_begin=1, it_end=None):
        self._call("cpml_batch_run", it_begin, self.nstep if it_end is None else it_end)

    def seismograms(self, q):
        n = max(self.nrec[q], 1)
        out = [np.zeros((n, self.nstep), self.dtype) for _ in range(2)]
        self._call("cpml_batch_get_seismograms", q, _ptr(out[0]), _ptr(out[1]))
        return tuple(o[: self.nrec[q]] for o in out)

    def field(self, q, name):
        out = np.zeros(self._shape(q), self.dtype)
        self._call("cpml_batch_get_field", q, FIELDS.index(name), _ptr(out))
        return out

    def elapsed_ms(self):
        v = C.c_double()
        self._call("cpml_batch_elapsed_ms", C.byref(v))
        return v.value
