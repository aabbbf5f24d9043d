"""Heterogeneous inputs for the parity tests: seeded layered media with a
smoothed random perturbation (the 3D benchmark medium at test size), C-PML
profiles with K_MAX_PML > 1, and a perturbation of the reduced 2D media.

Every kernel reads neighbouring material values (2-point and 4-point averages
in 2D, per-point DELTAT*lambda, DELTAT*mu, DELTAT/rho in 3D); on a homogeneous
medium a wrong row or plane offset on those arrays is invisible, here it is
not.  The profiles follow the C-PML formulas of
seismic_CPML_3D_isotropic_MPI_OpenMP.f90:399-667 with K_MAX_PML as a
parameter (the 3D viscoelastic program uses K_MAX_PML = 7); they are inputs
only: the oracle and the GPU take the same arrays.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.ndimage import gaussian_filter

import oracle_lib as orc

# dimensions of the 3D layered case (registered in golden/cases.py OTHER_CASES)
NX3, NY3, NZ3, NPML3 = 96, 80, 72, 10


def smooth_perturbation(shape, seed, sigma=3.0, amp=0.05):
    """Gaussian-smoothed white noise scaled to a maximum of +-amp."""
    rng = np.random.default_rng(seed)
    n = gaussian_filter(rng.standard_normal(shape), sigma, mode="wrap")
    return amp * n / np.abs(n).max()


def layered_cp(nlayers, n, seed):
    """Seeded layer velocities cp ~ U[1500, 5500] along one axis of n points."""
    rng = np.random.default_rng(seed)
    cuts = np.sort(rng.integers(1, n, size=nlayers - 1))
    cps = rng.uniform(1500.0, 5500.0, size=nlayers)
    return cps[np.searchsorted(cuts, np.arange(n), side="right")]


def elastic(cp):
    """cs = cp/sqrt(3), rho = 310 cp^0.25 (Gardner); returns lambda, mu, rho."""
    cs = cp / math.sqrt(3.0)
    rho = 310.0 * cp ** 0.25
    return rho * (cp * cp - 2.0 * cs * cs), rho * cs * cs, rho


def profile_kmax(n, delta, deltat, npml, cp, f0, kmax, dt, rcoef=1e-3, clamp_alpha=False):
    """(6, n) a, b, K, a_half, b_half, K_half with K_MAX_PML = kmax, following
    3D:399-667: d0 = -(NPOWER+1) cp ln(Rcoef) / (2 L), NPOWER = 2; in the
    slab d = d0 x^2, K = 1 + (kmax-1) x^2, alpha = pi f0 (1 - x);
    b = exp(-(d/K + alpha) DELTAT), a = d (b-1) / (K (d + K alpha))."""
    L = npml * delta
    d0 = -3.0 * cp * math.log(rcoef) / (2.0 * L)
    amax = math.pi * f0
    out = np.zeros((6, n))
    out[2] = out[5] = 1.0
    for half in (0, 1):
        x = (np.arange(n) + 0.5 * half) * delta
        xl = L - x
        xr = x - ((n - 1) * delta - L)
        xv = np.where(xl >= 0, xl, 0.0) + np.where(xr >= 0, xr, 0.0)
        inside = (xl >= 0) | (xr >= 0)
        xn = np.where(inside, xv / L, 0.0)
        d = d0 * xn ** 2
        K = np.where(inside, 1.0 + (kmax - 1.0) * xn ** 2, 1.0)
        alpha = np.where(inside, amax * (1.0 - xn), 0.0)
        if clamp_alpha:
            alpha = np.maximum(alpha, 0.0)
        b = np.exp(-(d / K + alpha) * deltat)
        with np.errstate(invalid="ignore", divide="ignore"):
            a = np.where(np.abs(d) > 1e-6, d * (b - 1.0) / (K * (d + K * alpha)), 0.0)
        out[3 * half], out[3 * half + 1], out[3 * half + 2] = a, b, K
    return out.astype(dt)


def layered3d_inputs(prec, nstep, kmax=1.0, seed=11, nx=NX3, ny=NY3, nz=NZ3, npml=NPML3,
                     dx=10.0, f0=12.0, courant=0.5):
    """The 3D benchmark medium at test size: 12 layers along z, a +-5%
    Gaussian-smoothed (sigma 3 cells) perturbation, DELTAT for Courant 0.5
    with cp_max, Ricker source at the centre, a 6 x 6 grid of receivers on
    the top interior plane (k = NZ - NPML).  Arrays (nz, ny, nx)."""
    dt = np.float64 if prec == "f64" else np.float32
    cp = layered_cp(12, nz, seed)[:, None, None] * (
        1.0 + smooth_perturbation((nz, ny, nx), seed + 1))
    lam, mu, rho = elastic(cp)
    cp_max = float(cp.max())
    deltat = courant * dx / (cp_max * math.sqrt(3.0))
    R = dt
    d = dict(dtype=dt, nstep=nstep, DELTAX=R(dx), DELTAY=R(dx), DELTAZ=R(dx), DELTAT=R(deltat))
    d["lam"], d["mu"], d["rho"] = lam.astype(dt), mu.astype(dt), rho.astype(dt)
    for ax, n in (("x", nx), ("y", ny), ("z", nz)):
        d[f"prof_{ax}"] = profile_kmax(n, dx, deltat, npml, cp_max, f0, kmax, dt,
                                       clamp_alpha=(ax == "x"))
    d["fx"], d["fy"] = orc.stf(nstep, R(deltat), R(f0), R(1.2 / f0), R(1e7), R(135.0), 2, prec)
    d["ISOURCE"], d["JSOURCE"], d["ksrc"] = nx // 2, ny // 2, nz // 2
    line_x = np.linspace(npml + 1, nx - npml, 6).round().astype(np.int32)
    line_y = np.linspace(npml + 1, ny - npml, 6).round().astype(np.int32)
    ix, iy = np.meshgrid(line_x, line_y, indexing="xy")
    d["ix_rec"], d["iy_rec"] = ix.ravel().astype(np.int32), iy.ravel().astype(np.int32)
    d["iz_rec"] = np.full(d["ix_rec"].size, nz - npml, np.int32)
    d["ewin"] = (npml + 1, nx - npml, npml + 1, ny - npml, npml + 1, nz - npml)
    return d


def perturbed(a, seed, amp=0.05):
    """a * (1 + smooth +-amp perturbation) in a's dtype (a 2D medium array)."""
    return (a.astype(np.float64) * (1.0 + smooth_perturbation(a.shape, seed, 2.0, amp))).astype(
        a.dtype)
