"""Synthetic workloads of the benchmark configurations.

* ``shipped_2d(order)``  -- seismic_CPML_2D_isotropic_{second,fourth}_order as
  shipped (101 x 641, homogeneous, first-derivative-of-Gaussian force).
* ``layered_2d``         -- 2D 4th-order, 4096^2, 5 m, NPML 20, FP32, 12 seeded
  horizontal layers (cp ~ U[1500,5500], cs = cp/sqrt(3), rho = 310 cp^0.25),
  Ricker 10 Hz at the centre, 256 receivers on the top interior row.
* ``layered_3d``         -- 3D 768^3, 10 m, NPML 20, FP32, layered along z plus
  a +-5% Gaussian-smoothed (sigma = 8 cells) perturbation, Ricker 8 Hz at the
  centre, 32 x 32 receivers on the top interior plane, DELTAT for Courant 0.5.

Large media are generated on the GPU with torch (device-memory plumbing only)
and handed to the library with set_medium_device; the time loop itself runs
in libcpml_gpu.  The damping profile needs one velocity: cp_max is used.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import solver as S


@dataclass
class Workload:
    name: str
    ndim: int
    order: int
    prec: int
    nx: int
    ny: int
    nz: int
    dx: float
    dt: float
    npml: int
    nstep: int
    f0: float
    src: tuple
    rec: tuple            # (ix, iy[, iz]) 1-based int32 arrays
    prof: tuple           # (prof_x, prof_y[, prof_z])
    force: tuple          # (force_x, force_y)
    medium_host: tuple | None = None     # (lam, mu, rho) numpy
    medium_dev: tuple | None = None      # torch tensors (lam, mu, rho)
    cp_max: float = 0.0
    extra: dict = field(default_factory=dict)

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    def make_solver(self, energy=False, graph_steps=16):
        s = S.CpmlGpu(self.ndim, self.order, self.prec, self.nx, self.ny, self.nz, self.dx,
                      self.dx, self.dx, self.dt, self.npml, self.nstep)
        if self.medium_dev is not None:
            lam, mu, rho = self.medium_dev
            s.set_medium_device(lam.data_ptr(), mu.data_ptr(), rho.data_ptr())
        else:
            s.set_medium(*self.medium_host)
        s.set_pml(*self.prof)
        s.set_source(*self.src, *self.force)
        s.set_receivers(*self.rec)
        if energy:
            s.set_energy(1)
        s.set_graph_steps(graph_steps)
        return s


def _layers(rng, n, nlayers=12):
    cuts = np.sort(rng.integers(1, n, size=nlayers - 1))
    cps = rng.uniform(1500.0, 5500.0, size=nlayers)
    return cps[np.searchsorted(cuts, np.arange(n), side="right")]


def _elastic(cp):
    cs = cp / math.sqrt(3.0)
    rho = 310.0 * cp ** 0.25
    mu = rho * cs * cs
    lam = rho * (cp * cp - 2.0 * cs * cs)
    return lam, mu, rho


def shipped_2d(order: int = 2, prec: int = 8) -> Workload:
    """2D2 / 2D4 as shipped (2D2:137-186, 2D4:131-187)."""
    dt_ = np.float64 if prec == 8 else np.float32
    nx, ny, npml = 101, 641, 10
    dx = dt_(10.0)
    deltat = dt_(2e-3) if order == 2 else dt_(2e-3) / dt_(2)
    nstep = 2000 if order == 2 else 4000
    cp, density = dt_(3300.0), dt_(2800.0)
    cs = cp / dt_(1.732)
    f0 = dt_(7.0)
    t0 = dt_(1.2) / f0
    isrc, jsrc = nx - 2 * npml - 1, 2 * ny // 3 + 1
    xs = dt_(isrc - 1) * dx
    px = S.pml_profile(nx, dx, deltat, npml, cp, f0, prec, clamp_alpha=True)
    py = S.pml_profile(ny, dx, deltat, npml, cp, f0, prec, origin_max_uses_n=(order == 4),
                       clamp_alpha=False)
    fx, fy = S.source_time_function(nstep, deltat, f0, t0, dt_(1e7), dt_(135.0), 1, prec)
    ix, iy = S.nearest_receivers(nx, ny, dx, dx, 2, xs - dt_(100.0), dt_(2300.0), xs,
                                 dt_(300.0), prec)
    rho = np.full((ny, nx), density, dt_)
    mu = np.full((ny, nx), density * cs * cs, dt_)
    lam = np.full((ny, nx), density * (cp * cp - dt_(2) * cs * cs), dt_)
    return Workload(f"2d{order}_shipped", 2, order, prec, nx, ny, 1, float(dx), float(deltat),
                    npml, nstep, float(f0), (isrc, jsrc, 1), (ix, iy), (px, py), (fx, fy),
                    medium_host=(lam, mu, rho), cp_max=float(cp))


def layered_2d(n=4096, nstep=4000, seed=2, order=4, prec=4, courant=0.5, npml=20,
               dx=5.0, f0=10.0, nrec=256) -> Workload:
    dt_ = np.float64 if prec == 8 else np.float32
    rng = np.random.default_rng(seed)
    cp = np.repeat(_layers(rng, n)[:, None], n, axis=1)  # layers along y (depth)
    lam, mu, rho = (a.astype(dt_) for a in _elastic(cp))
    cp_max = float(cp.max())
    deltat = courant * dx / (cp_max * math.sqrt(2.0))
    px = S.pml_profile(n, dx, deltat, npml, cp_max, f0, prec, clamp_alpha=True)
    py = S.pml_profile(n, dx, deltat, npml, cp_max, f0, prec, clamp_alpha=False)
    fx, fy = S.source_time_function(nstep, deltat, f0, 1.2 / f0, 1e7, 135.0, 2, prec)
    ix = np.linspace(npml + 1, n - npml, nrec).round().astype(np.int32)
    iy = np.full(nrec, n - npml, np.int32)
    return Workload(f"2d{order}_layered_{n}", 2, order, prec, n, n, 1, dx, deltat, npml, nstep,
                    f0, (n // 2, n // 2, 1), (ix, iy), (px, py), (fx, fy),
                    medium_host=(lam, mu, rho), cp_max=cp_max)


# Relaxation times of 3 standard linear solids fitted by the reference's
# SolvOpt routine for Qp = 65, Qs = 55 around 35 Hz (the shipped 2D
# viscoelastic program's values): tau_epsilon / tau_sigma for nu1 (Qp) and nu2
# (Qs).  The fit scales with 1/f0, so taus_for(f0) moves the band to f0.
SLS3_Q65_55_35HZ = (
    (0.024081581857537462, 0.0046996089908613505, 0.0009567997872435829),
    (0.022560146386368645, 0.004508471279712249, 0.000893787640376875),
    (0.024305444805272164, 0.0047281078292263955, 0.0009667252695863502),
    (0.022509197794294895, 0.0045013880073380965, 0.0008917332095369118),
)


def taus_for(f0: float):
    return tuple(np.array(t) * (35.0 / f0) for t in SLS3_Q65_55_35HZ)


def layered_3d(n=768, nstep=2000, seed=3, prec=4, courant=0.5, npml=20, dx=10.0, f0=8.0,
               nrec_side=32, sigma=8.0, amp=0.05, device="cuda", order=2) -> Workload:
    """3D layered + smoothed perturbation, generated on the GPU with torch.
    order 4 selects the 4th-order step (Courant 0.5 is inside its 6/7 limit)."""
    import torch

    tdt = torch.float32 if prec == 4 else torch.float64
    rng = np.random.default_rng(seed)
    cpz = torch.tensor(_layers(rng, n), dtype=torch.float64, device=device)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    noise = torch.randn((n, n, n), generator=g, device=device, dtype=torch.float32)
    # Gaussian smoothing with sigma cells, applied in the Fourier domain
    spec = torch.fft.rfftn(noise)
    del noise
    w_full = 2.0 * math.pi * torch.fft.fftfreq(n, device=device)
    w_half = 2.0 * math.pi * torch.fft.rfftfreq(n, device=device)
    gz = torch.exp(-0.5 * (sigma * w_full) ** 2).view(n, 1, 1)
    gy = torch.exp(-0.5 * (sigma * w_full) ** 2).view(1, n, 1)
    gx = torch.exp(-0.5 * (sigma * w_half) ** 2).view(1, 1, -1)
    spec *= gz
    spec *= gy
    spec *= gx
    noise = torch.fft.irfftn(spec, s=(n, n, n))
    del spec
    noise = noise / noise.abs().max()
    cp = (cpz.view(n, 1, 1).float() * (1.0 + amp * noise))  # (z, y, x) layers along z
    del noise
    cp_max = float(cp.max())
    cs = cp / math.sqrt(3.0)
    rho = 310.0 * cp ** 0.25
    mu = (rho * cs * cs).to(tdt).contiguous()
    lam = (rho * (cp * cp - 2.0 * cs * cs)).to(tdt).contiguous()
    rho = rho.to(tdt).contiguous()
    del cp, cs
    torch.cuda.synchronize()
    deltat = courant * dx / (cp_max * math.sqrt(3.0))
    px = S.pml_profile(n, dx, deltat, npml, cp_max, f0, prec, clamp_alpha=True)
    py = S.pml_profile(n, dx, deltat, npml, cp_max, f0, prec, clamp_alpha=False)
    pz = S.pml_profile(n, dx, deltat, npml, cp_max, f0, prec, clamp_alpha=False)
    fx, fy = S.source_time_function(nstep, deltat, f0, 1.2 / f0, 1e7, 135.0, 2, prec)
    line = np.linspace(npml + 1, n - npml, nrec_side).round().astype(np.int32)
    ix, iy = np.meshgrid(line, line, indexing="xy")
    ix = ix.ravel().astype(np.int32)
    iy = iy.ravel().astype(np.int32)
    iz = np.full(ix.size, n - npml, np.int32)
    return Workload(f"3d_layered_{n}", 3, order, prec, n, n, n, dx, deltat, npml, nstep, f0,
                    (n // 2, n // 2, n // 2), (ix, iy, iz), (px, py, pz), (fx, fy),
                    medium_dev=(lam, mu, rho), cp_max=cp_max)


def bytes_per_step(w: Workload, slabs, fused: bool = False, nsls: int = 0,
                   time_block: int = 0) -> dict:
    """Algorithmic HBM bytes of one time step, per kernel: every array touched
    once per kernel.  3D stress: read v(3), sigma(6), DELTAT*lambda, DELTAT*mu,
    write sigma(6) = 17 words; velocity: read sigma(6), v(3), DELTAT/rho, write
    v(3) = 13 words; each phase reads and writes 3 memory variables per
    absorbing slab point (6 words).  The fused 3D step reads v(3), sigma(6)
    and the 3 coefficients and writes v(3), sigma(6): 21 words, plus 12 words
    per slab point.  2D: stress 5 reads (vx, vy, sigma x3) + lambda, mu + 3
    writes = 10 words; velocity 5 + rho + 2 writes = 8 words; 2 memory
    variables per slab point per phase (4 words); the 2D viscoelastic stress
    kernel also reads and writes R_e1, R_e11, R_e13 per SLS (6 words per
    SLS).  A 2D time-blocked launch of `time_block` steps reads v(2), sigma(3),
    lambda, mu, rho and writes v(2), sigma(3) once (13 words) and reads and
    writes the 4 memory variables of each slab axis once (8 words per slab
    point); "time_block" is those bytes per launch, "step" per time step.  A
    3D time-blocked launch moves the fused step's 21 words per point and 12
    per slab point once per `time_block` steps."""
    es = 4 if w.prec == 4 else 8
    wx = slabs[0] + slabs[1]
    wy = slabs[2] + slabs[3]
    wz = slabs[4] + slabs[5]
    n = w.cells
    if w.ndim == 3:
        slab_pts = wx * w.ny * w.nz + w.nx * wy * w.nz + w.nx * w.ny * wz
        if time_block:
            f = es * (21 * n + 12 * slab_pts)
            return {"time_block": f, "step": f / time_block}
        if fused:
            f = es * (21 * n + 12 * slab_pts)
            return {"fused": f, "step": f}
        stress = es * (17 * n + 6 * slab_pts)
        vel = es * (13 * n + 6 * slab_pts)
    else:
        slab_pts = wx * w.ny + w.nx * wy
        if time_block:
            f = es * (13 * n + 8 * slab_pts)
            return {"time_block": f, "step": f / time_block}
        stress = es * ((10 + 6 * nsls) * n + 4 * slab_pts)
        vel = es * (8 * n + 4 * slab_pts)
    return {"stress": stress, "velocity": vel, "step": stress + vel}


@dataclass
class BatchWorkload:
    """Per-shot inputs of a batched 2D run (host arrays)."""
    order: int
    prec: int
    nstep: int
    shots: list  # dicts: NX NY DELTAX DELTAY DELTAT NPOINTS_PML lam mu rho prof_x prof_y
    #              ISOURCE JSOURCE fx fy ix_rec iy_rec f0

    @property
    def cells(self) -> int:
        return sum(s["NX"] * s["NY"] for s in self.shots)

    def make_batch(self, graph_steps=16):
        b = S.CpmlBatch(self.order, self.prec, self.shots, self.nstep)
        for q, s in enumerate(self.shots):
            b.set_medium(q, s["lam"], s["mu"], s["rho"])
            b.set_pml(q, s["prof_x"], s["prof_y"])
            b.set_source(q, s["ISOURCE"], s["JSOURCE"], s["fx"], s["fy"])
            b.set_receivers(q, s["ix_rec"], s["iy_rec"])
        b.set_tiling(4, 0, graph_steps)
        return b

    def make_solver(self, q, graph_steps=16):
        """The plain single-shot solver of shot q (the sequential comparator)."""
        s = self.shots[q]
        h = S.CpmlGpu(2, self.order, self.prec, s["NX"], s["NY"], 1, s["DELTAX"], s["DELTAY"],
                      1.0, s["DELTAT"], s["NPOINTS_PML"], self.nstep)
        h.set_medium(s["lam"], s["mu"], s["rho"])
        h.set_pml(s["prof_x"], s["prof_y"])
        h.set_source(s["ISOURCE"], s["JSOURCE"], 1, s["fx"], s["fy"])
        h.set_receivers(s["ix_rec"], s["iy_rec"])
        h.set_graph_steps(graph_steps)
        return h


def batch_2d(nshots=128, nmin=512, nmax=2048, nstep=3000, seed=4, order=4, prec=4,
             courant=0.5, npml=20, dx=5.0, nrec=64) -> BatchWorkload:
    """Multi-shot 2D survey: per shot a seeded grid size NX, NY ~ U{nmin..nmax}
    (uneven, non-square), its own 12-layer model, source position (at least
    NPML + 10 cells inside) and Ricker f0 ~ U[5, 15] Hz; DELTAT for Courant
    0.5 of the shot's cp_max; nrec receivers on the top interior row."""
    dt_ = np.float64 if prec == 8 else np.float32
    rng = np.random.default_rng(seed)
    shots = []
    for _ in range(nshots):
        nx, ny = (int(v) for v in rng.integers(nmin, nmax + 1, size=2))
        cpy = _layers(rng, ny)
        cp = np.broadcast_to(cpy[:, None], (ny, nx))
        lam, mu, rho = (np.ascontiguousarray(a, dt_) for a in _elastic(cp))
        cp_max = float(cpy.max())
        f0 = float(rng.uniform(5.0, 15.0))
        deltat = courant * dx / (cp_max * math.sqrt(2.0))
        m = npml + 10
        isrc, jsrc = int(rng.integers(m, nx - m)), int(rng.integers(m, ny - m))
        px = S.pml_profile(nx, dx, deltat, npml, cp_max, f0, prec, clamp_alpha=True)
        py = S.pml_profile(ny, dx, deltat, npml, cp_max, f0, prec, clamp_alpha=False)
        fx, fy = S.source_time_function(nstep, deltat, f0, 1.2 / f0, 1e7, 135.0, 2, prec)
        ix = np.linspace(npml + 1, nx - npml, nrec).round().astype(np.int32)
        iy = np.full(nrec, ny - npml, np.int32)
        shots.append(dict(NX=nx, NY=ny, DELTAX=dx, DELTAY=dx, DELTAT=deltat, NPOINTS_PML=npml,
                          lam=lam, mu=mu, rho=rho, prof_x=px, prof_y=py, ISOURCE=isrc,
                          JSOURCE=jsrc, fx=fx, fy=fy, ix_rec=ix, iy_rec=iy, f0=f0))
    return BatchWorkload(order, prec, nstep, shots)
