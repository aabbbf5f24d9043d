"""Learned thin absorbing boundary: training and evaluation on one GPU.

The corrector (libcpml_gpu, cpml_abc_*) replaces a 20-cell C-PML by a 4-cell
strip.  This module drives it through the C ABI:

* ``layered_model``   -- seeded random layered medium (cp ~ U[1500, 5500],
  cs = cp / sqrt(3), rho = 310 cp^0.25, layers along z), generated on the GPU.
* ``train``           -- online training on wide-C-PML teacher runs of this
  solver (interior n^3 + 20-cell C-PML; the student is the interior + 4 cells,
  i.e. the teacher window starting 16 cells in).
* ``reflection``      -- reflected energy of a boundary in dB: the test run
  and a reference run in a doubled domain (2n interior + 20-cell C-PML, the
  same medium extended by its edge values) record the three velocity
  components on an interior lattice; reflected energy = sum (u_test -
  u_ref)^2 over all receivers and steps, relative to sum u_ref^2.
* ``time_to_solution`` -- wall time of the learned form ((n + 8)^3 plus the
  corrector) against the plain 20-cell C-PML ((n + 40)^3), same steps.

Torch is used only to build media on the device (plumbing); every step and
every training kernel is in libcpml_gpu.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import solver as S

NPML_WIDE = 20
STRIP = 4
STRIP_CPML = 4  # C-PML cells of the strip under the corrector (training default)
OFFSET = NPML_WIDE - STRIP  # student/teacher window offset (cells)


def field_scales(z_ref: float = 8.0e6):
    """Velocities unscaled, stresses divided by a typical impedance rho*cp."""
    return np.array([1, 1, 1] + [z_ref] * 6, np.float32)


def init_weights(seed: int = 0, std: float = 0.05):
    """Random FP32 weights, padding channels zero."""
    rng = np.random.default_rng(seed)
    w = (rng.standard_normal((6, 3, 27, 16, 16)) * std).astype(np.float32)
    w[:, 0, :, 9:, :] = 0  # layer 1 inputs 9..15
    w[:, 2, :, :, 9:] = 0  # layer 3 outputs 9..15
    return w


@dataclass
class Model:
    cpz: np.ndarray      # cp per z index of the interior
    dx: float
    dt: float
    f0: float
    src: tuple           # 0-based interior cell of the source

    @property
    def n(self):
        return len(self.cpz)


def layered_model(n: int, seed: int, dx=10.0, courant=0.5, f0=None, nlayers=12,
                  src_margin=12) -> Model:
    rng = np.random.default_rng(seed)
    cuts = np.sort(rng.integers(1, n, size=nlayers - 1))
    cps = rng.uniform(1500.0, 5500.0, size=nlayers)
    cpz = cps[np.searchsorted(cuts, np.arange(n), side="right")]
    f0 = float(rng.uniform(5.0, 15.0)) if f0 is None else float(f0)
    margin = min(src_margin, (n - 1) // 2)
    src = tuple(int(v) for v in rng.integers(margin, n - margin, size=3))
    dt = courant * dx / (5500.0 * math.sqrt(3.0))  # Courant 0.5 of the fastest possible cp
    return Model(cpz, dx, dt, f0, src)


def _medium(cpz_ext, nxy, device="cuda"):
    """lam, mu, rho torch tensors (nz, ny, nx) for a z-layered cp profile."""
    import torch

    cp = torch.tensor(cpz_ext, dtype=torch.float64, device=device)
    cs = cp / math.sqrt(3.0)
    rho = 310.0 * cp ** 0.25
    mu = rho * cs * cs
    lam = rho * (cp * cp - 2.0 * cs * cs)
    out = []
    for a in (lam, mu, rho):
        out.append(a.float().view(-1, 1, 1).expand(-1, nxy[1], nxy[0]).contiguous())
    return out


# Learned thin C-PML: the profile family of 2D2:283-394 / 3D:399-667 with
# four trainable factors theta = (d0 scale, power, alpha_max scale, K_max):
#   d(x) = s_d d0 (x/L)^p, K(x) = 1 + (K_max - 1)(x/L)^p,
#   alpha(x) = s_a pi f0 (1 - x/L), d0 = -(2 + 1) cp ln(1e-3) / (2 L),
# x the depth into the layer of thickness L; a, b as 2D2:388-393.
THETA_REFERENCE = (1.0, 2.0, 1.0, 1.0)  # the reference's own profile


def profile_theta(n: int, dx: float, dt: float, width: int, cp: float, f0: float, theta,
                  prec: int = 4) -> np.ndarray:
    """(6, n) profile a, b, K, a_half, b_half, K_half of a C-PML of `width`
    cells at both ends of an axis of n points, with the factors theta."""
    s_d, p, s_a, kmax = (float(v) for v in theta)
    L = width * dx
    d0 = -(2.0 + 1.0) * cp * math.log(1e-3) / (2.0 * L)
    alpha_max = math.pi * f0
    x = np.arange(n, dtype=np.float64) * dx
    out = []
    for xv in (x, x + dx / 2):
        depth = np.maximum(L - xv, xv - ((n - 1) * dx - L))
        inside = depth >= 0  # as 2D2:333-381: the layer's inner edge is in it
        r = np.where(inside, depth / L, 0.0)
        d = np.where(inside, s_d * d0 * r ** p, 0.0)
        K = np.where(inside, 1.0 + (kmax - 1.0) * r ** p, 1.0)
        al = np.where(inside, np.maximum(s_a * alpha_max * (1.0 - r), 0.0), 0.0)
        b = np.exp(-(d / K + al) * dt)
        den = K * (d + K * al)
        a = np.where(np.abs(d) > 1e-6, d * (b - 1.0) / np.where(den != 0, den, 1.0), 0.0)
        out += [a, b, K]
    dt_ = np.float32 if prec == 4 else np.float64
    return np.stack([out[0], out[1], out[2], out[3], out[4], out[5]]).astype(dt_)


def make_run(model: Model, pad: int, absorb: int, nstep: int, ext: int = 0,
             receivers=None, source=True, theta=None):
    """A solver over the interior grown by ``ext`` cells of extended medium
    on every side, then ``pad`` boundary cells, of which the outer ``absorb``
    are a C-PML (0: no absorption; ``theta``: the learned profile factors
    instead of the reference's profile).  Returns (solver, offset of
    interior cell 0, media tensors kept alive)."""
    n_int = model.n + 2 * ext
    n = n_int + 2 * pad
    off = pad + ext
    z = np.clip(np.arange(n) - off, 0, model.n - 1)
    lam, mu, rho = _medium(model.cpz[z], (n, n))
    s = S.CpmlGpu(3, 2, 4, n, n, n, model.dx, model.dx, model.dx, model.dt,
                  max(absorb, 1), nstep)
    s.set_medium_device(lam.data_ptr(), mu.data_ptr(), rho.data_ptr())
    cp_max = float(model.cpz.max())
    if absorb > 0 and theta is not None:
        prof = profile_theta(n, model.dx, model.dt, absorb, cp_max, model.f0, theta)
        s.set_pml(prof, prof, prof)
    elif absorb > 0:
        # the C-PML of exactly `absorb` cells at the outer edge of the grid
        prof = S.pml_profile(n, model.dx, model.dt, absorb, cp_max, model.f0, 4,
                             clamp_alpha=False)
        s.set_pml(prof, prof, prof)
    else:
        p0 = S.no_absorption(n)
        s.set_pml(p0, p0, p0)
    if source:
        fx, fy = S.source_time_function(nstep, model.dt, model.f0, 1.2 / model.f0, 1e7, 135.0,
                                        2, 4)
        i, j, k = (c + off + 1 for c in model.src)
        s.set_source(i, j, k, fx, fy)
    else:
        s.set_source(0, 0, 0, np.zeros(nstep, np.float32), np.zeros(nstep, np.float32))
    if receivers is not None:
        ix, iy, iz = (np.asarray(r, np.int32) + off + 1 for r in receivers)
        s.set_receivers(ix, iy, iz)
    else:
        s.set_receivers(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32))
    return s, off, (lam, mu, rho)


def train(net: S.CpmlAbcNet, n: int = 128, runs: int = 16, nstep: int = 600,
          first_step: int = 30, stride: int = 1, seed: int = 1000, log=None,
          strip_cpml: int = 0):
    """Online training: for each seeded model a wide-C-PML teacher run
    (interior n^3 + 20 cells) and a student grid (n + 8)^3 whose strip has a
    C-PML of strip_cpml cells (0: none); the corrector learns what the
    student's step misses against the teacher.  Returns the mean loss per
    run."""
    net.strip_cpml = strip_cpml
    history = []
    for r in range(runs):
        m = layered_model(n, seed + r)
        teacher, _, keep_t = make_run(m, NPML_WIDE, NPML_WIDE, nstep)
        student, _, keep_s = make_run(m, STRIP, strip_cpml, 1, source=False)
        teacher.run(1, first_step - 1)
        losses = []
        for it in range(first_step, nstep + 1):
            if (it - first_step) % stride:
                teacher.run(it, it)
                continue
            losses.append(float(net.train_step(student, teacher, (OFFSET,) * 3, it).sum()))
        teacher.close()
        student.close()
        del keep_t, keep_s
        # relative error of the samples that carry a wave (the precursor
        # steps report 0), second half of the run
        tail = [v for v in losses[len(losses) // 2:] if v > 0]
        history.append(float(np.median(tail)) if tail else 0.0)
        if log:
            log(f"run {r}: f0 {m.f0:.1f} Hz, src {m.src}, median loss {history[-1]:.4e}")
    return history


def lattice(n: int, per_axis: int = 8, margin: int = 4):
    """Receiver lattice over the interior (0-based cells)."""
    c = np.linspace(margin, n - 1 - margin, per_axis).round().astype(np.int32)
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    return x.ravel(), y.ravel(), z.ravel()


def _seis(s):
    return np.stack(s.seismograms())  # (3, nrec, nstep)


def reflection(model: Model, nstep: int, net: S.CpmlAbcNet | None = None, per_axis: int = 8,
               reference=None):
    """Reflected energy (dB) and seismogram rel-L2 of the learned strip, the
    plain 20-cell C-PML, the rigid (no absorption) strip and the strip as a
    4-cell C-PML, each against the doubled-domain reference.  Returns a dict
    (and the reference seismograms for reuse)."""
    rec = lattice(model.n, per_axis)
    if reference is None:
        ref, _, keep = make_run(model, NPML_WIDE, NPML_WIDE, nstep, ext=model.n // 2,
                                receivers=rec)
        ref.run()
        reference = _seis(ref)
        ref.close()
        del keep
    e_ref = float((reference.astype(np.float64) ** 2).sum())
    out = {"reference_energy": e_ref}

    def score(u):
        err = float(((u.astype(np.float64) - reference) ** 2).sum())
        return {"db": 10.0 * math.log10(max(err, 1e-300) / e_ref),
                "rel_l2": math.sqrt(err / e_ref)}

    runs = [("cpml20", NPML_WIDE, NPML_WIDE, None), ("rigid4", STRIP, 0, None),
            ("cpml4", STRIP, STRIP, None)]
    if net is not None:
        runs.append(("learned4", STRIP, net.strip_cpml, net))
    for name, pad, absorb, nn in runs:
        s, _, keep = make_run(model, pad, absorb, nstep, receivers=rec)
        if nn is not None:
            s.set_boundary(nn)
        s.run()
        u = _seis(s)
        out[name] = score(u)
        out[name]["finite"] = bool(np.isfinite(u).all())
        s.close()
        del keep
    return out, reference


def reference_seismograms(model: Model, nstep: int, per_axis: int = 8):
    """Seismograms of the doubled-domain 20-cell C-PML reference run."""
    rec = lattice(model.n, per_axis)
    ref, _, keep = make_run(model, NPML_WIDE, NPML_WIDE, nstep, ext=model.n // 2, receivers=rec)
    ref.run()
    u = _seis(ref)
    ref.close()
    del keep
    return u


def strip_reflection_db(model: Model, nstep: int, width: int, theta, reference,
                        per_axis: int = 8, time_block: int = 0) -> float:
    """Reflected energy (dB) of a `width`-cell C-PML strip with profile
    factors theta against the reference seismograms."""
    rec = lattice(model.n, per_axis)
    s, _, keep = make_run(model, width, width, nstep, receivers=rec, theta=theta)
    if time_block:
        s.set_time_block(time_block)
    s.run()
    u = _seis(s)
    s.close()
    del keep
    if not np.isfinite(u).all():
        return 100.0
    e_ref = float((reference.astype(np.float64) ** 2).sum())
    err = float(((u.astype(np.float64) - reference) ** 2).sum())
    return 10.0 * math.log10(max(err, 1e-300) / e_ref)


def learn_profile(width: int = STRIP, n: int = 96, models: int = 4, seed: int = 2000,
                  maxfev: int = 160, nstep: int = 0, log=None):
    """Fit theta of a `width`-cell C-PML to seeded random layered models:
    minimise the mean reflected energy (dB) against their doubled-domain
    20-cell C-PML references (Nelder-Mead from the reference profile).
    nstep (default 6n) must let the reflections come back to the receivers
    through the slow layers.  Returns (theta, history)."""
    from scipy.optimize import minimize

    ms = [layered_model(n, seed + r) for r in range(models)]
    nstep = nstep or 6 * n
    refs = [reference_seismograms(m, nstep) for m in ms]
    hist = []

    def f(x):
        th = (math.exp(x[0]), 1.0 + math.exp(x[1]), math.exp(x[2]), 1.0 + math.exp(x[3]))
        db = float(np.mean([strip_reflection_db(m, nstep, width, th, r)
                            for m, r in zip(ms, refs)]))
        hist.append((th, db))
        if log:
            log(f"theta {tuple(round(v, 4) for v in th)}: {db:.2f} dB")
        return db

    x0 = np.array([0.0, 0.0, 0.0, math.log(1e-3)])  # (1, 2, 1, ~1): the reference profile
    res = minimize(f, x0, method="Nelder-Mead",
                   options={"maxfev": maxfev, "xatol": 1e-3, "fatol": 0.02,
                            "initial_simplex": x0 + np.vstack([np.zeros(4), 0.5 * np.eye(4)])})
    best = min(hist, key=lambda t: t[1])
    return best[0], hist


def time_to_solution(model: Model, nstep: int, net: S.CpmlAbcNet, repeats: int = 1,
                     time_block: int = 2):
    """Wall time of nstep steps: learned strip ((n+8)^3 + corrector), the
    strip as a 4-cell C-PML and the strip without absorption ((n+8)^3)
    against the 20-cell C-PML ((n+40)^3), no receivers.  Every form runs its
    fastest path: the C-PML and rigid grids are timed with single steps and
    with `time_block` steps per launch and the faster is kept (small grids,
    where most tiles touch a slab, favour single steps); the corrector acts
    after every step, so its form runs single steps.  Each entry records the
    steps per launch it used."""
    import torch

    blocks = (0, time_block) if time_block else (0,)
    res = {}
    for name, pad, absorb, nn, tbs in (("cpml20", NPML_WIDE, NPML_WIDE, None, blocks),
                                       ("learned4", STRIP, net.strip_cpml, net, (0,)),
                                       ("strip_single", STRIP, net.strip_cpml, None, (0,)),
                                       ("cpml4", STRIP, STRIP, None, blocks),
                                       ("rigid4", STRIP, 0, None, blocks)):
        for tb in tbs:
            s, _, keep = make_run(model, pad, absorb, nstep)
            if nn is not None:
                s.set_boundary(nn)
            elif tb:
                s.set_time_block(tb)
            s.run(1, min(16, nstep))  # build graphs
            best = None
            for _ in range(repeats):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                s.run(1, nstep)
                torch.cuda.synchronize()
                el = time.perf_counter() - t0
                best = el if best is None else min(best, el)
            if name not in res or best < res[name]["seconds"]:
                res[name] = {"seconds": best, "cells": (model.n + 2 * pad) ** 3,
                             "time_block": tb}
            s.close()
            del keep
    res["speedup"] = res["cpml20"]["seconds"] / res["learned4"]["seconds"]
    res["speedup_cpml4"] = res["cpml20"]["seconds"] / res["cpml4"]["seconds"]
    return res


def corrector_flops(n_grid: int) -> float:
    """Algorithmic FLOPs of one corrector pass on an n_grid^3 grid: per face
    column, layer 1 at 3 depths (27*9*16 MACs each), layer 2 at 2 depths
    (27*16*16), layer 3 at depth 0 (27*16*9); 2 FLOPs per MAC."""
    macs = 3 * 27 * 9 * 16 + 2 * 27 * 16 * 16 + 1 * 27 * 16 * 9
    return 2.0 * macs * 6 * n_grid * n_grid


WEIGHTS_FILE = Path(__file__).resolve().parent / "data" / "abc_weights.npz"


def save_weights(net: S.CpmlAbcNet, path=WEIGHTS_FILE, **meta):
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(path, weights=net.weights, scales=net.scales,
                        strip_cpml=np.array(net.strip_cpml, np.int32),
                        meta=np.array(repr(meta)))


def load_net(path=WEIGHTS_FILE):
    d = np.load(path, allow_pickle=False)
    net = S.CpmlAbcNet(d["scales"], d["weights"])
    net.strip_cpml = int(d["strip_cpml"]) if "strip_cpml" in d.files else 0
    return net


def main(argv=None):
    import argparse

    ap = argparse.ArgumentParser(description="train the learned thin absorbing boundary")
    ap.add_argument("--n", type=int, default=128, help="teacher interior cells per axis")
    ap.add_argument("--runs", type=int, default=96)
    ap.add_argument("--nstep", type=int, default=600)
    ap.add_argument("--lr", type=float, default=2e-4)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--noise", type=float, default=0.3,
                    help="relative input noise (stabilises long closed-loop runs)")
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--strip-cpml", type=int, default=STRIP_CPML,
                    help="C-PML cells of the student strip under the corrector (0: none)")
    ap.add_argument("--out", default=str(WEIGHTS_FILE))
    args = ap.parse_args(argv)
    net = S.CpmlAbcNet(field_scales(), init_weights(args.seed, std=0.02))
    net.train_init(lr=args.lr, batch=args.batch)
    net.train_noise(args.noise)
    t = time.perf_counter()
    hist = train(net, n=args.n, runs=args.runs, nstep=args.nstep, seed=args.seed,
                 log=lambda m: print(m, flush=True), strip_cpml=args.strip_cpml)
    save_weights(net, args.out, n=args.n, runs=args.runs, nstep=args.nstep, lr=args.lr,
                 noise=args.noise, seed=args.seed, strip_cpml=args.strip_cpml, loss=hist,
                 seconds=time.perf_counter() - t)
    print(f"saved {args.out} after {time.perf_counter() - t:.1f} s", flush=True)


if __name__ == "__main__":
    main()
