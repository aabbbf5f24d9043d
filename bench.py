#!/usr/bin/env python3
"""Benchmark of the C-PML time loop on one MI355X.

Workload (default): 3D isotropic elastic C-PML, 768^3 points, 10 m, NPML 20,
FP32, seeded layered medium (12 layers along z, cp ~ U[1500,5500] m/s,
cs = cp/sqrt(3), rho = 310 cp^0.25) with a +-5% Gaussian-smoothed (sigma = 8
cells) perturbation, Ricker 8 Hz at the centre, 32 x 32 receivers on the top
interior plane, DELTAT for Courant 0.5.  One step = one leapfrog time step
(stress + C-PML, velocity + C-PML, source, rigid faces, receivers) over the
whole grid.  Inputs are resident in HBM before the timed region.

Prints ONE JSON line: Gcell-updates/s, HIP-event kernel durations and the
HBM roofline of the dominant kernel, and a CPU baseline (the C oracle, the
reference loop structure with OpenMP) timed on a bounded sub-block.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0


def cpu_baseline(w, lam, mu, rho, nzs=128, seconds=15.0, threads=None):
    """Time the oracle's 3D loop (reference structure, full-grid memory
    variables) on the top nzs planes of the same medium."""
    import numpy as np

    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as orc

    threads = threads or min(16, os.cpu_count() or 1)
    n = w.nx
    z0 = w.nz - nzs
    lam_h = lam[z0:].cpu().numpy()
    mu_h = mu[z0:].cpu().numpy()
    rho_h = rho[z0:].cpu().numpy()
    px, py, pz = w.prof
    pzs = pz[:, z0:]
    prec = "f32" if w.prec == 4 else "f64"
    fx, fy = w.force
    state = None
    ixr = np.array([n // 2], np.int32)
    args = dict(prec=prec, nthreads=threads, energy_slabs=1)

    def run(nst, st):
        return orc.run3d(n, n, nzs, w.dx, w.dx, w.dx, w.dt, nst, lam_h, mu_h, rho_h, None, px,
                         py, pzs, n // 2, n // 2, nzs // 2, fx[:nst], fy[:nst], ixr, ixr,
                         np.array([nzs // 2], np.int32), state=st, **args)

    out = run(1, state)  # warm-up step (also pages the arrays in)
    t = time.perf_counter()
    out = run(2, out["state"])
    t1 = (time.perf_counter() - t) / 2
    steps = max(1, min(400, int(seconds / max(t1, 1e-3))))
    t = time.perf_counter()
    run(steps, out["state"])
    el = time.perf_counter() - t
    cells = n * n * nzs
    return {"value": round(cells * steps / el / 1e9, 4), "unit": "Gcell/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle C loop (reference structure, OpenMP) on the top {n}x{n}x{nzs} "
                      f"planes of the same medium, {steps} steps, {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=768, help="grid points per axis")
    ap.add_argument("--kchunk", type=int, default=0)
    ap.add_argument("--tile", type=int, default=0, help="rows per 3D tile (4, 8, 16)")
    ap.add_argument("--two-phase", action="store_true",
                    help="separate stress and velocity kernels instead of the fused step")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a device copy of 2 GiB (reference for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; runs on several GPUs are not part of it "
              f"(asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    import torch

    from seismic_cpml_amd import workloads

    torch.cuda.set_device(0)
    if args.copy_calibration:
        a = torch.ones(2 ** 29, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    nstep = args.warmup + args.steps
    w = workloads.layered_3d(n=args.n, nstep=nstep)
    s = w.make_solver()
    if args.kchunk:
        s.set_kchunk(args.kchunk)
    if not args.two_phase:
        s.set_fused(True)
    if args.tile:
        s.set_tile(args.tile)
    slabs = s.slab_widths()
    if args.warmup > 0:
        s.run(1, args.warmup)
    s.set_kernel_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.run(args.warmup + 1, nstep)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    cells = w.cells
    value = cells * args.steps / wall / 1e9
    nb = workloads.bytes_per_step(w, slabs)
    kt = {name: s.kernel_time(k) for k, name in ((0, "stress"), (1, "velocity"))}
    avg = {name: ms / max(n, 1) for name, (n, ms) in kt.items()}
    dom = max(avg, key=avg.get)
    achieved = nb[dom] / (avg[dom] * 1e-3) / 1e9
    step_gbs = nb["step"] / (wall / args.steps) / 1e9
    vnorm = s.velocity_norm()
    s.close()

    cpu = None
    if not args.no_cpu_baseline:
        lam, mu, rho = w.medium_dev
        cpu = cpu_baseline(w, lam, mu, rho, seconds=args.cpu_seconds)

    line = {
        "metric": "Gcell-updates/s (3D 768^3 FP32 C-PML time loop)",
        "value": round(value, 3),
        "unit": "Gcell/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"3D isotropic elastic C-PML {args.n}^3, NPML 20, 2nd order, layered "
                        "+-5% smoothed medium, Ricker 8 Hz, 1024 receivers",
            "grid": [w.nx, w.ny, w.nz], "npml": w.npml, "deltat": w.dt,
            "slab_widths": slabs,
            "bytes_per_cell_update": round(nb["step"] / cells, 3),
            "step_hbm_gbs": round(step_gbs, 1),
            "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 4),
            "kernel_ms": {k: round(v, 4) for k, v in avg.items()},
            "max_velocity_norm": vnorm,
        },
        "roofline": {
            "bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
