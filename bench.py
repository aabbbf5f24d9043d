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

Other workloads (--workload): ``2d`` -- 2D 4th order 4096^2, 5 m, NPML 20,
FP32, 12 seeded layers, Ricker 10 Hz at the centre, 256 receivers on the top
interior row, 4000 steps; ``batch`` -- 128 2D 4th-order shots with seeded
uneven grids 512..2048 per side, 3000 steps, batched solver against 128
sequential plain-solver runs; ``shipped`` -- 2D2 as shipped (101 x 641, FP64,
2000 steps); ``3d4`` -- the default 3D workload with the 4th-order step (the
elastic update of seismic_CPML_3D_viscoelastic_MPI.f90); ``2dve`` -- the 2D
workload in 2nd order with viscoelastic attenuation (3 SLS); ``ade8`` -- the
2D workload in 8th order stepped by the ADE-PML RK4 scheme.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0


def _orc():
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as orc

    return orc


def host_threads():
    """Host threads given to this GPU's process: OMP_NUM_THREADS where the
    environment sets it (the host's CPU share per GPU), else every logical
    CPU."""
    try:
        n = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        n = 0
    return n if n > 0 else (os.cpu_count() or 1)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_2d(nx, ny, order, dx, dt, lam, mu, rho, px, py, fx, fy, seconds=15.0,
                    threads=None, what=""):
    """Time the oracle's 2D loop (reference structure, full-grid memory
    variables, OpenMP) on a sample grid for about `seconds`."""
    import numpy as np

    orc = _orc()
    threads = threads or host_threads()
    prec = "f32" if lam.dtype == np.float32 else "f64"
    ix = np.array([nx // 2], np.int32)
    iy = np.array([ny // 2], np.int32)

    def run(nst, st):
        return orc.run2d(nx, ny, order, dx, dx, dt, nst, lam, mu, rho, px, py, nx // 2, ny // 2,
                         fx[:nst], fy[:nst], ix, iy, prec=prec, state=st, nthreads=threads)

    out = run(1, None)
    t = time.perf_counter()
    out = run(2, out["state"])
    t1 = (time.perf_counter() - t) / 2
    steps = max(1, min(len(fx), int(seconds / max(t1, 1e-4))))
    t = time.perf_counter()
    run(steps, out["state"])
    el = time.perf_counter() - t
    out = {"value": round(nx * ny * steps / el / 1e9, 4), "unit": "Gcell/s", "cores": threads,
           "kind": "port", "cpu_model": cpu_model(),
           "sample": f"oracle C 2D loop (reference structure, OpenMP) on {what}{nx}x{ny}, "
                     f"{steps} steps, {el:.1f} s"}
    if threads > 1 and ny >= 64:  # the shipped programs are serial: one thread on fewer rows
        r = max(64, ny // 16)
        one = cpu_baseline_2d(nx, r, order, dx, dt, lam[:r], mu[:r], rho[:r], px, py[:, :r], fx,
                              fy, seconds=seconds / 3, threads=1, what=what)
        out["value_1thread"] = one["value"]
        out["sample_1thread"] = one["sample"]
    return out


def cpu_baseline_ade8(w, tab, rows=256, seconds=15.0, threads=None):
    """Time the oracle's ADE-PML RK4 loop (the 8th-order program's
    structure: rates, stage copies and ADE memories as full arrays, OpenMP)
    on the first `rows` rows of the same medium and stage tables."""
    import numpy as np

    orc = _orc()
    threads = threads or host_threads()
    nx, ny = w.nx, min(rows, w.ny)
    lam, mu, rho = (np.pad(np.asarray(a[:ny], np.float32), ((5, 4), (5, 4)), mode="edge")
                    for a in w.medium_host)
    px, py = w.prof

    def ghost(a):  # (4, n) or (n,) -> the program's -4..n+4 layout
        a = np.asarray(a, np.float32)
        return np.pad(a, ((0, 0), (5, 4)) if a.ndim == 2 else (5, 4), mode="edge")

    prof = {"a_x": ghost(tab["a_x"]), "b_x": ghost(tab["b_x"]), "ah_x": ghost(tab["ah_x"]),
            "bh_x": ghost(tab["bh_x"]), "K_x": ghost(px[2]), "Kh_x": ghost(px[5]),
            "a_y": ghost(tab["a_y"][:, :ny]), "b_y": ghost(tab["b_y"][:, :ny]),
            "ah_y": ghost(tab["ah_y"][:, :ny]), "bh_y": ghost(tab["bh_y"][:, :ny]),
            "K_y": ghost(py[2][:ny]), "Kh_y": ghost(py[5][:ny])}
    ix = np.array([nx // 2], np.int32)
    iy = np.array([ny // 2], np.int32)

    def run(nst, st):
        return orc.runade8(nx, ny, w.dx, w.dx, w.dt, nst, w.npml, lam, mu, rho, prof, nx // 2,
                           ny // 2, 10.0, 0.12, 1e7, 0.0, ix, iy, prec="f32", state=st,
                           nthreads=threads)

    out = run(1, None)
    t = time.perf_counter()
    out = run(2, out["state"])
    t1 = (time.perf_counter() - t) / 2
    steps = max(1, int(seconds / max(t1, 1e-4)))
    t = time.perf_counter()
    run(steps, out["state"])
    el = time.perf_counter() - t
    return {"value": round(nx * ny * steps / el / 1e9, 4), "unit": "Gcell/s", "cores": threads,
            "kind": "port", "cpu_model": cpu_model(),
            "sample": f"oracle C ADE-PML RK4 loop (the program's structure, OpenMP) on the first "
                      f"{nx}x{ny} rows of the same medium and stage tables, {steps} steps, "
                      f"{el:.1f} s"}


def cpu_baseline(w, lam, mu, rho, nzs=128, seconds=15.0, threads=None):
    """Time the oracle's 3D loop (reference structure, full-grid memory
    variables; the 4th-order loop for an order-4 workload) on the top nzs
    planes of the same medium."""
    import numpy as np

    orc = _orc()
    threads = threads or host_threads()
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
        if w.order == 4:
            return orc.run3d4(n, n, nzs, w.dx, w.dx, w.dx, w.dt, nst, lam_h, mu_h, rho_h, px,
                              py, pzs, n // 2, n // 2, nzs // 2, fx[:nst], fy[:nst], ixr, ixr,
                              np.array([nzs // 2], np.int32), state=st, **args)
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
    out = {"value": round(cells * steps / el / 1e9, 4), "unit": "Gcell/s", "cores": threads,
           "kind": "port", "cpu_model": cpu_model(),
           "sample": f"oracle C loop (reference structure, OpenMP) on the top {n}x{n}x{nzs} "
                     f"planes of the same medium, {steps} steps, {el:.1f} s"}
    if threads > 1:  # the shipped programs are serial: the same loop on one thread
        one = cpu_baseline(w, lam, mu, rho, nzs=8, seconds=seconds / 3, threads=1)
        out["value_1thread"] = one["value"]
        out["sample_1thread"] = one["sample"]
    return out


REF_BIN = ROOT / "oracle" / "_ref"  # reference programs compiled from their sources
MPI_BIN = Path("/opt/conda/bin")       # the MPI the 3D reference is built against


def cpu_baseline_reference_3d(threads=None):
    """Time the reference's own 3D program (seismic_CPML_3D_isotropic_MPI_OpenMP,
    compiled from its sources by oracle/build_ref.py with OpenMP, float32) on a
    768 x 768 x 64-plane sample: 2 MPI ranks (the program needs an even
    count) of threads/2 OpenMP threads each, runs of 10 and 310 steps; the
    difference of their wall times is 300 steps without the set-up.  Its homogeneous model
    and its energy sum every step, as the program is.  None when the binaries
    or mpirun are not there."""
    lo, hi, nz = 10, 310, 64
    exes = [REF_BIN / f"ref3d_bench{n}_f32" for n in (lo, hi)]
    mpirun = MPI_BIN / "mpirun"
    if not mpirun.is_file() or not all(e.is_file() for e in exes):
        return None
    threads = threads or host_threads()
    per_rank = max(1, threads // 2)
    env = dict(os.environ, OMP_NUM_THREADS=str(per_rank),
               LD_LIBRARY_PATH=f"{MPI_BIN.parent / 'lib'}:" + os.environ.get("LD_LIBRARY_PATH", ""))
    walls = []
    with tempfile.TemporaryDirectory(prefix="cpml_ref3d_") as tmp:
        for exe in exes:
            t = time.perf_counter()
            r = subprocess.run([str(mpirun), "-np", "2", str(exe)], cwd=tmp, env=env,
                               capture_output=True, text=True, timeout=600)
            walls.append(time.perf_counter() - t)
            if r.returncode != 0:
                print(f"bench.py: reference baseline run failed ({r.returncode}): "
                      f"{(r.stdout + r.stderr)[-300:]}", file=sys.stderr)
                return None
    el = walls[1] - walls[0]
    if el <= 0:
        return None
    cells = 768 * 768 * nz
    return {"value": round(cells * (hi - lo) / el / 1e9, 4), "unit": "Gcell/s",
            "cores": 2 * per_rank,
            "kind": "reference", "cpu_model": cpu_model(),
            "sample": "seismic_CPML_3D_isotropic_MPI_OpenMP itself (amdflang -O3 -fopenmp, "
                      f"float32, its homogeneous model, energy every step) on 768x768x{nz}, "
                      f"2 MPI ranks x {per_rank} OpenMP threads: {hi - lo} steps = the "
                      f"{hi}-step run ({walls[1]:.1f} s) minus the {lo}-step run "
                      f"({walls[0]:.1f} s)"}


def _timed_run(s, warmup, nstep):
    import torch

    if warmup > 0:
        s.run(1, warmup)
    s.set_kernel_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.run(warmup + 1, nstep)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def _line(metric, value, args, steps, ms_per_step, dtype, config, roofline, cpu):
    return {
        "metric": metric, "value": round(value, 3), "unit": "Gcell/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic", "config": config, "roofline": roofline, "cpu_baseline": cpu,
    }


def _traffic(key, kernel):
    """Measured HBM bytes per dispatch of `kernel` in configuration `key`
    (profiles/traffic.json, from rocprofv3 FETCH_SIZE / WRITE_SIZE runs)."""
    p = ROOT / "profiles" / "traffic.json"
    if not p.is_file():
        return None
    d = json.loads(p.read_text()).get(key, {}).get(kernel)
    return None if d is None else int(d["read_bytes"] + d["write_bytes"])


BYTES_MODEL = ("two-phase compulsory bytes per cell-update (every array read or written once "
               "per phase, slab memory variables weighted by the slab fraction)")


def _roofline(nb, avg, key=None, canonical=None):
    """Roofline of the dominant kernel.  `canonical`: the bytes the
    canonical model (two-phase compulsory bytes per cell-update) assigns to
    one launch of it; a time-blocked launch of NT steps is NT steps of that
    model, and its own compulsory bytes (the state once per launch) are
    reported beside it."""
    dom = max(avg, key=avg.get)
    sec = avg[dom] * 1e-3
    algo = nb[dom] if canonical is None else canonical
    achieved = algo / sec / 1e9
    out = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
           "traffic": _traffic(key, dom) if key else None, "algorithmic_bytes": int(algo),
           "bytes_model": BYTES_MODEL}
    if canonical is not None:
        own = nb[dom] / sec / 1e9
        out.update({"own_algorithmic_bytes": int(nb[dom]), "own_achieved": round(own, 1),
                    "own_frac": round(own / HBM_PEAK_GBS, 4)})
    return out


def bench_3d(args, order=2):
    import torch

    from seismic_cpml_amd import workloads

    steps = args.steps or 2000
    # default: 2 steps per launch (time-blocked) for the 2nd-order step
    nt = args.time_block if args.time_block >= 0 else (2 if order == 2 and not args.fused else 0)
    if nt > 0:
        steps = max(nt, steps // nt * nt)  # whole blocks only in the timed region
    nstep = args.warmup + steps
    if order == 4 and (args.fused or nt):
        raise SystemExit("bench.py: the fused and time-blocked steps are second order only")
    w = workloads.layered_3d(n=args.n or 768, nstep=nstep, order=order)
    s = w.make_solver()
    if args.fused:  # the fused step's own defaults: 14-row blocks, 48-plane chunks
        args.tile = args.tile or 16
        args.kchunk = args.kchunk or 48
    if args.kchunk:
        s.set_kchunk(args.kchunk)
    if args.fused:
        s.set_fused(True)
    if args.tile:
        s.set_tile(args.tile)
    if nt:
        s.set_time_block(nt, args.strip, args.chunk_rows)
    slabs = s.slab_widths()
    wall = _timed_run(s, args.warmup, nstep)
    torch.cuda.synchronize()

    cells = w.cells
    value = cells * steps / wall / 1e9
    fused = args.fused
    nb = workloads.bytes_per_step(w, slabs, fused=fused, time_block=nt)
    names = (((0, "time_block"),) if nt else ((0, "fused"),) if fused
             else ((0, "stress"), (1, "velocity")))
    kt = {name: s.kernel_time(k) for k, name in names}
    avg = {name: ms / max(n, 1) for name, (n, ms) in kt.items()}
    step_gbs = nb["step"] / (wall / steps) / 1e9
    nb2 = workloads.bytes_per_step(w, slabs, fused=False)
    vnorm = s.velocity_norm()
    s.close()

    cpu = None
    if not args.no_cpu_baseline:
        lam, mu, rho = w.medium_dev
        port = cpu_baseline(w, lam, mu, rho, seconds=args.cpu_seconds)
        ref = cpu_baseline_reference_3d() if order == 2 else None
        cpu = port
        if ref is not None:  # the reference program itself; the port's figures beside it
            cpu = dict(ref, port_value=port["value"], port_sample=port["sample"],
                       port_cores=port["cores"])
            for k in ("value_1thread", "sample_1thread"):
                if k in port:
                    cpu[k] = port[k]
    config = {
        "workload": f"3D isotropic elastic C-PML {w.nx}^3, NPML 20, "
                    f"{'2nd' if order == 2 else '4th'} order, layered "
                    "+-5% smoothed medium, Ricker 8 Hz, 1024 receivers",
        "grid": [w.nx, w.ny, w.nz], "npml": w.npml, "deltat": w.dt,
        "kernel_path": f"time-blocked ({nt} steps per launch)" if nt else
        "fused" if fused else "two-phase",
        "time_block": {"steps": nt, "strip": args.strip, "chunk_planes": args.chunk_rows}
        if nt else None,
        "slab_widths": slabs,
        "bytes_per_cell_update": round(nb["step"] / cells, 3),
        "step_hbm_gbs": round(step_gbs, 1),
        "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 4),
        # the same rate priced with the two-phase byte count (each array once per phase)
        "two_phase_bytes_per_cell_update": round(nb2["step"] / cells, 3),
        "two_phase_model_frac": round(nb2["step"] / (wall / steps) / 1e9 / HBM_PEAK_GBS, 4),
        "kernel_ms": {k: round(v, 4) for k, v in avg.items()},
        "max_velocity_norm": vnorm,
    }
    if order == 4:
        key = f"3d4_{w.nx}_kchunk{args.kchunk or 8}"
        return _line("Gcell-updates/s (3D 768^3 FP32 4th-order C-PML time loop)", value, args,
                     steps, wall / steps * 1e3, "f32", config, _roofline(nb, avg, key), cpu)
    key = (f"3d_{w.nx}_tb{nt}" if nt else
           f"3d_{w.nx}_{'fused' if fused else 'two-phase'}_tile{args.tile or 4}_"
           f"kchunk{args.kchunk or 8}")
    canon = nb2["step"] * nt if nt else (nb2["step"] if fused else None)
    return _line("Gcell-updates/s (3D 768^3 FP32 C-PML time loop)", value, args, steps,
                 wall / steps * 1e3, "f32", config, _roofline(nb, avg, key, canon), cpu)


def bench_2d(args):
    import torch

    from seismic_cpml_amd import workloads

    steps = args.steps or 4000
    nt = args.time_block if args.time_block >= 0 else 2  # default: 2 steps per launch
    if nt > 0:
        steps = max(nt, steps // nt * nt)  # whole blocks only in the timed region
    nstep = args.warmup + steps
    n = args.n or 4096
    w = workloads.layered_2d(n=n, nstep=nstep)
    s = w.make_solver()
    if args.tile:
        s.set_tile(args.tile)
    if args.energy:  # the energy curve as the reference programs compute it
        s.set_energy(args.energy)
    if nt > 0:
        s.set_time_block(nt, args.strip, args.chunk_rows)
    slabs = s.slab_widths()
    wall = _timed_run(s, args.warmup, nstep)
    torch.cuda.synchronize()
    value = w.cells * steps / wall / 1e9
    nb = workloads.bytes_per_step(w, slabs, time_block=nt)
    if nt > 0:
        c, ms = s.kernel_time(0)
        avg = {"time_block": ms / max(c, 1)}
    else:
        kt = {name: s.kernel_time(k) for k, name in ((0, "stress"), (1, "velocity"))}
        avg = {name: ms / max(c, 1) for name, (c, ms) in kt.items()}
    step_gbs = nb["step"] / (wall / steps) / 1e9
    s.close()
    cpu = None
    if not args.no_cpu_baseline:
        lam, mu, rho = w.medium_host
        rows = 1024
        px, py = w.prof
        cpu = cpu_baseline_2d(n, rows, 4, w.dx, w.dt, lam[:rows], mu[:rows], rho[:rows], px,
                              py[:, :rows], *w.force, seconds=args.cpu_seconds,
                              what="the first rows of the same medium, ")
    config = {
        "workload": f"2D isotropic elastic C-PML {n}^2, 4th order, NPML 20, 12 seeded layers, "
                    "Ricker 10 Hz at the centre, 256 receivers on the top interior row",
        "grid": [w.nx, w.ny], "npml": w.npml, "deltat": w.dt, "slab_widths": slabs,
        "tile_rows": args.tile or 4,
        "time_block": {"steps": nt, "strip": args.strip, "chunk_rows": args.chunk_rows}
        if nt else None,
        "energy_cadence": args.energy,
        "bytes_per_cell_update": round(nb["step"] / w.cells, 3),
        "step_hbm_gbs": round(step_gbs, 1), "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 4),
        "kernel_ms": {k: round(v, 4) for k, v in avg.items()},
    }
    return _line("Gcell-updates/s (2D 4096^2 4th-order FP32 C-PML time loop)", value, args,
                  steps, wall / steps * 1e3, "f32", config,
                  _roofline(nb, avg, f"2d_{n}_tb{nt}" if nt else f"2d_{n}_tile{args.tile or 4}",
                            workloads.bytes_per_step(w, slabs)["step"] * nt if nt else None),
                  cpu)


def bench_2dve(args):
    """2D viscoelastic C-PML (2nd order, 3 SLS): the layered 2D workload with
    attenuation on."""
    import numpy as np
    import torch

    from seismic_cpml_amd import workloads

    steps = args.steps or 4000
    nstep = args.warmup + steps
    n = args.n or 4096
    w = workloads.layered_2d(n=n, nstep=nstep, order=2)
    taus = workloads.taus_for(w.f0)
    s = w.make_solver()
    s.set_attenuation(*taus)
    slabs = s.slab_widths()
    wall = _timed_run(s, args.warmup, nstep)
    torch.cuda.synchronize()
    value = w.cells * steps / wall / 1e9
    nb = workloads.bytes_per_step(w, slabs, nsls=3)
    kt = {name: s.kernel_time(k) for k, name in ((0, "stress"), (1, "velocity"))}
    avg = {name: ms / max(c, 1) for name, (c, ms) in kt.items()}
    s.close()
    cpu = None
    if not args.no_cpu_baseline:
        orc = _orc()
        lam, mu, rho = w.medium_host
        rows = 2048
        px, py = w.prof
        fx, fy = w.force
        threads = host_threads()
        ixr = np.array([n // 2], np.int32)
        iyr = np.array([rows // 2], np.int32)

        def run(nst, st):
            return orc.run2dve(n, rows, w.dx, w.dx, w.dt, nst, lam[:rows], mu[:rows],
                               rho[:rows], px, py[:, :rows], taus, n // 2, rows // 2,
                               fx[:nst], fy[:nst], ixr, iyr, prec="f32", state=st,
                               nthreads=threads)

        out = run(1, None)
        t = time.perf_counter()
        out = run(2, out["state"])
        t1 = (time.perf_counter() - t) / 2
        k = max(1, min(2000, int(args.cpu_seconds / max(t1, 1e-3))))
        t = time.perf_counter()
        run(k, out["state"])
        el = time.perf_counter() - t
        cpu = {"value": round(n * rows * k / el / 1e9, 4), "unit": "Gcell/s", "cores": threads,
               "kind": "port",
               "sample": f"oracle C loop (reference structure, OpenMP) on {n}x{rows} of the "
                         f"same medium, {k} steps, {el:.1f} s"}
    config = {
        "workload": f"2D viscoelastic C-PML {n}^2, 2nd order, 3 SLS (Qp 65, Qs 55), NPML 20, "
                    "12 seeded layers, Ricker 10 Hz at the centre, 256 receivers",
        "grid": [w.nx, w.ny], "npml": w.npml, "deltat": w.dt, "slab_widths": slabs,
        "bytes_per_cell_update": round(nb["step"] / w.cells, 3),
        "kernel_ms": {k: round(v, 4) for k, v in avg.items()},
    }
    return _line("Gcell-updates/s (2D 4096^2 viscoelastic FP32 C-PML time loop)", value, args,
                 steps, wall / steps * 1e3, "f32", config, _roofline(nb, avg, f"2dve_{n}"), cpu)


def bench_ade8(args):
    """ADE-PML RK4 8th-order step (the 8th-order program's scheme) on the
    layered 2D workload in 8th order.  The stage tables are the C-PML
    profiles of the workload repeated over the four stages (the timing does
    not depend on their values); the force at each stage time is the
    workload's Ricker."""
    import numpy as np
    import torch

    from seismic_cpml_amd import workloads

    steps = args.steps or 500
    nstep = args.warmup + steps
    n = args.n or 4096
    w = workloads.layered_2d(n=n, nstep=nstep, order=8)
    s = w.make_solver()
    px, py = w.prof
    tab = {"a_x": np.repeat(px[0][None], 4, 0), "b_x": np.repeat(px[1][None], 4, 0),
           "ah_x": np.repeat(px[3][None], 4, 0), "bh_x": np.repeat(px[4][None], 4, 0),
           "a_y": np.repeat(py[0][None], 4, 0), "b_y": np.repeat(py[1][None], 4, 0),
           "ah_y": np.repeat(py[3][None], 4, 0), "bh_y": np.repeat(py[4][None], 4, 0)}
    fx, fy = w.force
    s.set_ade_rk4(tab, np.repeat(np.asarray(fx)[:, None], 4, 1),
                  np.repeat(np.asarray(fy)[:, None], 4, 1))
    wall = _timed_run(s, args.warmup, nstep)
    torch.cuda.synchronize()
    s.close()
    value = w.cells * steps / wall / 1e9
    # canonical words per cell and step, the program's own array structure
    # (rates stored between the rate and accumulation phases, ADE8:692-906):
    # copy and first stage state 20; per stage velocity rates 6, stress
    # rates 7, accumulation with the next stage's state 25 (the last one 15)
    words = 20 + 4 * (6 + 7) + 3 * 25 + 15
    # this build's own compulsory words (X ping-pongs, no copy): stage 1
    # reads X0 5 (its state) and media 3 and writes X 5 and the next state
    # 5; stages 2-3 read the state 5, media 3, X 5, X0 5 and write X 5 and the
    # next state 5; stage 4 reads 13 and writes 5 (band memory variables not
    # counted)
    own_words = 18 + 2 * 28 + 18
    step_gbs = 4 * words * w.cells / (wall / steps) / 1e9
    own_gbs = 4 * own_words * w.cells / (wall / steps) / 1e9
    config = {
        "workload": f"2D ADE-PML RK4 {n}^2, 8th order (Holberg), NPML 20, 12 seeded layers, "
                    "Ricker 10 Hz at the centre, 256 receivers",
        "grid": [w.nx, w.ny], "npml": w.npml, "deltat": w.dt,
        "bytes_per_cell_update": 4 * words, "own_bytes_per_cell_update": 4 * own_words,
        "step_hbm_gbs": round(step_gbs, 1), "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 4),
        "launches_per_step": 4,
    }
    roof = {"bound": "hbm", "kernel": "step (4 launches)", "achieved": round(step_gbs, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(step_gbs / HBM_PEAK_GBS, 4),
            "traffic": None, "algorithmic_bytes": 4 * words * w.cells,
            "bytes_model": "the program's array structure (rates stored between phases)",
            "own_algorithmic_bytes": 4 * own_words * w.cells, "own_achieved": round(own_gbs, 1),
            "own_frac": round(own_gbs / HBM_PEAK_GBS, 4)}
    cpu = None if args.no_cpu_baseline else cpu_baseline_ade8(w, tab, seconds=args.cpu_seconds)
    return _line("Gcell-updates/s (2D 4096^2 ADE-PML RK4 8th-order FP32 time loop)", value, args,
                 steps, wall / steps * 1e3, "f32", config, roof, cpu)


def bench_batch(args):
    import numpy as np
    import torch

    from seismic_cpml_amd import workloads

    steps = args.steps or 3000
    w = workloads.batch_2d(nshots=args.shots, nstep=steps)
    t = time.perf_counter()
    b = w.make_batch()
    setup = time.perf_counter() - t
    cache = (Path(tempfile.gettempdir()) / "cpml_batch_tune.txt" if args.tune_cache is None
             else args.tune_cache)
    Path(cache).parent.mkdir(parents=True, exist_ok=True)
    # the autotuner leaves the batch at rest with the fastest launch shape set
    ty, bpc, tune_ms = b.autotune(trial_steps=min(30, steps), cache_path=str(cache))
    _, _, tb, jc = b.launch_shape()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run(1, steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cells = w.cells
    value = cells * steps / wall / 1e9
    seis = [b.seismograms(q) for q in range(len(w.shots))]
    b.close()
    # the same shots one after another with the plain solver: its default
    # single-step kernels, and each shot autotuned (single steps or
    # time-blocked, whichever its grid runs faster; tuning time counts in
    # with_setup_s only)
    nseq = len(w.shots) if args.seq_shots is None else min(args.seq_shots, len(w.shots))
    seq_cells = sum(s["NX"] * s["NY"] for s in w.shots[:nseq])

    def sequential(tuned):
        run, total, same, blocked = 0.0, 0.0, 0, 0
        for q in range(nseq):
            t0 = time.perf_counter()
            h = w.make_solver(q)
            if tuned:
                blocked += int(h.autotune(trial_steps=min(30, steps))[0] > 0)
            t1 = time.perf_counter()
            h.run()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            sx, sy = h.seismograms()
            h.close()
            run += t2 - t1
            total += t2 - t0
            same += int(np.array_equal(sx, seis[q][0]) and np.array_equal(sy, seis[q][1]))
        value = seq_cells * steps / run / 1e9 if nseq else None
        return {"shots": nseq, "autotuned": tuned, "gcell_s": value and round(value, 3),
                "shots_time_blocked": blocked,
                "run_s": round(run, 3), "with_setup_s": round(total, 3),
                "bitwise_equal_seismograms": same}

    seq = sequential(False)
    seq_tuned = sequential(True)
    seq_value = seq["gcell_s"]
    es = 4 if w.prec == 4 else 8
    # algorithmic bytes per cell-update (two-phase compulsory): 18 words of the
    # 2D interior; slabs add under 1%.  A time-blocked launch moves 13 words
    # per point per tb steps (its own compulsory bytes, reported beside it).
    bytes_per = es * 18
    own_per = es * (13.0 / tb if tb else 18)
    config = {
        "workload": f"{len(w.shots)} shots of 2D 4th-order C-PML, NX, NY ~ U{{512..2048}} "
                    "(seeded), NPML 20, per-shot layered model, source and Ricker f0 in "
                    "[5, 15] Hz, 64 receivers each",
        "cells_per_step": cells, "setup_s": round(setup, 2),
        "autotune": {"tile_rows": ty, "blocks_per_cu": bpc, "time_block": tb,
                     "chunk_rows": jc, "trial_ms": tune_ms},
        "sequential": seq,
        "sequential_autotuned": seq_tuned,
        "batched_vs_sequential": seq_value and round(value / seq_value, 3),
        "batched_vs_sequential_autotuned":
            seq_tuned["gcell_s"] and round(value / seq_tuned["gcell_s"], 3),
        "step_hbm_frac": round(bytes_per * value / HBM_PEAK_GBS, 4),
    }
    roof = {"bound": "hbm", "kernel": "batch step", "achieved": round(bytes_per * value, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(bytes_per * value / HBM_PEAK_GBS, 4),
            "traffic": _traffic(f"batch_{len(w.shots)}shots_tile{ty}_bpc{bpc}_tb{tb}", "step"),
            "bytes_model": BYTES_MODEL, "own_achieved": round(own_per * value, 1),
            "own_frac": round(own_per * value / HBM_PEAK_GBS, 4)}
    cpu = None
    if not args.no_cpu_baseline:
        s0 = w.shots[0]
        cpu = cpu_baseline_2d(s0["NX"], s0["NY"], 4, s0["DELTAX"], s0["DELTAT"], s0["lam"],
                              s0["mu"], s0["rho"], s0["prof_x"], s0["prof_y"], s0["fx"],
                              s0["fy"], seconds=args.cpu_seconds, what="shot 0, ")
    return _line("Gcell-updates/s (128-shot batched 2D 4th-order FP32 C-PML)", value, args,
                 steps, wall / steps * 1e3, "f32", config, roof, cpu)


def bench_shipped(args):
    import torch

    from seismic_cpml_amd import workloads

    w = workloads.shipped_2d(2, 8)
    s = w.make_solver()
    wall = _timed_run(s, 0, w.nstep)
    torch.cuda.synchronize()
    value = w.cells * w.nstep / wall / 1e9
    s.close()
    cpu = None
    if not args.no_cpu_baseline:
        lam, mu, rho = w.medium_host
        cpu = cpu_baseline_2d(w.nx, w.ny, 2, w.dx, w.dt, lam, mu, rho, *w.prof, *w.force,
                              seconds=args.cpu_seconds, threads=1, what="the shipped grid ")
    config = {"workload": "seismic_CPML_2D_isotropic_second_order as shipped: 101 x 641, "
                          "FP64, 2000 steps"}
    return _line("Gcell-updates/s (2D2 as shipped, FP64)", value, args, w.nstep,
                 wall / w.nstep * 1e3, "f64", config, None, cpu)


def bench_abc(args):
    """Learned thin boundary against the 20-cell C-PML: time to solution of
    the same steps on the 3D interior n^3, reflected energy against a
    doubled-domain reference, and the corrector's MFMA rate."""
    from seismic_cpml_amd import learned_abc as L

    n = args.n or 384
    if args.abc_weights and Path(args.abc_weights).is_file():
        net = L.load_net(args.abc_weights)
        trained = f"weights from {Path(args.abc_weights).name}"
    else:
        net = L.S.CpmlAbcNet(L.field_scales(), L.init_weights(args.abc_seed, std=0.02))
        net.train_init(lr=2e-4, batch=8)
        net.train_noise(0.3)
        t = time.perf_counter()
        L.train(net, n=128, runs=args.abc_runs, nstep=600, seed=args.abc_seed,
                strip_cpml=L.STRIP_CPML)
        trained = f"trained here: {args.abc_runs} teacher runs of 128^3, {time.perf_counter() - t:.0f} s"
    m = L.layered_model(n, seed=args.abc_seed + 7777, f0=10.0)
    nstep = args.steps or 3 * n
    tts = L.time_to_solution(m, nstep, net)
    # corrector alone: the learned grid's time without the corrector
    base = tts["strip_single"]["seconds"]  # the same grid and step path without the corrector
    corr_s = max(tts["learned4"]["seconds"] - base, 0.0) / nstep
    flops = L.corrector_flops(n + 2 * L.STRIP)
    refl, _ = L.reflection(m, nstep, net, per_axis=args.abc_lattice)
    tflops = flops / max(corr_s, 1e-12) / 1e12
    config = {
        "workload": f"3D {n}^3 interior FP32, layered seeded medium, Ricker 10 Hz, {nstep} steps:"
                    f" learned 4-cell strip ({n + 8}^3 + corrector"
                    f"{' over a 4-cell C-PML' if net.strip_cpml else ''}) vs 20-cell C-PML "
                    f"({n + 40}^3); reflection against a doubled-domain reference",
        "corrector": trained,
        "strip_cpml_cells": net.strip_cpml,
        "time_s": {k: round(tts[k]["seconds"], 4) for k in ("cpml20", "learned4", "cpml4",
                                                             "strip_single",
                                                             "rigid4")},
        "paths": "20-cell C-PML, 4-cell C-PML and rigid strip: single steps and 2 steps "
                 "per launch timed, the faster kept; the corrector form single steps (it "
                 "acts after every step)",
        "steps_per_launch": {k: tts[k]["time_block"] for k in ("cpml20", "learned4", "cpml4",
                                                              "strip_single", "rigid4")},
        "speedup_cpml4_without_corrector": round(tts["speedup_cpml4"], 4),
        "corrector_ms_per_step": round(corr_s * 1e3, 4),
        "reflection_db": {k: round(v["db"], 2) for k, v in refl.items() if isinstance(v, dict)},
        "seismogram_rel_l2": {k: round(v["rel_l2"], 5) for k, v in refl.items()
                              if isinstance(v, dict)},
    }
    roof = {"bound": "mfma", "kernel": "strip corrector (bf16 MFMA, fp32 accumulate)",
            "achieved": round(tflops, 2), "peak": 2500.0, "unit": "TFLOP/s",
            "frac": round(tflops / 2500.0, 4), "traffic": None}
    net.close()
    return {
        "metric": f"learned-ABC vs 20-cell C-PML time-to-solution (3D {n}^3 FP32)",
        "value": round(tts["speedup"], 4), "unit": "x", "n_gpus": 1, "steps": nstep,
        "warmup": 16, "ms_per_step": round(tts["learned4"]["seconds"] / nstep * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic", "config": config, "roofline": roof, "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload",
                    choices=("3d", "3d4", "2d", "2dve", "ade8", "batch", "shipped", "abc"),
                    default="3d")
    ap.add_argument("--steps", type=int, default=0, help="timed steps (0: the workload's own)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=0, help="grid points per axis (3d, 2d)")
    ap.add_argument("--kchunk", type=int, default=0)
    ap.add_argument("--tile", type=int, default=0, help="rows per tile (4, 8, 16)")
    ap.add_argument("--time-block", type=int, default=-1,
                    help="leapfrog steps per launch (time blocking; 0: off; -1: the "
                         "workload's default: 2 for 3d and 2d)")
    ap.add_argument("--strip", type=int, default=0, help="2d time blocking: strip width")
    ap.add_argument("--energy", type=int, default=0,
                    help="2d: energy curve every N steps (0: off; the reference programs: 1)")
    ap.add_argument("--chunk-rows", type=int, default=0, help="2d time blocking: chunk rows")
    ap.add_argument("--fused", action="store_true",
                    help="3d: one fused kernel per step instead of the stress and velocity "
                         "kernels")
    ap.add_argument("--shots", type=int, default=128, help="batch: number of shots")
    ap.add_argument("--seq-shots", type=int, default=None,
                    help="batch: shots to rerun sequentially (default all)")
    ap.add_argument("--tune-cache", default=None, help="batch: autotuner cache file")
    ap.add_argument("--abc-weights", default=str(ROOT / "seismic_cpml_amd" / "data" /
                                                 "abc_weights.npz"),
                    help="abc: trained corrector (trains a short one when missing)")
    ap.add_argument("--abc-runs", type=int, default=8, help="abc: teacher runs if training")
    ap.add_argument("--abc-seed", type=int, default=1000)
    ap.add_argument("--abc-lattice", type=int, default=8, help="abc: receivers per axis")
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

    torch.cuda.set_device(0)
    if args.copy_calibration:
        a = torch.ones(2 ** 29, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    fn = {"3d": bench_3d, "3d4": lambda a: bench_3d(a, order=4), "2d": bench_2d,
          "2dve": bench_2dve, "ade8": bench_ade8,
          "batch": bench_batch, "shipped": bench_shipped, "abc": bench_abc}
    print(json.dumps(fn[args.workload](args)), flush=True)


if __name__ == "__main__":
    main()
