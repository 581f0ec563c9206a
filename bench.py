#!/usr/bin/env python3
"""FDTD throughput on one MI355X: Mcell-updates/s and % of HBM peak.

Workload (default): the 400x400x200 dipole-over-random-slab model (32 M cells, 12-cell PML on all
faces, a 1-cell 50 ohm gap source, running DFTs at 51 frequencies on 6 planes), float32.  One bench
step is one full FDTD time step: E update, E sources/probes, H update, H probes, DFT accumulation.
Inputs (fields, coefficients, source series, DFT tables) are resident in HBM before timing starts.

Prints one JSON line with the throughput, the roofline of the dominant kernel (H update: 144
algorithmic bytes per cell: 6 split H read and written, 18 coefficients, 3 E totals read, 3 H totals
written), a measured copy-kernel bandwidth, and the CPU oracle (a restatement of the reference's
threaded C++ solver) timed on this host over a bounded window of the same workload.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
E_BYTES, H_BYTES = 120, 144  # algorithmic bytes per cell of the E and H half steps (sum 264)
METRIC = "Mcell-updates/s + % HBM peak, batched vs plain; adjoint gradients/s, 1 GPU"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_model(name, Nt):
    from rfnetwork_amd import workloads

    if name == "dipole_slab":
        return workloads.dipole_slab(Nt=Nt)
    if name == "microstrip":
        return workloads.microstrip(Nt=Nt)
    raise SystemExit(f"unknown workload {name}")


def dft_bytes_per_step(model):
    """Algorithmic DFT traffic per step: 4 B field read + 16 B complex64 read-modify-write per
    plane cell and frequency."""
    total = 0
    for m in model.monitors.values():
        shape = list(model.fshape[m["field"]])
        shape.pop(m["axis"])
        cells = shape[0] * shape[1]
        nf = 0 if m["frequency"] is None else len(m["frequency"])
        total += cells * (4 + 16 * nf)
    return total


def cpu_baseline(model, budget_s):
    """Time the CPU oracle on this host over a bounded number of steps of the same workload."""
    from oracle import cpu_solver

    threads = max(1, min(16, os.cpu_count() or 1, 20))
    cells = model.Nx * model.Ny * model.Nz

    def run(nsteps):
        c, probes, monitors, mem, Nx, Ny, Nz, Nt = model.pack(gpu=False, n_threads=threads)
        for p in probes:
            p["values"] = np.ascontiguousarray(p["values"][:nsteps])
        for m in monitors:
            if "dtft_phase" in m:
                m["dtft_phase"] = np.ascontiguousarray(m["dtft_phase"][: nsteps // m["n_step"] + 1])
            else:
                m["values"] = np.ascontiguousarray(m["values"][: nsteps // m["n_step"] + 1])
        t0 = time.perf_counter()
        cpu_solver.solver_run(c, probes, monitors, mem, Nx, Ny, Nz, nsteps, threads, 0)
        return time.perf_counter() - t0

    t_probe = run(2)
    per_step = t_probe / 2
    steps = int(max(4, min(200, budget_s / max(per_step, 1e-6))))
    dt = run(steps)
    return dict(value=round(cells * steps / dt / 1e6, 2), unit="Mcell-updates/s", cores=threads, kind="port",
                sample=f"{steps} steps of the full {model.Nx}x{model.Ny}x{model.Nz} workload from zero fields "
                       f"(oracle restating rfnetwork/core/src/solver.cpp, {threads} threads, x slabs, 2 barriers "
                       f"per step, -O2 -ffp-contract=off), {dt:.1f} s wall")


def batched_leg(n_steps):
    """The 4 uneven port variants x every port driven (12 excitations) advanced in one batched
    session, against the same 12 excitations run one after another through single-grid sessions.
    Both paths are prepared (coefficients, sources resident) before their timed runs; results must
    agree bit for bit."""
    from rfnetwork_amd import batch as rb, workloads
    from rfnetwork_amd.session import Session

    t0 = time.perf_counter()
    models = [workloads.port_variant(v, Nt=n_steps) for v in range(len(workloads.PORT_VARIANTS))]
    grids, inputs = [], []
    for m in models:
        ports = list(range(1, len(m.ports) + 1))
        exc, Nt = rb.pack_excitations(m, ports)
        coeff = m.coefficient_arrays(gpu=True)
        grids.append((coeff, m.Nx, m.Ny, m.Nz, len(ports)))
        inputs.append((m, coeff, exc, Nt))
    exc_cells = sum(m.Nx * m.Ny * m.Nz * len(e) for m, _, e, _ in inputs)
    log(f"batched: {len(models)} grids, {sum(len(e) for _, _, e, _ in inputs)} excitations, "
        f"{exc_cells} excitation-cells, built in {time.perf_counter() - t0:.1f}s")

    b = rb.Batch(grids)
    log(f"batched: autotune {b.autotune()}")
    k = 0
    for _, _, exc, Nt in inputs:
        for _, probes, monitors in exc:
            b.set_sources(k, probes, monitors, Nt)
            k += 1
    t0 = time.perf_counter()
    b.run(n_steps)
    b.synchronize()
    wall_b = time.perf_counter() - t0
    b.read()
    b.close()
    batched_series = [[np.array(p["values"]) for p in probes] for _, _, exc, _ in inputs for _, probes, _ in exc]
    log(f"batched: {n_steps} steps in {wall_b:.3f}s")

    sessions = []
    for m, coeff, exc, Nt in inputs:
        for port in range(len(exc)):
            exc2, _ = rb.pack_excitations(m, [port + 1])
            _, probes, monitors = exc2[0]
            s = Session(coeff, m.Nx, m.Ny, m.Nz)
            s.autotune()
            s.set_sources(probes, monitors, Nt)
            sessions.append((s, probes))
    t0 = time.perf_counter()
    for s, _ in sessions:
        s.run(n_steps)
        s.synchronize()
    wall_s = time.perf_counter() - t0
    identical = True
    for (s, probes), ref in zip(sessions, batched_series):
        s.read()
        identical &= all(np.array_equal(p["values"], r) for p, r in zip(probes, ref))
        s.close()
    log(f"sequential: {wall_s:.3f}s, identical {identical}")
    b_rate = exc_cells * n_steps / wall_b / 1e6
    s_rate = exc_cells * n_steps / wall_s / 1e6
    return dict(workload="4 port variants 64^3..192x128x96, 12 excitations", steps=n_steps,
                excitation_cells=exc_cells, batched_mcells=round(b_rate, 1), sequential_mcells=round(s_rate, 1),
                speedup=round(wall_s / wall_b, 3), batched_s=round(wall_b, 3), sequential_s=round(wall_s, 3),
                identical=bool(identical))


def adjoint_leg(n_steps, every, repeats=2):
    """dJ/deps_r of the 64x64x32 design box of the 200x160x80 two-port model, J = sum |S21|^2 at
    2..6 GHz: one gradient = checkpointed forward + reverse pass with recomputation; the best of
    `repeats` end-to-end evaluations (model assembly excluded, host objective included)."""
    from rfnetwork_amd import adjoint, workloads

    model = workloads.design_strip(Nt=n_steps)
    cbox = workloads.design_box()
    best = None
    for _ in range(repeats):
        t0 = time.perf_counter()
        J, grad, S, t = adjoint.eps_r_gradient(model, cbox, model.frequency, checkpoint_every=every,
                                               return_session_timing=True)
        wall = time.perf_counter() - t0
        if best is None or wall < best[0]:
            best = (wall, t, J, grad)
    wall, t, J, grad = best
    cells = model.Nx * model.Ny * model.Nz
    log(f"adjoint: {wall:.3f}s per gradient (forward {t['forward_ms']:.1f} ms, reverse {t['adjoint_ms']:.1f} ms)")
    return dict(workload=f"design_strip {model.Nx}x{model.Ny}x{model.Nz}, design box 64x64x32, {n_steps} steps, "
                         f"checkpoints every {every}", gradients_per_s=round(1.0 / wall, 4),
                s_per_gradient=round(wall, 3), forward_ms=round(t["forward_ms"], 1),
                reverse_ms=round(t["adjoint_ms"], 1), design_cells=int(grad.size), J=J,
                forward_mcells=round(cells * n_steps / (t["forward_ms"] * 1e-3) / 1e6, 1),
                finite=bool(np.isfinite(grad).all()))


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="dipole_slab")
    ap.add_argument("--coef", choices=("auto", "raw", "table"), default="auto")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--xplanes", type=int, default=0)
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU oracle work (0 = skip)")
    ap.add_argument("--no-copy-probe", action="store_true")
    ap.add_argument("--legs", default="plain,batched,adjoint",
                    help="comma list: plain (the headline), batched (12 excitations), adjoint (gradients/s)")
    ap.add_argument("--batch-steps", type=int, default=5000)
    ap.add_argument("--adjoint-steps", type=int, default=6000)
    ap.add_argument("--checkpoint-every", type=int, default=200)
    args = ap.parse_args()
    legs = set(args.legs.split(","))
    if args.gpus != 1:
        print(json.dumps(dict(error="this engine runs on one GPU; --gpus above 1 is not supported",
                              n_gpus=args.gpus)))
        sys.exit(2)

    from rfnetwork_amd import _lib
    from rfnetwork_amd.session import Session

    _lib.lib()  # fails loudly if the HIP library is not built
    Nt = args.warmup + args.steps
    t0 = time.perf_counter()
    model = build_model(args.workload, Nt)
    cells = model.Nx * model.Ny * model.Nz
    log(f"model {args.workload} {model.Nx}x{model.Ny}x{model.Nz} built in {time.perf_counter() - t0:.1f}s")
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = model.pack(gpu=True)
    flags = _lib.KERNEL_TIMING | {"auto": 0, "raw": _lib.COEF_RAW, "table": _lib.COEF_TABLE}[args.coef]
    t0 = time.perf_counter()
    s = Session(c, Nx, Ny, Nz, flags=flags)
    del c
    tuned = None
    if args.block or args.xplanes:
        s.set_launch(args.block, args.xplanes, 0)
    else:
        t1 = time.perf_counter()
        tuned = s.autotune()
        log(f"autotune {tuned} in {time.perf_counter() - t1:.1f}s")
    s.set_sources(probes, monitors, Nt)
    info = s.memory_info()
    log(f"session ready in {time.perf_counter() - t0:.1f}s: {info}")

    s.run(args.warmup)
    s.synchronize()
    log(f"warmup {args.warmup} steps done")
    t0 = time.perf_counter()
    s.run(args.steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    total_ms, e_ms, h_ms = s.timing()
    log(f"timed {args.steps} steps: wall {wall:.3f}s, events {total_ms / 1e3:.3f}s, E {e_ms:.4f} ms, H {h_ms:.4f} ms")
    s.read()
    finite = all(np.isfinite(p["values"][:Nt]).all() for p in probes)
    s.close()

    ms_step = wall * 1e3 / args.steps
    mcells = cells * args.steps / wall / 1e6
    dft_b = dft_bytes_per_step(model)
    step_bytes = cells * (E_BYTES + H_BYTES) + dft_b
    step_frac = step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS
    h_gbs = cells * H_BYTES / (h_ms * 1e-3) / 1e9 if h_ms > 0 else 0.0
    e_gbs = cells * E_BYTES / (e_ms * 1e-3) / 1e9 if e_ms > 0 else 0.0

    copy = None
    if not args.no_copy_probe:
        gbs, cms = _lib.copy_bandwidth(1 << 30, 20)
        copy = dict(gb_per_s=round(gbs, 1), bytes_per_copy=2 * (1 << 30), ms_per_copy=round(cms, 4))
        log(f"copy probe {gbs:.1f} GB/s")

    base = None
    if args.cpu_budget > 0:
        log("cpu baseline ...")
        base = cpu_baseline(model, args.cpu_budget)
        log(f"cpu baseline {base['value']} Mcell/s on {base['cores']} threads")

    batched = batched_leg(args.batch_steps) if "batched" in legs else None
    adj = adjoint_leg(args.adjoint_steps, args.checkpoint_every) if "adjoint" in legs else None

    out = dict(
        metric=METRIC, value=round(mcells, 1), unit="Mcell-updates/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=round(ms_step, 4), higher_is_better=True, scaling="weak",
        vs_baseline=None, dtype="f32", data="synthetic",
        config=dict(workload=f"{args.workload} {model.Nx}x{model.Ny}x{model.Nz}", cells=cells,
                    dft_planes=len(model.monitors), dft_frequencies=51 if model.monitors else 0,
                    coefficients=info["coef_kind"], coef_entries=[info["n_e_entries"], info["n_h_entries"]],
                    launch=tuned),
        roofline=dict(bound="hbm", kernel="k_update_h", achieved=round(h_gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=round(h_gbs / HBM_PEAK_GBS, 4), traffic=None,
                      algorithmic_bytes_per_launch=cells * H_BYTES, launch_ms=round(h_ms, 4)),
        step=dict(hbm_frac_algorithmic=round(step_frac, 4), algorithmic_bytes=step_bytes,
                  e_kernel_ms=round(e_ms, 4), e_kernel_gbs=round(e_gbs, 1), h_kernel_ms=round(h_ms, 4),
                  dft_bytes=dft_b),
        copy_probe=copy, cpu_baseline=base, finite=finite, batched=batched, adjoint=adj,
    )
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
