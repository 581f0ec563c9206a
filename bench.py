#!/usr/bin/env python3
"""FDTD throughput on one MI355X: Mcell-updates/s and % of HBM peak, batched vs plain, adjoint gradients/s.

Headline (the "plain" leg): the 400x400x200 dipole-over-random-slab model (32 M cells, 12-cell PML on
all faces, a 1-cell 50 ohm gap source, running DFTs at 51 frequencies on 6 planes), float32, 8000
timed steps.  One bench step is one full FDTD time step: E update, E sources/probes, H update, H
probes, DFT accumulation.  Inputs (fields, coefficients, source series, DFT tables) are resident in
HBM before timing starts.  The JSON line carries the roofline of the H update kernel (144 algorithmic
bytes per cell: 6 split H read and written, 18 coefficients, 3 E totals read, 3 H totals written;
the E update is 120 B, a step 264 B), its profiler-measured HBM bytes, a copy-kernel bandwidth, and
the CPU oracle (a restatement of the reference's threaded C++ solver) timed on this host over a
bounded window of the same workload.

Further legs in the same line: "c1" -- the 120x48x32 microstrip (4000 steps) on the GPU and, in
full, on the CPU oracle; "batched" -- 4 uneven multi-port grids x every port driven (12
excitations, 5000 steps) in one batched session vs the same excitations one after another;
"adjoint" -- gradients/s of dJ/deps_r over a 64x64x32 design box of a 200x160x80 model (6000 steps,
checkpoints every 200); "barrier" -- a device-wide barrier inside one cooperative launch against
the spacing of graph-replayed kernel boundaries (what a persistent launch would pay per step).  On
request: "c5" (8 design variants up to 512x512x256, batched forward and adjoint) and "adjoint_fd"
(that gradient against central differences on 32 cells).
"""
import argparse
import json
import os
import sys
import time
import traceback
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# The best streaming rate measured on MI355X with this library's own probes (the H update's mix of 27
# arrays read and 9 written, nontemporal, 128-thread workgroups): the practical ceiling beside the peak.
BEST_STREAMING_GBS = 5850.0
E_BYTES, H_BYTES = 120, 144  # algorithmic bytes per cell of the E and H half steps (sum 264)
METRIC = "Mcell-updates/s + % HBM peak, batched vs plain; adjoint gradients/s, 1 GPU"
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE on MI355X (KiB per
# dispatch x 1024, FETCH_SIZE x 2 and WRITE_SIZE x 1: the factors that make the counters of a plain
# 1 GiB copy kernel read 1 GiB each way); profiles/k_update_eh_dipole_slab_pmc.md
# HBM bytes per launch of the E and H kernels on the 400x400x200 workload, per launch shape
# (variant, block, x planes): rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE (KiB, calibrated on k_stream_copy).
MEASURED_TRAFFIC = {
    ("dipole_slab", "table", k, shape): v
    for shape, (e, h) in {
        (4, 512, 4): (2543600000, 2501600000),
        (1, 512, 4): (2563200000, 2504900000),
        (4, 256, 4): (2542900000, 2501400000),
        (1, 256, 4): (2572700000, 2501900000),
        (4, 128, 4): (2543100000, 2501400000),
        (1, 128, 4): (2571000000, 2502300000),
        (4, 512, 200): (2552000000, 2526600000),
        (4, 256, 200): (2493700000, 2452600000),
        (1, 256, 200): (2500000000, 2456700000),
        (4, 128, 200): (2496800000, 2452400000),
        (1, 128, 200): (2497900000, 2455000000),
    }.items()
    for k, v in (("k_update_e", e), ("k_update_h", h))
}
# The raw-coefficient form (the reference formulation's bytes) at the shapes the tuner picks for it
# (variant 10: nontemporal loads of the streamed arrays; variant 4 before it existed).
MEASURED_TRAFFIC.update({
    ("dipole_slab", "raw", "k_update_e", (4, 128, 4)): 3976700000,
    ("dipole_slab", "raw", "k_update_h", (4, 128, 4)): 4683300000,
    ("dipole_slab", "raw", "k_update_e", (10, 128, 4)): 3873700000,
    ("dipole_slab", "raw", "k_update_h", (10, 128, 4)): 4637100000,
    # variant 12: four z cells per thread, one x plane per workgroup (measured at 512 threads; the
    # block size moves the counters by under 0.1%, as the variant 4 rows above show)
    **{("dipole_slab", "raw", k, (12, blk, 1)): v for blk in (128, 256, 512)
       for k, v in (("k_update_e_w", 3925300000), ("k_update_h_w", 4648700000))},
    # the two-step pass at its tuned shape (12-row tiles, x chunks of 25 planes; 0 = the same automatic)
    ("dipole_slab", "table", "k_step2", (7, 256, 25)): 4175900000,
    ("dipole_slab", "table", "k_step2", (7, 256, 0)): 4175900000,
    # the table form of variant 10 in launch order, 4 x planes per workgroup
    ("dipole_slab", "table", "k_update_e", (10, 256, 4, 1)): 2766400000,
    ("dipole_slab", "table", "k_update_h", (10, 256, 4, 1)): 2722500000,
    # variant 10 with one x plane per workgroup and the workgroups in launch order (order 1): the
    # counters read 8-10% above the algorithmic bytes, the neighbour rows no longer sharing an L2
    ("dipole_slab", "raw", "k_update_e", (10, 256, 1, 1)): 4225500000,
    ("dipole_slab", "raw", "k_update_h", (10, 256, 1, 1)): 4993600000,
    # 26 x 28-cell tiles of 32-lane rows (variant 15), x chunks of 67 and 50 planes
    ("dipole_slab", "table", "k_step2", (15, 256, 67)): 3946000000,
    ("dipole_slab", "table", "k_step2", (15, 256, 50)): 3911700000,
    # the automatic chunking of the two-step pass (768 workgroups aimed at: 6 chunks of 128 tiles)
    # is the 67-plane shape on this grid
    ("dipole_slab", "table", "k_step2", (15, 256, 0)): 3946000000,
    # 28 x 28-cell tiles of 32-lane rows (variant 13), x chunks of 25 planes
    ("dipole_slab", "table", "k_step2", (13, 256, 25)): 4097500000,
    ("dipole_slab", "table", "k_step2", (13, 256, 0)): 4097500000,
})


def measured_traffic(workload, coef, kname, tuned):
    """HBM bytes per launch measured for the tuned shape: keys carry (variant, block, x planes) and,
    where it was measured apart, the separate kernels' workgroup order; entries without the order
    were measured with order 0 (the tiled passes ignore it)."""
    if not tuned:
        return None
    shape = (tuned["variant"], tuned["block"], tuned["x_planes"])
    order = tuned.get("order", 0)
    t = MEASURED_TRAFFIC.get((workload, coef, kname, shape + (order,)))
    if t is None and order == 0:
        t = MEASURED_TRAFFIC.get((workload, coef, kname, shape))
    return t


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


# A one-GPU job is allotted 16 host cores on the machines this runs on, whatever nproc reports.
HOST_CPU_SHARE = 16
REF_MAX_THREADS = 20  # the reference's thread cap (rfnetwork/core/inc/solver.h:17)
API_DEFAULT_THREADS = 4  # FDTD_Solver.solve(n_threads=4) (rfnetwork/solver.py:1410)


def host_cores():
    """Cores this process may run on: the affinity mask, capped by a cgroup CPU quota if one is set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if quota != "max":
            n = min(n, max(1, int(int(quota) / int(period))))
    except (OSError, ValueError):
        pass
    return max(1, n)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_thread_counts():
    """The reference's thread count on this host, min(cores, 20) (held to the host's share), and the
    API default of 4."""
    cores = host_cores()
    top = min(cores, REF_MAX_THREADS, HOST_CPU_SHARE)
    return cores, top, sorted({top, min(API_DEFAULT_THREADS, top)}, reverse=True)


def _oracle_inputs(model, threads, nsteps):
    """The model packed for the CPU oracle, its probe and monitor series cut to nsteps (the model's
    own series must be at least that long)."""
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = model.pack(gpu=False, n_threads=threads)
    if Nt < nsteps:
        raise ValueError(f"model has {Nt} steps of source samples, the window needs {nsteps}")
    for p in probes:
        p["values"] = np.ascontiguousarray(p["values"][:nsteps])
    for m in monitors:
        if "dtft_phase" in m:
            m["dtft_phase"] = np.ascontiguousarray(m["dtft_phase"][: nsteps // m["n_step"] + 1])
        else:
            m["values"] = np.ascontiguousarray(m["values"][: nsteps // m["n_step"] + 1])
    return c, probes, monitors, mem, Nx, Ny, Nz


def _time_oracle(model, threads, nsteps):
    """Wall time of one oracle run of nsteps from zero fields (fresh probe and monitor arrays)."""
    from oracle import cpu_solver

    c, probes, monitors, mem, Nx, Ny, Nz = _oracle_inputs(model, threads, nsteps)
    t0 = time.perf_counter()
    cpu_solver.solver_run(c, probes, monitors, mem, Nx, Ny, Nz, nsteps, threads, 0)
    return time.perf_counter() - t0


def cpu_baseline(workload, budget_s, max_window=200, runs=3):
    """The CPU oracle (a restatement of the reference's threaded solver, rfnetwork/core/src/solver.cpp,
    x slabs, a controller thread and two barriers per step) timed on this host over a window of the
    same workload, at the reference's thread count and at the API default: the median of `runs`
    runs each, the window sized so that each thread count takes about budget_s seconds.  The model
    is built here with source samples for the longest window, whatever the GPU leg's step count."""
    cores, top, counts = cpu_thread_counts()
    model = build_model(workload, max_window)
    cells = model.Nx * model.Ny * model.Nz
    per_count = {}
    for threads in counts:
        per_step = _time_oracle(model, threads, 2) / 2
        window = int(max(4, min(max_window, budget_s / runs / max(per_step, 1e-6))))
        walls = sorted(_time_oracle(model, threads, window) for _ in range(runs))
        med = walls[len(walls) // 2]
        per_count[threads] = dict(mcells=round(cells * window / med / 1e6, 2), window_steps=window,
                                  median_s=round(med, 3), runs_s=[round(w, 3) for w in walls])
        log(f"cpu baseline: {threads} threads, {window}-step window, median {med:.2f}s "
            f"-> {per_count[threads]['mcells']} Mcell/s")
    best = per_count[top]
    return dict(value=best["mcells"], unit="Mcell-updates/s", cores=top, kind="port",
                sample=f"{best['window_steps']} steps of the full {model.Nx}x{model.Ny}x{model.Nz} workload from zero "
                       f"fields, median of {runs} runs (oracle restating rfnetwork/core/src/solver.cpp: {top} threads "
                       f"= min(cores, 20) held to the host's {HOST_CPU_SHARE}-core share, x slabs, 2 barriers per step, "
                       f"-O2 -ffp-contract=off)",
                cpu_model=cpu_model(), host_cores=cores, by_threads={str(k): v for k, v in per_count.items()})


def c1_leg(flags=0):
    """The 120x48x32 microstrip (4000 steps, 2-port S at 101 frequencies): a launch-bound grid that
    fits the Infinity Cache.  GPU: resident session, tuned, all 4000 steps timed; CPU: the oracle
    over the same full run on this host's cores (the reference's CPU path restated)."""
    from oracle import cpu_solver
    from rfnetwork_amd import workloads
    from rfnetwork_amd.session import Session

    model = workloads.microstrip()
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = model.pack(gpu=True)
    cells = Nx * Ny * Nz
    with Session(c, Nx, Ny, Nz, flags=flags) as s:
        s.set_sources(probes, monitors, Nt)  # installed first: the tuner's trials then apply them too
        tuned = s.autotune()
        t0 = time.perf_counter()
        s.run(Nt)
        s.synchronize()
        wall = time.perf_counter() - t0
        s.read()
    model.unpack(probes, monitors, gpu=True)
    S_gpu = np.asarray(model.get_sparameters(model.frequency))
    # CPU: the full run on the oracle, median of 3 at the reference's thread count and at 4 threads
    _, top, counts = cpu_thread_counts()
    cpu = {}
    S_cpu = None
    for threads in counts:
        walls = []
        for _ in range(3):
            m2 = workloads.microstrip()
            c2, p2, mo2, mem2, *_ = m2.pack(gpu=False, n_threads=threads)
            t0 = time.perf_counter()
            cpu_solver.solver_run(c2, p2, mo2, mem2, Nx, Ny, Nz, Nt, threads, 0)
            walls.append(time.perf_counter() - t0)
        walls.sort()
        cpu[threads] = walls[1]
        if S_cpu is None:
            m2.unpack(p2, mo2, gpu=False)
            S_cpu = np.asarray(m2.get_sparameters(m2.frequency))
    wall_cpu = cpu[top]
    log(f"c1: GPU {wall * 1e3:.1f} ms, CPU oracle {wall_cpu:.2f} s on {top} threads "
        f"({', '.join(f'{k}: {v:.2f} s' for k, v in cpu.items())}), max |dS| {np.abs(S_gpu - S_cpu).max():.2e}")
    return dict(workload=f"microstrip {Nx}x{Ny}x{Nz}, {Nt} steps, 101 frequencies", gpu_ms=round(wall * 1e3, 2),
                gpu_mcells=round(cells * Nt / wall / 1e6, 1), launch=tuned,
                cpu_oracle_s=round(wall_cpu, 3), cpu_mcells=round(cells * Nt / wall_cpu / 1e6, 1), cpu_threads=top,
                cpu_oracle_s_by_threads={str(k): round(v, 3) for k, v in cpu.items()}, cpu_runs=3,
                speedup=round(wall_cpu / wall, 1), s_params_identical=bool(np.array_equal(S_gpu, S_cpu)))


def batched_leg(n_steps, flags=0):
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

    b = rb.Batch(grids, flags=flags)
    k = 0
    for _, _, exc, Nt in inputs:
        for _, probes, monitors in exc:
            b.set_sources(k, probes, monitors, Nt)
            k += 1
    log(f"batched: autotune {b.autotune()}")  # after the sources, as for the sequential sessions
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
            s = Session(coeff, m.Nx, m.Ny, m.Nz, flags=flags)
            s.set_sources(probes, monitors, Nt)
            s.autotune()
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

    # the same sequence through the reference formulation of the plain kernel: 30 raw coefficient
    # arrays per cell and one launch per kernel per step (no table, no graph replay)
    from rfnetwork_amd import _lib

    plain = []
    for m, coeff, exc, Nt in inputs:
        for port in range(len(exc)):
            exc2, _ = rb.pack_excitations(m, [port + 1])
            _, probes, monitors = exc2[0]
            s = Session(coeff, m.Nx, m.Ny, m.Nz, flags=flags | _lib.COEF_RAW | _lib.NO_GRAPHS)
            s.set_sources(probes, monitors, Nt)
            plain.append((s, probes))
    t0 = time.perf_counter()
    for s, _ in plain:
        s.run(n_steps)
        s.synchronize()
    wall_p = time.perf_counter() - t0
    for (s, probes), ref in zip(plain, batched_series):
        s.read()
        identical &= all(np.array_equal(p["values"], r) for p, r in zip(probes, ref))
        s.close()
    log(f"sequential, raw coefficients and per-step launches: {wall_p:.3f}s, identical {identical}")
    del sessions, plain, batched_series

    # end to end, as a user runs it: S-parameter columns of every port of every variant, either one
    # model.solve() per driven port (packing, coefficient upload, run, read back each time; the
    # reference's workflow) or solve_ports() (one batched session for all)
    freq = models[0].frequency
    jobs = [(m, list(range(1, len(m.ports) + 1))) for m in models]
    waves = [np.asarray(rb._waveform(m)) for m in models]
    t0 = time.perf_counter()
    res = rb.solve_ports(jobs, freq, flags=flags)
    e2e_b = time.perf_counter() - t0
    t0 = time.perf_counter()
    cols = []
    for (m, ports), wave in zip(jobs, waves):
        for port in ports:
            m.reset_excitations()
            m.assign_excitation(wave, port)
            m.solve(gpu=True, show_progress=False, flags=flags)
            cols.append(np.asarray(m.get_sparameters(freq, source_port=port))[..., 0])
    e2e_s = time.perf_counter() - t0
    ref = [S[..., c] for S, _ in res for c in range(S.shape[-1])]
    same = all(np.array_equal(a, b) for a, b in zip(ref, cols))
    log(f"end to end: solve_ports {e2e_b:.3f}s, sequential solve() {e2e_s:.3f}s, identical S {same}")
    # where batching pays: 12 excitations (ports 1 and 2, alternately) of the 120x48x32 microstrip,
    # a grid small enough that a single run is bound by launch latency, not bandwidth
    sm = workloads.microstrip(Nt=4000)
    sexc, sNt = rb.pack_excitations(sm, [1, 2] * 6)
    scoef = sm.coefficient_arrays(gpu=True)
    with rb.Batch([(scoef, sm.Nx, sm.Ny, sm.Nz, len(sexc))], flags=flags) as sb:
        for k, (_, probes, monitors) in enumerate(sexc):
            sb.set_sources(k, probes, monitors, sNt)
        sb.autotune()
        t0 = time.perf_counter()
        sb.run(sNt)
        sb.synchronize()
        small_b = time.perf_counter() - t0
        sb.read()
    small_series = [[np.array(p["values"]) for p in probes] for _, probes, _ in sexc]

    def small_sequential(extra_flags, tune):
        runs = []
        for port in [1, 2] * 6:
            (_, probes, monitors), = rb.pack_excitations(sm, [port])[0]
            ss = Session(scoef, sm.Nx, sm.Ny, sm.Nz, flags=flags | extra_flags)
            ss.set_sources(probes, monitors, sNt)
            if tune:
                ss.autotune()
            runs.append((ss, probes))
        t0 = time.perf_counter()
        for ss, _ in runs:
            ss.run(sNt)
            ss.synchronize()
        wall = time.perf_counter() - t0
        same = True
        for (ss, probes), ref in zip(runs, small_series):
            ss.read()
            same &= all(np.array_equal(p["values"], r) for p, r in zip(probes, ref))
            ss.close()
        return wall, same

    small_s, small_same = small_sequential(0, True)
    # the reference formulation of the plain kernel, as for the port variants above
    small_p, small_same_p = small_sequential(_lib.COEF_RAW | _lib.NO_GRAPHS, False)
    log(f"small grids: batched {small_b * 1e3:.1f} ms, sequential {small_s * 1e3:.1f} ms, sequential plain "
        f"formulation {small_p * 1e3:.1f} ms, identical {small_same and small_same_p}")
    b_rate = exc_cells * n_steps / wall_b / 1e6
    s_rate = exc_cells * n_steps / wall_s / 1e6
    return dict(workload="4 port variants 64^3..192x128x96, 12 excitations", steps=n_steps,
                excitation_cells=exc_cells, batched_mcells=round(b_rate, 1), sequential_mcells=round(s_rate, 1),
                speedup=round(wall_s / wall_b, 3), batched_s=round(wall_b, 3), sequential_s=round(wall_s, 3),
                sequential_plain_formulation_s=round(wall_p, 3), speedup_vs_plain_formulation=round(wall_p / wall_b, 3),
                identical=bool(identical),
                end_to_end=dict(solve_ports_s=round(e2e_b, 3), sequential_solve_s=round(e2e_s, 3),
                                speedup=round(e2e_s / e2e_b, 3), identical=bool(same)),
                small_grids=dict(workload="12 excitations of the 120x48x32 microstrip, 4000 steps",
                                 batched_ms=round(small_b * 1e3, 2), sequential_ms=round(small_s * 1e3, 2),
                                 speedup=round(small_s / small_b, 3),
                                 sequential_plain_formulation_ms=round(small_p * 1e3, 2),
                                 speedup_vs_plain_formulation=round(small_p / small_b, 3),
                                 identical=bool(small_same and small_same_p)))


def adjoint_leg(n_steps, every, repeats=2, flags=0):
    """dJ/deps_r of the 64x64x32 design box of the 200x160x80 two-port model, J = sum |S21|^2 at
    2..6 GHz.  Headline: one gradient = forward run with checkpoints every `every` steps + reverse pass
    recomputing each segment; beside it the form that records the design box's H every step during
    the forward run (no recomputation).  Best of `repeats` end-to-end evaluations each (model assembly
    excluded, host objective included); both gradients must agree bit for bit."""
    from rfnetwork_amd import adjoint, workloads

    model = workloads.design_strip(Nt=n_steps)
    cbox = workloads.design_box()
    res = {}
    for mode in ("history", "checkpoint"):
        best = None
        for _ in range(repeats):
            t0 = time.perf_counter()
            J, grad, S, t = adjoint.eps_r_gradient(model, cbox, model.frequency, checkpoint_every=every,
                                                   return_session_timing=True, flags=flags, mode=mode)
            wall = time.perf_counter() - t0
            if best is None or wall < best[0]:
                best = (wall, t, J, grad)
        res[mode] = best
        log(f"adjoint ({mode}): {best[0]:.3f}s per gradient (forward {best[1]['forward_ms']:.1f} ms, "
            f"reverse {best[1]['adjoint_ms']:.1f} ms)")
    wall, t, J, grad = res["history"]
    wc, tc, Jc, gc = res["checkpoint"]
    # a design loop on one resident session: the box's coefficients recomputed on the GPU per
    # evaluation instead of rebuilding the model and the session
    from rfnetwork_amd.design import DesignLoop

    eps_model = np.random.default_rng(4).uniform(1.0, 4.0, size=(64, 64, 32))  # design_strip's default design
    eps_next = np.random.default_rng(5).uniform(1.0, 4.0, size=eps_model.shape)
    with DesignLoop(model, cbox, model.frequency, flags=flags, mode="history") as loop:
        Jl, gl, _ = loop.evaluate(eps_model)
        loop_equal = bool(Jl == J and np.array_equal(gl, grad))
        walls = []
        for e in (eps_next, eps_model):
            t0 = time.perf_counter()
            loop.evaluate(e)
            walls.append(time.perf_counter() - t0)
        lt = loop.timing
    log(f"adjoint (design loop): {min(walls):.3f}s per gradient (update {lt['update_ms']:.2f} ms, forward "
        f"{lt['forward_ms']:.1f} ms, reverse {lt['adjoint_ms']:.1f} ms), equal to the rebuilt model's: {loop_equal}")
    cells = model.Nx * model.Ny * model.Nz
    # the headline is the configuration's own method: checkpoints every `every` steps, segments
    # recomputed in the reverse pass; the design-box history (no recomputation) beside it
    return dict(workload=f"design_strip {model.Nx}x{model.Ny}x{model.Nz}, design box 64x64x32, "
                         f"{n_steps} steps, checkpoints every {every} steps", checkpoint_every=every,
                gradients_per_s=round(1.0 / wc, 4), s_per_gradient=round(wc, 3),
                forward_ms=round(tc["forward_ms"], 1), reverse_ms=round(tc["adjoint_ms"], 1),
                design_cells=int(grad.size), J=J, adjoint_launch=t["adjoint_launch"],
                forward_mcells=round(cells * n_steps / (tc["forward_ms"] * 1e-3) / 1e6, 1),
                history=dict(gradients_per_s=round(1.0 / wall, 4), s_per_gradient=round(wall, 3),
                             forward_ms=round(t["forward_ms"], 1), reverse_ms=round(t["adjoint_ms"], 1),
                             history_bytes=int(adjoint.history_bytes([adjoint.gpu_box_for_cells(model, cbox)],
                                                                     n_steps))),
                modes_equal=bool(J == Jc and np.array_equal(grad, gc)),
                design_loop=dict(gradients_per_s=round(1.0 / min(walls), 4), s_per_gradient=round(min(walls), 3),
                                 coefficient_update_ms=round(lt["update_ms"], 2), equal=loop_equal),
                finite=bool(np.isfinite(grad).all()))


def c5_leg(n_steps, flags=0):
    """8 design variants from 96^3 to 512x512x256 (142 M cells), port 1 driven in each: the forward
    runs and the adjoint passes of all variants in batched launches, the checkpoint interval planned
    to fit device memory, the coefficients assembled on the GPU.  Reported: wall time of the whole
    gradient of all 8 (meshing excluded), the planned interval and bytes."""
    from rfnetwork_amd import _lib, adjoint, workloads

    jobs, cells = [], 0
    for v in range(len(workloads.DESIGN_VARIANTS)):
        t0 = time.perf_counter()
        m = workloads.design_variant(v, Nt=n_steps, coefficients="device")  # assembled on the GPU
        jobs.append((m, workloads.design_variant_box(v), m.frequency))
        cells += m.Nx * m.Ny * m.Nz
        log(f"c5: variant {v} {m.Nx}x{m.Ny}x{m.Nz} built in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    res, info = adjoint.eps_r_gradients_batch(jobs, flags=flags | _lib.VERBOSE, return_info=True)
    wall = time.perf_counter() - t0
    log(f"c5: {wall:.1f}s for all gradients, {info}")
    return dict(workload=f"8 design variants 96^3..512x512x256, design boxes 48^3..128^3, {n_steps} steps",
                cells=cells, wall_s=round(wall, 2), gradients_per_s=round(len(jobs) / wall, 4),
                forward_ms=round(info["forward_ms"], 1), reverse_ms=round(info["adjoint_ms"], 1),
                forward_mcells=round(cells * n_steps / (info["forward_ms"] * 1e-3) / 1e6, 1),
                mode=info["mode"], checkpoint_every=info.get("checkpoint_every"), planned_bytes=info.get("planned_bytes"),
                launch=info.get("launch"), J=[r[0] for r in res],
                finite=bool(all(np.isfinite(r[1]).all() for r in res)))


def monitors_leg(n_steps=1000, ring=64):
    """A field-visualisation run: the 400x400x200 workload with three time-domain plane monitors
    sampled every step (Ez on the z mid-plane, Hy on the x mid-plane, Ex on the y mid-plane;
    1.28 GB of samples over 1000 steps).  Timed from the start of the run until every sample is in
    host memory: the whole series kept on the GPU and copied back at the end, against the monitors
    streamed through a ring of `ring` planes into pinned host arrays during the run.  The two
    results must be identical."""
    from rfnetwork_amd.session import Session

    model = workloads_dipole(n_steps)
    cells = model.Nx * model.Ny * model.Nz
    out, results = {}, {}
    for mode in ("resident", "streamed"):
        c, probes, monitors, _, Nx, Ny, Nz, Nt = model.pack(gpu=True)
        with Session(c, Nx, Ny, Nz) as s:
            del c
            s.autotune()
            if mode == "streamed":
                s.set_monitor_streaming(ring)
            s.set_sources(probes, monitors, Nt)
            s.synchronize()
            t0 = time.perf_counter()
            s.run(Nt)
            s.read()
            wall = time.perf_counter() - t0
        results[mode] = [np.asarray(m["values"]) for m in monitors]
        mon_bytes = sum(m.nbytes for m in results[mode])
        out[mode] = dict(s_run_and_read=round(wall, 3), mcells=round(cells * Nt / wall / 1e6, 1))
        log(f"monitors ({mode}): {wall:.3f}s from run start to all {mon_bytes / 1e9:.2f} GB of samples on the host")
    same = all(np.array_equal(a, b) for a, b in zip(results["resident"], results["streamed"]))
    plane_bytes = sum(m[0].nbytes for m in results["streamed"])
    return dict(workload=f"dipole_slab 400x400x200, 3 time-domain planes every step, {n_steps} steps",
                sample_bytes=int(mon_bytes), ring=ring, device_ring_bytes=int(ring * plane_bytes),
                resident=out["resident"], streamed=out["streamed"],
                speedup=round(out["resident"]["s_run_and_read"] / out["streamed"]["s_run_and_read"], 3),
                identical=bool(same))


def dft64_leg(n_steps=1000):
    """The 400x400x200 workload with its 6 running-DFT planes (51 frequencies) accumulated in
    complex64, as the reference does, and in complex128 (float32 fields): ms per step of each, tuned,
    and how far the complex64 sums lie from the complex128 ones."""
    from rfnetwork_amd import workloads
    from rfnetwork_amd.session import Session

    out, sums = {}, {}
    for name, cdt in (("complex64", None), ("complex128", np.complex128)):
        model = workloads.dipole_slab(Nt=n_steps, dft_dtype=cdt)
        c, probes, monitors, _, Nx, Ny, Nz, Nt = model.pack(gpu=True)
        with Session(c, Nx, Ny, Nz) as s:
            del c
            s.set_sources(probes, monitors, Nt)
            tuned = s.autotune()
            s.synchronize()
            t0 = time.perf_counter()
            s.run(Nt)
            s.synchronize()
            wall = time.perf_counter() - t0
            s.read()
        sums[name] = [np.asarray(m["values"]) for m in monitors]
        out[name] = dict(ms_per_step=round(wall / Nt * 1e3, 4), launch=tuned,
                         accumulator_bytes=int(sum(v.nbytes for v in sums[name])))
        log(f"dft64 ({name}): {wall / Nt * 1e3:.4f} ms/step")
    rel = max(float(np.abs(a - b).max() / np.abs(b).max()) for a, b in zip(sums["complex64"], sums["complex128"]))
    return dict(workload=f"dipole_slab 400x400x200, 6 DFT planes x 51 frequencies, {n_steps} steps", **out,
                max_rel_diff_complex64_vs_complex128=rel)


def workloads_dipole(n_steps):
    from rfnetwork_amd import workloads

    model = workloads.dipole_slab(Nt=n_steps, monitors=False)
    model.add_field_monitor("t_ez", "ez", "z", index=model.Nz // 2, n_step=1)
    model.add_field_monitor("t_hy", "hy", "x", index=model.Nx // 2, n_step=1)
    model.add_field_monitor("t_ex", "ex", "y", index=model.Ny // 2, n_step=1)
    return model


def barrier_leg(iters=2000):
    """The cost of keeping time steps inside one persistent launch: workgroups of 256 threads (256,
    512, 1024) meeting at a device-wide barrier `iters` times in a cooperative launch, with no fences
    (the synchronisation alone, which does not make stores visible across the XCDs' L2 caches), with
    one wave per workgroup writing back and invalidating its L2 (the least that carries a step's fields
    to the next on other XCDs), and with every wave doing so; beside it, the spacing of dependent
    graph-replayed launches of an empty kernel of the same grid (a kernel boundary, which makes the
    stores visible by itself)."""
    from rfnetwork_amd import _lib

    rows = []
    for wg in (256, 512, 1024):
        for fences in (0, 1, 2):
            b, l = _lib.grid_barrier_cost(wg, fences, iters)
            rows.append(dict(workgroups=wg, fences=fences, barrier_us=round(b, 3), launch_us=round(l, 3)))
            log(f"barrier: {wg} workgroups, fences {fences}: {b:.3f} us per barrier, {l:.3f} us per graph launch")
    return dict(iters=iters, fences_meaning={"0": "none (no cross-XCD visibility)", "1": "one wave per workgroup",
                                             "2": "every wave"}, rows=rows)


def adjoint_fd_leg(n_steps, every, n_cells=32, seed=41, h=1e-3, oracle_cells=1):
    """Validation (not a throughput figure): the adjoint dJ/deps_r of the 200x160x80 design model in
    float64 and float32 against central differences (Delta eps_r = h) of float64 forward solves on
    n_cells random design cells.  The differences use the float64 HIP engine, which reproduces the
    float64 oracle bit for bit; for the first `oracle_cells` cells the oracle itself is run too and
    its J must equal the engine's exactly."""
    from oracle import cpu_solver
    from rfnetwork_amd import adjoint, workloads

    eps0 = np.random.default_rng(4).uniform(1.0, 4.0, size=(64, 64, 32))
    cbox = workloads.design_box()
    build = lambda e, dt=np.float64: workloads.design_strip(e, Nt=n_steps, coef_dtype=dt)
    m64 = build(eps0)
    t0 = time.perf_counter()
    J64, g64, _ = adjoint.eps_r_gradient(m64, cbox, m64.frequency, checkpoint_every=every, dtype=np.float64)
    log(f"adjoint_fd: float64 gradient in {time.perf_counter() - t0:.1f}s, J {J64:.9e}")
    m32 = build(eps0, np.float32)
    J32, g32, _ = adjoint.eps_r_gradient(m32, cbox, m32.frequency, checkpoint_every=every, dtype=np.float32)
    rng = np.random.default_rng(seed)
    cells = [tuple(int(rng.integers(0, n)) for n in (64, 64, 32)) for _ in range(n_cells)]
    fd, oracle_equal = [], []
    threads = max(1, min(16, os.cpu_count() or 1))
    for k, c in enumerate(cells):
        Js = []
        for sgn in (1, -1):
            e = eps0.copy()
            e[c] += sgn * h
            m = build(e)
            J = adjoint.objective_only(m, m.frequency, dtype=np.float64)[0]
            if k < oracle_cells:
                Jo = adjoint.objective_only(build(e), m.frequency, dtype=np.float64, backend=cpu_solver.solver_run,
                                            n_threads=threads)[0]
                oracle_equal.append(Jo == J)
            Js.append(J)
        fd.append((Js[0] - Js[1]) / (2 * h))
        log(f"adjoint_fd: cell {k} {c}: fd {fd[-1]:.9e} adjoint64 {g64[c]:.9e} adjoint32 {g32[c]:.9e}")
    fd = np.array(fd)
    a64 = np.array([g64[c] for c in cells])
    a32 = np.array([g32[c] for c in cells])
    r64 = np.abs(a64 - fd) / np.abs(fd)
    r32 = np.abs(a32 - fd) / np.abs(fd)
    return dict(workload=f"design_strip 200x160x80, design box 64x64x32, {n_steps} steps, checkpoints every {every}",
                cells=n_cells, seed=seed, delta_eps_r=h, J64=J64, J32=J32,
                max_rel_err_fp64=float(r64.max()), median_rel_err_fp64=float(np.median(r64)),
                max_rel_err_fp32=float(r32.max()), median_rel_err_fp32=float(np.median(r32)),
                normwise_rel_err_fp32=float(np.linalg.norm(a32 - fd) / np.linalg.norm(fd)),
                fp32_vs_fp64_gradient_rel_l2=float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64)),
                oracle_J_equal=oracle_equal, pass_1e3=bool(r64.max() <= 1e-3))


def guarded(name, fn, *a, **k):
    """Run one leg; a failure is logged and recorded as {"error": ...} so the JSON line still prints."""
    try:
        return fn(*a, **k)
    except Exception as e:  # noqa: BLE001 -- every leg's failure is reported, none ends the bench
        log(f"{name} failed: {type(e).__name__}: {e}")
        traceback.print_exc(file=sys.stderr)
        return dict(error=f"{type(e).__name__}: {e}")


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="dipole_slab")
    ap.add_argument("--coef", choices=("auto", "raw", "table"), default="auto")
    ap.add_argument("--variant", type=int, default=-1, help="kernel variant 0..15 (default: autotuned)")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--xplanes", type=int, default=0)
    ap.add_argument("--order", type=int, choices=(0, 1), default=0,
                    help="with a fixed launch shape: the separate kernels' workgroups in XCD ranges (0) or launch order (1)")
    ap.add_argument("--cpu-budget", type=float, default=12.0,
                    help="seconds of CPU oracle work per thread count (0 = skip)")
    ap.add_argument("--no-copy-probe", action="store_true")
    ap.add_argument("--legs", default="plain,raw,c1,batched,adjoint,barrier",
                    help="comma list: plain (the headline), raw (the same workload with the 30 raw coefficient "
                         "arrays: the reference formulation's bytes), c1 (the 120x48x32 microstrip, GPU and CPU "
                         "oracle), batched (12 excitations), adjoint (gradients/s), barrier (persistent-launch "
                         "barrier against graph-replayed launches), "
                         "c5 (8 design variants, batched forward+adjoint, minutes), "
                         "adjoint_fd (finite-difference validation, minutes), "
                         "monitors (streamed time-domain monitors against resident ones), "
                         "dft64 (the running DFTs accumulated in complex128 against complex64)")
    ap.add_argument("--batch-steps", type=int, default=5000)
    ap.add_argument("--adjoint-steps", type=int, default=6000)
    ap.add_argument("--checkpoint-every", type=int, default=200)
    ap.add_argument("--c5-steps", type=int, default=8000)
    ap.add_argument("--no-graphs", action="store_true", help="launch every step directly (no hipGraph replay)")
    args = ap.parse_args()
    legs = set(args.legs.split(","))
    if args.gpus != 1:
        print(json.dumps(dict(error="this engine runs on one GPU; --gpus above 1 is not supported",
                              n_gpus=args.gpus)))
        sys.exit(2)

    from rfnetwork_amd import _lib

    _lib.lib()  # fails loudly if the HIP library is not built
    out = dict(metric=METRIC, value=None, unit="Mcell-updates/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=None, higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f32",
               data="synthetic")
    if "plain" in legs:
        res = guarded("plain", plain_leg, args)
        if "error" in res:
            out["plain_error"] = res["error"]
        else:
            out.update(res)
    if "raw" in legs:
        res = guarded("raw", raw_leg, args)
        if "roofline" in res:
            out["roofline"] = res.pop("roofline")
        out["raw"] = res
    if "plain" in legs and args.cpu_budget > 0:
        log("cpu baseline ...")
        out["cpu_baseline"] = guarded("cpu_baseline", cpu_baseline, args.workload, args.cpu_budget)
    gflags = _lib.NO_GRAPHS if args.no_graphs else 0
    if "c1" in legs:
        out["c1"] = guarded("c1", c1_leg, gflags)
    if "batched" in legs:
        out["batched"] = guarded("batched", batched_leg, args.batch_steps, gflags)
    if "adjoint" in legs:
        out["adjoint"] = guarded("adjoint", adjoint_leg, args.adjoint_steps, args.checkpoint_every, flags=gflags)
    if "c5" in legs:
        out["c5"] = guarded("c5", c5_leg, args.c5_steps, gflags)
    if "adjoint_fd" in legs:
        out["adjoint_fd"] = guarded("adjoint_fd", adjoint_fd_leg, args.adjoint_steps, args.checkpoint_every)
    if "monitors" in legs:
        out["monitors"] = guarded("monitors", monitors_leg)
    if "dft64" in legs:
        out["dft64"] = guarded("dft64", dft64_leg)
    if "barrier" in legs:
        out["barrier"] = guarded("barrier", barrier_leg)
    print(json.dumps(out), flush=True)


def timed_run(args, coef, kernel_timing=True, separate=False):
    """W warm-up steps then K timed steps of the single-grid workload on a resident, tuned session
    (coefficients `coef`: "auto", "raw" or "table").  Per-launch E and H kernel times come from HIP
    events on the session's stream around up to 64 evenly spread steps of the timed region."""
    from rfnetwork_amd import _lib
    from rfnetwork_amd.session import Session

    Nt = args.warmup + args.steps
    t0 = time.perf_counter()
    model = build_model(args.workload, Nt)
    t_model = time.perf_counter() - t0
    log(f"model {args.workload} {model.Nx}x{model.Ny}x{model.Nz} built in {t_model:.1f}s")
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = model.pack(gpu=True)
    t_pack = time.perf_counter() - t0 - t_model
    flags = (_lib.KERNEL_TIMING if kernel_timing else 0) | {"auto": 0, "raw": _lib.COEF_RAW,
                                                             "table": _lib.COEF_TABLE}[coef]
    if separate:  # one step per pass: every field and coefficient array read once per step
        flags |= _lib.TUNE_SEPARATE
    t0 = time.perf_counter()
    s = Session(c, Nx, Ny, Nz, flags=flags)
    s.synchronize()
    t_session = time.perf_counter() - t0
    del c
    if args.block or args.xplanes or args.variant >= 0:
        s.set_launch(args.block, args.xplanes, 0)
        if args.variant >= 0:
            s.set_variant(args.variant)
        s.set_column_order(args.order)
        tuned = dict(variant=args.variant if args.variant >= 0 else 1, block=args.block or 256, x_planes=args.xplanes,
                     order=args.order, cached=None)
        s.set_sources(probes, monitors, Nt)
    else:
        s.set_sources(probes, monitors, Nt)  # installed first: the tuner's trials then apply them too
        t1 = time.perf_counter()
        tuned = s.autotune()
        log(f"autotune ({coef}) {tuned} in {time.perf_counter() - t1:.1f}s")
    info = s.memory_info()
    log(f"session ready in {time.perf_counter() - t0:.1f}s: {info}")
    s.run(args.warmup)
    s.synchronize()
    t0 = time.perf_counter()
    s.run(args.steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    total_ms, e_ms, h_ms = s.timing()
    log(f"timed {args.steps} steps ({coef}): wall {wall:.3f}s, events {total_ms / 1e3:.3f}s, E (or fused) "
        f"{e_ms:.4f} ms, H {h_ms:.4f} ms")
    s.read()
    finite = all(np.isfinite(p["values"][:Nt]).all() for p in probes)
    s.close()
    return dict(model=model, cells=model.Nx * model.Ny * model.Nz, wall=wall, e_ms=e_ms, h_ms=h_ms, tuned=tuned,
                info=info, finite=finite, host_s=t_model + t_pack + t_session)


def kernel_roofline(run, kname, kbytes, kms, traffic, bytes_model):
    """Roofline object of one kernel: algorithmic bytes per launch over its HIP-event launch time."""
    cells = run["cells"]
    k_gbs = cells * kbytes / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    return dict(bound="hbm", kernel=kname, coefficients=run["info"]["coef_kind"], achieved=round(k_gbs, 1),
                peak=HBM_PEAK_GBS, unit="GB/s", frac=round(k_gbs / HBM_PEAK_GBS, 4), traffic=traffic,
                bytes_per_cell=kbytes, bytes_model=bytes_model, algorithmic_bytes_per_launch=cells * kbytes,
                launch_ms=round(kms, 4),
                measured_hbm_gbs=round(traffic / (kms * 1e-3) / 1e9, 1) if traffic and kms else None,
                measured_hbm_frac=round(traffic / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic and kms else None,
                traffic_note="rocprofv3 FETCH_SIZE and WRITE_SIZE for this launch shape, scaled by a 1 GiB copy's "
                             "counts: the L2's fills from and write-backs to the data fabric, which the Infinity "
                             "Cache serves in part, so measured_hbm_gbs can exceed what HBM alone delivers")


# Table form: the deduplicated coefficient tuples live in L2, so a half step moves its fields (6 split
# values read and written, 3 neighbour totals read, 3 totals written: 72 B), the 4-byte entry index of
# the cell and 1/8 B of probe bitmask.
TABLE_BYTES = 72 + 4 + 0.125
# The two-step pass advances two steps per launch and reads and writes the 12 split values of a cell
# once (96 B), with the two entry indices (8 B) and two bitmask words per 32 cells, or the 30 raw
# coefficient arrays (120 B) instead of the indices.  Halo rows of its tiles are re-read on top.
TWO_STEP_TABLE_BYTES = 96 + 8 + 0.25
TWO_STEP_RAW_BYTES = 96 + 120
# kernel variants (rfdtd_set_variant): 0, 1, 4, 10 separate E and H kernels; 2, 3, 5, 6 the fused
# E+H pass; 7, 8, 9 the two-step pass
TWO_STEP_VARIANTS = (7, 8, 9, 13, 14, 15)
FUSED_VARIANTS = (2, 3, 5, 6) + TWO_STEP_VARIANTS
VECTOR_VARIANTS = (11, 12)  # separate kernels over 2 or 4 z cells per thread (k_update_e_w, k_update_h_w)


def plain_leg(args):
    """The headline: K timed steps of the single-grid workload after W warm-up steps, on the engine's
    default coefficient form (the deduplicated table wherever it is small)."""
    from rfnetwork_amd import _lib

    run = timed_run(args, args.coef)
    model, cells, wall, e_ms, h_ms, tuned, info = (run[k] for k in ("model", "cells", "wall", "e_ms", "h_ms",
                                                                     "tuned", "info"))
    ms_step = wall * 1e3 / args.steps
    mcells = cells * args.steps / wall / 1e6
    dft_b = dft_bytes_per_step(model)
    step_bytes = cells * (E_BYTES + H_BYTES) + dft_b
    step_frac = step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS
    variant = tuned.get("variant", 1) if tuned else 1
    fused = variant in FUSED_VARIANTS
    table = info["coef_kind"] == "table"
    # the dominant kernel: the two-step pass (two cell-updates per cell and launch), the fused E+H
    # pass or the H update
    if variant in TWO_STEP_VARIANTS:
        kname, kbytes, kms = "k_step2", (TWO_STEP_TABLE_BYTES if table else TWO_STEP_RAW_BYTES), e_ms
    elif fused:
        kname, kbytes, kms = "k_step_fused", (2 * TABLE_BYTES if table else E_BYTES + H_BYTES), e_ms
    else:
        kname = "k_update_h_w" if variant in VECTOR_VARIANTS else "k_update_h"
        kbytes, kms = (TABLE_BYTES if table else H_BYTES), h_ms
    shape = (tuned["variant"], tuned["block"], tuned["x_planes"]) if tuned else None
    traffic = measured_traffic(args.workload, info["coef_kind"], kname, tuned)
    if variant in TWO_STEP_VARIANTS:
        model_text = ("two-step pass, per cell and launch (2 cell-updates): 12 splits read and written once "
                      "(96 B) + " + ("E and H entry indices 8 B + 2 probe bitmask words per 32 cells" if table
                                     else "30 coefficients 120 B"))
    elif table:
        model_text = "table form: 72 B of fields + 4 B entry index + 1/8 B probe bitmask per cell"
    else:
        model_text = ("reference formulation: 144 B per cell (6 split H read and written, 18 coefficients, "
                      "3 E totals read, 3 H totals written)")
    copy = None
    if not args.no_copy_probe:
        gbs, cms = _lib.copy_bandwidth(1 << 30, 20)
        copy = dict(gb_per_s=round(gbs, 1), bytes_per_copy=2 * (1 << 30), ms_per_copy=round(cms, 4),
                    best_streaming_gb_per_s=BEST_STREAMING_GBS)
        log(f"copy probe {gbs:.1f} GB/s")
    setup = guarded("setup", setup_times, args.workload, args.warmup + args.steps, run["host_s"]) \
        if args.coef == "auto" else None
    e_gbs = cells * (TABLE_BYTES if table else E_BYTES) / (e_ms * 1e-3) / 1e9 if e_ms > 0 and not fused else None
    return dict(
        value=round(mcells, 1), ms_per_step=round(ms_step, 4),
        config=dict(workload=f"{args.workload} {model.Nx}x{model.Ny}x{model.Nz}", cells=cells,
                    dft_planes=len(model.monitors), dft_frequencies=51 if model.monitors else 0,
                    coefficients=info["coef_kind"], coef_entries=[info["n_e_entries"], info["n_h_entries"]],
                    launch=tuned),
        roofline_table=kernel_roofline(run, kname, kbytes, kms, traffic, model_text),
        step=dict(hbm_frac_reference_bytes=round(step_frac, 4), reference_bytes=step_bytes,
                  e_kernel_ms=None if fused else round(e_ms, 4), e_kernel_gbs=e_gbs and round(e_gbs, 1),
                  h_kernel_ms=None if fused else round(h_ms, 4), dft_bytes=dft_b),
        copy_probe=copy, finite=run["finite"], setup=setup,
    )


def raw_leg(args):
    """The same workload with the 30 coefficient arrays read per cell (the reference formulation,
    RFDTD_COEF_RAW) on the separate E and H kernels (one step per pass, RFDTD_TUNE_SEPARATE): its H
    update moves the 144 B per cell that the roofline's byte count assumes, so this is the
    like-for-like fraction of HBM peak."""
    run = timed_run(args, "raw", separate=True)
    tuned, h_ms, e_ms, cells = run["tuned"], run["h_ms"], run["e_ms"], run["cells"]
    fused = bool(tuned and tuned.get("variant", 1) in FUSED_VARIANTS)
    if fused:
        kname, kbytes, kms = "k_step_fused", E_BYTES + H_BYTES, e_ms
    else:
        kname = "k_update_h_w" if tuned and tuned.get("variant") in VECTOR_VARIANTS else "k_update_h"
        kbytes, kms = H_BYTES, h_ms
    shape = (tuned["variant"], tuned["block"], tuned["x_planes"]) if tuned else None
    traffic = measured_traffic(args.workload, "raw", kname, tuned)
    roof = kernel_roofline(run, kname, kbytes, kms, traffic,
                           "reference formulation: 144 B per cell (6 split H read and written, 18 coefficients, "
                           "3 E totals read, 3 H totals written); E update 120 B")
    e_frac = cells * E_BYTES / (e_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if e_ms > 0 and not fused else None
    return dict(roofline=roof, ms_per_step=round(run["wall"] * 1e3 / args.steps, 4),
                mcells=round(cells * args.steps / run["wall"] / 1e6, 1), launch=tuned,
                e_kernel_ms=None if fused else round(e_ms, 4), e_kernel_frac=e_frac and round(e_frac, 4),
                h_kernel_ms=round(h_ms, 4), finite=run["finite"])


def setup_times(workload, Nt, host_s):
    """Model to resident session, with the coefficients assembled on the host (mesh, 30 arrays,
    packing, copy, table) against the model meshed with coefficients="device" and its session
    assembling them on the GPU (the same arrays bit for bit)."""
    from rfnetwork_amd import workloads
    from rfnetwork_amd.session import Session

    t0 = time.perf_counter()
    build = dict(dipole_slab=workloads.dipole_slab, microstrip=workloads.microstrip)[workload]
    model = build(Nt=Nt, coefficients="device")
    t_model = time.perf_counter() - t0
    with Session.assembled(model) as s:
        s.synchronize()
        info = s.memory_info()
    dev_s = time.perf_counter() - t0
    log(f"setup: host assembly {host_s:.2f}s, device assembly {dev_s:.2f}s (model {t_model:.2f}s), {info['coef_kind']}")
    return dict(host_assembly_s=round(host_s, 3), device_assembly_s=round(dev_s, 3), device_model_s=round(t_model, 3),
                speedup=round(host_s / dev_s, 2))


if __name__ == "__main__":
    main()
