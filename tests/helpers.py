"""Shared helpers for the tests: run a model on the oracle and on the HIP engine."""
import copy

import numpy as np

from oracle import cpu_solver


def run_oracle(model, dtype=np.float32, n_threads=3, fields=True):
    """Pack the model in the CPU layout and run the oracle; returns (probes, monitors, fields)."""
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = model.pack(gpu=False, dtype=dtype, n_threads=n_threads)
    out = None
    if fields:
        out = {n: np.zeros(cpu_solver.global_field_shape(k, Nx, Ny, Nz), dtype)
               for n, k in zip(cpu_solver.FIELD_NAMES, cpu_solver.FIELD_COMPONENT)}
    cpu_solver.solver_run(c, probes, monitors, mem, Nx, Ny, Nz, Nt, n_threads, 0, dtype=dtype, fields_out=out)
    return probes, monitors, out


def cpu_monitor_to_global(model, monitors_cpu, monitors_gpu_shape=False):
    """Pad CPU-layout monitor values to full Yee planes (what FDTD_Solver.unpack does)."""
    m2 = copy.deepcopy(monitors_cpu)
    return m2


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0)
