"""Edge cases of the drop-in call against the oracle: one time step, runs without monitors,
without any probe or monitor, and a time monitor whose sampling interval is longer than the run
(one sample, the count `Nt / n_step + 1` of solver.cpp:383).  Results must be identical bit for
bit, as in the main parity tests."""
import copy

import numpy as np
import pytest

from oracle import cpu_solver
from rfnetwork_amd import workloads
from rfnetwork_amd.hip_func import solver_run_hip
from rfnetwork_amd.session import Session

from helpers import run_oracle

pytestmark = pytest.mark.gpu


def _drop_in(model, dtype):
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = model.pack(gpu=True, dtype=dtype)
    solver_run_hip(c, probes, monitors, mem, Nx, Ny, Nz, Nt, dtype=dtype)
    return probes, monitors


def _assert_same(model, gpu, cpu):
    (p_g, m_g), (p_c, m_c) = gpu, cpu
    assert len(p_g) == len(p_c)
    for a, b in zip(p_g, p_c):
        np.testing.assert_array_equal(a["values"], b["values"])
    model.unpack(p_g, m_g, gpu=True)
    g = {k: np.array(m["values"]) for k, m in model.monitors.items()}
    model.unpack(p_c, m_c, gpu=False)
    c = {k: np.array(m["values"]) for k, m in model.monitors.items()}
    assert g.keys() == c.keys()
    for k in g:
        assert g[k].shape == c[k].shape, k
        np.testing.assert_array_equal(g[k], c[k], err_msg=k)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_single_step(hip_ready, dtype):
    print(f"[case] single step {np.dtype(dtype).name}", flush=True)
    model = workloads.tiny_model(Nt=1)
    # a one-step Gaussian pulse normalises 0 / 0; drive the port with a plain nonzero series
    model.assign_excitation(np.array([3e-3, 7e-3]), 1)
    cpu = run_oracle(model, dtype, n_threads=2, fields=False)[:2]
    assert all(np.all(np.isfinite(p["values"])) for p in cpu[0])
    assert any(np.any(p["values"]) for p in cpu[0][1:])
    _assert_same(model, _drop_in(model, dtype), cpu)


def test_no_monitors(hip_ready):
    print("[case] no monitors", flush=True)
    model = workloads.tiny_model(Nt=40, time_monitors=False, phasor_monitors=False)
    cpu = run_oracle(model, np.float32, n_threads=2, fields=False)[:2]
    gpu = _drop_in(model, np.float32)
    assert gpu[1] == [] and cpu[1] == []
    _assert_same(model, gpu, cpu)


def test_monitor_interval_longer_than_run(hip_ready):
    print("[case] monitor interval > Nt", flush=True)
    Nt = 12
    model = workloads.tiny_model(Nt=Nt, time_monitors=False, phasor_monitors=False)
    model.add_field_monitor("t_once", "ez", "x", index=6, n_step=Nt + 5)
    model.add_field_monitor("t_last", "hy", "z", index=4, n_step=Nt)
    cpu = run_oracle(model, np.float32, n_threads=2, fields=False)[:2]
    gpu = _drop_in(model, np.float32)
    _assert_same(model, gpu, cpu)
    assert np.array(model.monitors["t_once"]["values"]).shape[0] == 1
    assert np.array(model.monitors["t_last"]["values"]).shape[0] == 2


def _extra_probes(model, n, Nt, rng, dtype):
    """n probe records at random interior cells and components, a third of them sources with
    random series; cells repeat, so several records share (component, cell) groups."""
    Nx, Ny, Nz = model.Nx, model.Ny, model.Nz
    out = []
    for _ in range(n):
        src = rng.random() < 1 / 3
        out.append({"values": (rng.uniform(-1e-3, 1e-3, Nt) if src else np.zeros(Nt)).astype(dtype),
                    "field": int(rng.integers(0, 6)),
                    "idx": [int(rng.integers(1, Nx)), int(rng.integers(1, Ny)), int(rng.integers(1, Nz))],
                    "is_source": int(src)})
    return out


def test_maximum_probe_and_monitor_counts(hip_ready):
    """Exactly the reference's limits (solver.h:15-16): 50 monitors and 5000 probe records."""
    from rfnetwork_amd import _lib

    print("[case] 5000 probes, 50 monitors", flush=True)
    Nt = 24
    model = workloads.tiny_model(Nt=Nt)
    f = np.array([12e9, 30e9])
    k = 0
    while len(model.monitors) < _lib.MAX_MONITORS:
        field = ("ex", "ey", "ez", "hx", "hy", "hz")[k % 6]
        axis = "xyz"[(k // 6) % 3]
        index = 2 + k % 5
        if k % 2:
            model.add_field_monitor(f"m{k}", field, axis, index=index, frequency=f, n_step=1 + k % 3)
        else:
            model.add_field_monitor(f"m{k}", field, axis, index=index, n_step=1 + k % 4)
        k += 1
    c, p_g, m_g, mem, Nx, Ny, Nz, _ = model.pack(gpu=True)
    c_c, p_c, m_c, mem_c, *_ = model.pack(gpu=False, n_threads=2)
    extra = _extra_probes(model, _lib.MAX_PROBES - len(p_g), Nt, np.random.default_rng(5), np.float32)
    p_g = p_g + copy.deepcopy(extra)
    p_c = p_c + copy.deepcopy(extra)
    assert len(p_g) == _lib.MAX_PROBES and len(m_g) == _lib.MAX_MONITORS
    cpu_solver.solver_run(c_c, p_c, m_c, mem_c, Nx, Ny, Nz, Nt, 2, 0, dtype=np.float32)
    solver_run_hip(c, p_g, m_g, mem, Nx, Ny, Nz, Nt)
    for i, (a, b) in enumerate(zip(p_g, p_c)):
        np.testing.assert_array_equal(a["values"], b["values"], err_msg=f"probe {i}")
    _assert_same(model, (p_g[:len(p_g) - len(extra)], m_g), (p_c[:len(p_c) - len(extra)], m_c))


def _thin_model(nx, ny, nz):
    """A grid with one axis of 2 cells (the smallest the solver takes): filled with a dielectric,
    PML on the two sides of the longest axis, no ports."""
    from rfnetwork_amd.geometry import Box
    from rfnetwork_amd.workloads import _grid_solver

    s, d = _grid_solver(nx, ny, nz, 1e-4)
    s.add_dielectric(Box((0, 0, 0), (nx * d, ny * d, nz * d)), er=2.2)
    axis = "xyz"[int(np.argmax([nx, ny, nz]))]
    s.assign_PML_boundaries(axis + "-", axis + "+", n_pml=3)
    s.generate_mesh(courant=0.99)
    return s


@pytest.mark.parametrize("dims", [(30, 20, 2), (30, 2, 20), (2, 20, 30), (2, 2, 40)])
def test_two_cell_axis(hip_ready, dims):
    """Source and record probes at random interior cells of a grid with a 2-cell axis: probes bit
    for bit equal to the oracle's."""
    print(f"[case] two-cell axis {dims}", flush=True)
    Nt = 30
    model = _thin_model(*dims)
    Nx, Ny, Nz = model.Nx, model.Ny, model.Nz
    assert (Nx, Ny, Nz) == dims
    # no port, so the arrays are taken directly (pack() wants an excitation); the oracle's work
    # buffer is sized as pack() sizes it
    c = model.coefficient_arrays(gpu=True, dtype=np.float32)
    c_c = model.coefficient_arrays(gpu=False, dtype=np.float32)
    mem_len = sum(3 * int(np.prod((Nx,) + s[1:])) for s in model.fshape.values()) + 4 * (Ny + 1) * (Nz + 1) * 2 * 8
    mem_c = np.zeros(mem_len, np.float32)
    extra = _extra_probes(model, 60, Nt, np.random.default_rng(sum(dims)), np.float32)
    p_g, p_c = copy.deepcopy(extra), copy.deepcopy(extra)
    cpu_solver.solver_run(c_c, p_c, [], mem_c, Nx, Ny, Nz, Nt, 2, 0, dtype=np.float32)
    solver_run_hip(c, p_g, [], np.zeros(1, np.float32), Nx, Ny, Nz, Nt)
    assert any(np.any(p["values"]) for p, e in zip(p_c, extra) if not e["is_source"])
    for i, (a, b) in enumerate(zip(p_g, p_c)):
        np.testing.assert_array_equal(a["values"], b["values"], err_msg=f"probe {i}")


def test_no_probes_no_monitors(hip_ready):
    """Nothing drives the grid: every field stays zero, on the GPU as in the oracle."""
    print("[case] no probes, no monitors", flush=True)
    model = workloads.tiny_model(Nt=20, time_monitors=False, phasor_monitors=False)
    c, _, _, mem, Nx, Ny, Nz, Nt = model.pack(gpu=True)
    with Session(c, Nx, Ny, Nz) as s:
        s.set_sources([], [], Nt)
        s.run(Nt)
        s.synchronize()
        s.read()
        f_g = s.fields()
    c_cpu, _, _, mem_cpu, *_ = model.pack(gpu=False, n_threads=2)
    f_c = {n: np.zeros(cpu_solver.global_field_shape(k, Nx, Ny, Nz), np.float32)
           for n, k in zip(cpu_solver.FIELD_NAMES, cpu_solver.FIELD_COMPONENT)}
    cpu_solver.solver_run(c_cpu, [], [], mem_cpu, Nx, Ny, Nz, Nt, 2, 0, dtype=np.float32, fields_out=f_c)
    assert f_g.keys() == f_c.keys()
    for k in f_g:
        assert not np.any(f_g[k]), k
        np.testing.assert_array_equal(f_g[k], f_c[k], err_msg=k)


def test_repeated_drop_in_calls_are_independent(hip_ready):
    """Two drop-in calls with the same arguments give the same outputs (no state carried over)."""
    print("[case] repeated drop-in calls", flush=True)
    model = workloads.tiny_model(Nt=30)
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = model.pack(gpu=True)
    c2, probes2, monitors2 = copy.deepcopy(c), copy.deepcopy(probes), copy.deepcopy(monitors)
    solver_run_hip(c, probes, monitors, mem, Nx, Ny, Nz, Nt)
    solver_run_hip(c2, probes2, monitors2, mem, Nx, Ny, Nz, Nt)
    for a, b in zip(probes, probes2):
        np.testing.assert_array_equal(a["values"], b["values"])
    for a, b in zip(monitors, monitors2):
        np.testing.assert_array_equal(np.asarray(a["values"]), np.asarray(b["values"]))
