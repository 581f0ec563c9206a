"""Batched sessions: uneven grids x excitations advanced in the same launches must give every instance
exactly what a separate single-grid run (and the CPU oracle) gives."""
import numpy as np
import pytest

from rfnetwork_amd import batch as rbatch, workloads
from rfnetwork_amd.batch import Batch
from rfnetwork_amd.session import Session

from helpers import run_oracle

pytestmark = pytest.mark.gpu


def _models():
    return [workloads.tiny_model(Nt=50, nx=14, ny=11, nz=9, seed=7),
            workloads.tiny_model(Nt=50, nx=19, ny=12, nz=10, seed=8),
            workloads.tiny_model(Nt=50, nx=16, ny=13, nz=70, seed=9)]


def test_batch_instances_equal_single_runs(hip_ready):
    print("[case] batch vs single", flush=True)
    models = _models()
    grids, per_inst = [], []
    for m in models:
        coeff = m.coefficient_arrays(gpu=True)
        grids.append((coeff, m.Nx, m.Ny, m.Nz, 2))
        for scale in (1.0, -0.5):  # two excitations of the same grid
            _, probes, monitors, _, Nx, Ny, Nz, Nt = m.pack(gpu=True, coefficients=False)
            for p in probes:
                p["values"] = (p["values"] * np.float32(scale)).astype(np.float32)
            per_inst.append((m, coeff, probes, monitors, Nt))
    with Batch(grids) as b:
        b.set_launch(128, 0, 16)
        for k, (_, _, probes, monitors, Nt) in enumerate(per_inst):
            b.set_sources(k, probes, monitors, Nt)
        b.run(37)
        b.run(13)
        b.synchronize()
        b.read()
        fields = [b.fields(k) for k in range(len(per_inst))]
    for k, (m, coeff, probes, monitors, Nt) in enumerate(per_inst):
        ref_p = [dict(p, values=p["values"].copy()) for p in probes]
        # the batch wrote results into probes/monitors; rerun the instance alone from its inputs
        _, src, mon, _, *_ = m.pack(gpu=True, coefficients=False)
        scale = 1.0 if k % 2 == 0 else -0.5
        for p in src:
            p["values"] = (p["values"] * np.float32(scale)).astype(np.float32)
        with Session(coeff, m.Nx, m.Ny, m.Nz) as s:
            s.set_sources(src, mon, Nt)
            s.run(Nt)
            s.synchronize()
            s.read()
            f = s.fields()
        for a, c in zip(ref_p, src):
            np.testing.assert_array_equal(a["values"], c["values"])
        for a, c in zip(monitors, mon):
            np.testing.assert_array_equal(a["values"], c["values"])
        for name in f:
            np.testing.assert_array_equal(fields[k][name], f[name], err_msg=f"inst {k} {name}")


def test_batch_matches_oracle(hip_ready):
    print("[case] batch vs oracle", flush=True)
    models = _models()
    grids, inputs = [], []
    for m in models:
        grids.append((m.coefficient_arrays(gpu=True), m.Nx, m.Ny, m.Nz, 1))
        _, probes, monitors, _, Nx, Ny, Nz, Nt = m.pack(gpu=True, coefficients=False)
        inputs.append((probes, monitors, Nt))
    with Batch(grids) as b:
        for k, (probes, monitors, Nt) in enumerate(inputs):
            b.set_sources(k, probes, monitors, Nt)
        b.run(inputs[0][2])
        b.synchronize()
        b.read()
        fields = [b.fields(k) for k in range(len(models))]
    for m, (probes, monitors, Nt), f in zip(models, inputs, fields):
        p_c, m_c, f_c = run_oracle(m, np.float32, n_threads=2)
        for a, c in zip(probes, p_c):
            np.testing.assert_array_equal(a["values"], c["values"])
        for k in f_c:
            np.testing.assert_array_equal(f[k], f_c[k], err_msg=k)


def test_solve_ports_equals_per_port_solves(hip_ready):
    """Every port of two uneven two-port models driven in one batch: the S columns equal those of
    separate solve() calls, one per driven port."""
    print("[case] solve_ports", flush=True)
    f = np.linspace(4e9, 12e9, 5)
    jobs = [(workloads.microstrip(Nt=400, nx=40, ny=20, nz=10, n_pml=4), [1, 2]),
            (workloads.microstrip(Nt=400, nx=52, ny=24, nz=12, n_pml=5), [2, 1])]
    results = rbatch.solve_ports(jobs, f)
    for (model, ports), (S, _) in zip(jobs, results):
        wave = np.asarray(model.ports[0]["src"])
        for col, port in enumerate(ports):
            ref = workloads.microstrip(Nt=400, nx=model.Nx, ny=model.Ny, nz=model.Nz, n_pml=4 if model.Nx == 40 else 5)
            ref.reset_excitations()
            ref.assign_excitation(wave, port)
            ref.solve(gpu=True, show_progress=False)
            S_ref = np.asarray(ref.get_sparameters(f, source_port=port))[..., 0]
            np.testing.assert_array_equal(S[..., col], S_ref)


def test_write_touchstone_full_matrix(hip_ready, tmp_path):
    """FDTD_Solver.write_touchstone: the file read back holds the S-matrix whose columns are the
    per-port solve() results."""
    print("[case] touchstone", flush=True)
    from rfnetwork_amd import touchstone

    f = np.linspace(4e9, 12e9, 5)
    model = workloads.microstrip(Nt=400, nx=40, ny=20, nz=10, n_pml=4)
    S = model.write_touchstone(tmp_path / "line.s2p", f)
    S2, _, _ = touchstone.read_snp(tmp_path / "line.s2p")
    np.testing.assert_allclose(np.asarray(S2), np.asarray(S), rtol=1e-10, atol=1e-13)
    ref = workloads.microstrip(Nt=400, nx=40, ny=20, nz=10, n_pml=4)
    ref.solve(gpu=True, show_progress=False)
    np.testing.assert_array_equal(S[..., 0], np.asarray(ref.get_sparameters(f, source_port=1))[..., 0])
