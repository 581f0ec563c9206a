"""Launch autotuner: it leaves the session's state untouched (results stay bit-exact with the oracle),
returns a valid launch shape, and caches it per shape."""
import numpy as np
import pytest

from rfnetwork_amd import autotune, workloads
from rfnetwork_amd.batch import Batch
from rfnetwork_amd.session import Session

from helpers import run_oracle

pytestmark = pytest.mark.gpu


def test_autotuned_session_matches_oracle(hip_ready, tmp_path, monkeypatch):
    print("[case] autotune session", flush=True)
    monkeypatch.setenv("RFNETWORK_AMD_TUNE_CACHE", str(tmp_path / "launch.json"))
    autotune._memory.clear()
    model = workloads.tiny_model(Nt=40)
    c, probes, monitors, _, Nx, Ny, Nz, Nt = model.pack(gpu=True)
    with Session(c, Nx, Ny, Nz) as s:
        s.set_sources(probes, monitors, Nt)
        cfg = s.autotune(trial_steps=5)
        assert cfg["cached"] is False
        assert cfg["variant"] in range(10) and cfg["block"] in (128, 256, 512) and cfg["x_planes"] >= 0
        assert cfg["ms_per_step"] > 0
        s.run(Nt)
        s.synchronize()
        s.read()
        f = s.fields()
    p_c, _, f_c = run_oracle(model, np.float32, n_threads=2)
    for a, b in zip(probes, p_c):
        np.testing.assert_array_equal(a["values"], b["values"])
    for k in f_c:
        np.testing.assert_array_equal(f[k], f_c[k], err_msg=k)
    # the same shape again: served from the cache file, not re-measured
    autotune._memory.clear()
    with Session(c, Nx, Ny, Nz) as s:
        cfg2 = s.autotune(trial_steps=5)
    assert cfg2["cached"] is True
    assert {k: cfg2[k] for k in ("variant", "block", "x_planes")} == {k: cfg[k] for k in ("variant", "block", "x_planes")}


def test_autotune_refuses_a_session_that_ran(hip_ready):
    model = workloads.tiny_model(Nt=10)
    c, probes, monitors, _, Nx, Ny, Nz, Nt = model.pack(gpu=True)
    with Session(c, Nx, Ny, Nz) as s:
        s.set_sources(probes, monitors, Nt)
        s.run(2)
        s.synchronize()
        with pytest.raises(Exception, match="before its first run"):
            s.autotune(use_cache=False)


def test_autotuned_batch_matches_single_runs(hip_ready, tmp_path, monkeypatch):
    print("[case] autotune batch", flush=True)
    monkeypatch.setenv("RFNETWORK_AMD_TUNE_CACHE", str(tmp_path / "launch.json"))
    models = [workloads.tiny_model(Nt=30, nx=14, ny=11, nz=9, seed=7),
              workloads.tiny_model(Nt=30, nx=21, ny=12, nz=10, seed=8)]
    grids, inputs = [], []
    for m in models:
        coeff = m.coefficient_arrays(gpu=True)
        grids.append((coeff, m.Nx, m.Ny, m.Nz, 1))
        _, probes, monitors, _, *_ = m.pack(gpu=True, coefficients=False)
        inputs.append((m, coeff, probes, monitors))
    with Batch(grids) as b:
        cfg = b.autotune(trial_steps=5, use_cache=False)
        assert cfg["block"] in (128, 256, 512)
        for k, (_, _, probes, monitors) in enumerate(inputs):
            b.set_sources(k, probes, monitors, 30)
        b.run(30)
        b.synchronize()
        b.read()
    for m, coeff, probes, _ in inputs:
        p_c, _, _ = run_oracle(m, np.float32, n_threads=2, fields=False)
        for a, c in zip(probes, p_c):
            np.testing.assert_array_equal(a["values"], c["values"])


def test_adjoint_shape_timed_once_and_cached(hip_ready, tmp_path, monkeypatch):
    """The first gradient of an autotuned session times the reverse-pass shapes; the chosen shape is
    cached under the grid's key and applied to the next session of that shape, with the same
    gradient."""
    print("[case] adjoint shape cache", flush=True)
    monkeypatch.setenv("RFNETWORK_AMD_TUNE_CACHE", str(tmp_path / "launch.json"))
    autotune._memory.clear()
    Nt = 30
    model = workloads.tiny_model(Nt=Nt, time_monitors=False, phasor_monitors=False)
    c, probes, _, _, Nx, Ny, Nz, _ = model.pack(gpu=True)
    g = [np.random.default_rng(2).standard_normal(Nt).astype(np.float32) for _ in probes]
    box = (0, Nx, 0, Ny, 0, Nz)
    out, shapes = [], []
    for _ in range(2):
        autotune._memory.clear()
        _, pr, _, _, *_ = model.pack(gpu=True, coefficients=False)
        with Session(c, Nx, Ny, Nz) as s:
            s.autotune(trial_steps=3)
            shapes.append(s.adjoint_launch())
            s.set_history(box, Nt)
            s.set_sources(pr, [], Nt)
            s.run(Nt)
            s.synchronize()
            out.append(s.gradient(box, g))
            assert s.adjoint_launch()["block"] in (128, 256, 512)
            shapes.append(s.adjoint_launch())
    assert shapes[0]["block"] == 0  # nothing cached yet: timed at the gradient
    assert shapes[2] == shapes[1] == shapes[3]  # cached, applied before the second gradient
    np.testing.assert_array_equal(out[0], out[1])


def test_autotune_with_sources_installed_leaves_them_intact(hip_ready, tmp_path, monkeypatch):
    """Tuning after set_sources times the trials with probes and monitors applied, then restores
    them: the run afterwards equals the oracle bit for bit."""
    print("[case] autotune with sources", flush=True)
    monkeypatch.setenv("RFNETWORK_AMD_TUNE_CACHE", str(tmp_path / "launch.json"))
    autotune._memory.clear()
    model = workloads.tiny_model(Nt=40)
    c, probes, monitors, _, Nx, Ny, Nz, Nt = model.pack(gpu=True)
    with Session(c, Nx, Ny, Nz) as s:
        s.set_sources(probes, monitors, Nt)
        s.autotune(trial_steps=5)
        s.run(Nt)
        s.synchronize()
        s.read()
        f = s.fields()
    p_c, m_c, f_c = run_oracle(model, np.float32, n_threads=2)
    for a, b in zip(probes, p_c):
        np.testing.assert_array_equal(a["values"], b["values"])
    model.unpack(probes, monitors, gpu=True)
    g = {k: np.array(m["values"]) for k, m in model.monitors.items()}
    model.unpack(p_c, m_c, gpu=False)
    for k, m in model.monitors.items():
        np.testing.assert_array_equal(g[k], np.array(m["values"]), err_msg=k)
    for k in f_c:
        np.testing.assert_array_equal(f[k], f_c[k], err_msg=k)
