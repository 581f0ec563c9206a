"""Streamed time-domain monitors (rfdtd_set_monitor_streaming): a small device ring per monitor,
the run cut into segments, each segment's sample planes copied to pinned host arrays on a copy
stream while the next segment runs.  Every monitor value, probe series and field must equal the
unstreamed run's bit for bit, whatever the ring size, chunking, graph replay, checkpointing or
resets."""
import numpy as np
import pytest

from rfnetwork_amd import _lib, workloads
from rfnetwork_amd.batch import Batch
from rfnetwork_amd.session import Session

pytestmark = pytest.mark.gpu


def _run(model, ring, chunks, flags=0, every=0, rerun=False):
    c, probes, monitors, _, Nx, Ny, Nz, Nt = model.pack(gpu=True)
    with Session(c, Nx, Ny, Nz, flags=flags) as s:
        if ring:
            s.set_monitor_streaming(ring)
        if every:
            s.set_checkpointing(every)
        for r in range(2 if rerun else 1):
            if r:  # the read filled the dictionaries with results: fresh sources for the rerun
                _, probes, monitors, _, *_ = model.pack(gpu=True, coefficients=False)
            s.reset()
            s.set_sources(probes, monitors, Nt)
            for n in chunks:
                s.run(n)
            s.synchronize()
            s.read()
        f = s.fields()
    return [np.array(p["values"]) for p in probes], [np.array(m["values"]) for m in monitors], f


def _assert_equal(a, b):
    for x, y in zip(a[0], b[0]):
        np.testing.assert_array_equal(x, y)
    for k, (x, y) in enumerate(zip(a[1], b[1])):
        np.testing.assert_array_equal(x, y, err_msg=f"monitor {k}")
    for k in a[2]:
        np.testing.assert_array_equal(a[2][k], b[2][k], err_msg=k)


@pytest.mark.parametrize("ring,flags", [(2, 0), (5, 0), (64, 0), (1000, 0), (6, _lib.NO_GRAPHS)])
def test_streamed_monitors_equal_unstreamed(hip_ready, ring, flags):
    print(f"[case] streaming ring={ring} flags={flags}", flush=True)
    model = workloads.tiny_model(Nt=150)
    ref = _run(model, 0, [150], flags)
    assert any(m.ndim == 3 and m.shape[0] > 10 for m in ref[1])
    _assert_equal(_run(model, ring, [37, 90, 23], flags), ref)


def test_streamed_monitors_with_checkpoints_and_reset(hip_ready):
    print("[case] streaming checkpoints reset", flush=True)
    model = workloads.tiny_model(Nt=120)
    ref = _run(model, 0, [120])
    _assert_equal(_run(model, 8, [50, 70], every=17, rerun=True), ref)


def test_streamed_monitors_in_a_batch(hip_ready):
    print("[case] streaming batch", flush=True)
    models = [workloads.tiny_model(Nt=80, seed=7), workloads.tiny_model(Nt=80, nx=19, ny=12, nz=10, seed=8)]
    refs = [_run(m, 0, [80]) for m in models]
    grids, inputs = [], []
    for m in models:
        coeff, probes, monitors, _, Nx, Ny, Nz, Nt = m.pack(gpu=True)
        grids.append((coeff, Nx, Ny, Nz, 1))
        inputs.append((probes, monitors))
    with Batch(grids) as b:
        b.set_monitor_streaming(4)
        for k, (probes, monitors) in enumerate(inputs):
            b.set_sources(k, probes, monitors, 80)
        b.run(33)
        b.run(47)
        b.synchronize()
        b.read()
    for (probes, monitors), ref in zip(inputs, refs):
        for p, r in zip(probes, ref[0]):
            np.testing.assert_array_equal(p["values"], r)
        for m, r in zip(monitors, ref[1]):
            np.testing.assert_array_equal(m["values"], r)


def test_synchronize_before_the_first_run(hip_ready):
    """A synchronize before any run (timing events not yet recorded) leaves no error behind for the
    run that follows, streamed or not."""
    print("[case] synchronize before run", flush=True)
    model = workloads.tiny_model(Nt=40)
    c, probes, monitors, _, Nx, Ny, Nz, Nt = model.pack(gpu=True)
    for ring in (0, 4):
        _, probes, monitors, _, *_ = model.pack(gpu=True, coefficients=False)
        with Session(c, Nx, Ny, Nz) as s:
            if ring:
                s.set_monitor_streaming(ring)
            s.set_sources(probes, monitors, Nt)
            s.synchronize()
            s.run(Nt)
            s.synchronize()
            s.read()


def test_streamed_total_field_monitor(hip_ready):
    """An "e_total" monitor (three component planes) streamed through a device ring and read back
    with get_monitor_data: the corner-averaged x, y, z stack equals the oracle run's bit for bit."""
    from helpers import run_oracle

    print("[case] streaming e_total", flush=True)
    ref = workloads.tiny_model(Nt=90)
    ref.add_field_monitor("tot", "e_total", "z", index=4, n_step=2)
    probes, monitors, _ = run_oracle(ref, fields=False)
    ref.unpack(probes, monitors, gpu=False)
    want = ref.get_monitor_data("tot")

    model = workloads.tiny_model(Nt=90)
    model.add_field_monitor("tot", "e_total", "z", index=4, n_step=2)
    c, probes, monitors, _, Nx, Ny, Nz, Nt = model.pack(gpu=True)
    with Session(c, Nx, Ny, Nz) as s:
        s.set_monitor_streaming(6)
        s.set_sources(probes, monitors, Nt)
        s.run(41)
        s.run(Nt - 41)
        s.synchronize()
        s.read()
    model.unpack(probes, monitors, gpu=True)
    got = model.get_monitor_data("tot")
    assert got.shape == want.shape and got.shape[0] == 3
    assert np.abs(np.asarray(want)).max() > 0
    np.testing.assert_array_equal(np.asarray(got), np.asarray(want))
