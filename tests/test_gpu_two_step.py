"""The two-step pass (variants 7 and 8: 12- and 4-row tiles of 64-lane rows; 9: its one-barrier form,
8-row tiles and the x planes skewed by three; 13, 15 and 14: 28-, 26- and 12-row tiles of 32-lane
rows, two rows per wavefront): time steps n and n + 1 of a tile in one launch, with two-cell halos
recomputed; step n's monitor samples are taken inside the pass.  Its
results must equal the oracle's and the other variants' bit for bit, whatever the tiling, x planes,
chunking (odd chunk lengths end on a single step), buffer parity, graph replay or batch, and a
forward run recorded for a gradient (checkpoints between pairs; a design-box history, which runs
the pass one step at a time) must give the same gradient."""
import numpy as np
import pytest

from rfnetwork_amd import _lib, workloads
from rfnetwork_amd.batch import Batch
from rfnetwork_amd.session import Session

from helpers import run_oracle
from test_gpu_fused import _assert_equal_to_oracle, _run

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant", [7, 8, 9, 13, 14, 15])
@pytest.mark.parametrize("flags", [0, _lib.COEF_RAW, _lib.NO_GRAPHS, _lib.COEF_RAW | _lib.NO_GRAPHS])
def test_two_step_tiny_fp32_bit_exact(hip_ready, flags, variant):
    """PML, dielectric, conductor, two ports, a voltage probe, time and phasor monitors on every axis;
    chunks of 37 and 23 steps (graph chunks of 8 steps, then pairs, then a single step)."""
    print(f"[case] two-step tiny flags={flags} variant={variant}", flush=True)
    model = workloads.tiny_model(Nt=60)
    p, m, f = _run(model, np.float32, flags, variant, chunks=[37, 23], launch=(0, 3, 10))
    _assert_equal_to_oracle(model, np.float32, p, m, f)


@pytest.mark.parametrize("variant", [7, 8, 9, 13, 14, 15])
def test_two_step_tiny_fp64_bit_exact(hip_ready, variant):
    print(f"[case] two-step tiny fp64 variant={variant}", flush=True)
    model = workloads.tiny_model(Nt=50)
    p, m, f = _run(model, np.float64, 0, variant, chunks=[21, 29], launch=(0, 4, 6))
    _assert_equal_to_oracle(model, np.float64, p, m, f)


def test_two_step_interleaves_with_other_variants(hip_ready):
    print("[case] two-step interleave", flush=True)
    model = workloads.tiny_model(Nt=60)
    p, m, f = _run(model, np.float32, 0, 7, chunks=[13, 14, 7, 11, 4, 5, 3, 2, 1],
                   switch=[7, 1, 9, 8, 0, 6, 13, 14, 15, 9])
    _assert_equal_to_oracle(model, np.float32, p, m, f)


def test_two_step_many_tiles_and_planes(hip_ready):
    """Several tiles along y (12 or 28 rows) and z (60 or 28 cells) and x chunks of 1 .. all planes:
    the probes, monitors and fields equal the separate kernels'."""
    print("[case] two-step tiles", flush=True)
    model = workloads.tiny_model(Nt=30, nx=23, ny=31, nz=131)
    p0, m0, ref = _run(model, variant=1)
    ref_p = [np.array(p["values"]) for p in p0]
    ref_m = [np.array(m["values"]) for m in m0]
    for variant in (7, 8, 9, 13, 14, 15):
        for xp in (1, 2, 3, 5, 23):
            p, m, f = _run(model, variant=variant, launch=(0, xp, 8), chunks=[11, 19])
            for k in ref:
                np.testing.assert_array_equal(f[k], ref[k], err_msg=f"{k} variant {variant} x planes {xp}")
            for a, b in zip(p, ref_p):
                np.testing.assert_array_equal(a["values"], b, err_msg=f"probe, variant {variant} x planes {xp}")
            for k, (a, b) in enumerate(zip(m, ref_m)):
                np.testing.assert_array_equal(a["values"], b, err_msg=f"monitor {k}, variant {variant} x planes {xp}")


@pytest.mark.parametrize("variant", [7, 8, 9, 13, 14, 15])
def test_two_step_batch_matches_oracle(hip_ready, variant):
    print(f"[case] two-step batch variant={variant}", flush=True)
    models = [workloads.tiny_model(Nt=40, nx=14, ny=11, nz=9, seed=7),
              workloads.tiny_model(Nt=40, nx=19, ny=26, nz=70, seed=8)]
    grids, inputs = [], []
    for m in models:
        grids.append((m.coefficient_arrays(gpu=True), m.Nx, m.Ny, m.Nz, 1))
        _, probes, monitors, _, *_ = m.pack(gpu=True, coefficients=False)
        inputs.append((probes, monitors))
    with Batch(grids) as b:
        b.set_variant(variant)
        for k, (probes, monitors) in enumerate(inputs):
            b.set_sources(k, probes, monitors, 40)
        b.run(25)
        b.run(15)
        b.synchronize()
        b.read()
        fields = [b.fields(k) for k in range(2)]
    for m, (probes, monitors), f in zip(models, inputs, fields):
        _assert_equal_to_oracle(m, np.float32, probes, monitors, f)


def test_two_step_streamed_time_monitors(hip_ready):
    """Time-domain monitors streamed through small device rings: the samples step n of each pair
    writes inside the pass land in the right ring slots."""
    print("[case] two-step streaming", flush=True)
    model = workloads.tiny_model(Nt=90)
    _, m_ref, _ = _run(model, variant=1)
    want = [np.array(m["values"]) for m in m_ref]
    c, probes, monitors, _, Nx, Ny, Nz, Nt = model.pack(gpu=True)
    with Session(c, Nx, Ny, Nz) as s:
        s.set_variant(9)
        s.set_monitor_streaming(6)
        s.set_sources(probes, monitors, Nt)
        for n in (33, 41, 16):
            s.run(n)
        s.synchronize()
        s.read()
    for k, (m, w) in enumerate(zip(monitors, want)):
        np.testing.assert_array_equal(np.asarray(m["values"]), w, err_msg=f"monitor {k}")


@pytest.mark.parametrize("mode", ["checkpoint", "history"])
def test_two_step_forward_gradient_equals_separate(hip_ready, mode):
    """A forward run on the two-step pass (checkpoints every 9 steps, or a design-box history that
    makes the pass run single steps) gives the separate kernels' gradient bit for bit, as do the
    separate kernels with their workgroups in launch order (forward run, the history's save blocks,
    the checkpoint recomputation)."""
    print(f"[case] two-step gradient {mode}", flush=True)
    Nt = 40
    model = workloads.tiny_model(Nt=Nt, time_monitors=False, phasor_monitors=False)
    coeff, probes, _, _, Nx, Ny, Nz, _ = model.pack(gpu=True)
    g = np.random.default_rng(3).standard_normal((len(probes), Nt)).astype(np.float32)
    box = (2, Nx - 2, 2, Ny - 2, 1, Nz - 1)
    out = []
    for variant, order in ((1, 0), (1, 1), (4, 1), (7, 0), (8, 0), (9, 0), (13, 0), (14, 0), (15, 0)):
        _, pr, _, _, *_ = model.pack(gpu=True, coefficients=False)
        with Session(coeff, Nx, Ny, Nz) as s:
            s.set_variant(variant)
            s.set_column_order(order)  # the separate kernels in launch order: the save blocks and the recomputation too
            if mode == "history":
                s.set_history(box, Nt)
            else:
                s.set_checkpointing(9)
            s.set_sources(pr, [], Nt)
            s.run(Nt)
            s.synchronize()
            out.append(s.gradient(box, list(g)))
    assert np.abs(out[0]).max() > 0
    for k in range(1, len(out)):
        np.testing.assert_array_equal(out[0], out[k])


@pytest.mark.parametrize("variant", [7, 8, 9, 13, 14, 15])
def test_two_step_large_grid_parity(hip_ready, variant):
    """The 400x400x200 workload with its 6 running-DFT planes, 60 steps on the two-step pass (graph
    replay): probes, monitors and fields bit-exact against the oracle."""
    print(f"[case] two-step 400x400x200 variant={variant}", flush=True)
    model = workloads.dipole_slab(Nt=60, monitors=True)
    p_c, m_c, f_c = run_oracle(model, np.float32, n_threads=16)
    p, m, f = _run(model, np.float32, 0, variant, chunks=[43, 17])
    for a, b in zip(p, p_c):
        np.testing.assert_array_equal(a["values"], b["values"])
    model.unpack(p, m, gpu=True)
    g = {k: np.array(v["values"]) for k, v in model.monitors.items()}
    model.unpack(p_c, m_c, gpu=False)
    for k, v in model.monitors.items():
        np.testing.assert_array_equal(g[k], np.array(v["values"]), err_msg=k)
    for k in f_c:
        np.testing.assert_array_equal(f[k], f_c[k], err_msg=k)


@pytest.fixture(scope="module")
def slab_440_oracle():
    """The oracle's probes, DFT planes and fields of the 400x400x200 workload after 440 steps,
    computed once for the variants below."""
    model = workloads.dipole_slab(Nt=440, monitors=True)
    p_c, m_c, f_c = run_oracle(model, np.float32, n_threads=16)
    model.unpack(p_c, m_c, gpu=False)
    return dict(probes=[np.array(p["values"]) for p in p_c],
                monitors={k: np.array(v["values"]) for k, v in model.monitors.items()}, fields=f_c)


@pytest.mark.parametrize("variant", [7, 13])
def test_two_step_large_grid_440_steps(hip_ready, slab_440_oracle, variant):
    """The two variants the tuner picks for the 400x400x200 workload, over 440 steps (graph chunks;
    both run lengths odd, so each ends on a single step), by which time the pulse has reached every
    PML slab and the z-normal DFT planes: probes, the 6 running-DFT planes and every field equal
    to the oracle's bit for bit."""
    print(f"[case] two-step 400x400x200 variant={variant}, 440 steps", flush=True)
    ref = slab_440_oracle
    model = workloads.dipole_slab(Nt=440, monitors=True)
    p, m, f = _run(model, np.float32, 0, variant, chunks=[263, 177])
    for a, b in zip(p, ref["probes"]):
        np.testing.assert_array_equal(a["values"], b)
    model.unpack(p, m, gpu=True)
    for k, v in model.monitors.items():
        got = np.array(v["values"])
        np.testing.assert_array_equal(got, ref["monitors"][k], err_msg=k)
        assert np.abs(got).max() > 0, k
    for k in ref["fields"]:
        np.testing.assert_array_equal(f[k], ref["fields"][k], err_msg=k)


def test_large_sessions_start_on_the_two_step_pass(hip_ready):
    """A float session on the coefficient table that advances at least 1.5 Mi cell-updates per step
    starts on the two-step pass: it holds the second field buffer of the pair, and its probes, DFT
    planes and fields equal those of the separate kernels (variant 1) bit for bit.  Below that size,
    and on the raw coefficient arrays, a session starts on the separate kernels (one buffer)."""
    import copy

    print("[case] starting shape of a 256x256x128 session", flush=True)
    model = workloads.dipole_slab(Nt=30, nx=256, ny=256, nz=128, monitors=True)
    c, probes, monitors, _, Nx, Ny, Nz, Nt = model.pack(gpu=True)
    cells = Nx * Ny * Nz
    assert cells >= 3 << 19
    out = []
    for variant in (None, 1):
        pr, mo = copy.deepcopy(probes), copy.deepcopy(monitors)
        with Session(c, Nx, Ny, Nz) as s:
            if variant is not None:
                s.set_variant(variant)
            s.set_sources(pr, mo, Nt)
            s.run(Nt)
            s.synchronize()
            s.read()
            out.append((pr, mo, s.fields(("ez", "hy")), s.memory_info()["device_bytes"]))
    (p0, m0, f0, b0), (p1, m1, f1, b1) = out
    assert b0 >= b1 + 18 * cells * 4
    for a, b in zip(p0, p1):
        np.testing.assert_array_equal(a["values"], b["values"])
    for a, b in zip(m0, m1):
        np.testing.assert_array_equal(a["values"], b["values"])
    for k in f0:
        np.testing.assert_array_equal(f0[k], f1[k], err_msg=k)
    assert np.abs(f0["ez"]).max() > 0
    held = []
    for flags, variant in ((_lib.COEF_RAW, None), (_lib.COEF_RAW, 1)):
        with Session(c, Nx, Ny, Nz, flags=flags) as s:
            if variant is not None:
                s.set_variant(variant)
            s.set_sources(copy.deepcopy(probes), [], Nt)
            s.run(2)
            s.synchronize()
            held.append(s.memory_info()["device_bytes"])
    assert held[0] == held[1]
    print("[case] starting shape of a 120x100x80 session", flush=True)
    small = workloads.dipole_slab(Nt=4, nx=120, ny=100, nz=80, monitors=False)
    c, probes, _, _, Nx, Ny, Nz, Nt = small.pack(gpu=True)
    held = []
    for variant in (None, 1):
        with Session(c, Nx, Ny, Nz) as s:
            if variant is not None:
                s.set_variant(variant)
            s.set_sources(copy.deepcopy(probes), [], Nt)
            s.run(2)
            s.synchronize()
            held.append(s.memory_info()["device_bytes"])
    assert held[0] == held[1]
