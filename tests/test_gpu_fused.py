"""The fused E+H pass (variants 2 and 3, and 5 and 6 with deeper load pipelining): one launch per
step that reads one field buffer and writes the other, with y/z neighbours through LDS and halo
cells recomputed.  Its results must equal the
oracle's and the separate kernels' bit for bit, whatever the tile, x planes, chunking and buffer
parity."""
import numpy as np
import pytest

from rfnetwork_amd import _lib, workloads
from rfnetwork_amd.batch import Batch
from rfnetwork_amd.session import Session

from helpers import run_oracle

pytestmark = pytest.mark.gpu


def _run(model, dtype=np.float32, flags=0, variant=2, chunks=None, launch=None, switch=None):
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = model.pack(gpu=True, dtype=dtype)
    with Session(c, Nx, Ny, Nz, dtype=dtype, flags=flags) as s:
        s.set_variant(variant)
        if launch:
            s.set_launch(*launch)
        s.set_sources(probes, monitors, Nt)
        for k, n in enumerate(chunks or [Nt]):
            if switch:
                s.set_variant(switch[k % len(switch)])
            s.run(n)
        s.synchronize()
        s.read()
        f = s.fields()
    return probes, monitors, f


def _assert_equal_to_oracle(model, dtype, probes, monitors, f):
    p_c, m_c, f_c = run_oracle(model, dtype, n_threads=3)
    for a, b in zip(probes, p_c):
        np.testing.assert_array_equal(a["values"], b["values"])
    model.unpack(probes, monitors, gpu=True)
    g = {k: np.array(m["values"]) for k, m in model.monitors.items()}
    model.unpack(p_c, m_c, gpu=False)
    for k, m in model.monitors.items():
        np.testing.assert_array_equal(g[k], np.array(m["values"]), err_msg=k)
    for k in f_c:
        np.testing.assert_array_equal(f[k], f_c[k], err_msg=k)


@pytest.mark.parametrize("variant", [2, 3, 5, 6])
@pytest.mark.parametrize("flags", [0, _lib.COEF_RAW, _lib.NO_GRAPHS])
def test_fused_tiny_fp32_bit_exact(hip_ready, variant, flags):
    """Every code path of the small model (PML, dielectric, conductor, two ports, a voltage probe,
    time and phasor monitors on every axis); odd chunk lengths leave the pair on either buffer."""
    print(f"[case] fused tiny variant={variant} flags={flags}", flush=True)
    model = workloads.tiny_model(Nt=60)
    p, m, f = _run(model, np.float32, flags, variant, chunks=[37, 23], launch=(0, 3, 10))
    _assert_equal_to_oracle(model, np.float32, p, m, f)


def test_fused_tiny_fp64_bit_exact(hip_ready):
    print("[case] fused tiny fp64", flush=True)
    model = workloads.tiny_model(Nt=50)
    p, m, f = _run(model, np.float64, 0, 6, chunks=[21, 29], launch=(0, 4, 6))
    _assert_equal_to_oracle(model, np.float64, p, m, f)


def test_fused_and_separate_kernels_interleave(hip_ready):
    """Switching between the fused pass and the separate kernels mid-run (which rebuilds the total
    arrays the fused pass leaves unwritten) changes nothing."""
    print("[case] fused/separate interleave", flush=True)
    model = workloads.tiny_model(Nt=60)
    p, m, f = _run(model, np.float32, 0, 2, chunks=[13, 20, 7, 11, 9], switch=[2, 1, 3, 0, 5, 6])
    _assert_equal_to_oracle(model, np.float32, p, m, f)


def test_fused_many_tiles_and_planes(hip_ready):
    """A grid with several tiles along y and z (rows 6 and 14, 62 cells per tile along z) and x
    chunks of every size: identical to the separate kernels."""
    print("[case] fused tiles", flush=True)
    model = workloads.tiny_model(Nt=30, nx=23, ny=31, nz=131)
    _, _, ref = _run(model, variant=1)
    for variant in (2, 3, 5, 6):
        for xp in (1, 2, 5, 23):
            _, _, f = _run(model, variant=variant, launch=(0, xp, 8))
            for k in ref:
                np.testing.assert_array_equal(f[k], ref[k], err_msg=f"{k} variant {variant} x planes {xp}")


def test_fused_batch_equals_single_runs(hip_ready):
    print("[case] fused batch", flush=True)
    models = [workloads.tiny_model(Nt=40, nx=14, ny=11, nz=9, seed=7),
              workloads.tiny_model(Nt=40, nx=19, ny=20, nz=70, seed=8)]
    grids, inputs = [], []
    for m in models:
        coeff = m.coefficient_arrays(gpu=True)
        grids.append((coeff, m.Nx, m.Ny, m.Nz, 1))
        _, probes, monitors, _, *_ = m.pack(gpu=True, coefficients=False)
        inputs.append((probes, monitors))
    with Batch(grids) as b:
        b.set_variant(3)
        for k, (probes, monitors) in enumerate(inputs):
            b.set_sources(k, probes, monitors, 40)
        b.run(25)
        b.run(15)
        b.synchronize()
        b.read()
        fields = [b.fields(k) for k in range(2)]
    for m, (probes, _), f in zip(models, inputs, fields):
        p_c, _, f_c = run_oracle(m, np.float32, n_threads=2)
        for a, c in zip(probes, p_c):
            np.testing.assert_array_equal(a["values"], c["values"])
        for k in f_c:
            np.testing.assert_array_equal(f[k], f_c[k], err_msg=k)


def test_fused_forward_gradient_equals_separate(hip_ready):
    """Checkpoints taken during fused steps (from whichever buffer is current) give the same
    adjoint gradient as a forward run with the separate kernels."""
    print("[case] fused gradient", flush=True)
    Nt = 40
    model = workloads.tiny_model(Nt=Nt, time_monitors=False, phasor_monitors=False)
    coeff, probes, _, _, Nx, Ny, Nz, _ = model.pack(gpu=True)
    g = np.random.default_rng(3).standard_normal((len(probes), Nt)).astype(np.float32)
    out = []
    for variant in (1, 2, 6):
        _, pr, _, _, *_ = model.pack(gpu=True, coefficients=False)
        with Session(coeff, Nx, Ny, Nz) as s:
            s.set_variant(variant)
            s.set_checkpointing(9)
            s.set_sources(pr, [], Nt)
            s.run(Nt)
            s.synchronize()
            out.append(s.gradient((0, Nx, 0, Ny, 0, Nz), list(g)))
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[0], out[2])


def test_fused_large_grid_parity(hip_ready):
    """The 400x400x200 workload, 12 fused steps: bit-exact against the oracle."""
    print("[case] fused 400x400x200", flush=True)
    model = workloads.dipole_slab(Nt=12, monitors=True)
    p_c, m_c, f_c = run_oracle(model, np.float32, n_threads=16)
    p, m, f = _run(model, np.float32, 0, 6)
    for a, b in zip(p, p_c):
        np.testing.assert_array_equal(a["values"], b["values"])
    for k in f_c:
        np.testing.assert_array_equal(f[k], f_c[k], err_msg=k)
