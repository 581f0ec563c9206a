"""The HIP engine at the benchmark configurations' own sizes (fewer steps where a full run is long):

* the batched forward workload: 4 uneven port variants (64^3 .. 192x128x96), every port driven
  (12 excitations), all in one batched session -- the 64^3 variant's excitations against the CPU
  oracle bit for bit, every instance against its own single-grid run;
* the adjoint workload: dJ/deps_r over the 64x64x32 design box of the 200x160x80 model, the history
  and checkpointed reverse passes equal bit for bit, and the float64 gradient against central
  finite differences of float64 forward solves;
* the batched forward+adjoint workload: the 8 design variants (96^3 .. 512x512x256) in one batched
  gradient, each of the small ones equal to its single-variant gradient, all finite.

A multi-excitation run is the reference's own scenario for coupled lines (even and odd solves with
reset_excitations between them, rfnetwork tests/test_solver/test_coupled_stripline.py:96-109)."""
import numpy as np
import pytest

from rfnetwork_amd import adjoint, batch as rbatch, workloads
from rfnetwork_amd.batch import Batch
from rfnetwork_amd.session import Session

from helpers import run_oracle

pytestmark = pytest.mark.gpu


def test_port_variants_batched_12_excitations(hip_ready):
    Nt = 300
    print(f"[case] port variants x 12 excitations, {Nt} steps", flush=True)
    models = [workloads.port_variant(v, Nt=Nt) for v in range(len(workloads.PORT_VARIANTS))]
    grids, inst = [], []
    for v, m in enumerate(models):
        ports = list(range(1, len(m.ports) + 1))
        exc, _ = rbatch.pack_excitations(m, ports)
        coeff = m.coefficient_arrays(gpu=True)
        grids.append((coeff, m.Nx, m.Ny, m.Nz, len(ports)))
        for port, probes, monitors in exc:
            inst.append((v, port, coeff, probes, monitors))
    assert len(inst) == 12
    with Batch(grids) as b:
        for k, (_, _, _, probes, monitors) in enumerate(inst):
            b.set_sources(k, probes, monitors, Nt)
        b.autotune()
        b.run(171)
        b.run(Nt - 171)
        b.synchronize()
        b.read()
        fields = [b.fields(k, names=("ez", "hx", "hy")) for k in range(len(inst))]
    for k, (v, port, coeff, probes, monitors) in enumerate(inst):
        m = models[v]
        (_, src, mon), = rbatch.pack_excitations(m, [port])[0]
        with Session(coeff, m.Nx, m.Ny, m.Nz) as s:
            s.set_sources(src, mon, Nt)
            s.run(Nt)
            s.synchronize()
            s.read()
            f = s.fields(("ez", "hx", "hy"))
        for a, c in zip(probes, src):
            np.testing.assert_array_equal(a["values"], c["values"], err_msg=f"instance {k}")
        for name in f:
            np.testing.assert_array_equal(fields[k][name], f[name], err_msg=f"instance {k} {name}")
        assert np.abs(f["ez"]).max() > 0
    # the 64^3 variant's two excitations on the CPU oracle
    m0 = models[0]
    wave = rbatch._waveform(m0)
    for k in (0, 1):
        ref = workloads.port_variant(0, Nt=Nt)
        ref.reset_excitations()
        ref.assign_excitation(wave, k + 1)
        p_c, _, f_c = run_oracle(ref, np.float32, n_threads=8)
        for a, c in zip(inst[k][3], p_c):
            np.testing.assert_array_equal(a["values"], c["values"], err_msg=f"oracle, port {k + 1}")
        for name in ("ez", "hx", "hy"):
            np.testing.assert_array_equal(fields[k][name], f_c[name], err_msg=f"oracle, port {k + 1} {name}")


def test_design_strip_gradient_full_size(hip_ready):
    """200x160x80 model, 64x64x32 design box: float32 history and checkpointed gradients equal bit
    for bit; the float64 gradient against central differences (Delta eps_r = 1e-3) of float64
    forward solves on 3 cells drawn with seed 41, within 1e-3 relative."""
    Nt = 2400
    print(f"[case] design_strip 200x160x80 gradient, {Nt} steps", flush=True)
    eps0 = np.random.default_rng(4).uniform(1.0, 4.0, size=(64, 64, 32))
    cbox = workloads.design_box()
    m32 = workloads.design_strip(eps0, Nt=Nt)
    J_h, g_h, _ = adjoint.eps_r_gradient(m32, cbox, m32.frequency, mode="history")
    J_c, g_c, _ = adjoint.eps_r_gradient(m32, cbox, m32.frequency, mode="checkpoint", checkpoint_every=200)
    assert J_h == J_c and J_h > 0
    np.testing.assert_array_equal(g_h, g_c)
    assert g_h.shape == (64, 64, 32) and np.isfinite(g_h).all() and np.abs(g_h).max() > 0

    build = lambda e: workloads.design_strip(e, Nt=Nt, coef_dtype=np.float64)
    m64 = build(eps0)
    J64, g64, _ = adjoint.eps_r_gradient(m64, cbox, m64.frequency, dtype=np.float64, mode="checkpoint",
                                         checkpoint_every=200)
    rng = np.random.default_rng(41)
    cells = [tuple(int(rng.integers(0, n)) for n in (64, 64, 32)) for _ in range(3)]
    h = 1e-3
    for c in cells:
        Js = []
        for sgn in (1, -1):
            e = eps0.copy()
            e[c] += sgn * h
            Js.append(adjoint.objective_only(build(e), m64.frequency, dtype=np.float64)[0])
        fd = (Js[0] - Js[1]) / (2 * h)
        err = abs(g64[c] - fd) / abs(fd)
        print(f"[adjoint] cell {c}: adjoint {g64[c]:.9e} fd {fd:.9e} rel err {err:.2e}", flush=True)
        assert err <= 1e-3
    # the float32 gradient tracks the float64 one
    assert np.linalg.norm(g_h - g64) / np.linalg.norm(g64) < 1e-3


def test_design_variants_batched_gradient(hip_ready):
    """The 8 design variants (142 M cells) in one batched forward+adjoint with the checkpoint
    interval planned to fit the device; the three smallest equal their single-variant gradients."""
    Nt = 300
    print(f"[case] 8 design variants batched gradient, {Nt} steps", flush=True)
    jobs = []
    for v in range(len(workloads.DESIGN_VARIANTS)):
        m = workloads.design_variant(v, Nt=Nt)
        jobs.append((m, workloads.design_variant_box(v), m.frequency))
    res, info = adjoint.eps_r_gradients_batch(jobs, return_info=True, mode="checkpoint")
    print(f"[c5] {info}", flush=True)
    assert info["checkpoint_every"] >= 1
    for v, ((J, g, _), (m, cbox, _)) in enumerate(zip(res, jobs)):
        b = cbox[1] - cbox[0]
        assert g.shape == (b, b, b) and np.isfinite(g).all() and np.isfinite(J)
        # in 300 steps the pulse reaches port 2 of the small variants only; where it has not, the
        # port voltage, S21, J and so the gradient are exactly zero
        assert (np.abs(g).max() > 0) == (J > 0), f"variant {v}: J {J}"
        if v < 3:
            assert J > 0, f"variant {v}"
    for v in range(3):
        m = workloads.design_variant(v, Nt=Nt)
        J, g, _ = adjoint.eps_r_gradient(m, workloads.design_variant_box(v), m.frequency, mode="history")
        assert J == res[v][0]
        np.testing.assert_array_equal(g, res[v][1], err_msg=f"variant {v}")
