"""The engine's adjoint (rfdtd_gradient) against an unrolled torch autograd of the same time loop,
and the eps_r gradient against central finite differences of the float64 oracle."""
import numpy as np
import pytest
import torch

from oracle import cpu_solver
from rfnetwork_amd import _lib, adjoint, workloads
from rfnetwork_amd.session import Session

import yee_torch
from helpers import rel_l2

pytestmark = pytest.mark.gpu

SPLITS = ("Cb_ex_y", "Cb_ex_z", "Cb_ey_z", "Cb_ey_x", "Cb_ez_x", "Cb_ez_y")


def _autograd_reference(model, g, Nt):
    c, probes, _, _, Nx, Ny, Nz, _ = model.pack(gpu=False, dtype=np.float64, n_threads=1)
    series, leaf = yee_torch.run(c, probes, Nx, Ny, Nz, Nt)
    torch.sum(series * torch.tensor(g)).backward()
    return {k: leaf[k].grad.numpy() for k in SPLITS}


def _engine_gradient(model, g, Nt, dtype, every):
    coeff, probes, _, _, Nx, Ny, Nz, _ = model.pack(gpu=True, dtype=dtype)
    with Session(coeff, Nx, Ny, Nz, dtype=dtype) as s:
        s.set_checkpointing(every)
        s.set_sources(probes, [], Nt)
        s.run(Nt)
        s.synchronize()
        return s.gradient((0, Nx, 0, Ny, 0, Nz), [row.astype(dtype) for row in g])


def _cpu_view(grad, k, Nx, Ny, Nz):
    """The part of the packed-layout gradient that the CPU-layout coefficient k covers."""
    comp = SPLITS[k][3:5]
    ny = Ny - 1 if comp in ("ex", "ez") else Ny
    nz = Nz - 1 if comp in ("ex", "ey") else Nz
    return grad[k][:, :ny, :nz]


@pytest.mark.parametrize("every", [7, 40])
def test_adjoint_fp64_matches_unrolled_autograd(hip_ready, every):
    """float64: the hand-written reverse pass (checkpoints every `every` steps, recomputation between
    them) against torch autograd through the unrolled loop, for a random dJ/dy on every probe."""
    print(f"[case] adjoint vs autograd fp64 every={every}", flush=True)
    Nt = 40
    model = workloads.tiny_model(Nt=Nt, time_monitors=False, phasor_monitors=False)
    Nx, Ny, Nz = model.n_cells
    n_probes = len(model.pack(gpu=True, coefficients=False)[1])
    g = np.random.default_rng(17).standard_normal((n_probes, Nt))
    ref = _autograd_reference(model, g, Nt)
    grad = _engine_gradient(model, g, Nt, np.float64, every)
    for k, name in enumerate(SPLITS):
        got = _cpu_view(grad, k, Nx, Ny, Nz)
        r = ref[name]
        if name[3:5] in ("ey", "ez"):  # the x = Nx-1 plane is the zeroed PEC wall
            got, r = got[:-1], r[:-1]
        scale = np.abs(r).max()
        assert scale > 0, name
        np.testing.assert_allclose(got, r, rtol=1e-9, atol=1e-12 * scale, err_msg=name)
    # entries of the packed layout that hold no updated component carry no gradient
    assert not np.any(grad[0][:, Ny - 1, :]) and not np.any(grad[0][:, :, Nz - 1])
    assert not np.any(grad[4][:, Ny - 1, :])


def test_adjoint_fp32_error_against_fp64(hip_ready):
    """float32 reverse pass: its gradient against the float64 autograd reference (stated)."""
    print("[case] adjoint fp32 vs fp64", flush=True)
    Nt = 40
    model = workloads.tiny_model(Nt=Nt, time_monitors=False, phasor_monitors=False)
    Nx, Ny, Nz = model.n_cells
    n_probes = len(model.pack(gpu=True, coefficients=False)[1])
    g = np.random.default_rng(18).standard_normal((n_probes, Nt))
    ref = _autograd_reference(model, g, Nt)
    grad = _engine_gradient(model, g, Nt, np.float32, 10)
    err = max(rel_l2(_cpu_view(grad, k, Nx, Ny, Nz)[:-1], ref[n][:-1]) for k, n in enumerate(SPLITS))
    print(f"[adjoint] fp32 vs fp64 gradient rel L2 {err:.3e}", flush=True)
    assert err < 1e-4


SMALL_F = (16e9, 20e9, 24e9)


def _small_design(eps, coef_dtype=np.float64, Nt=900, coefficients="host"):
    """The adjoint workload shrunk to 48x32x20 cells with an 8x8x6 design box and a 20 GHz pulse."""
    return workloads.design_strip(eps, Nt=Nt, nx=48, ny=32, nz=20, design=(8, 8, 6), n_pml=5, h=3,
                                  coef_dtype=coef_dtype, pulse=(20e9, 80e-12, 60e-12), frequency=SMALL_F,
                                  coefficients=coefficients)


def _fd_oracle(eps, cell, h, frequency):
    Js = []
    for sgn in (1, -1):
        e = eps.copy()
        e[cell] += sgn * h
        m = _small_design(e)
        Js.append(adjoint.objective_only(m, frequency, dtype=np.float64, backend=cpu_solver.solver_run)[0])
    return (Js[0] - Js[1]) / (2 * h)


def test_eps_gradient_matches_oracle_finite_differences(hip_ready):
    """dJ/deps_r, J = sum_f |S21|^2 at 16, 20, 24 GHz, from the float64 engine adjoint against central
    differences (Delta eps_r = 1e-3) of the float64 oracle, on cells across the design box."""
    print("[case] eps gradient vs oracle FD", flush=True)
    rng = np.random.default_rng(4)
    eps = rng.uniform(1.0, 4.0, size=(8, 8, 6))
    cbox = workloads.design_box(48, 32, (8, 8, 6))
    model = _small_design(eps)
    J, grad, _ = adjoint.eps_r_gradient(model, cbox, model.frequency, checkpoint_every=64, dtype=np.float64)
    J_o = adjoint.objective_only(_small_design(eps), model.frequency, dtype=np.float64,
                                 backend=cpu_solver.solver_run)[0]
    assert J == pytest.approx(J_o, rel=1e-12)
    cells = [(0, 0, 0), (4, 4, 2), (7, 3, 5), (2, 6, 1), (5, 1, 3)]
    fd = np.array([_fd_oracle(eps, c, 1e-3, model.frequency) for c in cells])
    ad = np.array([grad[c] for c in cells])
    err = np.abs(ad - fd) / np.abs(fd)
    print(f"[adjoint] J {J:.6e} grad {ad} fd {fd} rel err {err}", flush=True)
    assert np.all(err <= 1e-3)
    # float32 fields and coefficients: the same gradient, error stated
    m32 = _small_design(eps, coef_dtype=np.float32)
    J32, g32, _ = adjoint.eps_r_gradient(m32, cbox, m32.frequency, checkpoint_every=64, dtype=np.float32)
    e32 = rel_l2(g32, grad)
    print(f"[adjoint] fp32 gradient vs fp64 rel L2 {e32:.3e}, J {J32:.6e}", flush=True)
    assert e32 < 1e-2


def test_autograd_function_backward(hip_ready):
    """EpsRObjective.backward hands torch the adjoint gradient of the design cells."""
    print("[case] autograd function", flush=True)
    eps = torch.tensor(np.random.default_rng(4).uniform(1.0, 4.0, size=(8, 8, 6)), requires_grad=True)
    cbox = workloads.design_box(48, 32, (8, 8, 6))
    f = np.array(SMALL_F)
    J = adjoint.EpsRObjective.apply(eps, lambda e: _small_design(e, coef_dtype=np.float32), cbox, f, 100)
    (2.0 * J).backward()
    J_ref, g_ref, _ = adjoint.eps_r_gradient(_small_design(eps.detach().numpy(), coef_dtype=np.float32), cbox, f,
                                             checkpoint_every=100)
    assert float(J) == pytest.approx(J_ref, rel=1e-6)
    np.testing.assert_allclose(eps.grad.numpy(), 2.0 * g_ref, rtol=1e-6, atol=1e-9 * np.abs(g_ref).max())


def test_autograd_function_on_a_design_loop(hip_ready):
    """EpsRObjective on a resident DesignLoop: forward runs the forward solve only (no reverse pass
    until backward), a backward after a later forward is refused, and the gradient torch receives
    equals eps_r_gradient on a model rebuilt with that permittivity, bit for bit."""
    print("[case] autograd on design loop", flush=True)
    from rfnetwork_amd.design import DesignLoop

    rng = np.random.default_rng(11)
    cbox = workloads.design_box(48, 32, (8, 8, 6))
    f = np.array(SMALL_F)
    e1, e2 = rng.uniform(1.0, 4.0, size=(2, 8, 8, 6))
    base = _small_design(e1, coef_dtype=np.float32)
    with DesignLoop(base, cbox, f, mode="checkpoint", checkpoint_every=100) as loop:
        t1 = torch.tensor(e1, requires_grad=True)
        J1 = adjoint.EpsRObjective.apply(t1, loop)
        assert loop.timing["adjoint_ms"] == 0.0  # no reverse pass yet
        t2 = torch.tensor(e2, requires_grad=True)
        J2 = adjoint.EpsRObjective.apply(t2, loop)
        with pytest.raises(RuntimeError, match="must follow its own forward"):
            J1.backward()
        J2.backward()
        assert loop.timing["adjoint_ms"] > 0.0
    J_ref, g_ref, _ = adjoint.eps_r_gradient(_small_design(e2, coef_dtype=np.float32), cbox, f, checkpoint_every=100)
    assert float(J2.detach()) == J_ref
    np.testing.assert_array_equal(t2.grad.numpy(), g_ref)
    assert t1.grad is None


def test_batched_gradients_equal_single_sessions(hip_ready):
    """Reverse passes of uneven grids sharing launches give each instance exactly what a single-grid
    session gives; the checkpoint planner's interval fits and is used."""
    print("[case] batched gradients", flush=True)
    from rfnetwork_amd.batch import Batch

    Nt = 40
    models = [workloads.tiny_model(Nt=Nt, time_monitors=False, phasor_monitors=False, seed=7),
              workloads.tiny_model(Nt=Nt, nx=19, ny=12, nz=10, time_monitors=False, phasor_monitors=False, seed=8)]
    rng = np.random.default_rng(9)
    inputs = []
    for m in models:
        coeff, probes, _, _, Nx, Ny, Nz, _ = m.pack(gpu=True)
        inputs.append((coeff, probes, (Nx, Ny, Nz), rng.standard_normal((len(probes), Nt)).astype(np.float32)))
    boxes = [(0, 14, 0, 11, 0, 9), (3, 15, 2, 9, 1, 8)]
    with Batch([(c, *dims, 1) for c, _, dims, _ in inputs]) as b:
        every, need = b.plan_checkpoints(Nt, [(0, boxes[0]), (1, boxes[1])])
        assert 1 <= every <= Nt and need > 0
        with pytest.raises(Exception, match="available"):
            b.plan_checkpoints(Nt, [(0, boxes[0])], budget_bytes=1024)
        b.set_checkpointing(every)
        for k, (_, probes, _, _) in enumerate(inputs):
            b.set_sources(k, probes, [], Nt)
        b.run(Nt)
        b.synchronize()
        got = b.gradients([(k, boxes[k], inputs[k][3]) for k in range(2)])
    for k, (coeff, _, dims, g) in enumerate(inputs):
        _, probes, _, _, *_ = models[k].pack(gpu=True, coefficients=False)
        with Session(coeff, *dims) as s:
            s.set_checkpointing(every)
            s.set_sources(probes, [], Nt)
            s.run(Nt)
            s.synchronize()
            ref = s.gradient(boxes[k], list(g))
        np.testing.assert_array_equal(got[k], ref)
        assert np.abs(ref).max() > 0


def test_eps_gradients_batch_equal_single(hip_ready):
    """Two design models of different size through eps_r_gradients_batch: J and dJ/deps_r equal the
    single-model eps_r_gradient results."""
    print("[case] eps gradients batch", flush=True)
    eps_a = np.random.default_rng(4).uniform(1.0, 4.0, size=(8, 8, 6))
    eps_b = np.random.default_rng(5).uniform(1.0, 4.0, size=(6, 10, 5))
    ma = _small_design(eps_a, coef_dtype=np.float32, Nt=300)
    mb = workloads.design_strip(eps_b, Nt=300, nx=52, ny=36, nz=18, design=(6, 10, 5), n_pml=5, h=3,
                                pulse=(20e9, 80e-12, 60e-12), frequency=SMALL_F)
    ca, cb = workloads.design_box(48, 32, (8, 8, 6)), workloads.design_box(52, 36, (6, 10, 5))
    for mode in ("checkpoint", "history"):
        res, info = adjoint.eps_r_gradients_batch([(ma, ca, ma.frequency), (mb, cb, mb.frequency)],
                                                  checkpoint_every=70, return_info=True, mode=mode)
        for (J, g, _), (m, cbox) in zip(res, ((ma, ca), (mb, cb))):
            J1, g1, _ = adjoint.eps_r_gradient(m, cbox, m.frequency, checkpoint_every=70, mode="checkpoint")
            assert J == J1
            np.testing.assert_array_equal(g, g1)
        assert info["mode"] == mode
        if mode == "checkpoint":
            assert info["checkpoint_every"] == 70


def test_eps_gradients_batch_device_assembled_models(hip_ready):
    """The batched gradient of models meshed without host coefficient arrays (coefficients="device":
    the batch assembles them on the GPU, and the chain rule reads the design box's values back)
    equals that of the same models assembled on the host, J and dJ/deps_r bit for bit."""
    print("[case] eps gradients batch, device-assembled models", flush=True)
    eps_a = np.random.default_rng(4).uniform(1.0, 4.0, size=(8, 8, 6))
    eps_b = np.random.default_rng(5).uniform(1.0, 4.0, size=(6, 10, 5))
    kw_b = dict(Nt=300, nx=52, ny=36, nz=18, design=(6, 10, 5), n_pml=5, h=3, pulse=(20e9, 80e-12, 60e-12),
                frequency=SMALL_F)
    ca, cb = workloads.design_box(48, 32, (8, 8, 6)), workloads.design_box(52, 36, (6, 10, 5))
    out = []
    for coefficients in ("host", "device"):
        ma = _small_design(eps_a, coef_dtype=np.float32, Nt=300, coefficients=coefficients)
        mb = workloads.design_strip(eps_b, coefficients=coefficients, **kw_b)
        out.append(adjoint.eps_r_gradients_batch([(ma, ca, ma.frequency), (mb, cb, mb.frequency)],
                                                 checkpoint_every=70, mode="checkpoint"))
    for (Jh, gh, _), (Jd, gd, _) in zip(*out):
        assert Jh == Jd
        np.testing.assert_array_equal(gh, gd)
        assert np.abs(gh).max() > 0


def _session_gradient(model, g, box, variant, history, chunks, flags=0):
    coeff, probes, _, _, Nx, Ny, Nz, Nt = model.pack(gpu=True)
    with Session(coeff, Nx, Ny, Nz, flags=flags) as s:
        s.set_variant(variant)
        if history:
            s.set_history(box, Nt)
        else:
            s.set_checkpointing(9)
        s.set_sources(probes, [], Nt)
        for n in chunks:
            s.run(n)
        s.synchronize()
        return s.gradient(box, list(g))


@pytest.mark.parametrize("variant", [1, 2])
@pytest.mark.parametrize("flags", [0, _lib.NO_GRAPHS])
def test_history_gradient_equals_checkpointed(hip_ready, variant, flags):
    """Recording the design box's H at every forward step (inside captured chunks or step by step,
    with the separate kernels or the fused pass) gives the checkpointed gradient bit for bit."""
    print(f"[case] history gradient variant={variant} flags={flags}", flush=True)
    Nt = 60
    model = workloads.tiny_model(Nt=Nt, time_monitors=False, phasor_monitors=False)
    _, probes, _, _, Nx, Ny, Nz, _ = model.pack(gpu=True, coefficients=False)
    g = np.random.default_rng(3).standard_normal((len(probes), Nt)).astype(np.float32)
    box = (2, Nx - 3, 1, Ny - 2, 3, Nz - 1)
    ref = _session_gradient(model, g, box, 1, False, [Nt], flags)
    got = _session_gradient(model, g, box, variant, True, [37, 23], flags)
    np.testing.assert_array_equal(got, ref)
    assert np.abs(ref).max() > 0


def test_history_requests_checked(hip_ready):
    """The history is set before the first step, a gradient over another box needs checkpoints, and
    a run past the recorded length is refused."""
    print("[case] history errors", flush=True)
    Nt = 20
    model = workloads.tiny_model(Nt=Nt, time_monitors=False, phasor_monitors=False)
    coeff, probes, _, _, Nx, Ny, Nz, _ = model.pack(gpu=True)
    with Session(coeff, Nx, Ny, Nz) as s:
        with pytest.raises(Exception, match="outside"):
            s.set_history((0, Nx + 1, 0, Ny, 0, Nz), Nt)
        s.set_history((0, 4, 0, 4, 0, 4), Nt - 5)
        s.set_sources(probes, [], Nt)
        s.run(Nt - 5)
        with pytest.raises(Exception, match="history"):
            s.run(1)
        with pytest.raises(Exception, match="first step"):
            s.set_history((0, 4, 0, 4, 0, 4), Nt)
        s.synchronize()
        zero = [np.zeros(Nt, np.float32) for _ in probes]
        with pytest.raises(Exception, match="checkpointing"):
            s.gradient((0, 5, 0, 4, 0, 4), zero)
        assert s.gradient((0, 4, 0, 4, 0, 4), zero).shape == (6, 4, 4, 4)


def test_history_batch_gradients(hip_ready):
    """Histories of two uneven grids in one batch: each gradient equals its checkpointed single-grid
    gradient."""
    print("[case] history batch", flush=True)
    from rfnetwork_amd.batch import Batch

    Nt = 40
    models = [workloads.tiny_model(Nt=Nt, time_monitors=False, phasor_monitors=False, seed=7),
              workloads.tiny_model(Nt=Nt, nx=19, ny=12, nz=10, time_monitors=False, phasor_monitors=False, seed=8)]
    rng = np.random.default_rng(9)
    inputs = []
    for m in models:
        coeff, probes, _, _, Nx, Ny, Nz, _ = m.pack(gpu=True)
        inputs.append((coeff, probes, (Nx, Ny, Nz), rng.standard_normal((len(probes), Nt)).astype(np.float32)))
    boxes = [(0, 14, 0, 11, 0, 9), (3, 15, 2, 9, 1, 8)]
    with Batch([(c, *dims, 1) for c, _, dims, _ in inputs]) as b:
        b.set_history([(0, boxes[0]), (1, boxes[1])], Nt)
        for k, (_, probes, _, _) in enumerate(inputs):
            b.set_sources(k, probes, [], Nt)
        b.run(Nt)
        b.synchronize()
        got = b.gradients([(k, boxes[k], inputs[k][3]) for k in range(2)])
    for k, (_, _, _, g) in enumerate(inputs):
        ref = _session_gradient(models[k], g, boxes[k], 1, False, [Nt])
        np.testing.assert_array_equal(got[k], ref)


def test_adjoint_launch_shapes_give_identical_gradients(hip_ready):
    """The reverse pass's launch shapes (pipelined or plain columns, XCD ranges or launch order,
    workgroup sizes, x chunks down to one plane per workgroup) and the shape its timing picks all give
    the same gradient bits."""
    print("[case] adjoint launch shapes", flush=True)
    Nt = 40
    model = workloads.tiny_model(Nt=Nt, nx=17, ny=13, nz=70, time_monitors=False, phasor_monitors=False)
    coeff, probes, _, _, Nx, Ny, Nz, _ = model.pack(gpu=True)
    g = np.random.default_rng(11).standard_normal((len(probes), Nt)).astype(np.float32)
    box = (1, Nx - 1, 2, Ny, 0, Nz - 3)
    with Session(coeff, Nx, Ny, Nz) as s:
        s.set_history(box, Nt)
        s.set_sources(probes, [], Nt)
        s.run(Nt)
        s.synchronize()
        s.set_adjoint_launch(0, 256, 0)
        ref = s.gradient(box, list(g))
        for v, blk, ch in [(1, 256, 0), (1, 128, 1), (1, 512, 3), (0, 64, 7), (1, 64, Nx), (1, 1024, 2),
                           (3, 256, Nx), (2, 128, 1), (3, 512, 3)]:
            s.set_adjoint_launch(v, blk, ch)
            assert s.adjoint_launch() == dict(variant=v, block=blk, x_chunks=ch)
            np.testing.assert_array_equal(s.gradient(box, list(g)), ref, err_msg=f"{v} {blk} {ch}")
        s.set_adjoint_launch(1, 0, 0)
        np.testing.assert_array_equal(s.gradient(box, list(g)), ref)
        assert s.adjoint_launch()["block"] in (128, 256, 512)
        with pytest.raises(Exception, match="block_threads"):
            s.set_adjoint_launch(1, 100, 0)
        with pytest.raises(Exception, match="adjoint variant"):
            s.set_adjoint_launch(4, 256, 0)
    assert np.abs(ref).max() > 0
