"""Host side of the adjoint: the S-parameter objective, its derivative with respect to the probe
series, the chain rule from the packed Cb coefficients to cell permittivities, and the unrolled
torch stepper that serves as the adjoint kernels' reference.  All on the CPU."""
import copy

import numpy as np
import pytest
import torch

from oracle import cpu_solver
from rfnetwork_amd import adjoint, workloads
from rfnetwork_amd.geometry import Box, Rectangle
from rfnetwork_amd.solver import FDTD_Solver

import yee_torch
from helpers import run_oracle

IN_PER_M = 39.3701


def _small_strip(Nt=300):
    return workloads.microstrip(Nt=Nt, nx=40, ny=20, nz=10, n_pml=4, n_freq=7)


def test_objective_matches_get_sparameters():
    """S_21 from the packed series equals FDTD_Solver.get_sparameters after unpack
    (rfnetwork/solver.py:2442-2532 restated in rfnetwork_amd/solver.py)."""
    model = _small_strip()
    probes, monitors, _ = run_oracle(model, np.float64, n_threads=2, fields=False)
    series = np.array([p["values"] for p in probes])
    model.unpack(probes, monitors, gpu=False)
    S_ref = np.asarray(model.get_sparameters(model.frequency, source_port=1))[:, :, 0]
    J, S, _ = adjoint.sparam_objective(model, series, model.frequency, b_port=2, a_port=1)
    np.testing.assert_allclose(S, S_ref[:, 1], rtol=1e-12, atol=0)
    J11, S11, _ = adjoint.sparam_objective(model, series, model.frequency, b_port=1, a_port=1)
    np.testing.assert_allclose(S11, S_ref[:, 0], rtol=1e-12, atol=0)
    assert J == pytest.approx(np.sum(np.abs(S_ref[:, 1]) ** 2), rel=1e-12)


@pytest.mark.parametrize("b,a", [(2, 1), (1, 1)])
def test_objective_derivative_matches_central_differences(b, a):
    model = _small_strip()
    probes, _, _ = run_oracle(model, np.float64, n_threads=2, fields=False)
    series = np.array([p["values"] for p in probes])
    J, _, g = adjoint.sparam_objective(model, series, model.frequency, b_port=b, a_port=a)
    rng = np.random.default_rng(11)
    scale = np.abs(series).max()
    for _ in range(4):
        d = rng.standard_normal(series.shape)
        h = 1e-6 * scale
        Jp = adjoint.sparam_objective(model, series + h * d, model.frequency, b, a)[0]
        Jm = adjoint.sparam_objective(model, series - h * d, model.frequency, b, a)[0]
        fd = (Jp - Jm) / (2 * h)
        assert np.sum(g * d) == pytest.approx(fd, rel=1e-6)
    # probes outside the two ports' voltage and current sets carry no derivative
    used = set()
    for L in adjoint._port_probe_layout(model):
        used |= set(L["v_idx"].tolist()) | set(L["i_idx"])
    unused = [k for k in range(len(probes)) if k not in used]
    assert not np.any(g[unused])


def _design_model(er_design, origin, graded=True):
    """Non-uniform grid with a dielectric slab, a design box of per-cell eps_r, PML on two faces
    and a conductor and a port away from the box."""
    nx, ny, nz = 16, 13, 11
    d = 1e-4 * IN_PER_M
    # graded widths on every axis so that the half-width weights of the edge averaging differ
    wx = d * (1 + 0.5 * np.sin(np.arange(nx)) ** 2) if graded else np.full(nx, d)
    wy = d * (1 + 0.4 * np.cos(np.arange(ny) * 1.3) ** 2) if graded else np.full(ny, d)
    wz = d * (1 + 0.3 * np.sin(np.arange(nz) * 0.7 + 0.2) ** 2) if graded else np.full(nz, d)
    xs, ys, zs = [np.concatenate([[0], np.cumsum(w)]) for w in (wx, wy, wz)]
    s = FDTD_Solver(Box((0, 0, 0), (xs[-1], ys[-1], zs[-1])))
    s.set_grid(xs, ys, zs)
    s.add_dielectric(Box((0, 0, 0), (xs[-1], ys[-1], zs[2])), er=2.2)
    s.add_permittivity_volume(np.asarray(er_design, np.float64), origin)
    s.add_conductor(Rectangle([(xs[2], ys[3], zs[3]), (xs[2], ys[5], zs[3]), (xs[4], ys[5], zs[3])]))
    s.add_lumped_port(1, Rectangle([(xs[2], ys[3], 0), (xs[2], ys[5], 0), (xs[2], ys[5], zs[3])]), "z+")
    s.assign_PML_boundaries("x+", "z+", n_pml=3)
    s.coef_dtype = np.float64  # coefficients are assembled by generate_mesh
    s.generate_mesh(courant=0.99)
    return s


def test_chain_rule_matches_central_differences():
    """d(sum W * Cb)/d eps_r of the design cells, by the chain rule through Cb = 2 dt/(2 eps) and the
    half-width-weighted averaging onto edges (rfnetwork/solver.py:892-961), against central
    differences of the coefficient assembly itself (Delta eps_r = 1e-4, float64)."""
    origin, shape = (6, 5, 3), (4, 3, 3)
    rng = np.random.default_rng(5)
    er = rng.uniform(1.0, 4.0, size=shape)
    base = _design_model(er, origin)
    coeff = base.coefficient_arrays(gpu=True, dtype=np.float64)
    cbox = (origin[0], origin[0] + shape[0], origin[1], origin[1] + shape[1], origin[2], origin[2] + shape[2])
    gbox = adjoint.gpu_box_for_cells(base, cbox)
    nb = (gbox[1] - gbox[0], gbox[3] - gbox[2], gbox[5] - gbox[4])
    W = rng.standard_normal((6,) + nb)
    grad = adjoint.cb_to_eps_r_gradient(base, coeff, W, gbox, cbox)
    names = ("Cb_ex_y", "Cb_ex_z", "Cb_ey_z", "Cb_ey_x", "Cb_ez_x", "Cb_ez_y")
    sl = (slice(gbox[0], gbox[1]), slice(gbox[2], gbox[3]), slice(gbox[4], gbox[5]))

    def F(e):
        c = _design_model(e, origin).coefficient_arrays(gpu=True, dtype=np.float64)
        # the PEC wall planes are zeroed by the solver before use
        return sum(np.sum(W[k] * c[n][sl]) for k, n in enumerate(names))

    h = 1e-4
    for idx in [(0, 0, 0), (3, 2, 2), (1, 1, 1), (2, 0, 1), (3, 1, 0)]:
        ep, em = er.copy(), er.copy()
        ep[idx] += h
        em[idx] -= h
        fd = (F(ep) - F(em)) / (2 * h)
        assert grad[idx] == pytest.approx(fd, rel=1e-6), idx


def test_chain_rule_rejects_non_plain_edges():
    """A design box touching a PML, conductor or port edge has eps-dependent Ca or source terms
    that the chain rule does not cover: refused."""
    er = np.full((3, 3, 3), 2.0)
    m = _design_model(er, (1, 3, 1))  # touches the conductor and the port at x = 2
    coeff = m.coefficient_arrays(gpu=True, dtype=np.float64)
    cbox = (1, 4, 3, 6, 1, 4)
    gbox = adjoint.gpu_box_for_cells(m, cbox)
    nb = (gbox[1] - gbox[0], gbox[3] - gbox[2], gbox[5] - gbox[4])
    with pytest.raises(ValueError):
        adjoint.cb_to_eps_r_gradient(m, coeff, np.ones((6,) + nb), gbox, cbox)


def test_unrolled_torch_forward_equals_oracle():
    """The torch stepper (the adjoint's autograd reference) reproduces the oracle's float64 probe
    series; its time loop is the one torch differentiates."""
    model = workloads.tiny_model(Nt=30, time_monitors=False, phasor_monitors=False)
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = model.pack(gpu=False, dtype=np.float64, n_threads=2)
    c_t, p_t = copy.deepcopy(c), copy.deepcopy(probes)
    cpu_solver.solver_run(c, probes, monitors, mem, Nx, Ny, Nz, Nt, 2, 0, dtype=np.float64)
    series, _ = yee_torch.run(c_t, p_t, Nx, Ny, Nz, Nt)
    ref = np.array([p["values"][:Nt] for p in probes])
    np.testing.assert_allclose(series.detach().numpy(), ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())
    assert np.abs(ref).max() > 0


def test_unrolled_torch_gradient_matches_finite_differences():
    """Autograd through the torch stepper against central differences of the float64 oracle for a
    few Cb entries: pins the reference the GPU adjoint is compared with."""
    model = workloads.tiny_model(Nt=25, time_monitors=False, phasor_monitors=False)
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = model.pack(gpu=False, dtype=np.float64, n_threads=2)
    rng = np.random.default_rng(3)
    g = rng.standard_normal((len(probes), Nt))
    series, leaf = yee_torch.run(copy.deepcopy(c), copy.deepcopy(probes), Nx, Ny, Nz, Nt)
    torch.sum(series * torch.tensor(g)).backward()

    def L(cc):
        pr = copy.deepcopy(probes)
        cpu_solver.solver_run(cc, pr, [], mem, Nx, Ny, Nz, Nt, 2, 0, dtype=np.float64)
        return float(np.sum(g * np.array([p["values"][:Nt] for p in pr])))

    for name, idx in [("Cb_ez_x", (6, 5, 2)), ("Cb_ex_y", (5, 4, 1)), ("Cb_ey_x", (7, 3, 3))]:
        h = 1e-3 * abs(c[name][idx])
        cp, cm = copy.deepcopy(c), copy.deepcopy(c)
        cp[name][idx] += h
        cm[name][idx] -= h
        fd = (L(cp) - L(cm)) / (2 * h)
        ad = float(leaf[name].grad[idx])
        assert ad == pytest.approx(fd, rel=1e-6, abs=1e-9 * abs(ad)), name
        assert ad != 0


def test_tuner_trial_lengths():
    """Tuner trials: 20 steps on grids of up to 50 M cells, more on small ones (up to 200), fewer on
    larger ones (at least 6), never past the installed series."""
    from rfnetwork_amd.autotune import trial_steps_for

    assert trial_steps_for([(400, 400, 200, 1)]) == 20
    assert trial_steps_for([(120, 48, 32, 1)]) == 200
    assert trial_steps_for([(512, 512, 256, 1), (384, 384, 224, 1)]) == 9  # 100.1 M cells
    assert trial_steps_for([(1024, 1024, 512, 1)]) == 6
    assert trial_steps_for([(120, 48, 32, 1)], limit=50) == 50
