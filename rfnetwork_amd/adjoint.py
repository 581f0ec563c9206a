"""Gradients of port objectives with respect to per-cell relative permittivity.

The reference has no differentiation.  Here the objective is J = sum_f |S_ba(f)|^2 with S taken as
FDTD_Solver.get_sparameters does (rfnetwork/solver.py:2442-2532): port voltage from the E probes of
the port, current from its ampere loop, A = (V + r0 I)/2, B = (V - r0 I)/2, S_ba = B_b / A_a, and
plain (not decimated) DTFTs.  Its gradient with respect to eps_r of a box of cells is assembled from

1. dJ/dy[n] for every probe series y (exact, by Wirtinger calculus on the DTFT sums);
2. the engine's discrete adjoint (rfdtd_gradient): dJ/dCb for the six E splits of every cell, from a
   reverse-time pass over a checkpointed forward run;
3. the chain rule through Cb = 2 dt / (2 eps + sigma dt) (solver.py:954-961, sigma = 0 in the
   design region, so dCb/deps = -Cb/eps and Ca = 1 does not depend on eps) and through the
   half-width-weighted averaging of cell permittivities onto the E edges (solver.py:892-940).

``EpsRObjective`` wraps it as a torch.autograd.Function.
"""
import numpy as np

from .session import Session
from .units import const, conv

e0 = const.e0


# ---------------------------------------------------------------------------------------------
# objective and its derivative with respect to the probe series
# ---------------------------------------------------------------------------------------------

def _port_probe_layout(model):
    """Per port i (1-based order of model.ports): voltage weights per port probe index and the
    current-loop probe indices with their weights, as positions in the packed probe list
    (ports, inverted ports, then model.probes in order; rfnetwork/solver.py:1541-1575)."""
    offsets, k = [], 0
    for port in model.ports + model.ports_inv:
        if port is None:
            continue
        offsets.append((port, k))
        k += port["Vs_a"].size
    n_port_probes = k
    names = list(model.probes.keys())
    layout = []
    for i, port in enumerate(model.ports):
        start = [o for p, o in offsets if p is port][0]
        f_axis = "xyz".index(port["field"][1])
        idx = port["idx"][f_axis]
        cw = conv.m_in(model.d_cells[f_axis][idx]).astype(np.float64)
        shape = port["Vs_a"].shape
        b = [None] * 3
        b[f_axis] = slice(None)
        w = np.broadcast_to(cw[tuple(b)], shape).astype(np.float64) * (-port["direction"])
        # after the sum along the integration axis, the voltage is averaged over the remaining width
        rest = [s for a, s in enumerate(shape) if a != f_axis and s > 1]
        n_w = int(np.prod(rest)) if rest else 1
        vw = (w / n_w).ravel()
        prefix = f"port_{i + 1}"
        cur = [(n_port_probes + j, float(model.probes[nm]["d"])) for j, nm in enumerate(names)
               if nm[:len(prefix)] == prefix]
        layout.append(dict(v_idx=np.arange(start, start + port["Vs_a"].size), v_w=vw, i_idx=[c[0] for c in cur],
                           i_w=np.array([c[1] for c in cur]), direction=port["direction"], r0=port["r0"]))
    return layout


def sparam_objective(model, series, frequency, b_port=2, a_port=1):
    """J = sum_f |S_ba(f)|^2 and dJ/dy for every probe series.

    series: float array [P, Nt] of recorded probe values in packed order.  Returns (J, S_ba, g) with
    g float64 [P, Nt]."""
    series = np.asarray(series, dtype=np.float64)
    P, Nt = series.shape
    frequency = np.atleast_1d(np.asarray(frequency, dtype=np.float64))
    # the same sample rate get_sparameters hands to dtft: 1 / dt in the model's dtype
    omega = 2 * np.pi * (frequency / (1 / model.dt))
    ph = np.exp(-1j * np.outer(np.arange(Nt), omega))  # [Nt, Nf]
    layout = _port_probe_layout(model)
    V, I = [], []
    for L in layout:
        vp = L["v_w"] @ series[L["v_idx"]]
        ip = L["i_w"] @ series[L["i_idx"]] if len(L["i_idx"]) else np.zeros(Nt)
        V.append(vp @ ph)
        I.append(L["direction"] * (ip @ ph))
    a, b = layout[a_port - 1], layout[b_port - 1]
    A = (V[a_port - 1] + a["r0"] * I[a_port - 1]) / 2
    B = (V[b_port - 1] - b["r0"] * I[b_port - 1]) / 2
    S = B / A
    J = float(np.sum(np.abs(S) ** 2))
    # dJ = sum_f 2 Re(conj(S) dS), dS = dB / A - B dA / A^2
    dB_coef = 1 / A
    dA_coef = -B / A ** 2
    g = np.zeros((P, Nt))
    for pi, L in enumerate(layout):
        cv = np.zeros_like(A)  # d S / d V_i
        ci = np.zeros_like(A)  # d S / d I_i
        if pi == b_port - 1:
            cv = cv + dB_coef / 2
            ci = ci - dB_coef * L["r0"] / 2
        if pi == a_port - 1:
            cv = cv + dA_coef / 2
            ci = ci + dA_coef * L["r0"] / 2
        wv = 2 * np.conj(S) * cv  # per frequency
        wi = 2 * np.conj(S) * ci * L["direction"]
        base_v = np.real(ph @ wv)  # [Nt]: sum_f Re(w_f e^{-j w_f n})
        base_i = np.real(ph @ wi)
        for j, w in zip(L["v_idx"], L["v_w"]):
            g[j] += w * base_v
        for j, w in zip(L["i_idx"], L["i_w"]):
            g[j] += w * base_i
    return J, S, g


# ---------------------------------------------------------------------------------------------
# chain rule: packed Cb of the E splits -> edge permittivity -> cell permittivity
# ---------------------------------------------------------------------------------------------

def _pair_weights(d):
    """Half-width weights of the two cells adjacent to each interior edge along one axis."""
    d = np.asarray(d, dtype=np.float64)
    w0 = d[:-1] / (d[:-1] + d[1:])
    return w0, 1 - w0


def cb_to_eps_r_gradient(model, coeff, grad_cb, gbox, cbox, edges=None):
    """dJ/deps_r of the cells of cbox = (x0, x1, y0, y1, z0, z1) from dJ/dCb of the six E splits on
    the grid-cell box gbox (as returned by Session.gradient, GPU packing).  The edge permittivity
    and packed coefficients come from the model and its packed arrays `coeff`, or, given
    edges = design.edge_values(...), from those (values over gbox for the current design)."""
    gx0, gx1, gy0, gy1, gz0, gz1 = gbox
    x0, x1, y0, y1, z0, z1 = cbox
    Nx, Ny, Nz = model.n_cells
    dx, dy, dz = [conv.m_in(d).astype(np.float64) for d in model.d_cells]
    w = dict(x=_pair_weights(dx), y=_pair_weights(dy), z=_pair_weights(dz))
    out = np.zeros((x1 - x0, y1 - y0, z1 - z0))
    gx, gy, gz = np.meshgrid(np.arange(gx0, gx1), np.arange(gy0, gy1), np.arange(gz0, gz1), indexing="ij")
    sl = (slice(gx0, gx1), slice(gy0, gy1), slice(gz0, gz1))
    # component: (split Cb keys, split Ca keys, global index offset of the GPU cell, averaged axes)
    comps = (("ex", ("Cb_ex_y", "Cb_ex_z"), ("Ca_ex_y", "Ca_ex_z"), (0, 1, 1), ("y", "z")),
             ("ey", ("Cb_ey_z", "Cb_ey_x"), ("Ca_ey_z", "Ca_ey_x"), (1, 0, 1), ("x", "z")),
             ("ez", ("Cb_ez_x", "Cb_ez_y"), ("Ca_ez_x", "Ca_ez_y"), (1, 1, 0), ("x", "y")))
    eps_edge = dict(ex=model.eps_ex, ey=model.eps_ey, ez=model.eps_ez) if edges is None else None
    for c, (comp, cbk, cak, off, axes) in enumerate(comps):
        X, Y, Z = gx + off[0], gy + off[1], gz + off[2]
        # interior edges only: boundary edges hold e0 and take no cell permittivity
        valid = np.ones(gx.shape, bool)
        for ax_name, G, N in (("x", X, Nx), ("y", Y, Ny), ("z", Z, Nz)):
            if ax_name in axes:
                valid &= (G >= 1) & (G <= N - 1)
        if edges is None:
            Xc, Yc, Zc = np.clip(X, 0, Nx), np.clip(Y, 0, Ny), np.clip(Z, 0, Nz)
            eps = eps_edge[comp][Xc, Yc, Zc].astype(np.float64)
            cget = lambda key: coeff[key][sl]  # noqa: E731
        else:
            eps = edges[0][comp[1]].astype(np.float64)
            cget = lambda key: edges[1][key]  # noqa: E731
        cb1, cb2 = cget(cbk[0]).astype(np.float64), cget(cbk[1]).astype(np.float64)
        dJ_deps = np.where(valid, -(grad_cb[2 * c] * cb1 + grad_cb[2 * c + 1] * cb2) / eps, 0.0)
        plain = (cget(cak[0]) == 1.0) & (cget(cak[1]) == 1.0)
        for a in (0, 1):
            for b2 in (0, 1):
                # the cell: lower neighbour along each averaged axis, shifted by a / b
                cell = [X.copy(), Y.copy(), Z.copy()]
                wt = np.ones(gx.shape)
                for shift, ax_name in zip((a, b2), axes):
                    k = "xyz".index(ax_name)
                    idx = np.clip(cell[k] - 1, 0, [Nx, Ny, Nz][k] - 2)
                    wt = wt * w[ax_name][shift][idx]
                    cell[k] = cell[k] - 1 + shift
                inside = valid & (cell[0] >= x0) & (cell[0] < x1) & (cell[1] >= y0) & (cell[1] < y1) & \
                    (cell[2] >= z0) & (cell[2] < z1)
                if np.any(inside & ~plain):
                    raise ValueError("design region touches a lossy, PML, conductor or lumped-element edge")
                np.add.at(out, (cell[0][inside] - x0, cell[1][inside] - y0, cell[2][inside] - z0),
                          (wt * dJ_deps * e0)[inside])
    return out


def gpu_box_for_cells(model, cbox):
    """Grid-cell box whose E splits hold every edge adjacent to the cells of cbox."""
    x0, x1, y0, y1, z0, z1 = cbox
    Nx, Ny, Nz = model.n_cells
    return (max(0, x0 - 1), min(Nx, x1), max(0, y0 - 1), min(Ny, y1), max(0, z0 - 1), min(Nz, z1))


def history_bytes(boxes, Nt, dtype=np.float32):
    """Device bytes of the design-box H history of boxes (grid-cell boxes) over Nt steps."""
    esz = np.dtype(dtype).itemsize
    return sum(3 * (b[1] - b[0] + 1) * (b[3] - b[2] + 1) * (b[5] - b[4] + 1) * esz * Nt for b in boxes)


def _use_history(mode, boxes, Nt, dtype):
    if mode not in ("auto", "history", "checkpoint"):
        raise ValueError("mode must be auto, history or checkpoint")
    if mode != "auto":
        return mode == "history"
    from ._lib import device_memory

    return history_bytes(boxes, Nt, dtype) < 0.5 * device_memory()[0]


def _assemble_on_device(model, dtype):
    """Sessions of the gradient drivers assemble the coefficients on the GPU (no host packing or
    copy of the 30 arrays) when the model's coefficient dtype is the session's."""
    return np.dtype(model.coef_dtype) == np.dtype(dtype)


def _box_edges(model, gbox, box_coefficients, dtype):
    """The chain rule's inputs over gbox: each E component's edge permittivity (as the host
    assembly forms it) and the session's own packed coefficients there (read back from the GPU)."""
    from .design import edge_values

    eps, _ = edge_values(model, gbox, model.eps, model.sigma, dtype)
    return eps, box_coefficients


def eps_r_gradient(model, cbox, frequency, b_port=2, a_port=1, checkpoint_every=200, dtype=np.float32,
                   flags=0, return_session_timing=False, tune=True, mode="auto"):
    """Forward solve of the model on the GPU, then the adjoint pass.

    mode "history" records the design box's H at every forward step (no recomputation in the reverse
    pass), "checkpoint" saves the fields every checkpoint_every steps and recomputes between them,
    "auto" records the history when it takes less than half the free device memory.
    Returns (J, dJ/deps_r over cbox, S_ba).  The model must be meshed and excited."""
    gbox = gpu_box_for_cells(model, cbox)
    if _assemble_on_device(model, dtype):
        coeff = None
        _, probes, _, _, Nx, Ny, Nz, Nt = model.pack(gpu=True, dtype=dtype, coefficients=False)
        session = Session.assembled(model, dtype, flags)
    else:
        coeff, probes, _, _, Nx, Ny, Nz, Nt = model.pack(gpu=True, dtype=dtype)
        session = Session(coeff, Nx, Ny, Nz, dtype=dtype, flags=flags)
    history = _use_history(mode, [gbox], Nt, dtype)
    with session as s:
        if history:  # before tuning: a recorded history rules out the two-step pass
            s.set_history(gbox, Nt)
        if tune:
            s.autotune()
        if not history:
            s.set_checkpointing(checkpoint_every)
        s.set_sources(probes, [], Nt)
        s.run(Nt)
        s.synchronize()
        t_fwd = s.timing()[0]
        s.read()
        series = np.array([p["values"][:Nt] for p in probes])
        J, S, g = sparam_objective(model, series, frequency, b_port, a_port)
        grad_cb = s.gradient(gbox, [row.astype(dtype) for row in g])
        t_adj = s.timing()[0]
        adj_launch = s.adjoint_launch()
        edges = None if coeff is not None else _box_edges(model, gbox, s.coefficients(box=gbox), dtype)
    grad = cb_to_eps_r_gradient(model, coeff, grad_cb, gbox, cbox, edges=edges)
    if return_session_timing:
        return J, grad, S, dict(forward_ms=t_fwd, adjoint_ms=t_adj, mode="history" if history else "checkpoint",
                                adjoint_launch=adj_launch)
    return J, grad, S


def eps_r_gradients_batch(jobs, b_port=2, a_port=1, checkpoint_every=0, dtype=np.float32, flags=0, tune=True,
                          budget_bytes=0, return_info=False, mode="auto"):
    """Gradients of several models at once: one batched session holds every (meshed, excited)
    model; their forward runs and their reverse passes share every launch.

    jobs: [(model, cbox, frequency)], all models with the same number of steps.  checkpoint_every
    0 picks the interval that fits device memory (budget_bytes, 0: what is free).  Returns
    [(J, dJ/deps_r over cbox, S_ba)] per job (and a dict of timings and the interval)."""
    from .batch import Batch

    device = all(_assemble_on_device(model, dtype) for model, _, _ in jobs)
    grids, packed, Nt_all = [], [], None
    for model, cbox, _ in jobs:
        coeff, probes, _, _, Nx, Ny, Nz, Nt = model.pack(gpu=True, dtype=dtype, coefficients=not device)
        if Nt_all is None:
            Nt_all = Nt
        elif Nt != Nt_all:
            raise ValueError("all models of a batched gradient must run the same number of steps")
        grids.append((coeff, Nx, Ny, Nz, 1))
        packed.append((coeff, probes, gpu_box_for_cells(model, cbox)))
    info = {}
    history = _use_history(mode, [p[2] for p in packed], Nt_all, dtype)
    info["mode"] = "history" if history else "checkpoint"
    batch = Batch.assembled([j[0] for j in jobs], dtype=dtype, flags=flags) if device else \
        Batch(grids, dtype=dtype, flags=flags)
    with batch as b:
        if history:  # before tuning: a recorded history rules out the two-step pass
            b.set_history([(k, p[2]) for k, p in enumerate(packed)], Nt_all)
        if tune:
            info["launch"] = b.autotune()
        if not history:
            every = int(checkpoint_every)
            if every <= 0:
                every, info["planned_bytes"] = b.plan_checkpoints(Nt_all, [(k, p[2]) for k, p in enumerate(packed)],
                                                                  budget_bytes)
            info["checkpoint_every"] = every
            b.set_checkpointing(every)
        for k, (_, probes, _) in enumerate(packed):
            b.set_sources(k, probes, [], Nt_all)
        b.run(Nt_all)
        b.synchronize()
        info["forward_ms"] = b.timing()[0]
        b.read()
        requests, objectives = [], []
        for k, ((model, cbox, frequency), (_, probes, gbox)) in enumerate(zip(jobs, packed)):
            series = np.array([p["values"][:Nt_all] for p in probes])
            J, S, g = sparam_objective(model, series, frequency, b_port, a_port)
            objectives.append((J, S))
            requests.append((k, gbox, [row.astype(dtype) for row in g]))
        grads = b.gradients(requests)
        info["adjoint_ms"] = b.timing()[0]
        edges = [None if coeff is not None else _box_edges(model, gbox, b.coefficients(k, box=gbox), dtype)
                 for k, ((model, _, _), (coeff, _, gbox)) in enumerate(zip(jobs, packed))]
    out = []
    for (model, cbox, _), (coeff, _, gbox), (J, S), g, ed in zip(jobs, packed, objectives, grads, edges):
        out.append((J, cb_to_eps_r_gradient(model, coeff, g, gbox, cbox, edges=ed), S))
    return (out, info) if return_info else out


def objective_only(model, frequency, b_port=2, a_port=1, dtype=np.float32, backend=None, n_threads=4):
    """J of the model from a plain solve (GPU by default; backend: a solver_run-style callable run
    with n_threads)."""
    gpu = backend is None
    coeff, probes, monitors, mem, Nx, Ny, Nz, Nt = model.pack(gpu=gpu, dtype=dtype, n_threads=n_threads)
    if backend is None:
        from .hip_func import solver_run_hip
        solver_run_hip(coeff, probes, [], mem, Nx, Ny, Nz, Nt, 0, 0, dtype=dtype)
    else:
        backend(coeff, probes, [], mem, Nx, Ny, Nz, Nt, n_threads, 0, dtype=dtype)
    series = np.array([p["values"][:Nt] for p in probes])
    J, S, _ = sparam_objective(model, series, frequency, b_port, a_port)
    return J, S


try:  # torch is plumbing for autograd users; the engine itself does not need it
    import torch

    class EpsRObjective(torch.autograd.Function):
        """J(eps_r of a design box) with a hand-written adjoint backward, on one resident GPU session.

        apply(eps_r, loop): loop is a design.DesignLoop; forward recomputes the box's coefficients on
        the GPU (from eps_r's own device memory when it is a GPU tensor) and runs the forward solve,
        and the reverse pass runs only when backward is called.
        apply(eps_r, build, cbox, frequency[, checkpoint_every]): build(eps_r_numpy) returns a meshed,
        excited model whose cells in cbox have that permittivity; a loop is made for that model and
        released after backward.  J is sum_f |S_21|^2 (ports b=2, a=1)."""

        @staticmethod
        def forward(ctx, eps_r, loop, cbox=None, frequency=None, checkpoint_every=200):
            from .design import DesignLoop

            owned = not isinstance(loop, DesignLoop)
            if owned:
                eps = eps_r.detach().to("cpu", torch.float64).numpy()
                model = loop(eps)
                loop = DesignLoop(model, cbox, frequency, dtype=model.coef_dtype, checkpoint_every=checkpoint_every)
            # a tensor in device memory goes to the GPU's coefficient update as it is (no host copy)
            J, _ = loop.forward(eps_r.detach() if eps_r.is_cuda else eps_r.detach().to(torch.float64).numpy())
            ctx.loop, ctx.owned, ctx.generation = loop, owned, loop.generation
            return eps_r.new_tensor(J)

        @staticmethod
        def backward(ctx, grad_out):
            grad = ctx.loop.backward(ctx.generation)
            if ctx.owned:
                ctx.loop.close()
            g = torch.as_tensor(grad, dtype=grad_out.dtype, device=grad_out.device)
            return grad_out * g, None, None, None, None

except ImportError:  # pragma: no cover
    EpsRObjective = None
