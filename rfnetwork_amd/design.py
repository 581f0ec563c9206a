"""Design-region coefficients: what the GPU recomputes when the permittivity of a box of cells
changes (rfdtd_set_design_permittivity), the same formula on the host for the gradient's chain rule,
and a resident-session loop for repeated gradients of one model with changing design permittivity.

The formula is the host assembly's restricted to edges of plain dielectric (rfnetwork/solver.py:
892-961 for the half-width-weighted averaging of cell permittivity and conductivity onto the
edges and Ca/Cb, :1463-1523 for the packing): per E component the 4 cells around the edge are
averaged in float64 along the lower transverse axis first (Ex: y then z; Ey: x then z; Ez: x then
y), the result is rounded to the coefficient type, Ca = (2 eps - sig dt) / (2 eps + sig dt) and
Cb = 2 dt / (2 eps + sig dt) in that type; edges on the grid boundary hold eps0 without loss; the
packed Cb carries the inverse dual spacing (0 past the last) and the curl sign.
"""
import ctypes

import numpy as np

from ._lib import check
from .units import conv, const


def _avg(v0, v1, d0, d1):
    """Half-width-weighted average of two cell values (float64).  The widths are float32 values
    (held in float64 here): their halves are exact, and the sum of the halves is formed in float32,
    as rfnetwork/solver.py:892-912 forms it from float32 width arrays."""
    h0, h1 = np.asarray(d0 / 2), np.asarray(d1 / 2)
    return (v0 * h0 + v1 * h1) / (h0.astype(np.float32) + h1.astype(np.float32))


def edge_values(model, gbox, eps_cells, sig_cells, dtype):
    """For the packed cells of gbox: the edge permittivity of each E component (eps_x, eps_y, eps_z)
    and the 12 packed E coefficients (Ca_ex_y, Ca_ex_z, Cb_ex_y, Cb_ex_z, Ca_ey_z, ..., Cb_ez_y), in
    dtype, from full-grid cell permittivity and conductivity (float64)."""
    T = np.dtype(dtype).type
    gx0, gx1, gy0, gy1, gz0, gz1 = gbox
    Nx, Ny, Nz = model.n_cells
    dx, dy, dz = [np.asarray(conv.m_in(d), dtype=np.float64) for d in model.d_cells]
    ix, iy, iz = model.packing_inverses(dtype)
    i = np.arange(gx0, gx1)[:, None, None]
    j = np.arange(gy0, gy1)[None, :, None]
    k = np.arange(gz0, gz1)[None, None, :]
    xi, yi, zi = i <= Nx - 2, j <= Ny - 2, k <= Nz - 2
    i1, j1, k1 = np.minimum(i + 1, Nx - 1), np.minimum(j + 1, Ny - 1), np.minimum(k + 1, Nz - 1)
    dxi, dyj, dzk = dx[i], dy[j], dz[k]
    dxn, dyn, dzn = np.where(xi, dx[i1], 0.0), np.where(yi, dy[j1], 0.0), np.where(zi, dz[k1], 0.0)
    dt = T(model.dt)
    e0 = T(const.e0)
    out_eps, ca, cb = {}, {}, {}
    for comp, interior, first, second in (
            ("x", yi & zi, lambda v, a: _avg(v[i, j, a], v[i, j1, a], dyj, dyn), (k, k1, dzk, dzn)),
            ("y", xi & zi, lambda v, a: _avg(v[i, j, a], v[i1, j, a], dxi, dxn), (k, k1, dzk, dzn)),
            ("z", xi & yi, lambda v, a: _avg(v[i, a, k], v[i1, a, k], dxi, dxn), (j, j1, dyj, dyn))):
        a0, a1, d0, d1 = second
        vals = []
        for v in (eps_cells, sig_cells):
            vals.append(_avg(first(v, a0), first(v, a1), d0, d1))
        eps = np.where(interior, vals[0].astype(T), e0).astype(T)
        sig = np.where(interior, vals[1].astype(T), T(0)).astype(T)
        out_eps[comp] = eps
        with np.errstate(all="ignore"):
            ca[comp] = ((T(2) * eps - (sig * dt)) / (T(2) * eps + (sig * dt))).astype(T)
            cb[comp] = ((T(2) * dt) / (T(2) * eps + (sig * dt))).astype(T)
    ixv, iyv, izv = ix[i], iy[j], iz[k]
    tup = [ca["x"], ca["x"], cb["x"] * iyv, (-cb["x"]) * izv,
           ca["y"], ca["y"], cb["y"] * izv, (-cb["y"]) * ixv,
           ca["z"], ca["z"], cb["z"] * ixv, (-cb["z"]) * iyv]
    shape = (gx1 - gx0, gy1 - gy0, gz1 - gz0)
    tup = [np.broadcast_to(np.asarray(t, dtype=T), shape).copy() for t in tup]
    last = (i == Nx - 1)[:, 0, 0]
    for c in range(4, 12):
        tup[c][last] = T(0)
    names = ("Ca_ex_y", "Ca_ex_z", "Cb_ex_y", "Cb_ex_z", "Ca_ey_z", "Ca_ey_x", "Cb_ey_z", "Cb_ey_x",
             "Ca_ez_x", "Ca_ez_y", "Cb_ez_x", "Cb_ez_y")
    eps_b = {c: np.broadcast_to(out_eps[c], shape).copy() for c in out_eps}
    return eps_b, dict(zip(names, tup))


def design_permittivity_args(model, cbox, eps_now, sigma_now, eps_r, dtype):
    """Arguments of rfdtd_set_design_permittivity for replacing the relative permittivity of the
    cells of cbox = (x0, x1, y0, y1, z0, z1) by eps_r: the packed-cell box, the current and new
    permittivity (F/m) and the conductivity over that box grown by one cell on each upper side,
    the cell widths and the packing inverses.  eps_now, sigma_now: full-grid float64 arrays.
    eps_r None: no new permittivity (new is None; rfdtd_design_prepare's arguments)."""
    from .adjoint import gpu_box_for_cells

    x0, x1, y0, y1, z0, z1 = (int(v) for v in cbox)
    Nx, Ny, Nz = model.n_cells
    gbox = gpu_box_for_cells(model, cbox)
    gx0, gx1, gy0, gy1, gz0, gz1 = gbox
    ext = (slice(gx0, min(gx1 + 1, Nx)), slice(gy0, min(gy1 + 1, Ny)), slice(gz0, min(gz1 + 1, Nz)))
    now = np.ascontiguousarray(eps_now[ext], dtype=np.float64)
    new = None
    if eps_r is not None:
        eps_r = np.asarray(eps_r, dtype=np.float64)
        if eps_r.shape != (x1 - x0, y1 - y0, z1 - z0):
            raise ValueError("eps_r must have the shape of the design box")
        new = now.copy()
        new[x0 - gx0:x1 - gx0, y0 - gy0:y1 - gy0, z0 - gz0:z1 - gz0] = eps_r * const.e0
    sig = np.ascontiguousarray(sigma_now[ext], dtype=np.float64)
    widths = [np.ascontiguousarray(conv.m_in(d), dtype=np.float64) for d in model.d_cells]
    inverses = [np.ascontiguousarray(a) for a in model.packing_inverses(dtype)]
    return gbox, now, new, sig, widths, inverses, float(model.dt), float(const.e0)


def set_design_permittivity(handle, L, member, model, cbox, eps_r, state, dtype):
    """Replace eps_r of the cells of cbox in member `member` of a session on the GPU; state holds
    the session's current full-grid permittivity and conductivity per member (taken from the model
    the first time) and is updated on success."""
    if np.dtype(model.coef_dtype) != np.dtype(dtype):
        raise ValueError("the model's coefficient dtype must be the session's element type")
    if member not in state:
        state[member] = (np.array(model.eps, dtype=np.float64), np.array(model.sigma, dtype=np.float64))
    eps_now, sigma_now = state[member]
    gbox, now, new, sig, (dx, dy, dz), (ix, iy, iz), dt, e0 = design_permittivity_args(
        model, cbox, eps_now, sigma_now, eps_r, dtype)
    b = (ctypes.c_int32 * 6)(*gbox)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    check(L.rfdtd_set_design_permittivity(handle, int(member), ctypes.cast(b, ctypes.c_void_p), ptr(now), ptr(new),
                                          ptr(sig), ptr(dx), ptr(dy), ptr(dz), ptr(ix), ptr(iy), ptr(iz),
                                          ctypes.c_double(dt), ctypes.c_double(e0)), "rfdtd_set_design_permittivity")
    x0, x1, y0, y1, z0, z1 = (int(v) for v in cbox)
    eps_now[x0:x1, y0:y1, z0:z1] = np.asarray(eps_r, dtype=np.float64) * e0


def prepare_design_box(handle, L, member, model, cbox, state, dtype):
    """Keep the arrays of cbox's box of member `member` on the device for updates from device memory
    (rfdtd_design_prepare), checked against the session's coefficients; state as for
    set_design_permittivity (the permittivity the session holds now)."""
    if np.dtype(model.coef_dtype) != np.dtype(dtype):
        raise ValueError("the model's coefficient dtype must be the session's element type")
    if member not in state:
        state[member] = (np.array(model.eps, dtype=np.float64), np.array(model.sigma, dtype=np.float64))
    eps_now, sigma_now = state[member]
    gbox, now, _, sig, (dx, dy, dz), (ix, iy, iz), dt, e0 = design_permittivity_args(
        model, cbox, eps_now, sigma_now, None, dtype)
    b = (ctypes.c_int32 * 6)(*gbox)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    check(L.rfdtd_design_prepare(handle, int(member), ctypes.cast(b, ctypes.c_void_p), ptr(now), ptr(sig), ptr(dx),
                                 ptr(dy), ptr(dz), ptr(ix), ptr(iy), ptr(iz), dt, e0), "rfdtd_design_prepare")


def is_device_tensor(a):
    """True for a torch tensor in GPU memory."""
    return type(a).__module__.startswith("torch") and getattr(a, "is_cuda", False)


def set_design_permittivity_device(handle, L, member, cbox, eps_r):
    """Relative permittivity of the cells of cbox from a torch tensor in device memory (float32 or
    float64, the box's shape; other types are converted to float64), in a box prepared with
    prepare_design_box: no host copy.  torch's current stream (which produced eps_r) is waited for,
    then the update runs on the session's stream, before its later steps.  Returns the tensor the
    GPU reads: the caller keeps it alive until the session has synchronised.

    The library may run on a HIP runtime other than torch's (torch wheels bundle their own), so no
    torch stream handle is passed to it; device addresses are valid in either."""
    import torch

    x0, x1, y0, y1, z0, z1 = (int(v) for v in cbox)
    t = eps_r.detach()
    if t.dtype not in (torch.float32, torch.float64):
        t = t.to(torch.float64)
    t = t.contiguous()
    if tuple(t.shape) != (x1 - x0, y1 - y0, z1 - z0):
        raise ValueError("eps_r must have the shape of the design box")
    b = (ctypes.c_int32 * 6)(x0, x1, y0, y1, z0, z1)
    torch.cuda.current_stream(t.device).synchronize()
    check(L.rfdtd_set_design_permittivity_device(handle, int(member), ctypes.cast(b, ctypes.c_void_p),
                                                 ctypes.c_void_p(t.data_ptr()), int(t.dtype == torch.float64),
                                                 None), "rfdtd_set_design_permittivity_device")
    return t


class DesignLoop:
    """Repeated J and dJ/deps_r of one model with changing design-box permittivity, on one resident
    GPU session: each evaluation recomputes the box's coefficients on the device, reruns the forward
    solve and the reverse pass, and applies the chain rule with the design's edge values from
    edge_values (no model rebuild, packing or session set-up per evaluation).  Results equal
    adjoint.eps_r_gradient on a model rebuilt with the same permittivity, bit for bit.

    model: meshed and excited, with the design box cbox inside plain dielectric; frequency and the
    ports as for eps_r_gradient.  mode: "history", "checkpoint" or "auto" (see eps_r_gradient)."""

    def __init__(self, model, cbox, frequency, b_port=2, a_port=1, dtype=np.float32, flags=0, tune=True,
                 mode="auto", checkpoint_every=200):
        from . import adjoint
        from .session import Session

        if np.dtype(model.coef_dtype) != np.dtype(dtype):
            raise ValueError("the model's coefficient dtype must be the session's element type")
        self.model, self.cbox, self.frequency = model, tuple(int(v) for v in cbox), frequency
        self.b_port, self.a_port, self.dtype = b_port, a_port, np.dtype(dtype)
        _, probes, _, _, Nx, Ny, Nz, Nt = model.pack(gpu=True, dtype=dtype, coefficients=False)
        self.Nt = Nt
        self.gbox = adjoint.gpu_box_for_cells(model, cbox)
        self.history = adjoint._use_history(mode, [self.gbox], Nt, dtype)
        self.session = Session.assembled(model, dtype, flags)  # coefficients assembled on the GPU
        if self.history:  # recorded every evaluation; set before tuning, which it constrains
            self.session.set_history(self.gbox, self.Nt)
        if tune:
            self.session.autotune()
        self.checkpoint_every = checkpoint_every
        self.eps = np.array(model.eps, dtype=np.float64)  # the session's current cell permittivity
        self.sigma = np.array(model.sigma, dtype=np.float64)
        self.timing = {}
        self.generation = 0
        self._pending = None  # (generation, dJ/dy of that forward) until its backward runs
        self._prepared = False  # the box held on the device for updates from device memory
        self._eps_device = None  # permittivity set from device memory, not yet in self.eps

    def _host_eps(self):
        """Bring self.eps up to date with a permittivity last set from device memory (one copy of the
        box's values, made only when the host needs them)."""
        if self._eps_device is not None:
            x0, x1, y0, y1, z0, z1 = self.cbox
            vals = self._eps_device.detach().to("cpu").numpy().astype(np.float64)
            self.eps[x0:x1, y0:y1, z0:z1] = vals * const.e0
            self._eps_device = None

    def forward(self, eps_r):
        """(J, S_ba) for relative permittivity eps_r of the box cells: the box's coefficients
        recomputed on the GPU, then the forward solve (recording what a reverse pass needs).  The
        reverse pass runs only if backward() is called before the next forward().  eps_r: a numpy
        array, or a torch tensor in device memory, which the GPU reads directly (no host copy)."""
        import time

        from .adjoint import sparam_objective

        s, model = self.session, self.model
        t0 = time.perf_counter()
        s.reset()
        if is_device_tensor(eps_r):
            if not self._prepared:
                self._host_eps()
                prepare_design_box(s._h, s._L, 0, model, self.cbox, {0: (self.eps, self.sigma)}, self.dtype)
                self._prepared = True
            # t stays referenced until the solve below has synchronised; the device copy of its values
            # serves backward() (the caller may change eps_r before then)
            t = set_design_permittivity_device(s._h, s._L, 0, self.cbox, eps_r)
            self._eps_device = t.clone()
        else:
            self._host_eps()
            state = {0: (self.eps, self.sigma)}
            set_design_permittivity(s._h, s._L, 0, model, self.cbox, eps_r, state, self.dtype)
            self._prepared = False  # a host update discards the device-held box
        t1 = time.perf_counter()
        if self.history:
            s.set_history(self.gbox, self.Nt)
        else:
            s.set_checkpointing(self.checkpoint_every)
        _, probes, _, _, *_ = model.pack(gpu=True, dtype=self.dtype, coefficients=False)
        s.set_sources(probes, [], self.Nt)
        s.run(self.Nt)
        s.synchronize()
        t_fwd = s.timing()[0]
        s.read()
        series = np.array([p["values"][:self.Nt] for p in probes])
        J, S, g = sparam_objective(model, series, self.frequency, self.b_port, self.a_port)
        self.generation += 1
        self._pending = (self.generation, g)
        self.timing = dict(update_ms=(t1 - t0) * 1e3, forward_ms=t_fwd, adjoint_ms=0.0,
                           wall_ms=(time.perf_counter() - t0) * 1e3)
        return J, S

    def backward(self, generation=None):
        """dJ/deps_r over the box for the last forward() (generation: the forward it must belong to,
        as forward() numbers them; a later forward makes an earlier one's backward fail)."""
        import time

        from .adjoint import cb_to_eps_r_gradient

        if self._pending is None or (generation is not None and generation != self._pending[0]):
            raise RuntimeError("backward() must follow its own forward() on the design loop")
        t0 = time.perf_counter()
        _, g = self._pending
        self._pending = None
        s, model = self.session, self.model
        grad_cb = s.gradient(self.gbox, [row.astype(self.dtype) for row in g])
        self.timing["adjoint_ms"] = s.timing()[0]
        self._host_eps()  # the chain rule's edge values need the design permittivity on the host
        edges = edge_values(model, self.gbox, self.eps, self.sigma, self.dtype)
        grad = cb_to_eps_r_gradient(model, None, grad_cb, self.gbox, self.cbox, edges=edges)
        self.timing["wall_ms"] += (time.perf_counter() - t0) * 1e3
        return grad

    def evaluate(self, eps_r):
        """(J, dJ/deps_r over the box, S_ba) for relative permittivity eps_r of the box cells."""
        J, S = self.forward(eps_r)
        return J, self.backward(), S

    def close(self):
        self.session.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
