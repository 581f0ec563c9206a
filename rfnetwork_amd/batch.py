"""Batched solves: many excitations and uneven grids advanced together on one GPU.

A ``Batch`` session holds several grids; each grid is advanced for several excitations that share
its coefficients.  Every phase of a time step (E update, E sources, H update, H probes, DFT planes)
is one kernel launch over all instances, captured in hipGraph chunks, so a batch of small grids
costs the launches of one grid and the excitations of a grid share its coefficient reads.

``solve_ports(jobs, frequency)`` is the model-level entry: for each (model, ports) job every listed
port is driven in its own excitation (the others terminated in their resistance), exactly as
``model.assign_excitation(waveform, port); model.solve()`` would do one at a time, and the S-matrix
columns are returned.  The reference runs one excitation per ``solve()`` call
(rfnetwork/solver.py:1410; tests/test_solver/test_coupled_stripline.py:96-109 re-solves per mode).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import FIELD_COMPONENT, FIELD_NAMES, check
from .session import yee_shape


class Batch:
    """Several grids x excitations resident on the GPU."""

    def __init__(self, grids, dtype=np.float32, flags=0):
        """grids: list of (coefficients dict, Nx, Ny, Nz, n_excitations) in the GPU packing."""
        from .hip_func import _check_array, _real_dtype

        self.dtype = _real_dtype(dtype)
        self.grids = []
        dims, exc, ptrs = [], [], []
        for coeff, Nx, Ny, Nz, K in grids:
            for name in _lib.COEFF_NAMES:
                _check_array(coeff[name], name, (Nx, Ny, Nz), self.dtype)
                ptrs.append(coeff[name])
            dims += [int(Nx), int(Ny), int(Nz)]
            exc.append(int(K))
            self.grids.append((int(Nx), int(Ny), int(Nz), int(K)))
        self._L = _lib.lib()
        self._h = ctypes.c_void_p()
        d = (ctypes.c_int32 * len(dims))(*dims)
        e = (ctypes.c_int32 * len(exc))(*exc)
        parr = _lib.pointer_array(ptrs)
        f = int(flags) | (_lib.F64 if self.dtype == np.float64 else 0)
        check(self._L.rfdtd_create_batch(ctypes.byref(self._h), len(grids), ctypes.cast(d, ctypes.c_void_p),
                                         ctypes.cast(e, ctypes.c_void_p), ctypes.cast(parr, ctypes.c_void_p), f),
              "rfdtd_create_batch")
        self.instances = []  # (grid index, excitation index)
        for g, (_, _, _, K) in enumerate(self.grids):
            self.instances += [(g, k) for k in range(K)]
        self._keep = {}
        self.flags = int(flags)

    @classmethod
    def assembled(cls, models, excitations=None, dtype=np.float32, flags=0):
        """A batch of meshed models whose coefficient arrays are assembled on the GPU
        (rfdtd_create_batch_assembled; equal to the host assembly bit for bit); excitations: the
        instances per model (default 1 each)."""
        from . import assembly
        from .hip_func import _real_dtype

        self = cls.__new__(cls)
        self.dtype = _real_dtype(dtype)
        excitations = list(excitations) if excitations is not None else [1] * len(models)
        if len(excitations) != len(models):
            raise ValueError("one excitation count per model")
        descs = [assembly.Description(m, self.dtype) for m in models]
        arr = (_lib.Assembly * len(descs))(*[D.struct for D in descs])
        e = (ctypes.c_int32 * len(excitations))(*[int(k) for k in excitations])
        self.grids = [tuple(int(n) for n in m.n_cells) + (int(k),) for m, k in zip(models, excitations)]
        self._L = _lib.lib()
        self._h = ctypes.c_void_p()
        f = int(flags) | (_lib.F64 if self.dtype == np.float64 else 0)
        check(self._L.rfdtd_create_batch_assembled(ctypes.byref(self._h), len(descs), ctypes.cast(arr, ctypes.c_void_p),
                                                   ctypes.cast(e, ctypes.c_void_p), f), "rfdtd_create_batch_assembled")
        self.instances = []
        for g, (_, _, _, K) in enumerate(self.grids):
            self.instances += [(g, k) for k in range(K)]
        self.flags = int(flags)
        self._keep = {}
        return self

    def coefficients(self, member, box=None):
        """Member `member`'s 30 packed coefficient arrays read back from the GPU, whole or over box =
        (x0, x1, y0, y1, z0, z1) packed cells."""
        from . import assembly

        if box is not None:
            return assembly.read_coefficient_box(self._L, self._h, box, self.dtype, member)
        Nx, Ny, Nz, _ = self.grids[member]
        return assembly.read_coefficients(self._L, self._h, (Nx, Ny, Nz), self.dtype, member)

    def set_monitor_streaming(self, ring_samples):
        """As Session.set_monitor_streaming, for the instances' monitors installed afterwards."""
        check(self._L.rfdtd_set_monitor_streaming(self._h, int(ring_samples)), "rfdtd_set_monitor_streaming")
        self._stream_ring = int(ring_samples)

    def set_sources(self, instance, probes, monitors, Nt):
        from .hip_func import check_probe_cells, pack_monitors, pack_probes, pin_time_monitors

        if getattr(self, "_stream_ring", 0):
            pin_time_monitors(monitors)

        Nx, Ny, Nz, _ = self.grids[self.instances[instance][0]]
        check_probe_cells(probes, Nx, Ny, Nz)
        prb, kp = pack_probes(probes, int(Nt), self.dtype)
        mons, km = pack_monitors(monitors, Nx, Ny, Nz, int(Nt), self.dtype)
        check(self._L.rfdtd_set_instance_sources(self._h, int(instance), ctypes.cast(prb, ctypes.c_void_p),
                                                 len(probes), ctypes.cast(mons, ctypes.c_void_p), len(monitors),
                                                 int(Nt)), "rfdtd_set_instance_sources")
        self._keep[instance] = (prb, mons, kp, km)

    def run(self, n_steps, stream=None):
        s = None if stream is None else ctypes.c_void_p(int(stream))
        check(self._L.rfdtd_run(self._h, int(n_steps), s), "rfdtd_run")

    def synchronize(self):
        check(self._L.rfdtd_synchronize(self._h), "rfdtd_synchronize")

    def read(self):
        check(self._L.rfdtd_read_probes(self._h), "rfdtd_read_probes")
        check(self._L.rfdtd_read_monitors(self._h), "rfdtd_read_monitors")

    def fields(self, instance, names=FIELD_NAMES):
        Nx, Ny, Nz, _ = self.grids[self.instances[instance][0]]
        out, ptrs = {}, (ctypes.c_void_p * 18)()
        for a, name in enumerate(FIELD_NAMES):
            if name in names:
                out[name] = np.zeros(yee_shape(FIELD_COMPONENT[a], Nx, Ny, Nz), self.dtype)
                ptrs[a] = out[name].ctypes.data
        check(self._L.rfdtd_read_instance_fields(self._h, int(instance), ctypes.cast(ptrs, ctypes.c_void_p)))
        return out

    def timing(self):
        t, e, h = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(self._L.rfdtd_last_timing(self._h, ctypes.byref(t), ctypes.byref(e), ctypes.byref(h)))
        return t.value, e.value, h.value

    def set_launch(self, block_threads=0, x_planes=0, graph_steps=0):
        check(self._L.rfdtd_set_launch(self._h, int(block_threads), int(x_planes), int(graph_steps)))

    def autotune(self, trial_steps=20, use_cache=True):
        """Pick the fastest launch shape for this set of grids (before the first run); cached per
        set of shapes.  Returns {variant, block, x_planes, ms_per_step, cached}."""
        from . import autotune

        kind = ctypes.c_int32()
        check(self._L.rfdtd_memory_info(self._h, None, ctypes.byref(kind), None, None))
        return autotune.tune(self, self.grids, np.dtype(self.dtype).name, kind.value, trial_steps, use_cache)

    def set_checkpointing(self, every):
        """Save the split fields of every instance every `every` steps of later runs."""
        check(self._L.rfdtd_set_checkpointing(self._h, int(every)))

    def set_adjoint_launch(self, variant=1, block=0, x_chunks=0):
        """Reverse-pass launch shape: variant 1 pipelined adjoint columns, 0 plain, 2 and 3 the same with
        the workgroups in launch order; block threads per workgroup (0: timed at the next gradient);
        x_chunks workgroups along x (0: the forward's)."""
        check(self._L.rfdtd_set_adjoint_launch(self._h, int(variant), int(block), int(x_chunks)),
              "rfdtd_set_adjoint_launch")

    def adjoint_launch(self):
        """{variant, block, x_chunks} of the reverse pass (block 0 until the first gradient)."""
        cfg = (ctypes.c_int32 * 3)()
        check(self._L.rfdtd_get_adjoint_launch(self._h, ctypes.cast(cfg, ctypes.c_void_p)), "rfdtd_get_adjoint_launch")
        return dict(variant=int(cfg[0]), block=int(cfg[1]), x_chunks=int(cfg[2]))

    def set_design_permittivity(self, member, model, cbox, eps_r):
        """Relative permittivity eps_r of the cells of cbox of member `member` (a grid; every excitation
        of it sees the change), recomputed on the GPU as Session.set_design_permittivity does."""
        from .design import set_design_permittivity

        if not hasattr(self, "_design_state"):
            self._design_state = {}
        set_design_permittivity(self._h, self._L, int(member), model, cbox, eps_r, self._design_state, self.dtype)

    def set_history(self, requests, Nt):
        """Record the H totals of each request's box, requests [(instance, box), ...], at every step of
        later runs (up to Nt steps), for gradients without recomputation.  Before the first step."""
        n = len(requests)
        inst = (ctypes.c_int32 * max(n, 1))(*[int(r[0]) for r in requests])
        boxes = (ctypes.c_int32 * max(6 * n, 1))(*[int(v) for r in requests for v in r[1]])
        check(self._L.rfdtd_set_history(self._h, n, ctypes.cast(inst, ctypes.c_void_p), ctypes.cast(boxes, ctypes.c_void_p),
                                        int(Nt)), "rfdtd_set_history")
        self._history = bool(requests)

    def plan_checkpoints(self, Nt, requests, budget_bytes=0):
        """Checkpoint interval for gradients of requests [(instance, box), ...] after Nt steps that
        fits device memory (budget_bytes, 0: what is free).  Returns (every, bytes needed)."""
        n = len(requests)
        inst = (ctypes.c_int32 * n)(*[int(r[0]) for r in requests])
        boxes = (ctypes.c_int32 * (6 * n))(*[int(v) for r in requests for v in r[1]])
        every, need = ctypes.c_int32(), ctypes.c_uint64()
        check(self._L.rfdtd_plan_checkpoints(self._h, int(Nt), n, ctypes.cast(inst, ctypes.c_void_p),
                                             ctypes.cast(boxes, ctypes.c_void_p), int(budget_bytes),
                                             ctypes.byref(every), ctypes.byref(need)), "rfdtd_plan_checkpoints")
        return every.value, need.value

    def gradients(self, requests):
        """dJ/dCb of the six E splits in each request's box after a checkpointed forward run, for
        requests [(instance, box (x0, x1, y0, y1, z0, z1), adjoint_series per probe)], all reverse
        passes in the same launches.  Returns a float64 [6, nx, ny, nz] array per request."""
        n = len(requests)
        outs, keep = [], []
        inst = (ctypes.c_int32 * n)()
        boxes = (ctypes.c_int32 * (6 * n))()
        series_ptrs = (ctypes.c_void_p * n)()
        counts = (ctypes.c_int32 * n)()
        out_ptrs = (ctypes.c_void_p * n)()
        for k, (instance, box, adj) in enumerate(requests):
            box = [int(v) for v in box]
            inst[k] = int(instance)
            boxes[6 * k:6 * k + 6] = box
            arrs = [np.ascontiguousarray(a, dtype=self.dtype) for a in adj]
            ptrs = _lib.pointer_array(arrs)
            keep += [arrs, ptrs]
            series_ptrs[k] = ctypes.cast(ptrs, ctypes.c_void_p).value
            counts[k] = len(arrs)
            out = np.zeros((6, box[1] - box[0], box[3] - box[2], box[5] - box[4]), np.float64)
            outs.append(out)
            out_ptrs[k] = out.ctypes.data
        check(self._L.rfdtd_gradient_batch(self._h, n, ctypes.cast(inst, ctypes.c_void_p),
                                           ctypes.cast(boxes, ctypes.c_void_p), ctypes.cast(series_ptrs, ctypes.c_void_p),
                                           ctypes.cast(counts, ctypes.c_void_p), ctypes.cast(out_ptrs, ctypes.c_void_p)),
              "rfdtd_gradient_batch")
        from . import autotune

        autotune.remember_adjoint(self)
        return outs

    def set_variant(self, variant=1):
        check(self._L.rfdtd_set_variant(self._h, int(variant)))

    def set_column_order(self, order):
        """Workgroup numbering of the separate kernels (Session.set_column_order)."""
        check(self._L.rfdtd_set_column_order(self._h, int(order)), "rfdtd_set_column_order")

    def column_order(self):
        o = ctypes.c_int32()
        check(self._L.rfdtd_get_column_order(self._h, ctypes.byref(o)), "rfdtd_get_column_order")
        return int(o.value)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.rfdtd_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _waveform(model):
    for p in model.ports:
        if p is not None and p["src"] is not None:
            return np.asarray(p["src"])
    raise ValueError("assign an excitation waveform to the model before batching")


def pack_excitations(model, ports, dtype=np.float32):
    """Per-port excitation inputs of one model: a list of (port, probes, monitors), Nt."""
    wave = _waveform(model)
    saved = [None if p is None else p["src"] for p in model.ports]
    saved_inv = [None if p is None else p["src"] for p in model.ports_inv]
    out = []
    try:
        for port in ports:
            for p in model.ports + model.ports_inv:
                if p is not None:
                    p["src"] = None
            model.ports[port - 1]["src"] = wave.astype(np.float32)
            if model.ports_inv[port - 1] is not None:
                model.ports_inv[port - 1]["src"] = wave.astype(np.float32)
            model._time = np.arange(len(wave)) * model.dt
            _, probes, monitors, _, Nx, Ny, Nz, Nt = model.pack(gpu=True, dtype=dtype, coefficients=False)
            out.append((port, probes, monitors))
    finally:
        for p, s in zip(model.ports, saved):
            if p is not None:
                p["src"] = s
        for p, s in zip(model.ports_inv, saved_inv):
            if p is not None:
                p["src"] = s
    return out, Nt


def solve_ports(jobs, frequency, dtype=np.float32, flags=0, return_batch=False):
    """Drive every (model, ports) job's ports one excitation each, all in one batched session.

    Returns, per job, an array S[f, b, a] whose columns a are the driven ports (source_port order of
    ``ports``), computed with the model's get_sparameters for each excitation, and the models'
    monitor values per excitation (list of dicts)."""
    grids, inputs = [], []
    Nt_all = None
    # the coefficients are assembled on the GPU when every model's coefficient dtype is the batch's
    device = all(np.dtype(model.coef_dtype) == np.dtype(dtype) for model, _ in jobs)
    for model, ports in jobs:
        exc, Nt = pack_excitations(model, ports, dtype)
        if Nt_all is None:
            Nt_all = Nt
        elif Nt != Nt_all:
            raise ValueError("all batched models must run the same number of steps")
        coeff = None if device else model.coefficient_arrays(gpu=True, dtype=dtype)
        grids.append((coeff, model.Nx, model.Ny, model.Nz, len(ports)))
        inputs.append(exc)
    if device:
        batch = Batch.assembled([m for m, _ in jobs], [len(p) for _, p in jobs], dtype=dtype, flags=flags)
    else:
        batch = Batch(grids, dtype=dtype, flags=flags)
    k = 0
    for exc in inputs:
        for _, probes, monitors in exc:
            batch.set_sources(k, probes, monitors, Nt_all)
            k += 1
    batch.run(Nt_all)
    batch.synchronize()
    batch.read()
    results = []
    for (model, ports), exc in zip(jobs, inputs):
        cols, mons = [], []
        for port, probes, monitors in exc:
            model.unpack(probes, monitors, gpu=True)
            cols.append(np.asarray(model.get_sparameters(frequency, source_port=port))[..., 0])
            mons.append({n: np.array(m["values"]) for n, m in model.monitors.items()})
        results.append((np.stack(cols, axis=-1), mons))
    if return_batch:
        return results, batch
    batch.close()
    return results


def sparameter_matrix(model, frequency, dtype=np.float32, flags=0):
    """The model's full S-matrix [frequency, b, a]: every port driven once (the others terminated),
    all excitations in one batched session."""
    from .utils import LabeledArray

    ports = [i + 1 for i, p in enumerate(model.ports) if p is not None]
    (S, _), = solve_ports([(model, ports)], frequency, dtype, flags)
    return LabeledArray(S, coords=dict(frequency=np.atleast_1d(np.asarray(frequency, np.float64)),
                                       b=np.arange(1, S.shape[1] + 1), a=np.array(ports)))
