"""Resident GPU sessions over the session half of the C ABI (include/rfdtd.h).

A ``Session`` keeps the fields and coefficients of one grid in HBM between calls: set the probes and
monitors once, advance any number of steps, read probes, monitors or the full 18-array field state.
Arguments use the reference's calling convention (the GPU packing of rfnetwork/solver.py:1463-1523,
probe and monitor dictionaries of :1541-1622), validated by ``hip_func.pack_call``.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import FIELD_COMPONENT, FIELD_NAMES, check


def yee_shape(component, Nx, Ny, Nz):
    """Global Yee shape of field component 0..5 (ex ey ez hx hy hz)."""
    return [(Nx, Ny + 1, Nz + 1), (Nx + 1, Ny, Nz + 1), (Nx + 1, Ny + 1, Nz),
            (Nx + 1, Ny, Nz), (Nx, Ny + 1, Nz), (Nx, Ny, Nz + 1)][component]


class Session:
    """One grid resident on the GPU."""

    def __init__(self, coefficients, Nx, Ny, Nz, dtype=np.float32, flags=0):
        from .hip_func import _check_array, _real_dtype  # local: hip_func imports this module's peers

        self.dtype = _real_dtype(dtype)
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        arrays = []
        for name in _lib.COEFF_NAMES:
            a = coefficients[name]
            _check_array(a, name, (self.Nx, self.Ny, self.Nz), self.dtype)
            arrays.append(a)
        self._L = _lib.lib()
        self._h = ctypes.c_void_p()
        f = int(flags) | (_lib.F64 if self.dtype == np.float64 else 0)
        ptrs = _lib.pointer_array(arrays)
        check(self._L.rfdtd_create(ctypes.byref(self._h), self.Nx, self.Ny, self.Nz,
                                   ctypes.cast(ptrs, ctypes.c_void_p), f), "rfdtd_create")
        self._keep = []
        self.Nt = None
        self.flags = int(flags)

    @classmethod
    def assembled(cls, model, dtype=np.float32, flags=0):
        """A session of a meshed model whose 30 coefficient arrays are computed on the GPU from the
        model's description (cell permittivity and conductivity, widths, PML profiles, conductor and
        lumped-element edges, edge corrections) instead of being built on the host and copied: equal
        to the host assembly bit for bit (rfdtd_create_assembled)."""
        from . import assembly
        from .hip_func import _real_dtype

        self = cls.__new__(cls)
        self.dtype = _real_dtype(dtype)
        self.Nx, self.Ny, self.Nz = (int(n) for n in model.n_cells)
        self._L, self._h = assembly.create(model, self.dtype, flags)
        self._keep = []
        self.Nt = None
        self.flags = int(flags)
        return self

    def coefficients(self, box=None):
        """The session's 30 packed coefficient arrays read back from the GPU (from its table or raw
        arrays; the x = Nx-1 plane of arrays 4..17 reads zero, as the kernels see it), whole or over
        box = (x0, x1, y0, y1, z0, z1) packed cells."""
        from . import assembly

        if box is not None:
            return assembly.read_coefficient_box(self._L, self._h, box, self.dtype)
        return assembly.read_coefficients(self._L, self._h, (self.Nx, self.Ny, self.Nz), self.dtype)

    def set_monitor_streaming(self, ring_samples):
        """Time-domain monitors installed afterwards keep ring_samples sample planes on the GPU and
        stream the rest to their (page-locked) host arrays during the run; 0 turns it off."""
        check(self._L.rfdtd_set_monitor_streaming(self._h, int(ring_samples)), "rfdtd_set_monitor_streaming")
        self._stream_ring = int(ring_samples)

    def set_sources(self, probes, monitors, Nt):
        """Install probe/source records and monitors (dictionaries as in the reference's solve()).
        With monitor streaming on, time-domain monitor arrays are moved into pinned host memory."""
        from .hip_func import pack_monitors, pack_probes, pin_time_monitors

        if getattr(self, "_stream_ring", 0):
            pin_time_monitors(monitors)

        prb, keep_p = pack_probes(probes, int(Nt), self.dtype)
        mons, keep_m = pack_monitors(monitors, self.Nx, self.Ny, self.Nz, int(Nt), self.dtype)
        keep = keep_p + keep_m
        check(self._L.rfdtd_set_probes(self._h, ctypes.cast(prb, ctypes.c_void_p), len(probes), int(Nt)),
              "rfdtd_set_probes")
        check(self._L.rfdtd_set_monitors(self._h, ctypes.cast(mons, ctypes.c_void_p), len(monitors), int(Nt)),
              "rfdtd_set_monitors")
        self._keep = [prb, mons, keep]
        self.probes, self.monitors, self.Nt = probes, monitors, int(Nt)

    def run(self, n_steps, stream=None):
        """Advance n_steps (asynchronous on the session stream or a given HIP stream handle)."""
        s = None if stream is None else ctypes.c_void_p(int(stream))
        check(self._L.rfdtd_run(self._h, int(n_steps), s), "rfdtd_run")

    def synchronize(self):
        check(self._L.rfdtd_synchronize(self._h), "rfdtd_synchronize")

    def read(self):
        """Copy probe series and monitor values back into the dictionaries given to set_sources."""
        check(self._L.rfdtd_read_probes(self._h), "rfdtd_read_probes")
        check(self._L.rfdtd_read_monitors(self._h), "rfdtd_read_monitors")

    def fields(self, names=FIELD_NAMES):
        """Dict of Yee-shaped arrays for the requested field names."""
        out = {}
        ptrs = (ctypes.c_void_p * 18)()
        for a, name in enumerate(FIELD_NAMES):
            if name in names:
                out[name] = np.zeros(yee_shape(FIELD_COMPONENT[a], self.Nx, self.Ny, self.Nz), self.dtype)
                ptrs[a] = out[name].ctypes.data
        check(self._L.rfdtd_read_fields(self._h, ctypes.cast(ptrs, ctypes.c_void_p)), "rfdtd_read_fields")
        return out

    def reset(self):
        check(self._L.rfdtd_reset(self._h), "rfdtd_reset")

    @property
    def step(self):
        return int(self._L.rfdtd_step(self._h))

    def timing(self):
        """(total_ms, e_kernel_ms, h_kernel_ms) of the last run (after synchronize)."""
        t, e, h = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(self._L.rfdtd_last_timing(self._h, ctypes.byref(t), ctypes.byref(e), ctypes.byref(h)))
        return t.value, e.value, h.value

    def memory_info(self):
        b, k = ctypes.c_uint64(), ctypes.c_int32()
        ne, nh = ctypes.c_int64(), ctypes.c_int64()
        check(self._L.rfdtd_memory_info(self._h, ctypes.byref(b), ctypes.byref(k), ctypes.byref(ne), ctypes.byref(nh)))
        return dict(device_bytes=b.value, coef_kind=("raw", "table")[k.value], n_e_entries=ne.value,
                    n_h_entries=nh.value)

    def set_launch(self, block_threads=0, x_planes=0, graph_steps=0):
        check(self._L.rfdtd_set_launch(self._h, int(block_threads), int(x_planes), int(graph_steps)))

    def autotune(self, trial_steps=20, use_cache=True):
        """Pick the fastest launch shape for this grid (before the first run); cached per shape.  With
        sources already installed the trials apply them as a run does (then restored), so their
        cost is part of the timing.  Returns {variant, block, x_planes, ms_per_step, cached}."""
        from . import autotune

        kind = self.memory_info()["coef_kind"]
        return autotune.tune(self, [(self.Nx, self.Ny, self.Nz, 1)], np.dtype(self.dtype).name, kind,
                             trial_steps, use_cache)

    def set_checkpointing(self, every):
        """Save the split fields every `every` steps of later runs (needed by gradient())."""
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

    def set_design_permittivity(self, model, cbox, eps_r):
        """Relative permittivity eps_r of the cells of cbox = (x0, x1, y0, y1, z0, z1) of the model the
        session was built from: the E coefficients around the box are recomputed on the GPU (bit for
        bit what the model's own assembly gives for that permittivity).  The box must hold plain
        dielectric edges only (no PML, conductor, lumped element or edge correction); later calls
        start from the permittivity set last.  eps_r: a numpy array, or a torch tensor in device
        memory, which the GPU reads directly (the box is prepared on the device once; while device
        updates follow one another the host's copy of the permittivity is not kept)."""
        from . import design

        if not hasattr(self, "_design_state"):
            self._design_state = {}
            self._design_prepared = None  # the box held on the device
            self._design_stale = False  # the host's record of the permittivity is behind the device's
        key = tuple(int(v) for v in cbox)
        if design.is_device_tensor(eps_r):
            if self._design_prepared != key:
                if self._design_stale:
                    raise ValueError("another box after device updates: use a DesignLoop, which keeps track of "
                                     "the permittivity, to mix boxes and host and device updates")
                design.prepare_design_box(self._h, self._L, 0, model, key, self._design_state, self.dtype)
                self._design_prepared = key
            # the tensor read stays referenced until the next update (the session's work is ordered)
            self._design_keep = design.set_design_permittivity_device(self._h, self._L, 0, key, eps_r)
            self._design_stale = True
            return
        if self._design_stale:
            raise ValueError("the session's permittivity was last set from device memory: use a DesignLoop, "
                             "which keeps track of it, to mix host and device updates")
        design.set_design_permittivity(self._h, self._L, 0, model, cbox, eps_r, self._design_state, self.dtype)
        self._design_prepared = None

    def set_history(self, box, Nt):
        """Record the H totals of box at every step of later runs (up to Nt steps): a gradient over
        that box then needs no checkpoints and no recomputation.  Before the first step only."""
        inst = (ctypes.c_int32 * 1)(0)
        b = (ctypes.c_int32 * 6)(*[int(v) for v in box])
        check(self._L.rfdtd_set_history(self._h, 1, ctypes.cast(inst, ctypes.c_void_p), ctypes.cast(b, ctypes.c_void_p),
                                        int(Nt)), "rfdtd_set_history")
        self._history = True

    def plan_checkpoints(self, Nt, box, budget_bytes=0):
        """Checkpoint interval for a gradient over box after an Nt-step forward run that fits device
        memory (budget_bytes, 0: what is free).  Returns (every, bytes needed)."""
        inst = (ctypes.c_int32 * 1)(0)
        b = (ctypes.c_int32 * 6)(*[int(v) for v in box])
        every, need = ctypes.c_int32(), ctypes.c_uint64()
        check(self._L.rfdtd_plan_checkpoints(self._h, int(Nt), 1, ctypes.cast(inst, ctypes.c_void_p),
                                             ctypes.cast(b, ctypes.c_void_p), int(budget_bytes), ctypes.byref(every),
                                             ctypes.byref(need)), "rfdtd_plan_checkpoints")
        return every.value, need.value

    def gradient(self, box, adjoint_series, instance=0):
        """dJ/dCb of the 6 E splits in box (x0, x1, y0, y1, z0, z1 grid cells) after a checkpointed
        forward run; adjoint_series: per probe, dJ/dy[n] (n < steps run).  Returns float64
        [6, nx, ny, nz] in split order ex_y ex_z ey_z ey_x ez_x ez_y."""
        box = [int(v) for v in box]
        nx, ny, nz = box[1] - box[0], box[3] - box[2], box[5] - box[4]
        out = np.zeros((6, nx, ny, nz), np.float64)
        series = [np.ascontiguousarray(a, dtype=self.dtype) for a in adjoint_series]
        ptrs = _lib.pointer_array(series)
        b = (ctypes.c_int32 * 6)(*box)
        check(self._L.rfdtd_gradient(self._h, int(instance), ctypes.cast(b, ctypes.c_void_p),
                                     ctypes.cast(ptrs, ctypes.c_void_p), len(series), out.ctypes.data_as(ctypes.c_void_p)),
              "rfdtd_gradient")
        from . import autotune

        autotune.remember_adjoint(self)
        return out

    def set_variant(self, variant=1):
        """Kernel variant: 1 separate E and H kernels issuing the next x plane's loads early,
        0 the same one plane at a time, 2 and 3 the fused E+H pass with 6- or 14-row tiles (two field
        buffers per instance), 4 variant 1 with nontemporal field stores, 5 and 6 variants 2 and 3
        with deeper load pipelining, 7 and 8 the two-step pass (steps n and n + 1 of a 12- or 4-row x
        60-cell tile per launch, two field buffers per instance), 9 the two-step pass over 8-row tiles
        with the x planes skewed by three, 10 variant 4 with nontemporal loads of the arrays streamed
        once, 11 and 12 the separate kernels over 2 and 4 z cells per thread (Nz a multiple of 2 or 4),
        13, 14 and 15 the two-step pass over 28 x 28, 12 x 28 and 26 x 28-cell tiles of 32-lane rows.
        Results are identical.  A new session runs variant 15 (automatic x chunks) when it is float,
        on the coefficient table, with at least 1.5 Mi cell-updates per step, else variant 1."""
        check(self._L.rfdtd_set_variant(self._h, int(variant)))

    def set_column_order(self, order):
        """Workgroup numbering of the separate kernels: 0 contiguous tile ranges per XCD, 1 launch
        order (faster on large grids).  Results are identical."""
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

