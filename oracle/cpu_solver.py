"""TEST INFRASTRUCTURE ONLY -- Python entry point of the CPU oracle.

``solver_run(coefficients, probes, monitors, mem, Nx, Ny, Nz, Nt, n_threads, update_interval)``
mirrors ``rfnetwork.core.core_func.solver_run`` (rfnetwork/core/src/core_binding.cpp:350-394): the
CPU packing of the coefficients (rfnetwork/solver.py:1463-1523 with gpu=False), in-place probe and
monitor outputs, and the scratch array ``mem`` that holds the slab fields.  The numerical work is
done by liboracle_fdtd.so (oracle/fdtd_oracle.cpp), a plain-loop restatement of
rfnetwork/core/src/solver.cpp.

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may import this module.
"""
import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle_fdtd.so"

COEFF_NAMES = (
    "Ca_ex_y", "Ca_ex_z", "Cb_ex_y", "Cb_ex_z",
    "Ca_ey_z", "Ca_ey_x", "Cb_ey_z", "Cb_ey_x",
    "Ca_ez_x", "Ca_ez_y", "Cb_ez_x", "Cb_ez_y",
    "Da_hx_y", "Da_hx_z", "Db_hx_y1", "Db_hx_y2", "Db_hx_z1", "Db_hx_z2",
    "Da_hy_z", "Da_hy_x", "Db_hy_z1", "Db_hy_z2", "Db_hy_x1", "Db_hy_x2",
    "Da_hz_x", "Da_hz_y", "Db_hz_x1", "Db_hz_x2", "Db_hz_y1", "Db_hz_y2",
)
FIELD_NAMES = (
    "ex_y", "ex_z", "ey_z", "ey_x", "ez_x", "ez_y",
    "hx_y", "hx_z", "hy_z", "hy_x", "hz_x", "hz_y",
    "ex", "ey", "ez", "hx", "hy", "hz",
)
FIELD_COMPONENT = (0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 0, 1, 2, 3, 4, 5)


class _Probe(ctypes.Structure):
    _fields_ = [("field", ctypes.c_int32), ("idx", ctypes.c_int32 * 3), ("is_source", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("values", ctypes.c_void_p)]


class _Monitor(ctypes.Structure):
    _fields_ = [("axis", ctypes.c_int32), ("position", ctypes.c_int32), ("field", ctypes.c_int32),
                ("n_step", ctypes.c_int32), ("n_frequencies", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("values", ctypes.c_void_p), ("dtft_phase", ctypes.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} not built; run `make -C oracle`")
        L = ctypes.CDLL(str(LIB_PATH))
        vp, i32, u64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint64
        for fn in (L.oracle_run_f32, L.oracle_run_f64):
            fn.argtypes = [vp, vp, i32, vp, i32, vp, u64, i32, i32, i32, i32, i32, i32, vp]
            fn.restype = i32
        L.oracle_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def coefficient_shapes(Nx, Ny, Nz):
    """CPU-packing shapes of the 30 coefficient arrays (solver.cpp:247-346 with gpu = 0)."""
    ex, ey, ez = (Nx, Ny - 1, Nz - 1), (Nx, Ny, Nz - 1), (Nx, Ny - 1, Nz)
    hx, hy, hz = (Nx, Ny, Nz), (Nx, Ny + 1, Nz), (Nx, Ny, Nz + 1)
    return dict(zip(COEFF_NAMES, [ex] * 4 + [ey] * 4 + [ez] * 4 + [hx] * 6 + [hy] * 6 + [hz] * 6))


def global_field_shape(component, Nx, Ny, Nz):
    """Yee shape of a field component (rfnetwork/solver.py:685-692)."""
    return [(Nx, Ny + 1, Nz + 1), (Nx + 1, Ny, Nz + 1), (Nx + 1, Ny + 1, Nz),
            (Nx + 1, Ny, Nz), (Nx, Ny + 1, Nz), (Nx, Ny, Nz + 1)][component]


def monitor_plane_shape(axis, field, Nx, Ny, Nz):
    """(N1, N2) of a monitor plane in the CPU packing (solver.cpp:367-441)."""
    f_ny = [Ny + 1, Ny, Ny + 1, Ny, Ny + 1, Ny][field]
    f_nz = [Nz + 1, Nz + 1, Nz, Nz, Nz, Nz + 1][field]
    return ((f_ny, f_nz), (Nx, f_nz), (Nx, f_ny))[axis]


def mem_size(Nx, Ny, Nz, n_threads):
    """Scratch floats the reference's solve() allocates (rfnetwork/solver.py:1525-1534)."""
    shapes = [global_field_shape(c, Nx, Ny, Nz) for c in range(6)]
    n = sum(3 * Nx * s[1] * s[2] for s in shapes)
    return n + 4 * (Ny + 1) * (Nz + 1) * n_threads * 8


def solver_run(coefficients, probes, monitors, mem, Nx, Ny, Nz, Nt, n_threads, update_interval,
               dtype=np.float32, fields_out=None):
    """Run the CPU oracle in place.  fields_out: optional dict name -> Yee-shaped array to fill."""
    dtype = np.dtype(dtype)
    cdtype = np.dtype(np.complex64) if dtype == np.float32 else np.dtype(np.complex128)
    if not isinstance(coefficients, dict):
        raise TypeError("Expected a coefficients dictionary")
    if not isinstance(monitors, list) or not isinstance(probes, list):
        raise TypeError("Expected lists of probes and monitors")
    shapes = coefficient_shapes(Nx, Ny, Nz)
    carr = (ctypes.c_void_p * 30)()
    for k, name in enumerate(COEFF_NAMES):
        a = coefficients[name]
        if a.dtype != dtype or not a.flags["C_CONTIGUOUS"] or a.shape != shapes[name]:
            raise ValueError(f"Invalid data array {name}: expected {dtype} C-contiguous {shapes[name]}")
        carr[k] = a.ctypes.data
    if mem.dtype != dtype or mem.ndim != 1 or not mem.flags["C_CONTIGUOUS"]:
        raise ValueError("Invalid memory buffer array.")

    prb = (_Probe * max(1, len(probes)))()
    for p, d in enumerate(probes):
        v = d["values"]
        if v.dtype != dtype or v.ndim != 1 or not v.flags["C_CONTIGUOUS"] or v.shape[0] < Nt:
            raise ValueError("Invalid source array.")
        prb[p].field = int(d["field"])
        for a in range(3):
            prb[p].idx[a] = int(d["idx"][a])
        prb[p].is_source = 1 if "is_source" in d else 0
        prb[p].values = v.ctypes.data
    mons = (_Monitor * max(1, len(monitors)))()
    for m, d in enumerate(monitors):
        n_step, axis, field = int(d["n_step"]), int(d["axis"]), int(d["field"])
        Nm = Nt // n_step + 1
        N1, N2 = monitor_plane_shape(axis, field, Nx, Ny, Nz)
        rec = mons[m]
        rec.axis, rec.position, rec.field, rec.n_step = axis, int(d["position"]), field, n_step
        if "dtft_phase" in d:
            nf = int(d["n_frequencies"])
            ph, v = d["dtft_phase"], d["values"]
            if ph.dtype != cdtype or ph.shape != (Nm, nf) or v.dtype != cdtype or v.shape != (nf, N1, N2):
                raise ValueError("Invalid phasor monitor arrays.")
            rec.n_frequencies = nf
            rec.dtft_phase = ph.ctypes.data
        else:
            v = d["values"]
            if v.dtype != dtype or v.shape != (Nm, N1, N2):
                raise ValueError("Invalid monitor values array.")
            rec.n_frequencies = 0
        rec.values = d["values"].ctypes.data

    fptr = None
    if fields_out is not None:
        fptr = (ctypes.c_void_p * 18)()
        for a, name in enumerate(FIELD_NAMES):
            arr = fields_out.get(name)
            if arr is not None:
                want = global_field_shape(FIELD_COMPONENT[a], Nx, Ny, Nz)
                if arr.shape != want or arr.dtype != dtype or not arr.flags["C_CONTIGUOUS"]:
                    raise ValueError(f"fields_out[{name}] must be {dtype} {want}")
                arr[...] = 0
                fptr[a] = arr.ctypes.data
    fn = lib().oracle_run_f32 if dtype == np.float32 else lib().oracle_run_f64
    rc = fn(ctypes.cast(carr, ctypes.c_void_p), ctypes.cast(prb, ctypes.c_void_p), len(probes),
            ctypes.cast(mons, ctypes.c_void_p), len(monitors), mem.ctypes.data, mem.shape[0],
            int(Nx), int(Ny), int(Nz), int(Nt), int(n_threads), int(update_interval),
            None if fptr is None else ctypes.cast(fptr, ctypes.c_void_p))
    if rc != 0:
        raise ValueError(lib().oracle_last_error().decode())
    return 0
