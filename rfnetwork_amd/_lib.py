"""ctypes binding of librfdtd.so (the HIP engine's C ABI, include/rfdtd.h).

The library is built in-tree by ``rfnetwork_amd.build`` (``python -m rfnetwork_amd.build``).  There
is no fallback: if the library is missing or fails to load, ``lib()`` raises, so no solve can
silently run on anything but the HIP kernels.
"""
import ctypes
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "librfdtd.so"

NUM_COEFFS = 30
NUM_FIELDS = 18
MAX_MONITORS = 50
MAX_PROBES = 5000

F64 = 0x1
COEF_RAW = 0x2
COEF_TABLE = 0x4
NO_GRAPHS = 0x8
QUIET = 0x10
KERNEL_TIMING = 0x20
VERBOSE = 0x40
HOST_TABLE = 0x80
TUNE_SEPARATE = 0x100  # the autotuner tries the separate E and H kernels only
MONITOR_DFT64 = 0x1  # rfdtd_monitor.flags: running DFT accumulated in complex128

# Coefficient keys in the order of the C ABI (rfnetwork/solver.py:1473-1518).
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
# component (0 ex .. 5 hz) of each field array
FIELD_COMPONENT = (0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 0, 1, 2, 3, 4, 5)

# the C entry points declared in include/rfdtd.h
EXPORTS = (
    "rfdtd_last_error", "rfdtd_version", "rfdtd_solver_run", "rfdtd_create", "rfdtd_set_probes",
    "rfdtd_set_monitors", "rfdtd_run", "rfdtd_synchronize", "rfdtd_read_probes", "rfdtd_read_monitors",
    "rfdtd_read_fields", "rfdtd_reset", "rfdtd_step", "rfdtd_last_timing", "rfdtd_memory_info",
    "rfdtd_set_launch", "rfdtd_destroy", "rfdtd_copy_bandwidth", "rfdtd_create_batch",
    "rfdtd_instance_count", "rfdtd_set_instance_sources", "rfdtd_read_instance_fields", "rfdtd_set_variant",
    "rfdtd_set_checkpointing", "rfdtd_gradient", "rfdtd_autotune", "rfdtd_gradient_batch",
    "rfdtd_plan_checkpoints", "rfdtd_nf2ff", "rfdtd_set_history", "rfdtd_device_memory",
    "rfdtd_set_adjoint_launch", "rfdtd_get_adjoint_launch", "rfdtd_set_design_permittivity",
    "rfdtd_set_monitor_streaming", "rfdtd_host_alloc", "rfdtd_host_free",
    "rfdtd_create_assembled", "rfdtd_read_coefficients", "rfdtd_create_batch_assembled",
    "rfdtd_read_coefficient_box", "rfdtd_nf2ff_static", "rfdtd_design_prepare", "rfdtd_grid_barrier_cost",
    "rfdtd_set_column_order", "rfdtd_get_column_order",
    "rfdtd_set_design_permittivity_device",
)


class Probe(ctypes.Structure):
    _fields_ = [
        ("field", ctypes.c_int32),
        ("idx", ctypes.c_int32 * 3),
        ("is_source", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("values", ctypes.c_void_p),
    ]


class Monitor(ctypes.Structure):
    _fields_ = [
        ("axis", ctypes.c_int32),
        ("position", ctypes.c_int32),
        ("field", ctypes.c_int32),
        ("n_step", ctypes.c_int32),
        ("n_frequencies", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("values", ctypes.c_void_p),
        ("dtft_phase", ctypes.c_void_p),
    ]


class PmlSide(ctypes.Structure):  # rfdtd_pml_side
    _fields_ = [("axis", ctypes.c_int32), ("upper", ctypes.c_int32), ("n", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("sigma_edge", ctypes.c_void_p), ("sigma_half", ctypes.c_void_p)]


class EdgeFix(ctypes.Structure):  # rfdtd_edge_fix
    _fields_ = [("component", ctypes.c_int32), ("kind", ctypes.c_int32), ("wide", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("count", ctypes.c_int64), ("cells", ctypes.c_void_p),
                ("values", ctypes.c_void_p)]


class HScale(ctypes.Structure):  # rfdtd_h_scale
    _fields_ = [("coef", ctypes.c_int32), ("wide", ctypes.c_int32), ("cell", ctypes.c_int64),
                ("factor", ctypes.c_double)]


class Assembly(ctypes.Structure):  # rfdtd_assembly
    _fields_ = [("Nx", ctypes.c_int32), ("Ny", ctypes.c_int32), ("Nz", ctypes.c_int32), ("n_pml", ctypes.c_int32),
                ("eps", ctypes.c_void_p), ("sigma", ctypes.c_void_p), ("width", ctypes.c_void_p * 3),
                ("inv_dual", ctypes.c_void_p * 3), ("inv_width", ctypes.c_void_p * 3), ("dt", ctypes.c_double),
                ("e0", ctypes.c_double), ("u0", ctypes.c_double), ("pml", ctypes.c_void_p),
                ("n_fix", ctypes.c_int32), ("n_scale", ctypes.c_int32), ("fix", ctypes.c_void_p),
                ("scale", ctypes.c_void_p)]


class RfdtdError(RuntimeError):
    """A failed call into librfdtd (device error or invalid state)."""


_lib = None


def device_memory():
    """(free, total) device memory in bytes."""
    f, t = ctypes.c_uint64(), ctypes.c_uint64()
    check(lib().rfdtd_device_memory(ctypes.byref(f), ctypes.byref(t)), "rfdtd_device_memory")
    return f.value, t.value


def version() -> int:
    """Version number of the loaded engine library."""
    return int(lib().rfdtd_version())


def lib() -> ctypes.CDLL:
    """Load librfdtd.so once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m rfnetwork_amd.build` (hipcc --offload-arch=gfx950)"
        )
    L = ctypes.CDLL(str(LIB_PATH), mode=ctypes.RTLD_GLOBAL)
    vp, i32, u32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32, ctypes.c_int64
    L.rfdtd_last_error.restype = ctypes.c_char_p
    L.rfdtd_version.restype = i32
    L.rfdtd_solver_run.argtypes = [vp, vp, i32, vp, i32, i32, i32, i32, i32, i32, u32]
    L.rfdtd_create.argtypes = [ctypes.POINTER(vp), i32, i32, i32, vp, u32]
    L.rfdtd_set_probes.argtypes = [vp, vp, i32, i32]
    L.rfdtd_set_monitors.argtypes = [vp, vp, i32, i32]
    L.rfdtd_run.argtypes = [vp, i32, vp]
    L.rfdtd_synchronize.argtypes = [vp]
    L.rfdtd_read_probes.argtypes = [vp]
    L.rfdtd_read_monitors.argtypes = [vp]
    L.rfdtd_read_fields.argtypes = [vp, vp]
    L.rfdtd_reset.argtypes = [vp]
    L.rfdtd_step.argtypes = [vp]
    L.rfdtd_step.restype = i64
    L.rfdtd_last_timing.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double)]
    L.rfdtd_memory_info.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(i32),
                                    ctypes.POINTER(i64), ctypes.POINTER(i64)]
    L.rfdtd_set_launch.argtypes = [vp, i32, i32, i32]
    L.rfdtd_destroy.argtypes = [vp]
    L.rfdtd_destroy.restype = None
    L.rfdtd_create_batch.argtypes = [ctypes.POINTER(vp), i32, vp, vp, vp, u32]
    L.rfdtd_instance_count.argtypes = [vp]
    L.rfdtd_instance_count.restype = i32
    L.rfdtd_set_instance_sources.argtypes = [vp, i32, vp, i32, vp, i32, i32]
    L.rfdtd_read_instance_fields.argtypes = [vp, i32, vp]
    L.rfdtd_set_variant.argtypes = [vp, i32]
    L.rfdtd_set_column_order.argtypes = [vp, i32]
    L.rfdtd_get_column_order.argtypes = [vp, ctypes.POINTER(i32)]
    L.rfdtd_set_checkpointing.argtypes = [vp, i32]
    L.rfdtd_gradient.argtypes = [vp, i32, vp, vp, i32, vp]
    L.rfdtd_autotune.argtypes = [vp, i32, vp, ctypes.POINTER(ctypes.c_double)]
    L.rfdtd_gradient_batch.argtypes = [vp, i32, vp, vp, vp, vp, vp]
    L.rfdtd_set_history.argtypes = [vp, i32, vp, vp, i32]
    L.rfdtd_device_memory.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
    L.rfdtd_nf2ff.argtypes = [i32, vp, i32, vp, i32, vp, i32, vp, ctypes.c_float, vp]
    L.rfdtd_nf2ff_static.argtypes = [i32, vp, i32, vp, i32, vp, i32, vp, ctypes.c_float, vp]
    L.rfdtd_plan_checkpoints.argtypes = [vp, i32, i32, vp, vp, ctypes.c_uint64, ctypes.POINTER(i32),
                                         ctypes.POINTER(ctypes.c_uint64)]
    L.rfdtd_copy_bandwidth.argtypes = [ctypes.c_uint64, i32, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]
    L.rfdtd_grid_barrier_cost.argtypes = [i32, i32, i32, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double)]
    L.rfdtd_create_assembled.argtypes = [ctypes.POINTER(vp), vp, i32, u32]
    L.rfdtd_read_coefficients.argtypes = [vp, i32, vp]
    L.rfdtd_create_batch_assembled.argtypes = [ctypes.POINTER(vp), i32, vp, vp, u32]
    L.rfdtd_read_coefficient_box.argtypes = [vp, i32, vp, vp]
    L.rfdtd_design_prepare.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_double, ctypes.c_double]
    L.rfdtd_set_design_permittivity_device.argtypes = [vp, i32, vp, vp, i32, vp]
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    """Raise on a non-zero status: ValueError for invalid arguments, RfdtdError otherwise."""
    if rc == 0:
        return
    msg = lib().rfdtd_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == 1:
        raise ValueError(text)
    raise RfdtdError(text)


def pointer_array(arrays) -> "ctypes.Array":
    """void*[n] holding the data pointers of numpy arrays (None -> NULL)."""
    arr = (ctypes.c_void_p * len(arrays))()
    for k, a in enumerate(arrays):
        arr[k] = None if a is None else a.ctypes.data
    return arr


class FFFace(ctypes.Structure):
    _fields_ = [
        ("axis", ctypes.c_int32),
        ("n0", ctypes.c_int32),
        ("n1", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("position", ctypes.c_double),
        ("r0", ctypes.c_void_p),
        ("r1", ctypes.c_void_p),
        ("ds", ctypes.c_void_p),
        ("J", ctypes.c_void_p),
        ("M", ctypes.c_void_p),
    ]


def nf2ff(faces, beta, theta, phi, eta0, static_workspace=False):
    """Radiation integrals of far-field box faces on the GPU.  faces: dicts with axis, position,
    r0, r1, ds (float32 [n0, n1]) and J, M (complex64 [3, nf, n0, n1]); beta, theta, phi float32
    (radians).  Returns complex64 [2, nf, ntheta, nphi].  static_workspace: through
    rfdtd_nf2ff_static (no device allocation; requests up to 64 MiB of staging)."""
    import numpy as np

    arr = (FFFace * len(faces))()
    keep = []
    for k, f in enumerate(faces):
        r0 = np.ascontiguousarray(f["r0"], dtype=np.float32)
        r1 = np.ascontiguousarray(f["r1"], dtype=np.float32)
        ds = np.ascontiguousarray(f["ds"], dtype=np.float32)
        J = np.ascontiguousarray(f["J"], dtype=np.complex64)
        M = np.ascontiguousarray(f["M"], dtype=np.complex64)
        if J.shape != M.shape or J.shape[0] != 3 or J.shape[2:] != r0.shape or r1.shape != r0.shape or ds.shape != r0.shape:
            raise ValueError("inconsistent far-field face arrays")
        keep += [r0, r1, ds, J, M]
        arr[k].axis, arr[k].n0, arr[k].n1 = int(f["axis"]), r0.shape[0], r0.shape[1]
        arr[k].position = float(f["position"])
        arr[k].r0, arr[k].r1, arr[k].ds = r0.ctypes.data, r1.ctypes.data, ds.ctypes.data
        arr[k].J, arr[k].M = J.ctypes.data, M.ctypes.data
    beta = np.ascontiguousarray(beta, dtype=np.float32)
    theta = np.ascontiguousarray(theta, dtype=np.float32)
    phi = np.ascontiguousarray(phi, dtype=np.float32)
    if faces and keep[3].shape[1] != len(beta):
        raise ValueError("one beta per frequency of the currents is required")
    out = np.zeros((2, len(beta), len(theta), len(phi)), dtype=np.complex64)
    name = "rfdtd_nf2ff_static" if static_workspace else "rfdtd_nf2ff"
    check(getattr(lib(), name)(len(faces), ctypes.cast(arr, ctypes.c_void_p), len(beta), beta.ctypes.data, len(theta),
                               theta.ctypes.data, len(phi), phi.ctypes.data, float(eta0), out.ctypes.data), name)
    return out


def copy_bandwidth(nbytes: int = 1 << 30, iters: int = 20):
    """(GB/s counting read + write, ms per copy) of a plain device-to-device streaming copy."""
    g, ms = ctypes.c_double(), ctypes.c_double()
    check(lib().rfdtd_copy_bandwidth(int(nbytes), int(iters), ctypes.byref(g), ctypes.byref(ms)),
          "rfdtd_copy_bandwidth")
    return g.value, ms.value


def grid_barrier_cost(workgroups: int = 256, fences: int = 1, iters: int = 2000):
    """(us per device-wide barrier in one cooperative launch, us between dependent graph-replayed
    launches of an empty kernel) for `workgroups` workgroups of 256 threads; fences 0: none, 1: one
    wave per workgroup writes back and invalidates its L2 around the barrier, 2: every wave."""
    b, l = ctypes.c_double(), ctypes.c_double()
    check(lib().rfdtd_grid_barrier_cost(int(workgroups), int(fences), int(iters), ctypes.byref(b), ctypes.byref(l)),
          "rfdtd_grid_barrier_cost")
    return b.value, l.value


def gpu_visible() -> bool:
    """True when HIP reports at least one device (does not initialise torch)."""
    if os.environ.get("HIP_VISIBLE_DEVICES", None) == "":
        return False
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


def as_c(a: np.ndarray, dtype) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=dtype)


class _PinnedOwner:
    """Frees a page-locked host block of this library when the last array over it goes away."""

    def __init__(self, ptr):
        self.ptr = ptr

    def __del__(self):
        try:
            if self.ptr:
                lib().rfdtd_host_free(ctypes.c_void_p(self.ptr))
        except Exception:
            pass


def pinned_empty(shape, dtype):
    """A numpy array in page-locked host memory allocated by the library's runtime."""
    dtype = np.dtype(dtype)
    n = int(np.prod(shape)) * dtype.itemsize
    ptr = ctypes.c_void_p()
    check(lib().rfdtd_host_alloc(ctypes.c_uint64(max(n, 1)), ctypes.byref(ptr)), "rfdtd_host_alloc")
    buf = (ctypes.c_char * max(n, 1)).from_address(ptr.value)
    buf._owner = _PinnedOwner(ptr.value)
    return np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)
