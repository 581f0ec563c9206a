"""Drop-in replacement of the reference's GPU solver entry point.

``solver_run_hip(coefficients, probes, monitors, mem, Nx, Ny, Nz, Nt, n_threads, update_interval)``
takes exactly the arguments of ``rfnetwork.core.cuda_func.solver_run_cu`` (called at
rfnetwork/solver.py:1636-1641; parsed at rfnetwork/core/src/cuda_binding.cpp:20-68) and has the
same effects:

* coefficients: dict of the 30 float32 C-contiguous (Nx, Ny, Nz) arrays (solver.cpp:80-118 checks);
  like solver.cpp:294-328 the x = Nx-1 planes of the Ey, Ez and Hx coefficients are zeroed in the
  caller's arrays (the PEC wall at x = Nx);
* probes: list of dicts {values, field, idx, is_source}; idx is a global Yee index, shifted per
  component as in solver.cpp:503-520; ``values`` (float32, length >= Nt) holds the source samples on
  entry and the sampled total field on return.  As in the reference, a probe whose dict has an
  ``is_source`` key is injected whatever the key's value (solver.cpp:528 tests the key's presence);
* monitors: list of dicts {axis, position, field, n_step, values[, dtft_phase, n_frequencies]};
  values are filled in place: time samples (Nt/n_step+1, N1, N2) float32 or running DFTs
  (Nf, N1, N2) complex64 with (N1, N2) = (Ny, Nz), (Nx, Nz), (Nx, Ny) for axis 0, 1, 2;
* mem is validated (float32, 1-D, C-contiguous) and otherwise unused, n_threads is ignored, and
  update_interval > 0 prints the reference's progress bar.

Differences, all where the reference crashes: invalid input raises TypeError/ValueError instead of
aborting the process, probe or monitor indices outside the grid are rejected, and HIP errors raise
``RfdtdError``.  Monitors sample after both half steps as in the reference's CPU solver
(solver.cpp:1130), not before the E sources as its CUDA kernels do (solver.cu:196-275).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import COEFF_NAMES, check

_E_Y_PLANE = ("Cb_ey_z", "Cb_ey_x", "Ca_ey_z", "Ca_ey_x")
_E_Z_PLANE = ("Cb_ez_x", "Cb_ez_y", "Ca_ez_x", "Ca_ez_y")
_H_X_PLANE = ("Db_hx_y1", "Db_hx_y2", "Db_hx_z1", "Db_hx_z2", "Da_hx_y", "Da_hx_z")


def _real_dtype(dtype):
    dtype = np.dtype(dtype)
    if dtype not in (np.dtype(np.float32), np.dtype(np.float64)):
        raise ValueError("dtype must be float32 or float64")
    return dtype


def _check_array(a, name, shape, dtype):
    if not isinstance(a, np.ndarray):
        raise ValueError(f"Invalid data array {name}. Expected a numpy array.")
    if a.ndim != len(shape):
        raise ValueError(f"Invalid data array {name}. Wrong number of dimensions.")
    if a.dtype != dtype:
        raise ValueError(f"Invalid data array {name}. Must be {dtype} type.")
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError(f"Invalid data array {name}. Must be row ordered (C-style)")
    for ax, (got, want) in enumerate(zip(a.shape, shape)):
        if got != want:
            raise ValueError(f"Invalid data array {name}. Expected shape on axis {ax} of {want}")


def monitor_plane_shape(axis: int, Nx: int, Ny: int, Nz: int):
    """(N1, N2) of a monitor plane in the GPU packing (solver.cpp:377-460 with gpu = 1)."""
    return ((Ny, Nz), (Nx, Nz), (Nx, Ny))[axis]


def pack_monitors(monitors, Nx, Ny, Nz, Nt, dtype=np.float32):
    """Validate monitor dictionaries (solver.py:1597-1620) and build the C records."""
    dtype = _real_dtype(dtype)
    cdtype = np.dtype(np.complex64) if dtype == np.float32 else np.dtype(np.complex128)
    if not isinstance(monitors, list):
        raise TypeError("Expected a monitors list")
    if len(monitors) > _lib.MAX_MONITORS:
        raise ValueError("Exceeded maximum number of monitors.")
    mons = (_lib.Monitor * max(1, len(monitors)))()
    keep = []
    for m, d in enumerate(monitors):
        if not isinstance(d, dict):
            raise TypeError("Expected a list of monitor dictionaries")
        n_step, axis, field = int(d["n_step"]), int(d["axis"]), int(d["field"])
        if n_step < 1 or axis not in (0, 1, 2) or field not in range(6):
            raise ValueError("Invalid monitor configuration.")
        Nm = Nt // n_step + 1
        N1, N2 = monitor_plane_shape(axis, Nx, Ny, Nz)
        rec = mons[m]
        rec.axis, rec.field, rec.n_step = axis, field, n_step
        rec.position = int(d["position"])
        if "dtft_phase" in d:
            nf = int(d["n_frequencies"])
            # complex128 arrays in a float32 solve: the running DFT is accumulated in double
            wide = dtype == np.float32 and getattr(d["values"], "dtype", None) == np.complex128
            mdtype = np.dtype(np.complex128) if wide else cdtype
            _check_array(d["dtft_phase"], "dtft_phase", (Nm, nf), mdtype)
            _check_array(d["values"], "values", (nf, N1, N2), mdtype)
            rec.n_frequencies = nf
            rec.flags = _lib.MONITOR_DFT64 if wide else 0
            rec.dtft_phase = d["dtft_phase"].ctypes.data
        else:
            _check_array(d["values"], "values", (Nm, N1, N2), dtype)
            rec.n_frequencies = 0
        rec.values = d["values"].ctypes.data
        keep.append(d["values"])
    return mons, keep


def pin_time_monitors(monitors):
    """Move the value arrays of the time-domain monitors into page-locked host memory of the
    library's runtime (numpy arrays that free it when they go), so that streamed monitor copies run
    asynchronously into them."""
    for d in monitors:
        if "dtft_phase" in d:
            continue
        a = np.asarray(d["values"])
        v = _lib.pinned_empty(a.shape, a.dtype)
        v[...] = a
        d["values"] = v


def pack_probes(probes, Nt, dtype=np.float32):
    """Validate probe dictionaries (solver.py:1541-1575; solver.cpp:132-170) and build the C records."""
    dtype = _real_dtype(dtype)
    if not isinstance(probes, list):
        raise TypeError("Expected a probes list")
    if len(probes) > _lib.MAX_PROBES:
        raise ValueError("Exceeded maximum number of probes.")
    prb = (_lib.Probe * max(1, len(probes)))()
    keep = []
    for p, d in enumerate(probes):
        if not isinstance(d, dict):
            raise ValueError("Expected a list of probe dictionaries")
        v = d["values"]
        if not isinstance(v, np.ndarray) or v.ndim != 1:
            raise ValueError("Invalid source array. Wrong number of dimensions.")
        if v.dtype != dtype:
            raise ValueError("Invalid data array. Must be float type.")
        if not v.flags["C_CONTIGUOUS"]:
            raise ValueError("Invalid source array. Must be row ordered (C-style)")
        if v.shape[0] < Nt:
            raise ValueError(f"Invalid data source array shape. Expected shape greater or equal to {Nt}")
        idx = [int(i) for i in d["idx"]]
        if len(idx) != 3:
            raise ValueError("Probe index must have three entries.")
        rec = prb[p]
        rec.field = int(d["field"])
        rec.idx[0], rec.idx[1], rec.idx[2] = idx
        rec.is_source = 1 if "is_source" in d else 0
        rec.values = v.ctypes.data
        keep.append(v)
    return prb, keep


def check_probe_cells(probes, Nx, Ny, Nz):
    """Component shifts of solver.cpp:503-520 from global Yee indices to grid cells, with the
    reference's grid-edge error and an upper bound the reference does not check."""
    for d in probes:
        f = int(d["field"])
        x, y, z = (int(i) for i in d["idx"])
        x -= f in (1, 2, 3)
        y -= f in (0, 2, 4)
        z -= f in (0, 1, 5)
        if x < 0 or y < 0 or z < 0:
            raise ValueError("Source cannot be placed on grid edge")
        if x >= Nx or y >= Ny or z >= Nz:
            raise ValueError("probe index outside the grid")


def pack_call(coefficients, probes, monitors, mem, Nx, Ny, Nz, Nt, dtype=np.float32):
    """Validate the reference-format arguments and build the C ABI records.

    Returns (coefficient pointer array, probe records, monitor records, keep-alive list)."""
    dtype = _real_dtype(dtype)
    if not isinstance(coefficients, dict):
        raise TypeError("Expected a coefficients dictionary")
    if not isinstance(monitors, list):
        raise TypeError("Expected a monitors list")
    if not isinstance(probes, list):
        raise TypeError("Expected a probes list")
    Nx, Ny, Nz, Nt = int(Nx), int(Ny), int(Nz), int(Nt)
    if min(Nx, Ny, Nz) < 2 or Nt < 1:
        raise ValueError("Invalid grid size or number of time steps.")
    if not isinstance(mem, np.ndarray) or mem.dtype != dtype:
        raise ValueError("Invalid data array. Must be float type.")
    if not mem.flags["C_CONTIGUOUS"]:
        raise ValueError("Invalid memory data array. Must be row ordered (C-style)")
    if mem.ndim != 1:
        raise ValueError("Invalid memory buffer array. Must be 1 dimension.")
    coeff = []
    for name in COEFF_NAMES:
        if name not in coefficients:
            raise ValueError(f"Missing coefficient array {name}.")
        a = coefficients[name]
        _check_array(a, name, (Nx, Ny, Nz), dtype)
        coeff.append(a)
    # PEC wall at x = Nx: the reference zeroes these planes in the caller's arrays
    for name in _E_Y_PLANE + _E_Z_PLANE + _H_X_PLANE:
        coefficients[name][Nx - 1] = 0
    mons, keep_m = pack_monitors(monitors, Nx, Ny, Nz, Nt, dtype)
    prb, keep_p = pack_probes(probes, Nt, dtype)
    check_probe_cells(probes, Nx, Ny, Nz)
    return _lib.pointer_array(coeff), prb, mons, keep_m + keep_p + coeff


def solver_run_hip(coefficients, probes, monitors, mem, Nx, Ny, Nz, Nt, n_threads=0, update_interval=0,
                   dtype=np.float32, flags=0):
    """Run the FDTD time loop on the GPU; same arguments and in-place outputs as solver_run_cu."""
    dtype = _real_dtype(dtype)
    cptr, prb, mons, keep = pack_call(coefficients, probes, monitors, mem, Nx, Ny, Nz, Nt, dtype)
    L = _lib.lib()
    f = int(flags) | (_lib.F64 if dtype == np.float64 else 0)
    rc = L.rfdtd_solver_run(
        ctypes.cast(cptr, ctypes.c_void_p), ctypes.cast(prb, ctypes.c_void_p), len(probes),
        ctypes.cast(mons, ctypes.c_void_p), len(monitors), int(Nx), int(Ny), int(Nz), int(Nt),
        int(update_interval), f,
    )
    del keep
    check(rc, "solver_run_hip")
    return 0
