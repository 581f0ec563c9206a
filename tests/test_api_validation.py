"""Argument validation of the drop-in entry point (no GPU call is reached): inputs that crash the
reference (rfnetwork/core/src/solver.cpp:80-190, 465-532) raise TypeError/ValueError here."""
import copy

import numpy as np
import pytest

from rfnetwork_amd import hip_func, workloads


@pytest.fixture(scope="module")
def packed():
    return workloads.tiny_model(Nt=8).pack(gpu=True)


def _args(packed):
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = copy.deepcopy(packed)
    return c, probes, monitors, mem, Nx, Ny, Nz, Nt


def test_type_errors(packed):
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = _args(packed)
    with pytest.raises(TypeError, match="coefficients dictionary"):
        hip_func.pack_call(list(c.values()), probes, monitors, mem, Nx, Ny, Nz, Nt)
    with pytest.raises(TypeError, match="monitors list"):
        hip_func.pack_call(c, probes, tuple(monitors), mem, Nx, Ny, Nz, Nt)
    with pytest.raises(TypeError, match="probes list"):
        hip_func.pack_call(c, tuple(probes), monitors, mem, Nx, Ny, Nz, Nt)


@pytest.mark.parametrize("Nt,shrink", [(0, None), (-3, None), (8, 0), (8, 1), (8, 2)])
def test_step_count_and_grid_size(packed, Nt, shrink):
    """No time step, or an axis of fewer than 2 cells, is refused before any array is read."""
    c, probes, monitors, mem, Nx, Ny, Nz, _ = _args(packed)
    dims = [Nx, Ny, Nz]
    if shrink is not None:
        dims[shrink] = 1
    with pytest.raises(ValueError, match="grid size or number of time steps"):
        hip_func.pack_call(c, probes, monitors, mem, *dims, Nt)


@pytest.mark.parametrize("breakage", ["dtype", "shape", "order", "missing", "ndim"])
def test_coefficient_array_checks(packed, breakage):
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = _args(packed)
    a = c["Cb_ey_x"]
    if breakage == "dtype":
        c["Cb_ey_x"] = a.astype(np.float64)
    elif breakage == "shape":
        c["Cb_ey_x"] = np.zeros((Nx, Ny + 1, Nz), np.float32)
    elif breakage == "order":
        c["Cb_ey_x"] = np.asfortranarray(a)
    elif breakage == "missing":
        del c["Cb_ey_x"]
    else:
        c["Cb_ey_x"] = a.reshape(Nx, -1)
    with pytest.raises(ValueError):
        hip_func.pack_call(c, probes, monitors, mem, Nx, Ny, Nz, Nt)


def test_probe_and_monitor_checks(packed):
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = _args(packed)
    short = copy.deepcopy(probes)
    short[0]["values"] = short[0]["values"][: Nt - 1].copy()
    with pytest.raises(ValueError, match="greater or equal"):
        hip_func.pack_call(c, short, monitors, mem, Nx, Ny, Nz, Nt)
    many = [copy.deepcopy(probes[0]) for _ in range(5001)]
    with pytest.raises(ValueError, match="maximum number of probes"):
        hip_func.pack_call(c, many, monitors, mem, Nx, Ny, Nz, Nt)
    with pytest.raises(ValueError, match="maximum number of monitors"):
        hip_func.pack_call(c, probes, [monitors[0]] * 51, mem, Nx, Ny, Nz, Nt)
    bad = copy.deepcopy(monitors)
    bad[0]["values"] = np.zeros((1, 2, 3), np.float32)
    with pytest.raises(ValueError):
        hip_func.pack_call(c, probes, bad, mem, Nx, Ny, Nz, Nt)
    with pytest.raises(ValueError, match="1 dimension"):
        hip_func.pack_call(c, probes, monitors, np.zeros((2, 2), np.float32), Nx, Ny, Nz, Nt)


def test_probe_grid_edge_rejected(packed):
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = _args(packed)
    probes[0]["field"], probes[0]["idx"] = 2, [0, 3, 3]  # Ez at x = 0 lies on the grid edge
    with pytest.raises(ValueError, match="grid edge"):
        hip_func.pack_call(c, probes, monitors, mem, Nx, Ny, Nz, Nt)
    probes[0]["idx"] = [Nx + 1, 3, 3]
    with pytest.raises(ValueError, match="outside the grid"):
        hip_func.pack_call(c, probes, monitors, mem, Nx, Ny, Nz, Nt)


def test_pec_planes_zeroed_in_caller_arrays(packed):
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = _args(packed)
    assert np.any(c["Ca_ey_z"][-1] != 0)
    hip_func.pack_call(c, probes, monitors, mem, Nx, Ny, Nz, Nt)
    for k in ("Cb_ey_z", "Ca_ey_x", "Cb_ez_x", "Ca_ez_y", "Db_hx_y1", "Da_hx_z"):
        assert not np.any(c[k][-1]), k
    assert np.any(c["Ca_ex_y"][-1] != 0)


def test_double_accumulated_dft_monitors():
    """dft_dtype=complex128 in a float32 solve gives complex128 values and phase tables, packed with
    the double-accumulation flag; mixed dtypes are rejected; complex64 stays the default."""
    from rfnetwork_amd import _lib
    m = workloads.tiny_model(Nt=8, time_monitors=False, phasor_monitors=False)
    m.add_field_monitor("p32", "ez", "x", index=6, frequency=[1e10, 2e10])
    m.add_field_monitor("p64", "ez", "x", index=6, frequency=[1e10, 2e10], dft_dtype=np.complex128)
    with pytest.raises(ValueError, match="dft_dtype"):
        m.add_field_monitor("bad", "ez", "x", index=6, frequency=[1e10], dft_dtype=np.float64)
    c, probes, monitors, mem, Nx, Ny, Nz, Nt = m.pack(gpu=True)
    assert monitors[0]["values"].dtype == np.complex64 and monitors[0]["dtft_phase"].dtype == np.complex64
    assert monitors[1]["values"].dtype == np.complex128 and monitors[1]["dtft_phase"].dtype == np.complex128
    # the double table is the complex64 one evaluated in double
    np.testing.assert_allclose(monitors[1]["dtft_phase"], monitors[0]["dtft_phase"], rtol=0, atol=1e-6)
    recs, _ = hip_func.pack_monitors(monitors, Nx, Ny, Nz, Nt)
    assert recs[0].flags == 0 and recs[1].flags == _lib.MONITOR_DFT64
    recs, _ = hip_func.pack_monitors(m.pack(gpu=True, dtype=np.float64, coefficients=False)[2], Nx, Ny, Nz, Nt,
                                     dtype=np.float64)
    assert recs[0].flags == 0 and recs[1].flags == 0  # a float64 solve accumulates in double anyway
    mixed = copy.deepcopy(monitors)
    mixed[1]["dtft_phase"] = mixed[1]["dtft_phase"].astype(np.complex64)
    with pytest.raises(ValueError, match="dtft_phase"):
        hip_func.pack_monitors(mixed, Nx, Ny, Nz, Nt)
