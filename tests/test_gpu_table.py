"""The coefficient table built on the device (the default) against the one built on the host
(RFDTD_HOST_TABLE): the same number of E and H entries, and runs identical bit for bit; the
table-or-raw decision follows the same entry limit (cells / 8, or any count with RFDTD_COEF_TABLE).
The device pass hashes every cell's tuple, radix-sorts (hash, cell), splits runs of bit-equal
tuples and numbers them by first cell, so each cell gets the entry number the host pass gives it."""
import numpy as np
import pytest

from rfnetwork_amd import _lib, workloads
from rfnetwork_amd.session import Session

pytestmark = pytest.mark.gpu


def _run(model, flags, dtype=np.float32):
    c, probes, monitors, _, Nx, Ny, Nz, Nt = model.pack(gpu=True, dtype=dtype)
    with Session(c, Nx, Ny, Nz, dtype=dtype, flags=flags) as s:
        info = s.memory_info()
        s.set_sources(probes, monitors, Nt)
        s.run(Nt)
        s.synchronize()
        s.read()
        f = s.fields()
    return info, [np.array(p["values"]) for p in probes], [np.array(m["values"]) for m in monitors], f


def _assert_same_run(a, b):
    for x, y in zip(a[1], b[1]):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(a[2], b[2]):
        np.testing.assert_array_equal(x, y)
    for k in a[3]:
        np.testing.assert_array_equal(a[3][k], b[3][k], err_msg=k)


@pytest.mark.parametrize("name,dtype", [("tiny", np.float32), ("tiny", np.float64), ("ports", np.float32),
                                        ("design", np.float32)])
def test_device_table_equals_host_table(hip_ready, name, dtype):
    print(f"[case] device table {name} {np.dtype(dtype).name}", flush=True)
    if name == "tiny":
        model = workloads.tiny_model(Nt=40)
    elif name == "ports":
        model = workloads.port_variant(1, Nt=40)
    else:
        model = workloads.design_strip(np.random.default_rng(3).uniform(1.0, 4.0, size=(8, 8, 6)), Nt=40, nx=48,
                                       ny=32, nz=20, design=(8, 8, 6), n_pml=5, h=3, pulse=(20e9, 80e-12, 60e-12))
    dev = _run(model, 0, dtype)
    host = _run(model, _lib.HOST_TABLE, dtype)
    assert dev[0]["coef_kind"] == host[0]["coef_kind"] == "table"
    assert dev[0]["n_e_entries"] == host[0]["n_e_entries"] and dev[0]["n_h_entries"] == host[0]["n_h_entries"]
    _assert_same_run(dev, host)


def test_entry_limit_and_forced_table(hip_ready):
    """Random permittivity in every cell: more distinct E tuples than cells / 8 (and than the floor
    of 4096 entries), so the arrays stay raw by default; RFDTD_COEF_TABLE forces the table (one
    entry per distinct tuple).  All three runs are identical."""
    print("[case] device table limit", flush=True)
    model = workloads.tiny_model(Nt=30, nx=40, ny=32, nz=30, time_monitors=False)
    model.add_permittivity_volume(np.random.default_rng(9).uniform(1.0, 4.0, size=(40, 32, 30)), (0, 0, 0))
    model.generate_mesh(courant=0.99)
    src = model.gaussian_modulated_source(f0=20e9, width=60e-12, t0=30e-12, t_len=30.5 * model.dt)
    model.assign_excitation(1e-2 * np.asarray(src), 1)
    default = _run(model, 0)
    forced = _run(model, _lib.COEF_TABLE)
    forced_host = _run(model, _lib.COEF_TABLE | _lib.HOST_TABLE)
    cells = model.Nx * model.Ny * model.Nz
    assert default[0]["coef_kind"] == "raw"
    assert forced[0]["coef_kind"] == "table" and forced[0]["n_e_entries"] > cells // 8
    assert forced[0]["n_e_entries"] == forced_host[0]["n_e_entries"]
    assert forced[0]["n_h_entries"] == forced_host[0]["n_h_entries"]
    _assert_same_run(default, forced)
    _assert_same_run(default, forced_host)
