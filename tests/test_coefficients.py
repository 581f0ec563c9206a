"""Closed-form checks of the coefficient assembly and the solve() packing.

Expected values come from Maxwell's equations and the update formulas' textbook forms, not from the
code under test: vacuum and dielectric Ca/Cb/Da/Db, the curl signs folded into the packed
coefficients, the time step, the cubic PML grading (rfnetwork/solver.py:1239-1325), the lumped-port
termination (:1086-1236), PEC conductors (:817-868) and the Gaussian sources (:1328-1362).
"""
import numpy as np

from rfnetwork_amd import FDTD_Solver, workloads
from rfnetwork_amd.geometry import Box
from rfnetwork_amd.units import const, conv

IN_PER_M = 39.3701
e0, u0, c0 = const.e0, const.u0, const.c0


def _box(n=(12, 10, 8), er_slab=None, pml=None, courant=0.95):
    d = 1e-4 * IN_PER_M
    s = FDTD_Solver(Box((0, 0, 0), tuple(k * d for k in n)))
    s.set_grid(*[np.arange(k + 1) * d for k in n])
    if er_slab is not None:
        s.add_dielectric(Box((0, 0, 0), (n[0] * d, n[1] * d, 4 * d)), er=er_slab)
    if pml:
        s.assign_PML_boundaries(*pml, n_pml=3)
    s.generate_mesh(courant=courant)
    return s, float(conv.m_in(np.float32(d)))


def test_time_step_courant():
    for courant in (0.95, 0.99, 0.5):
        s, d = _box(courant=courant)
        assert np.isclose(float(s.dt), courant * d / (c0 * np.sqrt(3)), rtol=2e-7)


def test_vacuum_coefficients_and_curl_signs():
    s, d = _box()
    dt = float(s.dt)
    c = s.coefficient_arrays(gpu=True)
    i = (5, 4, 3)
    ce, ch = dt / (e0 * d), dt / (u0 * d)
    # dEx/dt = (dHz/dy - dHy/dz)/eps ; dEy/dt = (dHx/dz - dHz/dx)/eps ; dEz/dt = (dHy/dx - dHx/dy)/eps
    for name, sign in (("Cb_ex_y", 1), ("Cb_ex_z", -1), ("Cb_ey_z", 1), ("Cb_ey_x", -1), ("Cb_ez_x", 1),
                       ("Cb_ez_y", -1)):
        assert np.isclose(c[name][i], sign * ce, rtol=1e-6), name
    # dHx/dt = -(dEz/dy - dEy/dz)/mu ; dHy/dt = -(dEx/dz - dEz/dx)/mu ; dHz/dt = -(dEy/dx - dEx/dy)/mu
    for name, sign in (("Db_hx_y1", -1), ("Db_hx_y2", -1), ("Db_hx_z1", 1), ("Db_hx_z2", 1), ("Db_hy_z1", -1),
                       ("Db_hy_z2", -1), ("Db_hy_x1", 1), ("Db_hy_x2", 1), ("Db_hz_x1", -1), ("Db_hz_x2", -1),
                       ("Db_hz_y1", 1), ("Db_hz_y2", 1)):
        assert np.isclose(c[name][i], sign * ch, rtol=1e-6), name
    for name in ("Ca_ex_y", "Ca_ey_x", "Ca_ez_y", "Da_hx_y", "Da_hy_x", "Da_hz_y"):
        assert c[name][i] == 1.0
    # the last half spacing is zero: upper-edge E components get no curl term (PEC walls)
    assert c["Cb_ex_y"][3, -1, 3] == 0 and c["Cb_ex_z"][3, 3, -1] == 0 and c["Cb_ey_x"][-1, 3, 3] == 0


def test_dielectric_and_interface_averaging():
    er = 4.0
    s, d = _box(er_slab=er)
    dt = float(s.dt)
    Cb = s.Cb["ex_y"]
    assert np.isclose(Cb[5, 4, 2], dt / (e0 * er), rtol=1e-6)   # inside the slab
    assert np.isclose(Cb[5, 4, 6], dt / e0, rtol=1e-6)          # above it
    # Ex on the z = 4 interface sees two cells of each material: eps = e0 (er + 1) / 2
    assert np.isclose(Cb[5, 4, 4], dt / (e0 * (er + 1) / 2), rtol=1e-6)
    # Ez is normal to the interface: its edge averages cells in x and y only (all in one layer)
    assert np.isclose(s.Cb["ez_x"][5, 4, 3], dt / (e0 * er), rtol=1e-6)


def test_pml_cubic_grading():
    s, d = _box(pml=("x+",))
    dt, n = float(s.dt), 3
    sig_max = 0.8 * 4 / (np.sqrt(u0 / e0) * d)
    Nx = s.Nx
    for i in range(n):
        sig = sig_max * (i / n) ** 3
        a = sig * dt / (2 * e0)
        # Ey edges x = Nx - n + i (sigma = 0 on the PML interface); x = Nx is the PEC wall
        assert np.isclose(s.Ca["ey_x"][Nx - n + i, 4, 3], (1 - a) / (1 + a), rtol=1e-6)
        assert np.isclose(s.Cb["ey_x"][Nx - n + i, 4, 3], dt / e0 / (1 + a), rtol=1e-6)
        sig_h = sig_max * ((i + 0.5) / n) ** 3
        am = sig_h * dt / (2 * e0)  # sigma_m dt / (2 mu0) with sigma_m = sigma mu0 / eps0
        assert np.isclose(s.Da["hz_x"][Nx - n + i, 4, 3], (1 - am) / (1 + am), rtol=1e-6)
        assert np.isclose(s.Db["hz_x1"][Nx - n + i, 4, 3], dt / u0 / (1 + am), rtol=1e-6)
    # only the split along the PML axis is graded
    assert s.Ca["ey_z"][Nx - 1, 4, 3] == 1.0 and s.Da["hz_y"][Nx - 1, 4, 3] == 1.0
    assert s.Ca["ey_x"][Nx - n - 1, 4, 3] == 1.0 and s.Da["hz_x"][Nx - n - 1, 4, 3] == 1.0


def test_lumped_port_termination():
    s = workloads.microstrip(Nt=10, nx=40, ny=20, nz=10, n_pml=4)
    port = s.ports[0]
    dt = float(s.dt)
    d = float(conv.m_in(s.d_cells[0][0]))
    eps = float(np.asarray(s.eps_ez[port["idx"]]).ravel()[0])
    nw, nf = 5, 4  # 6-cell face minus its two end components, 4 cells of substrate
    r_cell = 50 * nw / nf
    g = d / (2 * r_cell * d * d)  # df / (2 r_cell dn dw)
    ca = (eps / dt - g) / (eps / dt + g)
    cb = 1 / (eps / dt + g)
    idx = tuple(np.array(v).ravel()[0] if not isinstance(v, slice) else v.start for v in port["idx"])
    assert np.isclose(s.Ca["ez_x"][idx], ca, rtol=1e-5) and np.isclose(s.Ca["ez_y"][idx], ca, rtol=1e-5)
    assert np.isclose(s.Cb["ez_x"][idx], cb, rtol=1e-5)
    assert np.allclose(port["Vs_a"], -1 / ((eps / dt + g) * r_cell * d * d * nf), rtol=1e-5)
    assert port["Vs_a"].shape == (1, 5, 4)


def test_pec_conductor_coefficients():
    s = workloads.microstrip(Nt=10, nx=40, ny=20, nz=10, n_pml=4)
    # strip on z = 4 between the ports: tangential E flips sign each step and gets a vanishing Cb
    assert s.Ca["ex_y"][20, 10, 4] == -1.0 and abs(s.Cb["ex_y"][20, 10, 4]) < 1e-15
    assert s.Ca["ey_x"][20, 10, 4] == -1.0
    assert s.Ca["ez_x"][20, 10, 4] == 1.0  # normal component untouched


def test_gaussian_sources():
    s, _ = _box()
    dt = float(s.dt)
    n = 400
    g = np.asarray(s.gaussian_source(width=20e-12, t0=30e-12, t_len=n * dt + 0.5 * dt))
    t = np.linspace(0, s.dt * n, n)  # the reference samples on linspace: spacing dt * n / (n - 1)
    want = np.exp(-((t - 30e-12) / 5e-12) ** 2)
    np.testing.assert_allclose(g, want / want.max(), rtol=1e-12, atol=1e-15)
    m = np.asarray(s.gaussian_modulated_source(f0=50e9, width=20e-12, t0=30e-12, t_len=n * dt + 0.5 * dt))
    assert np.isclose(np.abs(m).max(), 1.0) and len(m) == n
