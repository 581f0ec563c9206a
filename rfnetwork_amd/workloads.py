"""Model builders for the FDTD benchmark and test workloads.

Each builder returns a meshed, excited ``FDTD_Solver`` on a uniform grid of ``cell_m`` metre cells
(the grid is given in inches, the solver's unit: cell_in = cell_m * 39.3701).

microstrip      120x48x32 cells of 0.1 mm: eps_r 3.66 substrate 4 cells thick on the z = 0 ground,
                PEC strip 6 cells wide on top, 10-cell PML on x-, x+, y-, y+, z+, a 50 ohm port at
                each end 1 cell in from the x PML, Gaussian-modulated sine (5.5 GHz, 430 ps) on
                port 1, 4000 steps at 0.99 of the Courant limit, S-parameters at 101 frequencies
                1-10 GHz.
dipole_slab     400x400x200 cells of 0.1 mm, 12-cell PML on every face, eps_r ~ U[1, 4.5) per cell
                (seed 2) in the centred 200x200x20 slab, two 40-cell PEC Ez rods fed by a 1-cell
                50 ohm gap source with a 50 ps Gaussian pulse, 8000 steps, running DFTs at 51
                frequencies 5-30 GHz on 6 planes 2 cells inside the PML.
tiny_model      small grid with PML, a dielectric block, a conductor, a port, a probe and monitors,
                for parity tests.
design_strip    200x160x80 cells of 0.1 mm: a 64x64x32 box of per-cell eps_r (U[1, 4), seed 4) on
                the z = 0 ground, centred in x and y, in the gap between two microstrip feed lines
                (6 cells wide, 8-cell eps_r 3.66 substrate) that run from a 50 ohm port 1 cell in
                from each x PML to one cell short of the box; 10-cell PML on x-, x+, y-, y+, z+;
                Gaussian-modulated sine at 4 GHz; objective frequencies 2, 3, 4, 5, 6 GHz; 6000
                steps.  design_box() gives the box as (x0, x1, y0, y1, z0, z1) cells.
port_variant    variant v of the batched workload: V0 64^3, V1 128x96x64, V2 160x128x96, V3
                192x128x96 cells of 0.1 mm with an eps_r ~ U[1, 6) block (seed 300 + v) centred in x
                and y over half the interior extent, from the ground to half height; 6-cell strips on
                a 4-cell eps_r 3.66 substrate from 2/3/3/4 ports (x-, x+, y-, y+ sides, 1 cell in from
                the PML) to the centre; 10-cell PML on every face but the z = 0 ground; the C1 pulse;
                101 frequencies 1-10 GHz; 5000 steps.
design_variant  variant v of the batched forward+adjoint workload: grids from 96^3 to 512x512x256
                cells of 0.1 mm, a cubic eps_r ~ U[1, 4) design box (seed 500 + v) from 48^3 to 128^3
                on the ground, centred in x and y, between 2/2/3/3/3/4/4/4 microstrip feeds (x-, x+,
                y-, y+ sides) that end one cell short of the box; 10-cell PML except on the ground;
                port 1 driven (4 GHz Gaussian-modulated sine); objective |S21|^2 at 2..6 GHz; 8000
                steps.
half_wave_dipole  rfnetwork's own 10 GHz dipole test case (tests/test_solver/test_dipole.py) on a
                uniform grid of 0.0075 in cells: two 4-cell-wide PEC strips of 37 cells each side of
                a 2-cell gap port, 5-cell PML on every face, far-field monitor at 10 and 20 GHz,
                edge-singularity correction on the strip edges, 50 ps Gaussian pulse.
"""
import numpy as np

from .geometry import Box, Line, Rectangle
from .solver import FDTD_Solver

IN_PER_M = 39.3701


def _grid_solver(nx, ny, nz, cell_m):
    d = cell_m * IN_PER_M
    s = FDTD_Solver(Box((0, 0, 0), (nx * d, ny * d, nz * d)))
    s.set_grid(np.arange(nx + 1) * d, np.arange(ny + 1) * d, np.arange(nz + 1) * d)
    return s, d


def microstrip(Nt=4000, nx=120, ny=48, nz=32, cell_m=1e-4, courant=0.99, n_pml=10, n_freq=101, coef_dtype=np.float32,
               coefficients="host"):
    """Two-port microstrip line (the 120x48x32 plain-FDTD workload)."""
    s, d = _grid_solver(nx, ny, nz, cell_m)
    s.coef_dtype = coef_dtype
    h = 4  # substrate thickness, cells
    y0, y1 = ny // 2 - 3, ny // 2 + 3  # 6-cell strip centred in y
    xa, xb = n_pml + 1, nx - n_pml - 1  # ports 1 cell in from the x PML
    s.add_dielectric(Box((0, 0, 0), (nx * d, ny * d, h * d)), er=3.66)
    s.add_conductor(Rectangle([(xa * d, y0 * d, h * d), (xa * d, y1 * d, h * d), (xb * d, y1 * d, h * d)]))
    s.add_lumped_port(1, Rectangle([(xa * d, y0 * d, 0), (xa * d, y1 * d, 0), (xa * d, y1 * d, h * d)]), "z+")
    s.add_lumped_port(2, Rectangle([(xb * d, y0 * d, 0), (xb * d, y1 * d, 0), (xb * d, y1 * d, h * d)]), "z+")
    s.assign_PML_boundaries("x-", "x+", "y-", "y+", "z+", n_pml=n_pml)
    s.generate_mesh(courant=courant, coefficients=coefficients)
    src = s.gaussian_modulated_source(f0=5.5e9, width=430e-12, t0=320e-12, t_len=(Nt + 0.5) * s.dt)
    s.assign_excitation(1e-2 * np.asarray(src), 1)
    s.frequency = np.linspace(1e9, 1e10, n_freq)
    return s


def dipole_slab(Nt=8000, nx=400, ny=400, nz=200, cell_m=1e-4, courant=0.99, n_pml=12, n_freq=51, seed=2,
                monitors=True, dft_dtype=None, coefficients="host"):
    """Dipole over a random dielectric slab (the 400x400x200 bandwidth workload).  dft_dtype:
    accumulator type of the 6 running-DFT planes (default complex64 in a float32 solve)."""
    s, d = _grid_solver(nx, ny, nz, cell_m)
    cx, cy, cz = nx // 2, ny // 2, nz // 2
    rng = np.random.default_rng(seed)
    sx, sy, sz = nx // 2, ny // 2, 20
    er = rng.uniform(1.0, 4.5, size=(sx, sy, sz))
    s.add_permittivity_volume(er, (cx - sx // 2, cy - sy // 2, cz - sz // 2))
    # rods: Ez components on the (cx, cy) grid line, 40 cells each, 1-cell gap at z = cz
    s.add_conductor(Line((cx * d, cy * d, (cz - 40) * d), (cx * d, cy * d, cz * d)))
    s.add_conductor(Line((cx * d, cy * d, (cz + 1) * d), (cx * d, cy * d, (cz + 41) * d)))
    # gap source: a port face 2 cells wide in y (one interior Ez component) and 1 cell tall
    s.add_lumped_port(1, Rectangle([(cx * d, (cy - 1) * d, cz * d), (cx * d, (cy + 1) * d, cz * d),
                                    (cx * d, (cy + 1) * d, (cz + 1) * d)]), "z+")
    s.assign_PML_boundaries("x-", "x+", "y-", "y+", "z-", "z+", n_pml=n_pml)
    s.generate_mesh(courant=courant, coefficients=coefficients)
    src = s.gaussian_source(width=50e-12, t0=40e-12, t_len=(Nt + 0.5) * s.dt)
    s.assign_excitation(np.asarray(src), 1)
    freqs = np.linspace(5e9, 3e10, n_freq)
    s.frequency = freqs
    if monitors:
        k = n_pml + 2
        s.add_field_monitor("ez_x0", "ez", "x", index=k, frequency=freqs, dft_dtype=dft_dtype)
        s.add_field_monitor("ez_x1", "ez", "x", index=nx - k - 1, frequency=freqs, dft_dtype=dft_dtype)
        s.add_field_monitor("ez_y0", "ez", "y", index=k, frequency=freqs, dft_dtype=dft_dtype)
        s.add_field_monitor("ez_y1", "ez", "y", index=ny - k - 1, frequency=freqs, dft_dtype=dft_dtype)
        s.add_field_monitor("ex_z0", "ex", "z", index=k, frequency=freqs, dft_dtype=dft_dtype)
        s.add_field_monitor("ex_z1", "ex", "z", index=nz - k - 1, frequency=freqs, dft_dtype=dft_dtype)
    return s


def tiny_model(Nt=60, nx=14, ny=11, nz=9, cell_m=1e-4, courant=0.99, n_pml=3, seed=7, all_sides=True,
               time_monitors=True, phasor_monitors=True, coef_dtype=np.float32, coefficients="host"):
    """Small model that touches every code path: PML, random dielectric, conductor, port, probes,
    time-domain and phasor monitors on every axis."""
    s, d = _grid_solver(nx, ny, nz, cell_m)
    s.coef_dtype = coef_dtype
    rng = np.random.default_rng(seed)
    s.add_permittivity_volume(rng.uniform(1.0, 4.0, size=(nx - 6, ny - 6, 3)), (3, 3, 1))
    s.add_dielectric(Box((0, 0, 0), (nx * d, ny * d, 1 * d)), er=2.2)
    s.add_conductor(Rectangle([(4 * d, 4 * d, 4 * d), (4 * d, 7 * d, 4 * d), (10 * d, 7 * d, 4 * d)]))
    s.add_lumped_port(1, Rectangle([(4 * d, 4 * d, 0), (4 * d, 7 * d, 0), (4 * d, 7 * d, 4 * d)]), "z+")
    s.add_lumped_port(2, Rectangle([(10 * d, 4 * d, 0), (10 * d, 7 * d, 0), (10 * d, 7 * d, 4 * d)]), "z+")
    sides = ("x-", "x+", "y-", "y+", "z-", "z+") if all_sides else ("x+", "z+")
    s.assign_PML_boundaries(*sides, n_pml=n_pml)
    s.generate_mesh(courant=courant, coefficients=coefficients)
    src = s.gaussian_modulated_source(f0=20e9, width=60e-12, t0=30e-12, t_len=(Nt + 0.5) * s.dt)
    s.assign_excitation(1e-2 * np.asarray(src), 1)
    s.add_voltage_probe("v1", Line((7 * d, 5 * d, 0), (7 * d, 5 * d, 4 * d)))
    if time_monitors:
        s.add_field_monitor("t_ez_x", "ez", "x", index=6, n_step=3)
        s.add_field_monitor("t_hy_y", "hy", "y", index=5, n_step=2)
        s.add_field_monitor("t_ex_z", "ex", "z", index=2, n_step=1)
        s.add_field_monitor("t_hz_z", "hz", "z", index=4, n_step=5)
    if phasor_monitors:
        f = np.array([10e9, 20e9, 35e9])
        s.add_field_monitor("p_ey_x", "ey", "x", index=8, frequency=f)
        s.add_field_monitor("p_hx_y", "hx", "y", index=6, frequency=f, n_step=2)
        s.add_field_monitor("p_ez_z", "ez", "z", index=3, frequency=f)
    s.frequency = np.linspace(5e9, 40e9, 8)
    return s


def design_box(nx=200, ny=160, design=(64, 64, 32)):
    """Cell box (x0, x1, y0, y1, z0, z1) of the design region of design_strip."""
    dx, dy, dz = design
    x0, y0 = (nx - dx) // 2, (ny - dy) // 2
    return (x0, x0 + dx, y0, y0 + dy, 0, dz)


def design_strip(eps_design=None, Nt=6000, nx=200, ny=160, nz=80, design=(64, 64, 32), cell_m=1e-4,
                 courant=0.99, n_pml=10, h=8, seed=4, coef_dtype=np.float32, pulse=(4e9, 600e-12, 420e-12),
                 frequency=(2e9, 3e9, 4e9, 5e9, 6e9), coefficients="host"):
    """Two microstrip feeds coupled through a box of per-cell permittivity (the adjoint workload).

    eps_design: eps_r of the box cells, shape ``design``; default U[1, 4) from ``seed``.
    pulse: (f0, width, t0) of the Gaussian-modulated sine on port 1; frequency: objective points."""
    if eps_design is None:
        eps_design = np.random.default_rng(seed).uniform(1.0, 4.0, size=design)
    eps_design = np.asarray(eps_design, np.float64)
    if eps_design.shape != tuple(design):
        raise ValueError(f"eps_design must have shape {tuple(design)}")
    s, d = _grid_solver(nx, ny, nz, cell_m)
    s.coef_dtype = coef_dtype
    x0, x1, y0, y1, z0, z1 = design_box(nx, ny, design)
    w0, w1 = ny // 2 - 3, ny // 2 + 3  # 6-cell strip centred in y
    xa, xb = n_pml + 1, nx - n_pml - 1
    s.add_dielectric(Box((0, 0, 0), (nx * d, ny * d, h * d)), er=3.66)
    s.add_permittivity_volume(eps_design, (x0, y0, z0))
    # feeds end one node short of the box so that no conductor edge borders a design cell
    s.add_conductor(Rectangle([(xa * d, w0 * d, h * d), (xa * d, w1 * d, h * d), ((x0 - 1) * d, w1 * d, h * d)]))
    s.add_conductor(Rectangle([((x1 + 1) * d, w0 * d, h * d), ((x1 + 1) * d, w1 * d, h * d), (xb * d, w1 * d, h * d)]))
    s.add_lumped_port(1, Rectangle([(xa * d, w0 * d, 0), (xa * d, w1 * d, 0), (xa * d, w1 * d, h * d)]), "z+")
    s.add_lumped_port(2, Rectangle([(xb * d, w0 * d, 0), (xb * d, w1 * d, 0), (xb * d, w1 * d, h * d)]), "z+")
    s.assign_PML_boundaries("x-", "x+", "y-", "y+", "z+", n_pml=n_pml)
    s.generate_mesh(courant=courant, coefficients=coefficients)
    f0, width, t0 = pulse
    src = s.gaussian_modulated_source(f0=f0, width=width, t0=t0, t_len=(Nt + 0.5) * s.dt)
    s.assign_excitation(1e-2 * np.asarray(src), 1)
    s.frequency = np.asarray(frequency, dtype=np.float64)
    return s


PORT_VARIANTS = ((64, 64, 64, 2), (128, 96, 64, 3), (160, 128, 96, 3), (192, 128, 96, 4))


def port_variant(v, Nt=5000, cell_m=1e-4, courant=0.99, n_pml=10, h=4, n_freq=101, coef_dtype=np.float32):
    """Variant v (0..3) of the batched multi-port workload; port 1 carries the excitation."""
    nx, ny, nz, n_ports = PORT_VARIANTS[v]
    s, d = _grid_solver(nx, ny, nz, cell_m)
    s.coef_dtype = coef_dtype
    rng = np.random.default_rng(300 + v)
    ix, iy = nx - 2 * n_pml, ny - 2 * n_pml  # interior extent
    bx, by, bz = ix // 2, iy // 2, nz // 2
    bx0, by0 = (nx - bx) // 2, (ny - by) // 2
    s.add_dielectric(Box((0, 0, 0), (nx * d, ny * d, h * d)), er=3.66)
    s.add_permittivity_volume(rng.uniform(1.0, 6.0, size=(bx, by, bz)), (bx0, by0, 0))
    cx, cy = nx // 2, ny // 2
    # (side, port position): strips along x from the x ports, along y from the y ports, to the centre
    ends = (("x", n_pml + 1), ("x", nx - n_pml - 1), ("y", n_pml + 1), ("y", ny - n_pml - 1))[:n_ports]
    for k, (axis, p) in enumerate(ends):
        if axis == "x":
            a, b = sorted((p, cx + 3 if p < cx else cx - 3))
            s.add_conductor(Rectangle([(a * d, (cy - 3) * d, h * d), (a * d, (cy + 3) * d, h * d),
                                       (b * d, (cy + 3) * d, h * d)]))
            s.add_lumped_port(k + 1, Rectangle([(p * d, (cy - 3) * d, 0), (p * d, (cy + 3) * d, 0),
                                                (p * d, (cy + 3) * d, h * d)]), "z+")
        else:
            a, b = sorted((p, cy + 3 if p < cy else cy - 3))
            s.add_conductor(Rectangle([((cx - 3) * d, a * d, h * d), ((cx + 3) * d, a * d, h * d),
                                       ((cx + 3) * d, b * d, h * d)]))
            s.add_lumped_port(k + 1, Rectangle([((cx - 3) * d, p * d, 0), ((cx + 3) * d, p * d, 0),
                                                ((cx + 3) * d, p * d, h * d)]), "z+")
    s.assign_PML_boundaries("x-", "x+", "y-", "y+", "z+", n_pml=n_pml)
    s.generate_mesh(courant=courant)
    src = s.gaussian_modulated_source(f0=5.5e9, width=430e-12, t0=320e-12, t_len=(Nt + 0.5) * s.dt)
    s.assign_excitation(1e-2 * np.asarray(src), 1)
    s.frequency = np.linspace(1e9, 1e10, n_freq)
    return s


DESIGN_VARIANTS = ((96, 96, 96, 2), (128, 128, 128, 2), (160, 160, 128, 3), (192, 192, 160, 3),
                   (256, 256, 160, 3), (320, 320, 192, 4), (384, 384, 224, 4), (512, 512, 256, 4))


def design_variant_box(v):
    """(x0, x1, y0, y1, z0, z1) of the design box of design_variant(v): a cube 48 + 80 v / 7 cells."""
    nx, ny, nz, _ = DESIGN_VARIANTS[v]
    b = int(round(48 + 80 * v / (len(DESIGN_VARIANTS) - 1)))
    x0, y0 = (nx - b) // 2, (ny - b) // 2
    return (x0, x0 + b, y0, y0 + b, 0, b)


def design_variant(v, eps_design=None, Nt=8000, cell_m=1e-4, courant=0.99, n_pml=10, h=8, coef_dtype=np.float32,
                   coefficients="host"):
    """Variant v (0..7) of the batched forward+adjoint workload (coefficients="device": no full-grid
    coefficient arrays on the host; a session assembles them on the GPU)."""
    nx, ny, nz, n_ports = DESIGN_VARIANTS[v]
    x0, x1, y0, y1, z0, z1 = design_variant_box(v)
    b = x1 - x0
    if eps_design is None:
        eps_design = np.random.default_rng(500 + v).uniform(1.0, 4.0, size=(b, b, b))
    eps_design = np.asarray(eps_design, np.float64)
    s, d = _grid_solver(nx, ny, nz, cell_m)
    s.coef_dtype = coef_dtype
    cx, cy = nx // 2, ny // 2
    s.add_dielectric(Box((0, 0, 0), (nx * d, ny * d, h * d)), er=3.66)
    s.add_permittivity_volume(eps_design, (x0, y0, z0))
    ends = (("x", n_pml + 1, x0 - 1), ("x", nx - n_pml - 1, x1 + 1), ("y", n_pml + 1, y0 - 1),
            ("y", ny - n_pml - 1, y1 + 1))[:n_ports]
    for k, (axis, p, stop) in enumerate(ends):
        a, e = sorted((p, stop))
        if axis == "x":
            s.add_conductor(Rectangle([(a * d, (cy - 3) * d, h * d), (a * d, (cy + 3) * d, h * d),
                                       (e * d, (cy + 3) * d, h * d)]))
            s.add_lumped_port(k + 1, Rectangle([(p * d, (cy - 3) * d, 0), (p * d, (cy + 3) * d, 0),
                                                (p * d, (cy + 3) * d, h * d)]), "z+")
        else:
            s.add_conductor(Rectangle([((cx - 3) * d, a * d, h * d), ((cx + 3) * d, a * d, h * d),
                                       ((cx + 3) * d, e * d, h * d)]))
            s.add_lumped_port(k + 1, Rectangle([((cx - 3) * d, p * d, 0), ((cx + 3) * d, p * d, 0),
                                                ((cx + 3) * d, p * d, h * d)]), "z+")
    s.assign_PML_boundaries("x-", "x+", "y-", "y+", "z+", n_pml=n_pml)
    s.generate_mesh(courant=courant, coefficients=coefficients)
    src = s.gaussian_modulated_source(f0=4e9, width=600e-12, t0=420e-12, t_len=(Nt + 0.5) * s.dt)
    s.assign_excitation(1e-2 * np.asarray(src), 1)
    s.frequency = np.array([2e9, 3e9, 4e9, 5e9, 6e9])
    return s


def graded_edges(fixed, lo, hi, d_fine, d_max, ratio):
    """Grid edges from lo to hi (inches) through every edge in `fixed`: cells of at most d_fine between
    consecutive fixed edges (equal within each span), growing by `ratio` per cell up to d_max outside
    them; a last cell shorter than half its neighbour is merged into it."""
    fe = sorted(float(v) for v in fixed)
    edges = [fe[0]]
    for a, b in zip(fe[:-1], fe[1:]):
        n = max(1, int(np.ceil((b - a) / d_fine - 1e-9)))
        edges += list(a + (b - a) * np.arange(1, n + 1) / n)

    def grow(start, stop, sign):
        out, d = [start], d_fine
        while sign * (stop - out[-1]) > 1e-12:
            d = min(d * ratio, d_max)
            out.append(out[-1] + sign * d if sign * (stop - out[-1] - sign * d) > 0 else stop)
        if len(out) > 2 and abs(out[-1] - out[-2]) < 0.5 * abs(out[-2] - out[-3]):
            out.pop(-2)
        return out

    up, down = grow(fe[-1], hi, 1), grow(fe[0], lo, -1)
    return np.array(down[::-1][:-1] + edges + up[1:])


def half_wave_dipole_exact(d_fine=0.0075, d_max=0.0075, ratio=1.0, coefficients="host"):
    """The strip dipole of rfnetwork's dipole test (tests/test_solver/test_dipole.py:14-92) with its
    exact dimensions (inches: strip 0.030 wide, gap 0.015, 0.546 end to end, solve box 1.5 x 1.1 x
    1.1) on a grid whose lines pass through the strip edges, the gap and the arm ends: cells of at
    most d_fine along the dipole (equal within each span), growing by `ratio` up to d_max towards the
    box, as the reference's mesher grades between its d_min and d_max (the default: d_fine
    everywhere).  PML of 5 cells on every side, a far-field box at 10 and 20 GHz, edge corrections
    along the strip edges, the reference's Gaussian excitation."""
    ms_w, gap, dl = 0.030, 0.015, 0.546
    xs = graded_edges([-d_fine, 0.0, d_fine], -0.75, 0.75, d_fine, d_max, ratio)
    ys = graded_edges([-ms_w / 2, ms_w / 2], -0.55, 0.55, d_fine, d_max, ratio)
    zs = graded_edges([-dl / 2, -gap / 2, gap / 2, dl / 2], -0.55, 0.55, d_fine, d_max, ratio)
    s = FDTD_Solver(Box((xs[0], ys[0], zs[0]), (xs[-1], ys[-1], zs[-1])))
    s.set_grid(xs, ys, zs)
    ms_y = (-ms_w / 2, ms_w / 2)
    z1, z2 = (-dl / 2, -gap / 2), (gap / 2, dl / 2)
    s.add_conductor(Rectangle([(0, ms_y[0], z1[0]), (0, ms_y[1], z1[0]), (0, ms_y[1], z1[1])]))
    s.add_conductor(Rectangle([(0, ms_y[0], z2[0]), (0, ms_y[1], z2[0]), (0, ms_y[1], z2[1])]))
    s.add_lumped_port(1, Rectangle([(0, ms_y[0], gap / 2), (0, ms_y[1], gap / 2), (0, ms_y[1], -gap / 2)]), "z-")
    s.assign_PML_boundaries("x-", "x+", "y-", "y+", "z+", "z-", n_pml=5)
    s.generate_mesh(coefficients=coefficients)
    s.add_farfield_monitor(frequency=[10e9, 20e9])
    for i, mz in enumerate((z1, z2)):
        s.edge_correction((0, ms_y[0], mz[0]), (0, ms_y[0], mz[1]), integration_line="y-")
        s.edge_correction((0, ms_y[1], mz[0]), (0, ms_y[1], mz[1]), integration_line="y+")
        s.edge_correction((0, ms_y[0], mz[i]), (0, ms_y[1], mz[i]), integration_line=("z-" if i == 0 else "z+"))
    s.assign_excitation(s.gaussian_source(width=50e-12, t0=40e-12, t_len=600e-12), 1)
    return s


def coupled_stripline(sl_w, sl_sp, d=0.0025, coefficients="host"):
    """The edge-coupled stripline of rfnetwork's coupled-line test (tests/test_solver/
    test_coupled_stripline.py:22-96), inches: strips sl_w wide, sl_sp apart, in a 0.06 thick
    substrate of er 3.66 between PEC walls (the solve box, 0.25 x 0.4 x 0.06), running from 0.1 past
    the x- wall into a 5-cell PML at x+, each fed by a lumped port across the substrate whose
    integration line runs from the lower wall to its strip; edge corrections on the first strip's
    edges applied twice, as that test applies them; uniform d cells, whose lines pass through every
    strip edge for the test's three geometries.  Returns the model and the test's 10 GHz
    Gaussian-modulated voltage (1e-2 amplitude)."""
    b, er = 0.06, 3.66
    sbox_w, sbox_len = 0.4, 0.25
    y1, y2 = -(sl_w / 2) - (sl_sp / 2), (sl_w / 2) + (sl_sp / 2)
    ms_x = (-sbox_len / 2 + 0.1, sbox_len / 2)
    xs = -sbox_len / 2 + np.arange(int(round(sbox_len / d)) + 1) * d
    ys = -sbox_w / 2 + np.arange(int(round(sbox_w / d)) + 1) * d
    zs = -b / 2 + np.arange(int(round(b / d)) + 1) * d
    s = FDTD_Solver(Box((xs[0], ys[0], zs[0]), (xs[-1], ys[-1], zs[-1])))
    s.set_grid(xs, ys, zs)
    s.add_dielectric(Box((xs[0], ys[0], zs[0]), (xs[-1], ys[-1], zs[-1])), er=er)
    for i, y in enumerate((y1, y2)):
        s.add_conductor(Rectangle([(ms_x[0], y - sl_w / 2, 0), (ms_x[0], y + sl_w / 2, 0),
                                   (ms_x[1], y + sl_w / 2, 0)]))
        s.add_lumped_port(i + 1, Rectangle([(ms_x[0], y - sl_w / 2, -b / 2), (ms_x[0], y + sl_w / 2, -b / 2),
                                             (ms_x[0], y + sl_w / 2, b / 2)]),
                          Line((ms_x[0], y, -b / 2), (ms_x[0], y, 0)))
    s.assign_PML_boundaries("x+", n_pml=5)
    s.generate_mesh(coefficients=coefficients)
    for y in (y1, y1):
        s.edge_correction((ms_x[0], y + sl_w / 2, 0), (ms_x[1], y + sl_w / 2, 0), integration_line="y+")
        s.edge_correction((ms_x[0], y - sl_w / 2, 0), (ms_x[1], y - sl_w / 2, 0), integration_line="y-")
    vsrc = 1e-2 * np.asarray(s.gaussian_modulated_source(f0=10e9, width=200e-12, t0=100e-12, t_len=500e-12))
    return s, vsrc


def half_wave_dipole(d=0.0075, coefficients="host"):
    """The strip dipole of rfnetwork's dipole test on a uniform grid of d inches (the strip width, gap
    and arm length snapped to it: 4, 2 and 37 cells)."""
    nx, ny, nz = 200, 148, 148
    xs = (np.arange(nx + 1) - nx // 2) * d
    ys = (np.arange(ny + 1) - ny // 2) * d
    zs = (np.arange(nz + 1) - nz // 2) * d
    s = FDTD_Solver(Box((xs[0], ys[0], zs[0]), (xs[-1], ys[-1], zs[-1])))
    s.set_grid(xs, ys, zs)
    ms_w, gap, half = 4 * d, 2 * d, 37 * d
    ms_y = (-ms_w / 2, ms_w / 2)
    z1, z2 = (-half, -gap / 2), (gap / 2, half)
    s.add_conductor(Rectangle([(0, ms_y[0], z1[0]), (0, ms_y[1], z1[0]), (0, ms_y[1], z1[1])]))
    s.add_conductor(Rectangle([(0, ms_y[0], z2[0]), (0, ms_y[1], z2[0]), (0, ms_y[1], z2[1])]))
    s.add_lumped_port(1, Rectangle([(0, ms_y[0], gap / 2), (0, ms_y[1], gap / 2), (0, ms_y[1], -gap / 2)]), "z-")
    s.assign_PML_boundaries("x-", "x+", "y-", "y+", "z+", "z-", n_pml=5)
    s.generate_mesh(coefficients=coefficients)
    s.add_farfield_monitor(frequency=[10e9, 20e9])
    for i, mz in enumerate((z1, z2)):
        s.edge_correction((0, ms_y[0], mz[0]), (0, ms_y[0], mz[1]), integration_line="y-")
        s.edge_correction((0, ms_y[1], mz[0]), (0, ms_y[1], mz[1]), integration_line="y+")
        s.edge_correction((0, ms_y[0], mz[i]), (0, ms_y[1], mz[i]), integration_line=("z-" if i == 0 else "z+"))
    s.assign_excitation(s.gaussian_source(width=50e-12, t0=40e-12, t_len=600e-12), 1)
    return s


def microstrip_line(len_axis=0, width_axis=1, normal_axis=2, d_max=0.02, coefficients="host"):
    """The microstrip of rfnetwork's microstrip test (tests/test_solver/test_microstrip.py:15-112),
    inches, with its length, width and substrate normal along any permutation of the axes: a strip
    0.04 wide on a substrate 0.02 thick of er 3.66, in a 1.3 x 0.6 x 0.5 solve box; the strip runs from
    0.5 before the middle into a 10-cell PML, fed there by a lumped port across the substrate; edge
    corrections on both strip edges; a voltage probe from the strip to the ground and an ampere loop
    around the strip halfway along it; a monitor of the normal E component on the substrate top; the
    mesher's cells at most d_max.  Returns the model and the test's Gaussian voltage (1e-2 amplitude,
    80 ps wide, 300 ps long)."""
    if sorted((len_axis, width_axis, normal_axis)) != [0, 1, 2]:
        raise ValueError("the three axes must be a permutation of 0, 1, 2")
    ms_w, ms_len, sub_h = 0.04, 1.0, 0.02
    sbox_h, sbox_w, sbox_len = 0.5, 0.6, 1.3
    ends = (-ms_len / 2, sbox_len / 2)

    def dims(length, width, height):
        out = [0.0] * 3
        out[len_axis], out[width_axis], out[normal_axis] = length, width, height
        return out

    def cube(center, size):
        return Box(np.subtract(center, np.divide(size, 2)), np.add(center, np.divide(size, 2)))

    na = "xyz"[normal_axis]
    s = FDTD_Solver(cube(dims(0, 0, sbox_h / 2), dims(sbox_len, sbox_w, sbox_h)))
    s.add_dielectric(cube(dims(0, 0, sub_h / 2), dims(sbox_len, sbox_w, sub_h)), er=3.66)
    s.add_conductor(Rectangle([dims(ends[0], -ms_w / 2, sub_h), dims(ends[0], ms_w / 2, sub_h),
                               dims(ends[1], ms_w / 2, sub_h)]))
    s.add_lumped_port(1, Rectangle([dims(ends[0], -ms_w / 2, sub_h), dims(ends[0], ms_w / 2, sub_h),
                                    dims(ends[0], ms_w / 2, 0)]), integration_line=f"{na}+")
    s.assign_PML_boundaries(f"{'xyz'[len_axis]}+", n_pml=10)
    s.generate_mesh(d_max=d_max, coefficients=coefficients)
    wa = "xyz"[width_axis]
    for sgn in (1, -1):
        s.edge_correction(dims(ends[0], sgn * ms_w / 2, sub_h), dims(ends[1], sgn * ms_w / 2, sub_h),
                          f"{wa}{'+' if sgn > 0 else '-'}")
    s.add_field_monitor("mon1", f"e{na}", na, sub_h, 5)
    s.add_current_probe("c1", Rectangle([dims(0, -ms_w / 2 - 0.001, sub_h + 0.001),
                                         dims(0, ms_w / 2 + 0.001, sub_h + 0.001),
                                         dims(0, ms_w / 2 + 0.001, sub_h - 0.001)]))
    s.add_voltage_probe("v1", Line(dims(0, 0, sub_h), dims(0, 0, 0)))
    vsrc = 1e-2 * np.asarray(s.gaussian_source(width=80e-12, t0=80e-12, t_len=300e-12))
    return s, vsrc


def coupled_microstrip(ms_w, ms_sp, d_max=0.01, d_min=0.005, coefficients="host"):
    """The edge-coupled microstrip of rfnetwork's coupled-microstrip test (tests/test_solver/
    test_coupled_ustrip.py:22-96), inches: strips ms_w wide and ms_sp apart on a substrate 0.03 thick
    of er 3.66 in a 0.3 x 0.3 x 0.25 solve box, from 0.1 past the x- wall into the PML at x+, each fed
    by a lumped port across the substrate (integration line z+); PML of 5 cells on x-, x+, z+, y- and
    y+; edge corrections along both edges of both strips; the mesher's cells at most d_max (and
    d_min, which only bounds the graded cells of the reference's mesher).  Returns the model and the
    test's Gaussian voltage (1e-2 amplitude, 80 ps wide, t0 50 ps, 300 ps long)."""
    sub_h, er = 0.03, 3.66
    sbox_h, sbox_w, sbox_len = 0.25, 0.3, 0.3
    ys = (-(ms_w / 2) - (ms_sp / 2), (ms_w / 2) + (ms_sp / 2))
    ms_x = (-sbox_len / 2 + 0.1, sbox_len / 2)
    s = FDTD_Solver(Box((-sbox_len / 2, -sbox_w / 2, 0), (sbox_len / 2, sbox_w / 2, sbox_h)))
    s.add_dielectric(Box((-sbox_len / 2, -sbox_w / 2, 0), (sbox_len / 2, sbox_w / 2, sub_h)), er=er)
    for i, y in enumerate(ys):
        s.add_conductor(Rectangle([(ms_x[0], y - ms_w / 2, sub_h), (ms_x[0], y + ms_w / 2, sub_h),
                                   (ms_x[1], y + ms_w / 2, sub_h)]))
        s.add_lumped_port(i + 1, Rectangle([(ms_x[0], y - ms_w / 2, sub_h), (ms_x[0], y + ms_w / 2, sub_h),
                                            (ms_x[0], y + ms_w / 2, 0)]), integration_line="z+")
    s.assign_PML_boundaries("x+", "x-", "z+", "y-", "y+", n_pml=5)
    s.generate_mesh(d_max=d_max, d_min=d_min, coefficients=coefficients)
    for y in ys:
        s.edge_correction((ms_x[0], y + ms_w / 2, sub_h), (ms_x[1], y + ms_w / 2, sub_h), integration_line="y+")
        s.edge_correction((ms_x[0], y - ms_w / 2, sub_h), (ms_x[1], y - ms_w / 2, sub_h), integration_line="y-")
    vsrc = 1e-2 * np.asarray(s.gaussian_source(width=80e-12, t0=50e-12, t_len=300e-12))
    return s, vsrc
