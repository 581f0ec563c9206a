"""The design-region coefficient formula (design.edge_values, the statement the GPU kernel of
rfdtd_set_design_permittivity follows) against the host assembly of a whole model
(rfnetwork/solver.py:892-961, :1463-1523): the edge permittivities and the 12 packed E coefficients
over the design box are equal bit for bit, in float32 and float64, including the grid edges where
the box reaches the boundary."""
import numpy as np
import pytest

from rfnetwork_amd import workloads
from rfnetwork_amd.adjoint import gpu_box_for_cells
from rfnetwork_amd.design import edge_values

DESIGN = (8, 8, 6)


def _model(coef_dtype, seed=3):
    eps = np.random.default_rng(seed).uniform(1.0, 4.0, size=DESIGN)
    return workloads.design_strip(eps, Nt=50, nx=48, ny=32, nz=20, design=DESIGN, n_pml=5, h=3,
                                  coef_dtype=coef_dtype, pulse=(20e9, 80e-12, 60e-12))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_edge_values_equal_host_assembly(dtype):
    model = _model(dtype)
    cbox = workloads.design_box(48, 32, DESIGN)
    gbox = gpu_box_for_cells(model, cbox)
    eps_g, coef_g = edge_values(model, gbox, model.eps, model.sigma, dtype)
    coeff = model.coefficient_arrays(gpu=True, dtype=dtype)
    gx0, gx1, gy0, gy1, gz0, gz1 = gbox
    sl = (slice(gx0, gx1), slice(gy0, gy1), slice(gz0, gz1))
    for key, v in coef_g.items():
        np.testing.assert_array_equal(v, coeff[key][sl], err_msg=key)
        assert v.dtype == np.dtype(dtype)
    # edge permittivity: packed cell (i, j, k) holds Ex of (i, j+1, k+1), Ey of (i+1, j, k+1), Ez of (i+1, j+1, k)
    np.testing.assert_array_equal(eps_g["x"], model.eps_ex[gx0:gx1, gy0 + 1:gy1 + 1, gz0 + 1:gz1 + 1])
    np.testing.assert_array_equal(eps_g["y"], model.eps_ey[gx0 + 1:gx1 + 1, gy0:gy1, gz0 + 1:gz1 + 1])
    np.testing.assert_array_equal(eps_g["z"], model.eps_ez[gx0 + 1:gx1 + 1, gy0 + 1:gy1 + 1, gz0:gz1])


def test_edge_values_at_the_grid_boundary():
    """A box reaching the upper grid faces: boundary edges hold eps0, the last packing inverse is 0
    and the Ey/Ez entries of the last x plane read zero.  The whole grid is compared where its
    coefficients are plain (no PML, conductor or port)."""
    model = workloads.tiny_model(Nt=10, all_sides=False, time_monitors=False, phasor_monitors=False)
    gbox = (0, model.Nx, 0, model.Ny, 0, model.Nz)
    eps_g, coef_g = edge_values(model, gbox, model.eps, model.sigma, np.float32)
    coeff = model.coefficient_arrays(gpu=True, dtype=np.float32)
    # zero-plane rule of the GPU upload (arrays 4..11 on x = Nx - 1) applied to the host arrays
    for n, key in enumerate(coef_g):
        ref = coeff[key].copy()
        if n >= 4:
            ref[-1] = 0
        plain = (coeff["Ca_ex_y"] == 1) & (coeff["Ca_ey_z"] == 1) & (coeff["Ca_ez_x"] == 1)
        plain &= (coeff["Ca_ex_z"] == 1) & (coeff["Ca_ey_x"] == 1) & (coeff["Ca_ez_y"] == 1)
        assert plain.sum() > 50
        assert plain[-1].any() and plain[:, -1].any() and plain[..., -1].any()  # the upper faces are covered
        np.testing.assert_array_equal(coef_g[key][plain], ref[plain], err_msg=key)


def test_edge_values_on_a_graded_grid():
    """Cell widths that differ from cell to cell: the averaging weights' sum is formed in float32
    as the host assembly forms it (the widths are float32), so the packed coefficients still agree
    bit for bit."""
    from rfnetwork_amd.solver import FDTD_Solver
    from rfnetwork_amd.geometry import Box

    rng = np.random.default_rng(11)
    d = 1e-4 * 39.3701
    edges = [np.concatenate([[0.0], np.cumsum(d * rng.uniform(0.6, 1.7, size=n))]) for n in (12, 10, 9)]
    s = FDTD_Solver(Box((0, 0, 0), tuple(float(e[-1]) for e in edges)))
    s.set_grid(*edges)
    s.add_permittivity_volume(rng.uniform(1.0, 5.0, size=(12, 10, 9)), (0, 0, 0),
                              sigma=rng.uniform(0.0, 0.3, size=(12, 10, 9)))
    s.generate_mesh(courant=0.9)
    for dtype in (np.float32, np.float64):
        s.coef_dtype = dtype
        s.generate_mesh(courant=0.9)
        gbox = (0, s.Nx, 0, s.Ny, 0, s.Nz)
        _, coef_g = edge_values(s, gbox, s.eps, s.sigma, dtype)
        coeff = s.coefficient_arrays(gpu=True, dtype=dtype)
        for n, key in enumerate(coef_g):
            ref = coeff[key].copy()
            if n >= 4:
                ref[-1] = 0
            np.testing.assert_array_equal(coef_g[key], ref, err_msg=f"{key} {np.dtype(dtype)}")
