"""Whole-grid coefficient assembly on the GPU (rfdtd_create_assembled).

The host assembly of FDTD_Solver (rfnetwork/solver.py:871-1003 edge averaging and Ca/Cb,
:1239-1325 PML, :817-868 conductors, :1086-1236 lumped elements, :2534-2700 edge corrections) and
the packing of :1463-1523 build 30 arrays of Nx*Ny*Nz values on the host and copy them to the
device.  Here the host describes the model instead -- the cell permittivity and conductivity, the
cell widths, the time step, the PML profiles of each side, and the few edges that conductors and
lumped elements override and the H coefficients that edge corrections scale -- and the device
computes every array from it, with the host's floating-point types, so that the arrays are equal
bit for bit; then it builds the deduplicated table on the device.  The description's size is
16 bytes per cell (the two cell arrays) against the 120 of the raw coefficients.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check
from .units import conv, const

# packed coefficient index of each unpacked Db array (the order of _lib.COEFF_NAMES)
_DB_INDEX = {"hx_y1": 14, "hx_y2": 15, "hx_z1": 16, "hx_z2": 17, "hy_z1": 20, "hy_z2": 21,
             "hy_x1": 22, "hy_x2": 23, "hz_x1": 26, "hz_x2": 27, "hz_y1": 28, "hz_y2": 29}
# the GPU packing's cell (x, y, z) holds Ex(x, y+1, z+1), Ey(x+1, y, z+1), Ez(x+1, y+1, z),
# Hx(x+1, y, z), Hy(x, y+1, z), Hz(x, y, z+1): global index minus this shift
_E_SHIFT = ((0, 1, 1), (1, 0, 1), (1, 1, 0))
_H_SHIFT = ((1, 0, 0), (0, 1, 0), (0, 0, 1))


def _wide(value):
    """True when numpy arithmetic with value promotes a float32 operand to float64: a float64 array
    or numpy scalar (a Python float is weak and takes the other operand's type)."""
    if type(value) is float or type(value) is int:
        return False
    return np.asarray(value).dtype == np.float64


def _packed_cells(global_idx, shift, n_cells):
    """Linear packed-cell indices of global Yee indices (3 arrays) shifted by `shift`; entries that
    fall outside the packed grid (components the packing drops) are removed, with a mask of the
    kept ones."""
    p = [np.asarray(g, np.int64) - s for g, s in zip(global_idx, shift)]
    keep = np.ones(p[0].shape, bool)
    for a in range(3):
        keep &= (p[a] >= 0) & (p[a] < n_cells[a])
    Nx, Ny, Nz = n_cells
    lin = (p[0] * Ny + p[1]) * Nz + p[2]
    return lin[keep], keep


def _block_indices(block, shape):
    """Global indices (3 flat arrays, C order over the block) of a tuple of slices / ints."""
    axes = []
    for b, n in zip(block, shape):
        axes.append(np.arange(*b.indices(n)) if isinstance(b, slice) else np.array([int(b)]))
    grid = np.meshgrid(*axes, indexing="ij")
    return [g.ravel() for g in grid]


class Description:
    """The inputs of rfdtd_create_assembled for one meshed model and element type; keeps every
    array alive while the C structure points at it."""

    def __init__(self, model, dtype=np.float32):
        model.check_mesh()
        T = np.dtype(dtype)
        if np.dtype(model.coef_dtype) != T:
            raise ValueError("the model's coefficient dtype (coef_dtype) must be the session's element type")
        self.dtype = T
        Nx, Ny, Nz = model.n_cells
        n_cells = (Nx, Ny, Nz)
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a.ctypes.data

        A = _lib.Assembly()
        A.Nx, A.Ny, A.Nz = Nx, Ny, Nz
        A.eps = arr(model.eps, np.float64)
        A.sigma = arr(model.sigma, np.float64)
        inv_dual = model.packing_inverses(T)
        for d in range(3):
            w = conv.m_in(model.d_cells[d])
            A.width[d] = arr(np.asarray(w, dtype=np.float64), np.float64)
            A.inv_dual[d] = arr(inv_dual[d], T)
            A.inv_width[d] = arr(1 / w.astype(T), T)
        A.dt = float(model.dt)
        A.e0, A.u0 = float(const.e0), float(const.u0)
        # PML sides in the order the host applies them
        sides = []
        if model._n_pml is not None:
            for side in model.pml_boundaries:
                ai, upper, n, sig_edge, sig_half = model._pml_profiles(side)
                P = _lib.PmlSide()
                P.axis, P.upper, P.n = ai, int(upper), n
                P.sigma_edge = arr(sig_edge, np.float64)
                P.sigma_half = arr(sig_half, np.float64)
                sides.append(P)
        if len(sides) > 6:
            raise ValueError("at most 6 PML sides")
        pml = (_lib.PmlSide * max(1, len(sides)))(*sides)
        A.n_pml, A.pml = len(sides), ctypes.addressof(pml)
        # E overrides: conductors, then lumped elements, in the host's order
        fixes = []
        comp = {"ex": 0, "ey": 1, "ez": 2}
        for f, blk, inside, sig in model._conductor_edges():
            c = comp[f]
            shape = model.fshape[f]
            g = _block_indices(blk, shape)
            sel = np.asarray(inside).ravel()
            cells, _ = _packed_cells([x[sel] for x in g], _E_SHIFT[c], n_cells)
            sdt = sig * model.dt
            F = _lib.EdgeFix()
            F.component, F.kind, F.wide, F.count = c, 0, int(_wide(sdt)), len(cells)
            F.cells = arr(cells, np.int64)
            F.values = arr(np.array([float(sdt)]), np.float64)
            fixes.append(F)
        for f_axis, eidx, q in model._lumped_fix:
            shape = model.fshape[("ex", "ey", "ez")[f_axis]]
            g = _block_indices(eidx, shape)
            cells, kept = _packed_cells(g, _E_SHIFT[f_axis], n_cells)
            qv = np.asarray(q).ravel()[kept]
            F = _lib.EdgeFix()
            F.component, F.kind, F.wide, F.count = f_axis, 1, int(_wide(q)), len(cells)
            F.cells = arr(cells, np.int64)
            F.values = arr(qv.astype(np.float64), np.float64)
            fixes.append(F)
        fix = (_lib.EdgeFix * max(1, len(fixes)))(*fixes)
        A.n_fix, A.fix = len(fixes), ctypes.addressof(fix)
        # edge-correction factors, in order
        scales = []
        for name, ix, factor in getattr(model, "_edge_ops", []):
            h = "xyz".index(name[1])
            shape = model.fshape[f"h{name[1]}"]
            cells, _ = _packed_cells(_block_indices(ix, shape), _H_SHIFT[h], n_cells)
            for c in cells:
                S = _lib.HScale()
                S.coef, S.wide, S.cell, S.factor = _DB_INDEX[name], int(_wide(factor)), int(c), float(factor)
                scales.append(S)
        scale = (_lib.HScale * max(1, len(scales)))(*scales)
        A.n_scale, A.scale = len(scales), ctypes.addressof(scale)
        self.struct = A
        self._keep = (keep, pml, fix, scale)
        self.n_fix_edges = sum(F.count for F in fixes)
        self.n_scales = len(scales)


def create(model, dtype=np.float32, flags=0, excitations=1):
    """A session handle whose coefficients are assembled on the GPU from `model` (meshed)."""
    D = Description(model, dtype)
    L = _lib.lib()
    h = ctypes.c_void_p()
    f = int(flags) | (_lib.F64 if D.dtype == np.float64 else 0)
    check(L.rfdtd_create_assembled(ctypes.byref(h), ctypes.byref(D.struct), int(excitations), f),
          "rfdtd_create_assembled")
    return L, h


def read_coefficients(L, handle, n_cells, dtype, member=0):
    """The member's 30 packed coefficient arrays as a dict (the x = Nx-1 plane rule applied)."""
    out = {k: np.empty(n_cells, dtype=dtype) for k in _lib.COEFF_NAMES}
    ptrs = _lib.pointer_array([out[k] for k in _lib.COEFF_NAMES])
    check(L.rfdtd_read_coefficients(handle, int(member), ctypes.cast(ptrs, ctypes.c_void_p)), "rfdtd_read_coefficients")
    return out


def read_coefficient_box(L, handle, box, dtype, member=0):
    """The member's 30 packed coefficient arrays over box (x0, x1, y0, y1, z0, z1, packed cells)."""
    x0, x1, y0, y1, z0, z1 = (int(v) for v in box)
    shape = (x1 - x0, y1 - y0, z1 - z0)
    out = {k: np.empty(shape, dtype=dtype) for k in _lib.COEFF_NAMES}
    ptrs = _lib.pointer_array([out[k] for k in _lib.COEFF_NAMES])
    b = (ctypes.c_int32 * 6)(x0, x1, y0, y1, z0, z1)
    check(L.rfdtd_read_coefficient_box(handle, int(member), ctypes.cast(b, ctypes.c_void_p), ctypes.cast(ptrs, ctypes.c_void_p)),
          "rfdtd_read_coefficient_box")
    return out
