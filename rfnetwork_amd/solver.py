"""FDTD model API of rfnetwork's field solver, driving the MI355X engine.

``FDTD_Solver`` keeps the public interface of rfnetwork.solver.FDTD_Solver (rfnetwork/solver.py):
geometry registration (add_dielectric / add_conductor / add_lumped_port / add_resistor /
assign_PML_boundaries), generate_mesh, probes and monitors, Gaussian sources, solve(),
get_sparameters(), vi_probe_values() and get_monitor_data().  Geometry objects are anything with
pyvista-style ``points``/``n_cells``/``faces`` (rfnetwork_amd.geometry provides axis-aligned boxes,
rectangles and lines).

What is restated from the reference (same formulas, same float32/float64 promotion order, so the
30 coefficient arrays come out with the values the reference computes):
  _init_grid                 solver.py:640-707
  _init_dielectrics          solver.py:794-815
  _init_coefficients         solver.py:871-1003 (dt = courant * d_min / c0, courant 0.95 by default)
  _init_PML                  solver.py:1239-1325 (cubic sigma grading)
  _init_conductors           solver.py:817-868 (axis-aligned boxes and rectangles)
  _init_lumped_elements      solver.py:1086-1236
  add_current_probe          solver.py:1893-2018
  add_voltage_probe          solver.py:2021-2079
  add_field_monitor          solver.py:1703-1775
  sources / excitations      solver.py:1328-1408
  solve (packing, unpacking) solver.py:1410-1700
  get_sparameters            solver.py:2442-2532
  edge_correction            solver.py:2534-2700
Additions for the engine: an explicit-grid constructor (``set_grid``) and a uniform mesher standing
in for the reference's graded pyvista mesher, per-cell permittivity volumes
(``add_permittivity_volume``), a ``courant`` factor, and ``solve(dtype=...)`` for float64 runs.
``solve`` runs on the GPU through ``hip_func.solver_run_hip``; there is no CPU solver in this
package.
"""
import itertools
import sys
import time
from copy import copy

import numpy as np

from . import geometry, hip_func, utils
from .units import conv, const
from .utils import LabeledArray

u0 = const.u0
e0 = const.e0
c0 = const.c0

_AXES = ("x", "y", "z")
_FIELDS = ("ex", "ey", "ez", "hx", "hy", "hz")


class FDTD_Solver:
    """FDTD EM solver for rectilinear (PCB-style) geometries."""

    def __init__(self, bounding_box):
        self.bounding_box = bounding_box
        self.dielectrics = dict()
        self.conductors = dict()
        self.lumped_elements = dict()
        self.eps_volumes = []
        self.pml_boundaries = []
        self.monitors = dict()
        self.probes = dict()
        self.ports = []
        self.ports_inv = []
        self._n_pml = None
        self._auto_name_counter = 0
        pts = np.asarray(bounding_box.points, float)
        self.sbox_max = np.max(pts, axis=0)
        self.sbox_min = np.min(pts, axis=0)
        self._places = 5
        self._tol = 1 / (10 ** self._places)
        self._meshed = False
        self._solved = False
        self._time = None
        self._edges = None
        self.courant = 0.95
        # element type of the coefficient assembly: float32 as in the reference; float64 gives exact
        # coefficients for finite-difference checks of gradients
        self.coef_dtype = np.float32
        self.farfield = dict()

    # ------------------------------------------------------------------ state
    @property
    def time(self):
        return self._time

    def invalidate_mesh(self):
        self._meshed = False
        self.invalidate_solution()

    def invalidate_solution(self):
        self._solved = False

    def check_mesh(self):
        if not self._meshed:
            raise RuntimeError("Mesh has been invalidated or has not been created yet.")

    def check_solution(self):
        if not self._solved:
            raise RuntimeError("Solution has been invalidated or model has not been solved yet.")

    # ------------------------------------------------------------------ model objects
    def _add_object(self, group, objects, properties, name=None):
        self.invalidate_mesh()
        if not isinstance(objects, (tuple, list, np.ndarray)):
            objects = [objects]
        if name is not None and name in self.conductors.keys():
            raise ValueError(f"Duplicate conductor name: {name}")
        for i, obj in enumerate(objects):
            if name is None:
                key = f"obj_{self._auto_name_counter}"
                self._auto_name_counter += 1
            else:
                key = name + "_{i}" if len(objects) > 1 else name
            group[key] = dict(obj=obj, **properties)

    def add_dielectric(self, *objects, er: float, loss_tan: float = 0, f0: float = 0, style: dict = None,
                       name: str = None):
        """Rectangular dielectric boxes with relative permittivity er (non-dispersive loss_tan at f0)."""
        self._add_object(self.dielectrics, objects, dict(er=er, loss_tan=loss_tan, f0=f0, style=style or {}), name)

    def add_conductor(self, *objects, sigma: float = 1e16, style: dict = None, name: str = None):
        """Conductors (boxes or rectangles); sigma defaults to the reference's PEC value 1e16."""
        self._add_object(self.conductors, objects, dict(sigma=sigma, style=style or {}), name)

    def add_lumped_port(self, number: int, face, integration_line, r0: float = 50, name: str = None):
        """Lumped port on a rectangular face; integration_line is "x+", ..., "z-" or a Line on the face."""
        if isinstance(integration_line, str):
            self._add_object(self.lumped_elements, face,
                             dict(integration_line=integration_line, r=r0, port=number), name)
            return
        # a line that starts on one face edge and ends inside the face: split the face into the
        # primary port (along the line) and an inverted secondary port, each 2*r0 (solver.py:185-270)
        pts = np.asarray(integration_line.points, float)
        if len(pts) != 2:
            raise ValueError("Integration line must contain only 2 points.")
        f_p0, f_p1 = pts
        f_axis = int(np.argmax(np.abs(f_p1 - f_p0)))
        if not np.all(geometry.inside_axis_aligned(pts, face, self._tol)):
            raise ValueError("Integration line must lie on the lumped element or port face.")
        if np.count_nonzero(np.abs(f_p1 - f_p0) < self._tol) != 2:
            raise ValueError("Integration line must be aligned with the cartesian axes.")
        f_dir = 1 if f_p1[f_axis] > f_p0[f_axis] else -1
        axis_str = f"{_AXES[f_axis]}{'+' if f_dir == 1 else '-'}"
        fpts = np.asarray(face.points, float)
        f_face0, f_face1 = np.min(fpts[:, f_axis]), np.max(fpts[:, f_axis])
        if f_dir == -1:
            f_face0, f_face1 = f_face1, f_face0
        if np.abs(f_face0 - f_p0[f_axis]) > self._tol:
            raise ValueError("Integration line must start on face edge.")
        if not np.abs(f_face0 - f_p1[f_axis]) > self._tol:
            raise ValueError("Use axis notation for integration lines that span the full face")
        fmin, fmax = np.min(fpts, axis=0), np.max(fpts, axis=0)

        def sub_face(lo, hi):
            c = np.array([fmin, fmax, fmax])
            c[0, f_axis] = lo
            c[1, f_axis] = lo
            c[2, f_axis] = hi
            return geometry.Rectangle(c)

        if f_dir == -1:
            pri, sec = sub_face(f_p1[f_axis], f_p0[f_axis]), sub_face(fmin[f_axis], f_p1[f_axis])
        else:
            pri, sec = sub_face(f_p0[f_axis], f_p1[f_axis]), sub_face(f_p1[f_axis], fmax[f_axis])
        self._add_object(self.lumped_elements, pri, dict(integration_line=axis_str, r=r0 * 2, port=number), name)
        inv = axis_str[0] + "-" if f_dir == 1 else "+"
        self._add_object(self.lumped_elements, sec, dict(integration_line=inv, r=r0 * 2, port=-number), name)

    def add_resistor(self, face, value: float, integration_line: str, name: str = None):
        """Resistive sheet of total resistance ``value`` on a rectangular face."""
        self._add_object(self.lumped_elements, face, dict(integration_line=integration_line, r=value, port=None), name)

    def add_permittivity_volume(self, er: np.ndarray, origin_idx, sigma: np.ndarray = None):
        """Per-cell relative permittivity (and optional conductivity, S/m) over a block of cells whose
        lowest cell index is origin_idx; it overrides the dielectric boxes in that block."""
        self.invalidate_mesh()
        er = np.asarray(er, dtype=np.float64)
        if er.ndim != 3:
            raise ValueError("er must be a 3-D array of cells")
        self.eps_volumes.append(dict(er=er, origin=tuple(int(i) for i in origin_idx),
                                     sigma=None if sigma is None else np.asarray(sigma, np.float64)))

    def assign_PML_boundaries(self, *sides: str, n_pml: int = 10):
        """Split-field PML on the given sides ("x+", "x-", ...), n_pml cells on each."""
        self.invalidate_mesh()
        valid = ("x+", "x-", "y+", "y-", "z+", "z-")
        if any(s not in valid for s in sides):
            raise ValueError(f"PML side not recognized. Expecting one of {valid}")
        self.pml_boundaries = copy(list(sides))
        self._n_pml = n_pml

    # ------------------------------------------------------------------ grid
    def set_grid(self, x_edges, y_edges, z_edges):
        """Use explicit grid edges (inches) instead of the mesher."""
        self.invalidate_mesh()
        self._edges = [np.asarray(e, dtype=np.float64) for e in (x_edges, y_edges, z_edges)]
        for e in self._edges:
            if e.ndim != 1 or len(e) < 3 or np.any(np.diff(e) <= 0):
                raise ValueError("grid edges must be increasing 1-D arrays with at least 3 points")

    def pos_to_idx(self, position, mode="edge"):
        """Index of the grid edge (or cell centre) on or just above position on each axis."""
        self.check_mesh()
        modes = [mode] * 3 if isinstance(mode, str) else mode
        out = []
        for i, m in enumerate(modes):
            g = self.g_edges[i] if m == "edge" else self.g_cells[i]
            diff = g - position[i]
            out.append(int(np.argmax(diff >= -self._tol)) if diff[-1] > -self._tol else len(g))
        return tuple(out)

    def field_pos_to_idx(self, position, field):
        """Index of the field component on or just above position on each axis."""
        self.check_mesh()
        out = []
        for i, g in enumerate(self.floc[field]):
            diff = g - position[i]
            out.append(int(np.argmax(diff >= -self._tol)) if diff[-1] > -self._tol else len(g))
        return tuple(out)

    def _mesh_edges(self, d_max):
        """Uniform mesher: object extents are hard edges; each gap is split evenly into cells no wider
        than d_max (the reference grades cells with a pyvista-based mesher, solver.py:505-792)."""
        objs = [self.bounding_box] + [d["obj"] for d in self.dielectrics.values()]
        objs += [d["obj"] for d in self.conductors.values()] + [d["obj"] for d in self.lumped_elements.values()]
        edges = []
        for a in range(3):
            pts = np.concatenate([np.asarray(o.points, float)[:, a] for o in objs])
            pts = np.clip(pts, self.sbox_min[a], self.sbox_max[a])
            hard = np.unique(np.around(pts, decimals=self._places))
            e = [hard[0]]
            for lo, hi in zip(hard[:-1], hard[1:]):
                n = max(1, int(np.ceil((hi - lo) / d_max - 1e-9)))
                e.extend(lo + (hi - lo) * np.arange(1, n + 1) / n)
            edges.append(np.around(np.array(e), decimals=self._places))
        return edges

    def generate_mesh(self, d_max: float = None, d_min: float = None, surface_tolerance: float = 0.1,
                      courant: float = None, coefficients: str = "host"):
        """Build the grid and the FDTD coefficients (order of solver.py:413-449).

        coefficients="device" skips the host's full-grid coefficient arrays: the model keeps the cell
        permittivity and conductivity, the time step, the ports (their source gains from the edge
        permittivity of the port edges alone) and the edge corrections as a list, and a session of it
        assembles the arrays on the GPU (Session.assembled; solve() does so).  The arrays are the same
        bit for bit; coefficient_arrays() is then unavailable."""
        if coefficients not in ("host", "device"):
            raise ValueError('coefficients must be "host" or "device"')
        if courant is not None:
            self.courant = float(courant)
        if self._edges is None:
            if d_max is None:
                raise ValueError("d_max is required without an explicit grid (set_grid)")
            if d_min is not None and d_min > d_max / 1.5:
                raise ValueError("d_min must be less than d_max.")
            edges = self._mesh_edges(d_max)
        else:
            edges = self._edges
        self._meshed = True
        self._coef_mode = coefficients
        self._edge_ops = []  # edge-correction factors applied since this mesh: (Db name, index, factor)
        self._init_grid(edges)
        self._init_dielectrics()
        if coefficients == "host":
            self._init_coefficients()
            for side in self.pml_boundaries:
                self._init_PML(side)
            self._init_conductors()
        else:
            self._init_time_step()
        self._init_lumped_elements()
        self.invalidate_solution()

    def _init_grid(self, edges):
        dtype_ = np.float32
        gx, gy, gz = edges
        dx, dy, dz = np.diff(gx).astype(dtype_), np.diff(gy).astype(dtype_), np.diff(gz).astype(dtype_)
        self.n_cells = len(dx), len(dy), len(dz)
        self.Nx, self.Ny, self.Nz = self.n_cells
        self.dx, self.dy, self.dz = dx, dy, dz
        gx_h, gy_h, gz_h = [(g[1:] + g[:-1]) / 2 for g in (gx, gy, gz)]
        dx_h, dy_h, dz_h = [(d[1:] + d[:-1]) / 2 for d in (dx, dy, dz)]
        self.g_edges = gx, gy, gz
        self.g_cells = gx_h, gy_h, gz_h
        self.d_cells = dx, dy, dz
        self.dh_cells = dx_h, dy_h, dz_h
        self.dx_h, self.dy_h, self.dz_h = dx_h, dy_h, dz_h
        self.eps = np.ones(self.n_cells) * e0
        self.sigma = np.zeros(self.n_cells)
        self.floc = dict(ex=(gx_h, gy, gz), ey=(gx, gy_h, gz), ez=(gx, gy, gz_h),
                         hx=(gx, gy_h, gz_h), hy=(gx_h, gy, gz_h), hz=(gx_h, gy_h, gz))
        Nx, Ny, Nz = self.n_cells
        self.fshape = dict(ex=(Nx, Ny + 1, Nz + 1), ey=(Nx + 1, Ny, Nz + 1), ez=(Nx + 1, Ny + 1, Nz),
                           hx=(Nx + 1, Ny, Nz), hy=(Nx, Ny + 1, Nz), hz=(Nx, Ny, Nz + 1))
        dx_hp, dy_hp, dz_hp = np.pad(dx_h, (1, 1)), np.pad(dy_h, (1, 1)), np.pad(dz_h, (1, 1))
        self.fcell_w = dict(ex=(dx, dy_hp, dz_hp), ey=(dx_hp, dy, dz_hp), ez=(dx_hp, dy_hp, dz),
                            hx=(dx_hp, dy, dz), hy=(dx, dy_hp, dz), hz=(dx, dy, dz_hp))

    def _init_dielectrics(self):
        for sub in self.dielectrics.values():
            if sub["obj"].n_cells > 6:
                raise NotImplementedError("Only rectangular dielectrics are supported.")
            pts = np.asarray(sub["obj"].points, float)
            x0, y0, z0 = self.pos_to_idx(np.min(pts, axis=0), mode="cell")
            x1, y1, z1 = self.pos_to_idx(np.max(pts, axis=0), mode="cell")
            blk = (slice(x0, x1), slice(y0, y1), slice(z0, z1))
            # the larger permittivity and conductivity win where boxes overlap
            self.eps[blk] = np.where(self.eps[blk] > e0 * sub["er"], self.eps[blk], e0 * sub["er"])
            s = sub["loss_tan"] * e0 * sub["er"] * 2 * np.pi * sub["f0"]
            self.sigma[blk] = np.where(self.sigma[blk] > s, self.sigma[blk], s)
        for vol in self.eps_volumes:
            er, (x0, y0, z0) = vol["er"], vol["origin"]
            nx, ny, nz = er.shape
            if x0 < 0 or y0 < 0 or z0 < 0 or x0 + nx > self.Nx or y0 + ny > self.Ny or z0 + nz > self.Nz:
                raise ValueError("permittivity volume outside the grid")
            blk = (slice(x0, x0 + nx), slice(y0, y0 + ny), slice(z0, z0 + nz))
            self.eps[blk] = er * e0
            if vol["sigma"] is not None:
                self.sigma[blk] = vol["sigma"]

    def _init_time_step(self):
        dtype_ = self.coef_dtype
        dx, dy, dz = [conv.m_in(d) for d in self.d_cells]
        # time step from the smallest cell on each axis (free-space Courant limit)
        lmin = np.array([np.min(dx), np.min(dy), np.min(dz)], dtype=dtype_)
        dmin = 1 / np.sqrt(((1 / lmin) ** 2).sum())
        dt = self.courant * (dmin / const.c0)
        if isinstance(dt, float):  # a python float courant keeps dt in float32 (numpy promotion)
            dt = np.float32(dt)
        self.dt = dt

    def _edge_eps(self, f, blk):
        """Permittivity of E component f on a block of its global edges, in the coefficient dtype,
        formed as _init_coefficients forms eps_ex / eps_ey / eps_ez (a float64 average of the 4
        adjacent cells weighted by float32 half widths; e0 on the boundary), without the full-grid
        arrays."""
        T = np.dtype(self.coef_dtype).type
        dx, dy, dz = [conv.m_in(d) for d in self.d_cells]
        d = (dx, dy, dz)
        Nx, Ny, Nz = self.n_cells
        axes = []
        for b, n in zip(blk, self.fshape[f]):
            axes.append(np.arange(*b.indices(n)) if isinstance(b, slice) else np.array([int(b)]))
        # the two axes averaged, in the host's order (Ex: y then z, Ey: x then z, Ez: x then y)
        first, second = {"ex": (1, 2), "ey": (0, 2), "ez": (0, 1)}[f]
        v = self.eps
        out = np.empty([len(a) for a in axes], dtype=T)

        def avg(v0, v1, w0, w1):
            return (v0 * (w0 / 2) + v1 * (w1 / 2)) / (w0 / 2 + w1 / 2)

        for n, g in enumerate(np.ndindex(*out.shape)):
            gi = [int(axes[a][g[a]]) for a in range(3)]
            lim = (Nx, Ny, Nz)
            if not (1 <= gi[first] <= lim[first] - 1 and 1 <= gi[second] <= lim[second] - 1):
                out[g] = T(e0)
                continue
            c = list(gi)

            def cell(a_off, b_off):
                cc = list(c)
                cc[first] = gi[first] - 1 + a_off
                cc[second] = gi[second] - 1 + b_off
                return v[tuple(cc)]

            w_a0, w_a1 = d[first][gi[first] - 1], d[first][gi[first]]
            w_b0, w_b1 = d[second][gi[second] - 1], d[second][gi[second]]
            u0 = avg(cell(0, 0), cell(1, 0), w_a0, w_a1)
            u1 = avg(cell(0, 1), cell(1, 1), w_a0, w_a1)
            out[g] = T(avg(u0, u1, w_b0, w_b1))
        return out

    def _init_coefficients(self):
        dtype_ = self.coef_dtype
        dx, dy, dz = [conv.m_in(d) for d in self.d_cells]
        Nx, Ny, Nz = self.n_cells
        self._init_time_step()
        dt = self.dt

        def edge_average(v):
            # cells adjacent to an edge, weighted by their half widths along each axis
            ax0, ax1 = dx[:-1][:, None, None], dx[1:][:, None, None]
            ay0, ay1 = dy[:-1][None, :, None], dy[1:][None, :, None]
            az0, az1 = dz[:-1][None, None], dz[1:][None, None]
            vx = (v[:-1] * (ax0 / 2) + v[1:] * (ax1 / 2)) / (ax0 / 2 + ax1 / 2)
            vy = (v[:, :-1] * (ay0 / 2) + v[:, 1:] * (ay1 / 2)) / (ay0 / 2 + ay1 / 2)
            vz = (v[..., :-1] * (az0 / 2) + v[..., 1:] * (az1 / 2)) / (az0 / 2 + az1 / 2)
            vyz = (vy[..., :-1] * (az0 / 2) + vy[..., 1:] * (az1 / 2)) / (az0 / 2 + az1 / 2)
            vxz = (vx[..., :-1] * (az0 / 2) + vx[..., 1:] * (az1 / 2)) / (az0 / 2 + az1 / 2)
            vxy = (vx[:, :-1] * (ay0 / 2) + vx[:, 1:] * (ay1 / 2)) / (ay0 / 2 + ay1 / 2)
            return vx, vy, vz, vxy, vxz, vyz

        ex_, ey_, ez_, exy, exz, eyz = edge_average(self.eps)
        sx_, sy_, sz_, sxy, sxz, syz = edge_average(self.sigma)
        fs = self.fshape
        self.eps_ex = np.ones(fs["ex"], dtype=dtype_) * e0
        self.eps_ey = np.ones(fs["ey"], dtype=dtype_) * e0
        self.eps_ez = np.ones(fs["ez"], dtype=dtype_) * e0
        self.eps_ex[:, 1:-1, 1:-1] = eyz
        self.eps_ey[1:-1, :, 1:-1] = exz
        self.eps_ez[1:-1, 1:-1] = exy
        self.eps_hx = np.ones(fs["hx"], dtype=dtype_) * e0
        self.eps_hy = np.ones(fs["hy"], dtype=dtype_) * e0
        self.eps_hz = np.ones(fs["hz"], dtype=dtype_) * e0
        self.eps_hx[1:-1] = ex_
        self.eps_hy[:, 1:-1] = ey_
        self.eps_hz[..., 1:-1] = ez_
        self.sig_ex = np.zeros(fs["ex"], dtype=dtype_)
        self.sig_ey = np.zeros(fs["ey"], dtype=dtype_)
        self.sig_ez = np.zeros(fs["ez"], dtype=dtype_)
        self.sig_ex[:, 1:-1, 1:-1] = syz
        self.sig_ey[1:-1, :, 1:-1] = sxz
        self.sig_ez[1:-1, 1:-1] = sxy

        def ca(eps, sig):
            return (2 * eps - (sig * dt)) / (2 * eps + (sig * dt))

        def cb(eps, sig):
            return (2 * dt) / ((2 * eps + (sig * dt)))

        Ca = {c: ca(e, s) for c, e, s in (("x", self.eps_ex, self.sig_ex), ("y", self.eps_ey, self.sig_ey),
                                          ("z", self.eps_ez, self.sig_ez))}
        Cb = {c: cb(e, s) for c, e, s in (("x", self.eps_ex, self.sig_ex), ("y", self.eps_ey, self.sig_ey),
                                          ("z", self.eps_ez, self.sig_ez))}
        splits = dict(ex=("ex_y", "ex_z"), ey=("ey_z", "ey_x"), ez=("ez_x", "ez_y"))
        self.Ca, self.Cb = {}, {}
        for f, (s1, s2) in splits.items():
            for s in (s1, s2):
                self.Ca[s] = Ca[f[1]].copy()
                self.Cb[s] = Cb[f[1]].copy()
        Db0 = dt / u0
        hshape = dict(hx=fs["hx"], hy=fs["hy"], hz=fs["hz"])
        hsplit = dict(hx=("y", "z"), hy=("z", "x"), hz=("x", "y"))
        self.Da, self.Db = {}, {}
        for h, (a, b) in hsplit.items():
            for ax in (a, b):
                self.Da[f"{h}_{ax}"] = np.ones(hshape[h], dtype=dtype_) * 1
                self.Db[f"{h}_{ax}1"] = np.ones(hshape[h], dtype=dtype_) * Db0
                self.Db[f"{h}_{ax}2"] = np.ones(hshape[h], dtype=dtype_) * Db0

    def _pml_profiles(self, side: str):
        """(axis index, upper side, layers, sigma at the E edges, sigma at the H half cells) of one
        PML side (solver.py:1255-1274): float64 arrays over the layers, counted from the inner
        edge."""
        n = self._n_pml
        ai = _AXES.index(side[0])
        order = 3
        cells = self.d_cells[ai]
        d_border = conv.m_in(cells[-1]) if side[1] == "+" else conv.m_in(cells[0])
        eta0 = np.sqrt(u0 / e0)
        sigma_max = 0.8 * (order + 1) / (eta0 * d_border)
        i = np.arange(0, n)
        sig_edge = sigma_max * (i / n) ** order          # sigma at the E edges of the layer
        sig_half = sigma_max * ((i + 0.5) / n) ** order  # sigma at the H half-cell points
        return ai, side[1] == "+", n, sig_edge, sig_half

    def _init_PML(self, side: str):
        n = self._n_pml
        if n is None:
            return
        axis = side[0]
        ai, _, _, sig_edge, sig_half = self._pml_profiles(side)
        dt = self.dt
        shape = [None] * 3
        shape[ai] = slice(None)
        sig_edge = sig_edge[tuple(shape)]
        sig_half = sig_half[tuple(shape)]
        e_idx = [slice(None)] * 3
        h_idx = [slice(None)] * 3
        e_idx[ai] = slice(n, 0, -1) if side[1] == "-" else slice(-n - 1, -1)
        h_idx[ai] = slice(n - 1, None, -1) if side[1] == "-" else slice(-n, None)
        e_idx, h_idx = tuple(e_idx), tuple(h_idx)
        feps = dict(ex=self.eps_ex, ey=self.eps_ey, ez=self.eps_ez, hx=self.eps_hx, hy=self.eps_hy, hz=self.eps_hz)
        e_graded = (("ey", "ez"), ("ez", "ex"), ("ex", "ey"))[ai]
        h_graded = (("hy", "hz"), ("hz", "hx"), ("hx", "hy"))[ai]
        for e in e_graded:
            eps = feps[e][e_idx]
            sig = np.broadcast_to(sig_edge, eps.shape).copy()
            sig *= eps / e0  # constant sigma/eps across the layer
            self.Ca[f"{e}_{axis}"][e_idx] = (2 * eps - (sig * dt)) / (2 * eps + (sig * dt))
            self.Cb[f"{e}_{axis}"][e_idx] = (2 * dt) / ((2 * eps + (sig * dt)))
        for h in h_graded:
            eps = feps[h][h_idx]
            sig = np.broadcast_to(sig_half, eps.shape).copy()
            sig *= eps / e0
            sig_m = sig * u0 / eps
            self.Da[f"{h}_{axis}"][h_idx] = (2 * u0 - (sig_m * dt)) / (2 * u0 + (sig_m * dt))
            self.Db[f"{h}_{axis}1"][h_idx] = (2 * dt) / ((2 * u0 + (sig_m * dt)))
            self.Db[f"{h}_{axis}2"][h_idx] = (2 * dt) / ((2 * u0 + (sig_m * dt)))

    def _conductor_edges(self):
        """Per conductor and E component, in the order the host applies them: (component, block of
        global edge indices, inside mask over the block, sigma)."""
        for cond in self.conductors.values():
            obj, sig = cond["obj"], cond["sigma"]
            pts = np.asarray(obj.points, float)
            for f in ("ex", "ey", "ez"):
                p0 = self.field_pos_to_idx(np.min(pts, axis=0), f)
                p1 = self.field_pos_to_idx(np.max(pts, axis=0), f)
                blk = tuple(slice(p0[i], p1[i] + 1) for i in range(3))
                grid = np.meshgrid(*[self.floc[f][i][blk[i]] for i in range(3)], indexing="ij")
                points = np.stack(grid, axis=-1)
                inside = geometry.inside_axis_aligned(points, obj, self._tol)
                yield f, blk, inside, sig

    def _init_conductors(self):
        feps = dict(ex=self.eps_ex, ey=self.eps_ey, ez=self.eps_ez)
        splits = dict(ex=("ex_y", "ex_z"), ey=("ey_z", "ey_x"), ez=("ez_x", "ez_y"))
        for f, blk, inside, sig in self._conductor_edges():
            eps = feps[f][blk]
            ca_c = (2 * eps - (sig * self.dt)) / (2 * eps + (sig * self.dt))
            cb_c = (2 * self.dt) / ((2 * eps + (sig * self.dt)))
            for s in splits[f]:
                self.Ca[s][blk] = np.where(inside, ca_c, self.Ca[s][blk])
                self.Cb[s][blk] = np.where(inside, cb_c, self.Cb[s][blk])

    def _init_lumped_elements(self):
        self._lumped_fix = []  # (component axis, global edge block, df / (2 r dn dw)) per element
        ports = [el for el in self.lumped_elements.values() if el["port"] is not None]
        if ports:
            n_ports = max(abs(el["port"]) for el in ports)
            self.ports = [None] * n_ports
            self.ports_inv = [None] * n_ports
        d_cells = [conv.m_in(d) for d in self.d_cells]
        dh_cells = [conv.m_in(d) for d in self.dh_cells]
        for el in self.lumped_elements.values():
            face, r, port = el["obj"], el["r"], el["port"]
            line = el["integration_line"]
            if face.n_cells > 1 or face.faces[0] != 4:
                raise ValueError("Only rectangular port faces are supported.")
            fpts = np.asarray(face.points, float)
            p1 = np.array(self.pos_to_idx(np.min(fpts, axis=0), mode="edge"))
            p2 = np.array(self.pos_to_idx(np.max(fpts, axis=0), mode="edge"))
            span = p2 - p1
            n_axis = int(np.argmin(span))
            if np.count_nonzero(span) != 2:
                raise ValueError("Port face must be on cartesian grid, and must be 2D.")
            f_axis = _AXES.index(line[0])
            f_dir = {"+": 1, "-": -1}[line[1]]
            w_axis = [a for a in (0, 1, 2) if a != f_axis and a != n_axis][0]
            if port is not None and span[w_axis] < 2:
                raise NotImplementedError("Port faces must span at least 2 grid cells.")
            if port is not None:  # port faces drop the two end components along the width
                p1[w_axis] += 1
                p2[w_axis] -= 1
            if f_axis == n_axis:
                raise ValueError("Integration direction must be in the port face.")
            na, wa, fa = _AXES[n_axis], _AXES[w_axis], _AXES[f_axis]

            def place(normal, width, along):
                idx = [slice(None)] * 3
                idx[w_axis], idx[f_axis], idx[n_axis] = width, along, normal
                return tuple(idx)

            dw = dh_cells[w_axis][p1[w_axis] - 1: p2[w_axis]]
            dn = dh_cells[n_axis][p1[n_axis] - 1: p2[n_axis]]
            df = d_cells[f_axis][p1[f_axis]: p2[f_axis]]
            gx, gy, gz = np.meshgrid(*place(dn, dw, df), indexing="ij")
            dn, dw, df = [(gx, gy, gz)[a] for a in (n_axis, w_axis, f_axis)]
            eidx = place(slice(p1[n_axis], p2[n_axis] + 1), slice(p1[w_axis], p2[w_axis] + 1),
                         slice(p1[f_axis], p2[f_axis]))
            if getattr(self, "_coef_mode", "host") == "host":
                eps_r = (self.eps_ex, self.eps_ey, self.eps_ez)[f_axis][eidx]
            else:
                eps_r = self._edge_eps(f"e{fa}", eidx)
            nw = (p2 - p1 + 1)[w_axis]
            nf = (p2 - p1)[f_axis]
            r_cell = r * nw / nf  # the parallel/series split of r over the face's components
            rterm = r_cell * dn * dw
            q = df / (2 * rterm)
            self._lumped_fix.append((f_axis, eidx, q))
            denom = (eps_r / self.dt) + q
            if getattr(self, "_coef_mode", "host") == "host":
                for s in (f"e{fa}_{wa}", f"e{fa}_{na}"):
                    self.Ca[s][eidx] = ((eps_r / self.dt) - q) / denom
                    self.Cb[s][eidx] = 1 / denom
            if port is None:
                continue
            group = self.ports if port > 0 else self.ports_inv
            group[abs(port) - 1] = dict(idx=eidx, Vs_a=-f_dir / (denom * rterm * nf), field=f"e{fa}", axis=n_axis,
                                        r0=r, direction=f_dir, src=None)
            # ampere loop around the termination, normal to the integration axis
            fp1, fp2 = np.min(fpts, axis=0), np.max(fpts, axis=0)
            loc = self.floc[f"e{fa}"][f_axis]
            fpos_n = loc[int((p2 + p1)[f_axis] / 2)]
            cface = geometry.Rectangle([
                place(fp1[n_axis] - 0.001, fp1[w_axis] - 0.001, fpos_n),
                place(fp1[n_axis] - 0.001, fp2[w_axis] + 0.001, fpos_n),
                place(fp1[n_axis] + 0.001, fp2[w_axis] + 0.001, fpos_n),
            ])
            self.add_current_probe(f"port_{port}", cface)

    # ------------------------------------------------------------------ probes and monitors
    def add_field_monitor(self, name: str, field: str, axis: str, position: float = None, index: int = None,
                          n_step: int = 1, frequency=None, dft_dtype=None):
        """Plane monitor of one component (or "e_total") normal to axis, time samples every n_step
        steps or running DFTs at ``frequency``.  dft_dtype=np.complex128 accumulates the running DFT
        (and evaluates its phase table) in double also in a float32 solve; by default it is complex64
        in a float32 solve, as in the reference (solver.py:1604-1617)."""
        if dft_dtype is not None and np.dtype(dft_dtype) not in (np.dtype(np.complex64), np.dtype(np.complex128)):
            raise ValueError("dft_dtype must be complex64 or complex128")
        supported = tuple(self.fshape.keys()) + ("e_total",)
        if field not in supported:
            raise ValueError(f"Unsupported field: {field}. Expecting one of: {supported}")
        if axis not in _AXES:
            raise ValueError(f"Unsupported axis: {axis}")
        ai = _AXES.index(axis)
        if field == "e_total":
            fields, names = ("ex", "ey", "ez"), [f"{name}_{a}" for a in _AXES]
        else:
            fields, names = [field], [name]
        for nm, f in zip(names, fields):
            shape = list(self.fshape[f])
            axis_len = shape.pop(ai)
            idx, pos = index, position
            if pos is not None:
                full = [0] * 3
                full[ai] = pos
                idx = int(self.field_pos_to_idx(full, f)[ai])
            elif idx is not None:
                pos = self.floc[f][ai][idx]
            if idx < 0 or idx >= axis_len - 1:
                raise ValueError("Field position out of bounds")
            self.monitors[nm] = dict(field=f, axis=ai, position=pos, index=idx, n_step=n_step, shape=tuple(shape),
                                     frequency=np.atleast_1d(frequency) if frequency is not None else None,
                                     dft_dtype=None if dft_dtype is None else np.dtype(dft_dtype))

    def add_farfield_monitor(self, frequency, padding: int = 2):
        """Phasor monitors on the faces of a box ``padding`` cells inside the PML, for the far field
        (rfnetwork/solver.py:1777-1891): on every face with PML outside it, the two tangential E
        components on the face and the two tangential H components on the planes either side of it.
        Faces without PML are PEC walls and get no monitors (they radiate nothing)."""
        self.check_mesh()
        if not len(self.pml_boundaries):
            raise ValueError("No radiation or PML boundaries found. Unable to create farfield monitor.")
        frequency = np.atleast_1d(np.asarray(frequency, dtype=np.float64))
        ff_idx = np.zeros((3, 2), dtype=np.int64)
        self.farfield["surf_pos"] = np.zeros((3, 2), dtype=np.float64)
        self.farfield["cell_pos"] = [None] * 3
        self.farfield["cell_w"] = [None] * 3
        for axis in range(3):
            a_s = _AXES[axis]
            sf0, sf1 = [i for i in (0, 1, 2) if i != axis]
            for j, side in enumerate(("n", "p")):
                pml = a_s + ("-" if side == "n" else "+")
                skipped = pml not in self.pml_boundaries
                if skipped:
                    ff_idx[axis, j] = 0 if j == 0 else len(self.g_edges[axis]) - 1
                elif side == "n":
                    ff_idx[axis, j] = self._n_pml + padding
                else:
                    ff_idx[axis, j] = len(self.g_edges[axis]) - self._n_pml - 1 - padding
                surf = int(ff_idx[axis, j])
                self.farfield["surf_pos"][axis, j] = conv.m_in(self.g_edges[axis][surf])
                if skipped:
                    continue
                for f in (sf0, sf1):
                    f_s = _AXES[f]
                    self.add_field_monitor(f"ff_e{f_s}_{side}{a_s}", f"e{f_s}", a_s, index=surf, frequency=frequency)
                    self.add_field_monitor(f"ff_h{f_s}1_{side}{a_s}", f"h{f_s}", a_s, index=surf, frequency=frequency)
                    other = surf - 1 if j == 0 else surf + 1
                    self.add_field_monitor(f"ff_h{f_s}2_{side}{a_s}", f"h{f_s}", a_s, index=other, frequency=frequency)
        for axis in range(3):
            sf0, sf1 = [i for i in (0, 1, 2) if i != axis]
            c0s = slice(ff_idx[sf0, 0], ff_idx[sf0, 1])
            c1s = slice(ff_idx[sf1, 0], ff_idx[sf1, 1])
            self.farfield["cell_pos"][axis] = np.meshgrid(conv.m_in(self.g_cells[sf0][c0s]),
                                                          conv.m_in(self.g_cells[sf1][c1s]), indexing="ij")
            self.farfield["cell_w"][axis] = np.meshgrid(conv.m_in(self.d_cells[sf0][c0s]),
                                                        conv.m_in(self.d_cells[sf1][c1s]), indexing="ij")
        self.farfield["frequency"] = frequency
        self.farfield["idx"] = ff_idx

    def _farfield_faces(self):
        """Equivalent surface currents on the far-field box faces (rfnetwork/solver.py:2907-3002): E
        and H averaged to the cell centres of each face, J = n x H, M = -n x E, complex64."""
        ff_idx = self.farfield["idx"]
        nf = len(self.farfield["frequency"])
        faces = []
        for axis in range(3):
            a_s = _AXES[axis]
            sf0, sf1 = [i for i in (0, 1, 2) if i != axis]
            shape = self.farfield["cell_pos"][axis][0].shape
            r0 = np.ascontiguousarray(self.farfield["cell_pos"][axis][0], dtype=np.float32)
            r1 = np.ascontiguousarray(self.farfield["cell_pos"][axis][1], dtype=np.float32)
            ds = np.ascontiguousarray(self.farfield["cell_w"][axis][0] * self.farfield["cell_w"][axis][1],
                                      dtype=np.float32)
            for j, side in enumerate(("n", "p")):
                e_xyz = np.zeros((3, nf) + shape, dtype=np.complex64)
                h_xyz = np.zeros((3, nf) + shape, dtype=np.complex64)
                present = False
                for f in (sf0, sf1):
                    f_s = _AXES[f]
                    emon, hm1, hm2 = f"ff_e{f_s}_{side}{a_s}", f"ff_h{f_s}1_{side}{a_s}", f"ff_h{f_s}2_{side}{a_s}"
                    if emon not in self.monitors:
                        continue
                    present = True
                    edata = np.asarray(self.get_monitor_data(emon))
                    h1, h2 = np.asarray(self.get_monitor_data(hm1)), np.asarray(self.get_monitor_data(hm2))
                    w1 = self.d_cells[axis][self.monitors[hm1]["index"]]
                    w2 = self.d_cells[axis][self.monitors[hm2]["index"]]
                    hdata = (h1 * (w1 / 2) + h2 * (w2 / 2)) / ((w1 / 2) + (w2 / 2))
                    # E averaged along the other in-face axis, H along its own axis, to the cell centres
                    ax_e = 1 if f == sf1 else 2
                    ax_h = 1 if f == sf0 else 2
                    lo, hi = [slice(None)] * 3, [slice(None)] * 3
                    lo[ax_e], hi[ax_e] = slice(1, None), slice(None, -1)
                    ecell = (edata[tuple(lo)] + edata[tuple(hi)]) / 2
                    lo, hi = [slice(None)] * 3, [slice(None)] * 3
                    lo[ax_h], hi[ax_h] = slice(1, None), slice(None, -1)
                    hcell = (hdata[tuple(lo)] + hdata[tuple(hi)]) / 2
                    win = (slice(None), slice(ff_idx[sf0, 0], ff_idx[sf0, 1]), slice(ff_idx[sf1, 0], ff_idx[sf1, 1]))
                    e_xyz[f] = ecell[win]
                    h_xyz[f] = hcell[win]
                if not present:
                    continue
                n = np.zeros(3)
                n[axis] = -1 if j == 0 else 1
                J = np.ascontiguousarray(np.cross(n, h_xyz, axis=0), dtype=np.complex64)
                M = np.ascontiguousarray(np.cross(-n, e_xyz, axis=0), dtype=np.complex64)
                faces.append(dict(axis=axis, position=float(self.farfield["surf_pos"][axis, j]), r0=r0, r1=r1, ds=ds,
                                  J=J, M=M))
        return faces

    def get_farfield_rE(self, theta, phi):
        """Far-field r E (exp(-j beta r)/r removed) for theta and phi in degrees: labelled array
        [polarization (thetapol, phipol), frequency, theta, phi] (rfnetwork/solver.py:2870-3020).
        The radiation integrals run on the GPU (rfdtd_nf2ff)."""
        from . import _lib

        self.check_solution()
        if not len(self.farfield.keys()):
            raise RuntimeError("Far-field monitor not found.")
        theta = np.atleast_1d(theta)
        phi = np.atleast_1d(phi)
        frequency = self.farfield["frequency"]
        faces = self._farfield_faces()
        data = _lib.nf2ff(faces, np.asarray(2 * np.pi * frequency / const.c0, dtype=np.float32),
                          np.asarray(np.deg2rad(theta), dtype=np.float32), np.asarray(np.deg2rad(phi), dtype=np.float32),
                          float(const.eta0))
        return LabeledArray(data, coords=dict(polarization=["thetapol", "phipol"], frequency=frequency, theta=theta,
                                              phi=phi))

    def get_farfield_gain(self, theta, phi, polarization=None):
        """Realised gain [polarization, frequency, theta, phi] against the power the sources deliver into
        50 ohm (rfnetwork/solver.py:2804-2868); polarization: thetapol, phipol, rhcp, lhcp."""
        rE = np.asarray(self.get_farfield_rE(theta, phi))
        if polarization is None:
            polarization = ["thetapol", "phipol"]
        polarization = list(np.atleast_1d(polarization))
        vectors = dict(thetapol=[1, 0], phipol=[0, 1], rhcp=[1 / np.sqrt(2), 1j / np.sqrt(2)],
                       lhcp=[1 / np.sqrt(2), -1j / np.sqrt(2)])
        proj = np.array([vectors[k] for k in polarization])
        rE_pol = np.einsum("ij,j...->i...", proj, rE)
        frequency = self.farfield["frequency"]
        sources = [p["src"] for p in self.ports if p is not None and p["src"] is not None]
        Vs = np.array([utils.dtft(v, frequency, 1 / self.dt, downsample=False) for v in sources])
        Pin = np.sum((1 / 2) * (np.abs(Vs) ** 2 / 50), axis=0)
        U = (1 / (2 * const.eta0)) * np.abs(rE_pol) ** 2
        gain = (4 * np.pi / Pin[None, :, None, None]) * U
        return LabeledArray(gain, coords=dict(polarization=polarization, frequency=frequency,
                                              theta=np.atleast_1d(theta), phi=np.atleast_1d(phi)))

    def add_current_probe(self, name: str, face):
        """Ampere-loop probes of H around a rectangular face; vi_probe_values(name) gives the current
        along the positive normal axis."""
        dx_h, dy_h, dz_h = [conv.m_in(d) for d in self.dh_cells]
        if face.n_cells > 1 or face.faces[0] != 4:
            raise ValueError("Only rectangular current faces are supported.")
        fpts = np.asarray(face.points, float)
        pmin, pmax = np.min(fpts, axis=0), np.max(fpts, axis=0)
        size = pmax - pmin
        axis = int(np.argmin(size))
        if np.count_nonzero(size) != 2:
            raise ValueError("Port face must be on cartesian grid, and must be 2D.")
        i = 0

        def put(f, index, d):
            nonlocal i
            self.probes[f"{name}_{i}"] = dict(field=f, index=tuple(int(v) for v in index), d=d)
            i += 1

        if axis == 0:
            x0 = self.field_pos_to_idx(pmin, "hy")[0]
            y0 = self.field_pos_to_idx(pmin, "hz")[1] - 1
            y1 = self.field_pos_to_idx(pmax, "hz")[1]
            z0 = self.field_pos_to_idx(pmin, "hy")[2] - 1
            z1 = self.field_pos_to_idx(pmax, "hy")[2]
            for z in np.arange(z0 + 1, z1 + 1):
                put("hz", (x0, y0, z), -dz_h[z - 1])
                put("hz", (x0, y1, z), dz_h[z - 1])
            for y in np.arange(y0 + 1, y1 + 1):
                put("hy", (x0, y, z0), dy_h[y - 1])
                put("hy", (x0, y, z1), -dy_h[y - 1])
        elif axis == 1:
            y0 = self.field_pos_to_idx(pmin, "hz")[1]
            z0 = self.field_pos_to_idx(pmin, "hx")[2] - 1
            z1 = self.field_pos_to_idx(pmax, "hx")[2]
            x0 = self.field_pos_to_idx(pmin, "hz")[0] - 1
            x1 = self.field_pos_to_idx(pmax, "hz")[0]
            for z in np.arange(z0 + 1, z1 + 1):
                put("hz", (x0, y0, z), dz_h[z - 1])
                put("hz", (x1, y0, z), -dz_h[z - 1])
            for x in np.arange(x0 + 1, x1 + 1):
                put("hx", (x, y0, z0), -dx_h[x - 1])
                put("hx", (x, y0, z1), dx_h[x - 1])
        else:
            z0 = self.field_pos_to_idx(pmin, "hx")[2]
            y0 = self.field_pos_to_idx(pmin, "hx")[1] - 1
            y1 = self.field_pos_to_idx(pmax, "hx")[1]
            x0 = self.field_pos_to_idx(pmin, "hy")[0] - 1
            x1 = self.field_pos_to_idx(pmax, "hy")[0]
            for y in np.arange(y0 + 1, y1 + 1):
                put("hy", (x0, y, z0), -dy_h[y - 1])
                put("hy", (x1, y, z0), dy_h[y - 1])
            for x in np.arange(x0 + 1, x1 + 1):
                put("hx", (x, y0, z0), dx_h[x - 1])
                put("hx", (x, y1, z0), -dx_h[x - 1])

    def add_voltage_probe(self, name: str, line):
        """E probes along an axis-aligned line; vi_probe_values(name) is the line integral p1 -> p2."""
        pts = np.asarray(line.points, float)
        seg = np.diff(pts, axis=0)
        if np.count_nonzero(seg) != 1:
            raise ValueError("Voltage probe must be parallel to the caresian axes.")
        axis = int(np.argmax(np.any(seg, axis=0)))
        field = ("ex", "ey", "ez")[axis]
        if len(pts) != 2:
            raise ValueError("Voltage probe must be a line.")
        start, end = np.min(pts, axis=0), np.max(pts, axis=0)
        i0 = list(self.field_pos_to_idx(start, field))
        i1 = list(self.field_pos_to_idx(end, field))
        if self.floc[field][axis][i1[axis]] > end[axis] + self._tol:
            i1[axis] -= 1
        n = i1[axis] - i0[axis] + 1
        ijk = np.broadcast_to(i0, (n, 3)).copy()
        ijk[:, axis] = np.arange(i0[axis], i1[axis] + 1)
        widths = self.fcell_w[field]
        direction = -1 if pts[0][axis] > pts[1][axis] else 1
        for k, idx in enumerate(ijk):
            cw = widths[axis][idx[axis]]
            self.probes[f"{name}_{k}"] = dict(field=field, index=tuple(int(v) for v in idx),
                                              d=direction * conv.m_in(cw))

    # ------------------------------------------------------------------ sources
    def gaussian_source(self, width: float, t0: float, t_len: float):
        """Gaussian pulse exp(-((t - t0)/(width/4))^2) normalised to peak 1, t_len seconds long."""
        n = int(t_len / self.dt)
        t = np.linspace(0, self.dt * n, n)
        v = np.exp(-((t - t0) / (width / 4)) ** 2)
        return LabeledArray(v / np.max(np.abs(v)), coords=dict(time=t))

    def gaussian_modulated_source(self, f0: float, width: float, t0: float, t_len: float):
        """Gaussian-modulated sine at f0, normalised to peak 1."""
        n = int(t_len / self.dt)
        t = np.linspace(0, self.dt * n, n)
        v = np.exp(-((t - t0) / (width / 4)) ** 2) * np.sin(2 * np.pi * f0 * t)
        return LabeledArray(v / np.max(np.abs(v)), coords=dict(time=t))

    def assign_excitation(self, waveform: np.ndarray, ports):
        """Drive ports with waveform (one sample per time step; padded or cut to the run length)."""
        self.invalidate_solution()
        self.check_mesh()
        waveform = np.asarray(waveform)
        if self._time is None:
            self._time = np.arange(0, self.dt * len(waveform), self.dt)
        if len(self._time) > len(waveform):
            pad = np.zeros_like(self._time)
            pad[:len(waveform)] = waveform
            waveform = pad
        if len(self._time) < len(waveform):
            waveform = waveform[:len(self._time)]
        for p in np.atleast_1d(ports):
            self.ports[p - 1]["src"] = waveform.astype(np.float32)
            if self.ports_inv[p - 1] is not None:
                self.ports_inv[p - 1]["src"] = waveform.astype(np.float32)

    def reset_excitations(self):
        self._time = None
        for p in self.ports:
            p["src"] = None
        for p in self.ports_inv:
            if p is not None:
                p["src"] = None

    # ------------------------------------------------------------------ solve
    def packing_inverses(self, dtype=np.float32):
        """The inverse dual spacings the packing folds into Cb, one per cell along each axis (the
        last one 0: that component's update coefficient vanishes), in dtype: as coefficient_arrays
        forms them."""
        dtype_ = np.dtype(dtype)
        out = []
        for d in self.dh_cells:
            inv = 1 / conv.m_in(d).astype(dtype_)
            out.append(np.concatenate([inv, np.zeros(1, dtype=dtype_)]).astype(dtype_))
        return out

    def coefficient_arrays(self, gpu: bool = True, dtype=np.float32):
        """The 30 solver coefficient arrays with cell spacings and curl signs folded in
        (solver.py:1439-1523): all (Nx, Ny, Nz) for gpu=True, the CPU solver's shapes otherwise."""
        self.check_mesh()
        if getattr(self, "_coef_mode", "host") != "host":
            raise ValueError('the model was meshed with coefficients="device": its arrays are assembled on the '
                             "GPU (Session.assembled)")
        dtype_ = np.dtype(dtype)
        dx, dy, dz = [conv.m_in(d).astype(dtype_) for d in self.d_cells]
        dx_inv, dy_inv, dz_inv = 1 / dx[:, None, None], 1 / dy[None, :, None], 1 / dz[None, None, :]
        # one more (zero) spacing for the last component: its update coefficient vanishes (PEC edge)
        ix, iy, iz = self.packing_inverses(dtype_)
        dx1, dy1, dz1 = ix[:, None, None], iy[None, :, None], iz[None, None, :]
        end = None if gpu else -1
        hs = 1 if gpu else 0
        Ca, Cb, Da, Db = self.Ca, self.Cb, self.Da, self.Db
        c = dict(
            Ca_ex_y=Ca["ex_y"][:, 1:end, 1:end], Ca_ex_z=Ca["ex_z"][:, 1:end, 1:end],
            Cb_ex_y=Cb["ex_y"][:, 1:end, 1:end] * dy1[:, :end], Cb_ex_z=-Cb["ex_z"][:, 1:end, 1:end] * dz1[..., :end],
            Ca_ey_z=Ca["ey_z"][1:, :, 1:end], Ca_ey_x=Ca["ey_x"][1:, :, 1:end],
            Cb_ey_z=Cb["ey_z"][1:, :, 1:end] * dz1[..., :end], Cb_ey_x=-Cb["ey_x"][1:, :, 1:end] * dx1,
            Ca_ez_x=Ca["ez_x"][1:, 1:end, :], Ca_ez_y=Ca["ez_y"][1:, 1:end, :],
            Cb_ez_x=Cb["ez_x"][1:, 1:end, :] * dx1, Cb_ez_y=-Cb["ez_y"][1:, 1:end, :] * dy1[:, :end],
            Da_hx_y=Da["hx_y"][1:], Da_hx_z=Da["hx_z"][1:],
            Db_hx_y1=-Db["hx_y1"][1:] * dy_inv, Db_hx_y2=-Db["hx_y2"][1:] * dy_inv,
            Db_hx_z1=Db["hx_z1"][1:] * dz_inv, Db_hx_z2=Db["hx_z2"][1:] * dz_inv,
            Da_hy_z=Da["hy_z"][:, hs:], Da_hy_x=Da["hy_x"][:, hs:],
            Db_hy_z1=-Db["hy_z1"][:, hs:] * dz_inv, Db_hy_z2=-Db["hy_z2"][:, hs:] * dz_inv,
            Db_hy_x1=Db["hy_x1"][:, hs:] * dx_inv, Db_hy_x2=Db["hy_x2"][:, hs:] * dx_inv,
            Da_hz_x=Da["hz_x"][:, :, hs:], Da_hz_y=Da["hz_y"][:, :, hs:],
            Db_hz_x1=-Db["hz_x1"][:, :, hs:] * dx_inv, Db_hz_x2=-Db["hz_x2"][:, :, hs:] * dx_inv,
            Db_hz_y1=Db["hz_y1"][:, :, hs:] * dy_inv, Db_hz_y2=Db["hz_y2"][:, :, hs:] * dy_inv,
        )
        return {k: np.array(v, order="C", dtype=dtype_) for k, v in c.items()}

    def pack(self, gpu: bool = True, dtype=np.float32, n_threads: int = 4, coefficients: bool = True):
        """Solver inputs in the reference's calling convention (solver.py:1439-1622).

        Returns (coefficients, probes, monitors, mem, Nx, Ny, Nz, Nt).  gpu=True gives the GPU packing
        (all arrays (Nx, Ny, Nz)); gpu=False the CPU packing of the reference's threaded solver.
        coefficients=False skips building the coefficient arrays (returned as None)."""
        self.check_mesh()
        srcs = [p["src"] for p in self.ports if p is not None and p["src"] is not None]
        if not srcs:
            raise ValueError("No port excitations found.")
        Nt = len(srcs[0])
        if not all(Nt == len(s) for s in srcs):
            raise ValueError("Excitations must have identical lengths.")
        dtype_ = np.dtype(dtype)
        coefficients = self.coefficient_arrays(gpu=gpu, dtype=dtype_) if coefficients else None
        Nx, Ny, Nz = self.n_cells
        mem_len = sum(3 * int(np.prod((Nx,) + s[1:])) for s in self.fshape.values())
        mem_len += (4 * (Ny + 1) * (Nz + 1)) * n_threads * 8
        mem = np.zeros(mem_len, dtype=dtype_) if not gpu else np.zeros(1, dtype=dtype_)

        probes = []
        for port in self.ports + self.ports_inv:
            if port is None:
                continue
            src = port["src"] if port["src"] is not None else np.zeros(Nt, dtype=dtype_)
            idx_list = [list(np.arange(v.start, v.stop)) if isinstance(v, slice) else [v] for v in port["idx"]]
            vs = port["Vs_a"].flatten()
            for j, ijk in enumerate(itertools.product(*idx_list)):
                probes.append(dict(values=np.array(vs[j] * src, dtype=dtype_, order="C"),
                                   field=_FIELDS.index(port["field"]), idx=[int(v) for v in ijk],
                                   is_source=int(port["src"] is not None)))
        for p in self.probes.values():
            probes.append(dict(values=np.zeros(Nt, dtype=dtype_, order="C"), field=_FIELDS.index(p["field"]),
                               idx=[int(v) for v in p["index"]], is_source=0))

        cdt = np.complex64 if dtype_ == np.float32 else np.complex128
        monitors = []
        for m in self.monitors.values():
            n_m = int(Nt / m["n_step"]) + 1
            fi = _FIELDS.index(m["field"])
            f_ny = [self.Ny + 1, self.Ny, self.Ny + 1, self.Ny, self.Ny + 1, self.Ny][fi]
            f_nz = [self.Nz + 1, self.Nz + 1, self.Nz, self.Nz, self.Nz, self.Nz + 1][fi]
            if m["axis"] == 0:
                shape = (Ny, Nz) if gpu else (f_ny, f_nz)
            elif m["axis"] == 1:
                shape = (Nx, Nz) if gpu else (self.Nx, f_nz)
            else:
                shape = (Nx, Ny) if gpu else (self.Nx, f_ny)
            cfg = dict(axis=int(m["axis"]), position=int(m["index"]), field=fi, n_step=int(m["n_step"]))
            if m["frequency"] is not None:
                f = m["frequency"]
                omega = 2 * np.pi * (f / (1 / (self.dt * m["n_step"])))
                mcdt = cdt if m.get("dft_dtype") is None else np.promote_types(cdt, m["dft_dtype"])
                cfg["values"] = np.zeros((len(f),) + shape, dtype=mcdt, order="C")
                # DTFT phase per sample and frequency, evaluated in the monitor's complex dtype
                cfg["dtft_phase"] = np.exp(-1j * omega[None] * np.arange(n_m)[:, None], dtype=mcdt, order="C")
                cfg["n_frequencies"] = int(len(f))
            else:
                cfg["values"] = np.zeros((n_m,) + shape, dtype=dtype_, order="C")
            monitors.append(cfg)
        return coefficients, probes, monitors, mem, Nx, Ny, Nz, Nt

    def unpack(self, probes, monitors, gpu: bool = True):
        """Move solver outputs into the model (solver.py:1646-1698)."""
        Nt = len(probes[0]["values"]) if probes else 0
        for i, (k, m) in enumerate(self.monitors.items()):
            v = monitors[i]["values"]
            # rows for the components on the x = 0 edge, which the solver grids do not hold
            if m["axis"] in (1, 2) and m["field"] in ("hx", "ey", "ez"):
                v = np.pad(v, ((0, 0), (1, 0), (0, 0)))
            if gpu:
                if m["field"] in ("hy", "ex", "ez"):
                    if m["axis"] == 0:
                        v = np.pad(v, ((0, 0), (1, 0), (0, 0)))
                    if m["axis"] == 2:
                        v = np.pad(v, ((0, 0), (0, 0), (1, 0)))
                if m["field"] in ("hz", "ex", "ey"):
                    if m["axis"] == 0:
                        v = np.pad(v, ((0, 0), (0, 0), (1, 0)))
                    if m["axis"] == 1:
                        v = np.pad(v, ((0, 0), (0, 0), (1, 0)))
            self.monitors[k]["values"] = v
        series = [p["values"][:Nt] for p in probes]
        i0 = 0
        for port in self.ports + self.ports_inv:
            if port is None:
                continue
            n = port["Vs_a"].size
            port["values"] = np.array(series[i0: i0 + n]).reshape(port["Vs_a"].shape + (Nt,))
            i0 += n
        for i, k in enumerate(self.probes.keys()):
            self.probes[k]["values"] = probes[i + i0]["values"]
        self._solved = True

    def solve(self, n_threads: int = 4, show_progress: bool = True, gpu: bool = True, dtype=np.float32,
              backend=None, flags: int = 0):
        """Run the FDTD time loop on the GPU (HIP engine) and store probe and monitor results.

        backend: optional callable with the signature of solver_run_hip, used instead of it (the
        reference dispatches the same way between its CPU and CUDA entry points, solver.py:1636-1641).
        """
        if backend is None:
            if not gpu:
                raise NotImplementedError("rfnetwork_amd runs on the GPU only: call solve(gpu=True)")
            if getattr(self, "_coef_mode", "host") == "device":
                return self._solve_assembled(dtype, flags, show_progress)
            backend = hip_func.solver_run_hip
        coefficients, probes, monitors, mem, Nx, Ny, Nz, Nt = self.pack(gpu=gpu, dtype=dtype, n_threads=n_threads)
        n_cells = self.Nx * self.Ny * self.Nz
        if show_progress:
            sys.stdout.flush()
            print(f"Running solver with {n_cells / 1e3:.1f}k cells, and {Nt} time steps...")
            pct = np.clip(n_cells * len(self._time) / 100e6, 20, 100)
            update_interval = int(Nt / pct)
            t0 = time.time()
        else:
            update_interval = 0
        kwargs = dict(dtype=dtype)
        if flags:
            kwargs["flags"] = flags
        backend(coefficients, probes, monitors, mem, Nx, Ny, Nz, Nt, n_threads, update_interval, **kwargs)
        if show_progress:
            sys.stdout.write(f"\rDone in {time.time() - t0:.3f}s" + " " * 20 + "\n")
        self.unpack(probes, monitors, gpu=gpu)

    def _solve_assembled(self, dtype, flags, show_progress):
        """solve() of a model meshed with coefficients="device": the session assembles the
        coefficient arrays on the GPU; probes and monitors as in solve()."""
        from .session import Session

        _, probes, monitors, _, Nx, Ny, Nz, Nt = self.pack(gpu=True, dtype=dtype, coefficients=False)
        t0 = time.time()
        if show_progress:
            print(f"Running solver with {Nx * Ny * Nz / 1e3:.1f}k cells, and {Nt} time steps...")
        with Session.assembled(self, dtype, flags) as s:
            s.set_sources(probes, monitors, Nt)
            s.run(Nt)
            s.synchronize()
            s.read()
        if show_progress:
            sys.stdout.write(f"\rDone in {time.time() - t0:.3f}s" + " " * 20 + "\n")
        self.unpack(probes, monitors, gpu=True)

    # ------------------------------------------------------------------ results
    def line_probe_values(self, name: str) -> np.ndarray:
        self.check_solution()
        return np.array([p["values"] for k, p in self.probes.items() if k[:len(name)] == name])

    def vi_probe_values(self, name: str) -> np.ndarray:
        """Voltage (line probe) or current (loop probe): sum of the probe series weighted by d."""
        self.check_solution()
        return np.sum([p["values"] * p["d"] for k, p in self.probes.items() if k[:len(name)] == name], axis=0)

    def get_sparameters(self, frequency, source_port: int = 1, downsample: bool = False):
        """Column ``source_port`` of the S-matrix from port voltages and loop currents (Pozar 4.58)."""
        self.check_solution()
        frequency = np.atleast_1d(frequency)
        nports, nf = len(self.ports), len(frequency)
        Vp = np.zeros((nports, nf), dtype=np.complex128)
        Ip = np.zeros((nports, nf), dtype=np.complex128)
        for i, port in enumerate(self.ports):
            fidx, field, direction = port["idx"], port["field"], port["direction"]
            f_axis = _AXES.index(field[1])
            cw = conv.m_in(self.d_cells[f_axis][fidx[f_axis]])
            b = [None] * 4
            b[f_axis] = slice(None)
            vp = (-direction) * np.sum(port["values"] * cw[tuple(b)], axis=f_axis).squeeze()
            if len(vp.shape) > 1:
                vp = np.mean(vp, axis=0)
            Vp[i] = utils.dtft(vp, frequency, 1 / self.dt, downsample)
            ip = self.vi_probe_values(f"port_{i + 1}")
            Ip[i] = direction * utils.dtft(ip, frequency, 1 / self.dt, downsample)
        r0 = np.array([p["r0"] for p in self.ports])[..., None]
        Ap = (Vp + r0 * Ip) / 2
        Bp = (Vp - r0 * Ip) / 2
        S = Bp / Ap[source_port - 1][None]
        return LabeledArray(S.T[..., None], coords=dict(frequency=frequency, b=np.arange(1, nports + 1),
                                                       a=[source_port]))

    def write_touchstone(self, filepath, frequency, fmt: str = "ri", freq_unit: str = "hz", comments=()):
        """Solve every port of the model (one batched GPU session) and write the full S-matrix to an
        .sNp file (the circuit side reads it with rfnetwork/touchstone.py:7).  Returns the matrix."""
        from . import batch, touchstone

        S = batch.sparameter_matrix(self, frequency)
        r0 = {p["r0"] for p in self.ports if p is not None}
        if len(r0) != 1:
            raise ValueError("Touchstone files hold one reference impedance; the ports differ")
        touchstone.write_snp(filepath, S, frequency, fmt=fmt, ref_z=float(r0.pop()), freq_unit=freq_unit,
                             comments=comments)
        return S

    def get_monitor_data(self, name: str):
        """Monitor values with time (or frequency) and the two in-plane coordinates (inches).

        A name whose three monitors ``name_x``, ``name_y``, ``name_z`` exist (an "e_total" monitor)
        returns the total field instead (rfnetwork/solver.py:2720-2722)."""
        self.check_solution()
        if all(f"{name}_{a}" in self.monitors for a in _AXES):
            return self.get_total_monitor_data(name)
        m = self.monitors[name]
        axes = [a for a in range(3) if a != m["axis"]]
        coords = {_AXES[a]: self.floc[m["field"]][a] for a in axes}
        if m["frequency"] is not None:
            return LabeledArray(m["values"], coords=dict(frequency=m["frequency"], **coords))
        t = np.arange(len(m["values"])) * (self.dt * m["n_step"])
        return LabeledArray(m["values"], coords=dict(time=t, **coords))

    def get_total_monitor_data(self, name: str):
        """The three E components of an "e_total" monitor on the corners of the plane's cells,
        stacked x, y, z: shape (3, samples, N1 - 1, N2 - 1) (rfnetwork/solver.py:2750-2802).

        With (a1, a2) the in-plane axes in increasing order and a3 the normal, the a1 component is
        averaged over its two neighbours along a1 (its edge rows along a2 dropped), the a2 component
        over its two neighbours along a2 (edge rows along a1 dropped), and the normal component keeps
        its interior rows, so all three sit on the interior corners.  The averages are formed in the
        monitor's dtype, (u[1:] + u[:-1]) / 2.  The coordinates are those the x component carries
        after its first slice, as the reference labels them.  Phasor monitors are refused, as in the
        reference."""
        names = [f"{name}_{a}" for a in _AXES]
        if not all(n in self.monitors for n in names):
            raise ValueError(f"Monitor {name} does not contain data for all three field components.")
        m0 = self.monitors[names[0]]
        if m0["frequency"] is not None:
            raise NotImplementedError("Phasor monitors not implemented yet for total fields.")
        axis = m0["axis"]
        surf = [a for a in range(3) if a != axis]
        order = surf + [axis]  # in-plane components first, then the normal one
        parts = [self.get_monitor_data(names[a]) for a in order]
        vals = [np.asarray(p) for p in parts]
        v1 = (vals[0][:, 1:, 1:-1] + vals[0][:, :-1, 1:-1]) / 2
        v2 = (vals[1][:, 1:-1, 1:] + vals[1][:, 1:-1, :-1]) / 2
        v3 = vals[2][:, 1:-1, 1:-1]
        # the slice each component's labels follow (the first operand of its average)
        cuts = [(slice(1, None), slice(1, -1)), (slice(1, -1), slice(1, None)), (slice(1, -1), slice(1, -1))]
        by_axis = {a: (v, parts[k], cuts[k]) for k, (a, v) in enumerate(zip(order, (v1, v2, v3)))}
        vx, px, cx = by_axis[0]
        d1, d2 = _AXES[surf[0]], _AXES[surf[1]]
        coords = dict(component=("x", "y", "z"), time=px.coords["time"],
                      **{d1: np.asarray(px.coords[d1])[cx[0]], d2: np.asarray(px.coords[d2])[cx[1]]})
        return LabeledArray(np.array([by_axis[a][0] for a in range(3)]), coords=coords)

    # ------------------------------------------------------------------ edge correction
    def edge_correction(self, p1, p2, integration_line: str, CFe: float = None, CFh: float = 1):
        """Scale the H update coefficients next to a PEC edge (singular field correction)."""
        self.invalidate_solution()
        self.check_mesh()
        if CFe is None:
            CFe = 2 * np.sqrt(1 / 2)
        p1, p2 = np.array(p1, float), np.array(p2, float)
        span = np.abs(p2 - p1)
        if np.count_nonzero(span) != 1:
            raise NotImplementedError("Edge correction is only supported for edges along the cartesian axis.")
        e_axis = int(np.argmax(span))
        f_axis = _AXES.index(integration_line[0])
        f_dir = {"+": 1, "-": 0}[integration_line[1]]
        if e_axis == f_axis:
            raise ValueError("Integration axis must be perpendicular to the edge.")
        if p1[e_axis] > p2[e_axis]:
            p1, p2 = p2, p1
        i1 = self.pos_to_idx(p1, mode="edge")
        i2 = self.pos_to_idx(p2, mode="edge")
        n_axis = [a for a in (0, 1, 2) if a != e_axis and a != f_axis][0]
        ea, fa, na = _AXES[e_axis], _AXES[f_axis], _AXES[n_axis]

        def at(edge, along, normal):
            idx = [slice(None)] * 3
            idx[e_axis], idx[f_axis], idx[n_axis] = edge, along, normal
            return tuple(idx)

        centers = slice(i1[e_axis], i2[e_axis])
        edges = slice(i1[e_axis], i2[e_axis] + 1)
        fh = i1[f_axis] if f_dir else i1[f_axis] - 1
        ni = i1[n_axis]
        host = getattr(self, "_coef_mode", "host") == "host"
        fs = self.fshape
        ops = []  # (Db name, index, factor), applied in this order
        idx = at(centers, fh, ni)
        ops.append((f"h{na}_{fa}2", idx, 1 / CFe))
        ops.append((f"h{na}_{fa}1", idx, 1 / CFe))
        if i1[e_axis] > 0:
            ops.append((f"h{na}_{ea}2", at(i1[e_axis] - 1, fh, ni), CFe))
        if i2[e_axis] < fs[f"h{na}"][e_axis]:
            ops.append((f"h{na}_{ea}1", at(i2[e_axis], fh, ni), CFe))
        ops.append((f"h{ea}_{na}2", at(edges, fh, ni - 1), CFe))
        ops.append((f"h{ea}_{na}1", at(edges, fh, ni), CFe))
        for k in (ni - 1, ni):
            ops.append((f"h{ea}_{fa}2", at(edges, i1[f_axis] - 1, k), CFe))
            ops.append((f"h{ea}_{fa}1", at(edges, i1[f_axis], k), CFe))
            ops.append((f"h{fa}_{na}1", at(centers, i1[f_axis], k), 1 / CFe))
            ops.append((f"h{fa}_{na}2", at(centers, i1[f_axis], k), 1 / CFe))
            if i1[e_axis] > 0:
                ops.append((f"h{fa}_{ea}2", at(i1[e_axis] - 1, i1[f_axis], k), CFe))
            if i2[e_axis] < fs[f"h{fa}"][e_axis]:
                ops.append((f"h{fa}_{ea}1", at(i2[e_axis], i1[f_axis], k), CFe))
        if host:
            for name, ix, factor in ops:
                self.Db[name][ix] *= factor
        self._edge_ops.extend(ops)
