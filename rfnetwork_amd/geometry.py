"""Axis-aligned geometry objects for the solver API.

rfnetwork describes geometry with pyvista PolyData; the solver only reads ``.points``,
``.n_cells`` and ``.faces`` of them (rfnetwork/solver.py:794-868, :1086-1236, :1893-2079).  These
small classes expose the same attributes for boxes, rectangles and lines, so models can be set
up without pyvista.  pyvista objects with the same attributes are accepted unchanged.
"""
import numpy as np


class _Shape:
    points: np.ndarray
    n_cells: int
    faces: np.ndarray

    @property
    def bounds(self):
        p0, p1 = self.points.min(axis=0), self.points.max(axis=0)
        return (p0[0], p1[0], p0[1], p1[1], p0[2], p1[2])


class Box(_Shape):
    """Rectangular box between two corners (six faces, like ``pyvista.Box``)."""

    def __init__(self, p0, p1):
        p0, p1 = np.asarray(p0, float), np.asarray(p1, float)
        lo, hi = np.minimum(p0, p1), np.maximum(p0, p1)
        self.points = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
        self.n_cells = 6
        self.faces = np.array([4, 0, 1, 3, 2] * 6)


def Cube(center=(0.0, 0.0, 0.0), x_length=1.0, y_length=1.0, z_length=1.0):
    """Box from a centre and edge lengths (``pyvista.Cube`` signature)."""
    c = np.asarray(center, float)
    h = np.array([x_length, y_length, z_length], float) / 2
    return Box(c - h, c + h)


class Rectangle(_Shape):
    """Planar rectangle from three corners; the fourth is p0 + p2 - p1 (``pyvista.Rectangle``)."""

    def __init__(self, points):
        p = np.asarray(points, float)
        if p.shape != (3, 3):
            raise ValueError("Rectangle takes three corner points")
        self.points = np.array([p[0], p[1], p[2], p[0] + p[2] - p[1]])
        self.n_cells = 1
        self.faces = np.array([4, 0, 1, 2, 3])


class Line(_Shape):
    """Straight line between two points (``pyvista.Line``)."""

    def __init__(self, a, b):
        self.points = np.array([np.asarray(a, float), np.asarray(b, float)])
        self.n_cells = 1
        self.faces = np.array([2, 0, 1])


def is_surface(obj) -> bool:
    """True when the object's extent is zero along exactly one axis."""
    ext = np.ptp(np.asarray(obj.points, float), axis=0)
    return np.count_nonzero(ext > 1e-12) == 2


def inside_axis_aligned(points: np.ndarray, obj, tol: float) -> np.ndarray:
    """Points (..., 3) lying inside or on the axis-aligned bounding box of obj, within tol.

    For the boxes and rectangles accepted here this is the reference's inclusion test: implicit
    distance <= tol for solids (rfnetwork/solver.py:856-859) and the polygon test with edge
    tolerance for planar faces (rfnetwork/utils_mesh.py:383), both of which include the boundary.
    """
    p = np.asarray(obj.points, float)
    lo, hi = p.min(axis=0) - tol, p.max(axis=0) + tol
    return np.all((points >= lo) & (points <= hi), axis=-1)
