"""Regular-grid SDF lookups, restated from rainbow/geometry/grid3.py.

* ``Grid``            grid3.py:5-123 (spacing = (max-min)/(res-1), linear index
                      i + I*(j + J*k), enclosing cell clipped to [0, dim-2]).
* ``get_value``       grid3.py:126-177, trilinear interpolation.
* ``get_gradient``    grid3.py:180-254, analytic gradient of the trilinear
                      interpolant, with the reference's local coordinates
                      ``(p - cell_origin) / min_coord`` (grid3.py:209-211; the
                      value path divides by ``spacing``, grid3.py:155-157).
* ``is_inside``       grid3.py:275-300, strict comparisons against the box
                      shrunk by ``boundary`` (default 0.5).

Every function has a scalar form (a line-by-line restatement) and a vectorised
form over an (M, 3) point array; tests check the two against each other.

Pinned by tests/test_oracle_grid.py: the reference's own cases from
unit_tests/test_geometry_grid3.py (enclosing cells :56-91, values and gradients
of linear fields :108-191 and :255-508, is_inside with boundary 0 and 0.5
:193-239) and, for non-linear fields, by the closed-form trilinear polynomial and
its derivative (any tabulated field), including the min_coord-scaled local
coordinates of the gradient.
"""
import numpy as np


class Grid:
    """grid3.py:5-45."""

    def __init__(self, min_coord, max_coord, I, J, K, values=None):
        min_coord = np.asarray(min_coord, dtype=np.float64).ravel()
        max_coord = np.asarray(max_coord, dtype=np.float64).ravel()
        if np.any(min_coord > max_coord) or I < 2 or J < 2 or K < 2:
            raise ValueError()
        self.min_coord = min_coord
        self.max_coord = max_coord
        self.I, self.J, self.K = int(I), int(J), int(K)
        res = np.array([I - 1, J - 1, K - 1], dtype=np.float64)
        self.spacing = np.divide(max_coord - min_coord, res)
        self.values = None if values is None else np.asarray(values, dtype=np.float64).ravel()

    def get_linear_index(self, i, j, k):
        return i + self.I * (j + self.J * k)

    def get_node_value(self, i, j, k):
        return self.values[self.get_linear_index(i, j, k)]

    def get_node_coord(self, i, j, k):
        return self.min_coord + np.multiply(self.spacing, np.array([i, j, k], dtype=np.float64))

    def get_enclosing_cell_idx(self, p):
        idx = np.floor((p - self.min_coord) / self.spacing)
        i = int(np.clip(int(idx[0]), 0, self.I - 2))
        j = int(np.clip(int(idx[1]), 0, self.J - 2))
        k = int(np.clip(int(idx[2]), 0, self.K - 2))
        return i, j, k


def nodes_array(grid):
    """grid3.py:303-317, in linear-index order."""
    k, j, i = np.meshgrid(np.arange(grid.K), np.arange(grid.J), np.arange(grid.I), indexing="ij")
    idx = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    return grid.min_coord + grid.spacing * idx


def eval_on_grid(grid, func):
    """grid3.py:257-272."""
    nodes = nodes_array(grid)
    grid.values = np.array([func(p) for p in nodes], dtype=np.float64)


def _corners(grid, i, j, k):
    g = grid.get_node_value
    return (g(i, j, k), g(i, j, k + 1), g(i, j + 1, k), g(i, j + 1, k + 1),
            g(i + 1, j, k), g(i + 1, j, k + 1), g(i + 1, j + 1, k), g(i + 1, j + 1, k + 1))


def get_value(grid, p):
    """grid3.py:126-177."""
    i, j, k = grid.get_enclosing_cell_idx(p)
    d000, d001, d010, d011, d100, d101, d110, d111 = _corners(grid, i, j, k)
    h, m = grid.spacing, grid.min_coord
    s = (p[0] - ((i * h[0]) + m[0])) / h[0]
    t = (p[1] - ((j * h[1]) + m[1])) / h[1]
    u = (p[2] - ((k * h[2]) + m[2])) / h[2]
    x00 = (d100 - d000) * s + d000
    x01 = (d101 - d001) * s + d001
    x10 = (d110 - d010) * s + d010
    x11 = (d111 - d011) * s + d011
    y0 = (x10 - x00) * t + x00
    y1 = (x11 - x01) * t + x01
    return (y1 - y0) * u + y0


def get_gradient(grid, p):
    """grid3.py:180-254 (local coordinates divided by min_coord)."""
    i, j, k = grid.get_enclosing_cell_idx(p)
    d000, d001, d010, d011, d100, d101, d110, d111 = _corners(grid, i, j, k)
    h, m = grid.spacing, grid.min_coord
    s = (p[0] - (i * h[0] + m[0])) / m[0]
    t = (p[1] - (j * h[1] + m[1])) / m[1]
    u = (p[2] - (k * h[2] + m[2])) / m[2]
    x00 = (d100 - d000) * s + d000
    x01 = (d101 - d001) * s + d001
    x10 = (d110 - d010) * s + d010
    x11 = (d111 - d011) * s + d011
    y0 = (x10 - x00) * t + x00
    y1 = (x11 - x01) * t + x01
    dx00_ds = d100 - d000
    dx01_ds = d101 - d001
    dx10_ds = d110 - d010
    dx11_ds = d111 - d011
    dy0_ds = (dx10_ds - dx00_ds) * t + dx00_ds
    dy1_ds = (dx11_ds - dx01_ds) * t + dx01_ds
    dy0_dt = x10 - x00
    dy1_dt = x11 - x01
    dp_ds = (dy1_ds - dy0_ds) * u + dy0_ds
    dp_dt = (dy1_dt - dy0_dt) * u + dy0_dt
    dp_du = y1 - y0
    return np.array([dp_ds * (1.0 / h[0]), dp_dt * (1.0 / h[1]), dp_du * (1.0 / h[2])],
                    dtype=np.float64)


def is_inside(grid, p, boundary=0.5):
    """grid3.py:275-300."""
    if boundary < 0:
        raise ValueError()
    M, m = grid.max_coord, grid.min_coord
    if p[0] > (M[0] - boundary) or p[1] > (M[1] - boundary) or p[2] > (M[2] - boundary):
        return False
    if p[0] < (m[0] + boundary) or p[1] < (m[1] + boundary) or p[2] < (m[2] + boundary):
        return False
    return True


# ---------------------------------------------------------------- vectorised

def cell_idx_v(grid, P):
    idx = np.floor((P - grid.min_coord) / grid.spacing)
    hi = np.array([grid.I - 2, grid.J - 2, grid.K - 2])
    return np.clip(idx.astype(np.int64), 0, hi)


def _corners_v(grid, ijk):
    i, j, k = ijk[:, 0], ijk[:, 1], ijk[:, 2]
    I, J = grid.I, grid.J
    v = grid.values

    def g(a, b, c):
        return v[a + I * (b + J * c)]
    return (g(i, j, k), g(i, j, k + 1), g(i, j + 1, k), g(i, j + 1, k + 1),
            g(i + 1, j, k), g(i + 1, j, k + 1), g(i + 1, j + 1, k), g(i + 1, j + 1, k + 1))


def get_value_v(grid, P):
    P = np.atleast_2d(P)
    ijk = cell_idx_v(grid, P)
    d000, d001, d010, d011, d100, d101, d110, d111 = _corners_v(grid, ijk)
    h, m = grid.spacing, grid.min_coord
    s = (P[:, 0] - ((ijk[:, 0] * h[0]) + m[0])) / h[0]
    t = (P[:, 1] - ((ijk[:, 1] * h[1]) + m[1])) / h[1]
    u = (P[:, 2] - ((ijk[:, 2] * h[2]) + m[2])) / h[2]
    x00 = (d100 - d000) * s + d000
    x01 = (d101 - d001) * s + d001
    x10 = (d110 - d010) * s + d010
    x11 = (d111 - d011) * s + d011
    y0 = (x10 - x00) * t + x00
    y1 = (x11 - x01) * t + x01
    return (y1 - y0) * u + y0


def get_gradient_v(grid, P):
    P = np.atleast_2d(P)
    ijk = cell_idx_v(grid, P)
    d000, d001, d010, d011, d100, d101, d110, d111 = _corners_v(grid, ijk)
    h, m = grid.spacing, grid.min_coord
    s = (P[:, 0] - (ijk[:, 0] * h[0] + m[0])) / m[0]
    t = (P[:, 1] - (ijk[:, 1] * h[1] + m[1])) / m[1]
    u = (P[:, 2] - (ijk[:, 2] * h[2] + m[2])) / m[2]
    x00 = (d100 - d000) * s + d000
    x01 = (d101 - d001) * s + d001
    x10 = (d110 - d010) * s + d010
    x11 = (d111 - d011) * s + d011
    y0 = (x10 - x00) * t + x00
    y1 = (x11 - x01) * t + x01
    dx00_ds = d100 - d000
    dx01_ds = d101 - d001
    dx10_ds = d110 - d010
    dx11_ds = d111 - d011
    dy0_ds = (dx10_ds - dx00_ds) * t + dx00_ds
    dy1_ds = (dx11_ds - dx01_ds) * t + dx01_ds
    dy0_dt = x10 - x00
    dy1_dt = x11 - x01
    dp_ds = (dy1_ds - dy0_ds) * u + dy0_ds
    dp_dt = (dy1_dt - dy0_dt) * u + dy0_dt
    dp_du = y1 - y0
    return np.stack([dp_ds * (1.0 / h[0]), dp_dt * (1.0 / h[1]), dp_du * (1.0 / h[2])], axis=1)


def is_inside_v(grid, P, boundary=0.5):
    P = np.atleast_2d(P)
    hi = grid.max_coord - boundary
    lo = grid.min_coord + boundary
    return np.all(P <= hi, axis=1) & np.all(P >= lo, axis=1)
