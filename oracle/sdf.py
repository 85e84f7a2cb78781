"""Signed-distance grid construction (the role of libigl's signed_distance).

The reference builds every shape's SDF with ``igl.signed_distance(nodes, V, F)``
(grid3.py:320-339) over a res^3 node grid spanning the mesh bounding box grown
by ``boundary`` (api.py:232-241).  libigl (a third-party dependency, version not
pinned by the reference's CI: ``pip install libigl``,
.github/workflows/python-package-conda.yml:37) is not installed here, so its
published algorithm is restated: the magnitude is the exact Euclidean distance
to the closest point of the closest triangle; the sign is negative inside,
decided by the generalized winding number (sum of signed solid angles / 4 pi,
Van Oosterom & Strackee 1983), which equals libigl's winding-number sign for
closed meshes at points off the surface.

Closest point on a triangle follows Ericson, "Real-Time Collision Detection",
section 5.1.5 (Voronoi-region tests), vectorised over nodes x triangles.

Pinned by tests/test_oracle_sdf.py: the analytic signed distance of an
axis-aligned box (the box mesh is the box), and brute-force point-to-triangle
distances computed by dense sampling of barycentric coordinates.  Parity with
libigl's own output is unpinned (libigl is absent).
"""
import numpy as np

from .grid import Grid, nodes_array


def closest_point_sqdist(P, A, B, C):
    """Squared distance from points P (M,3) to triangles (A,B,C) each (T,3).

    Returns an (M, T) array.  Ericson 5.1.5, region tests in barycentric form.
    """
    P = P[:, None, :]
    A = A[None, :, :]
    B = B[None, :, :]
    C = C[None, :, :]
    ab = B - A
    ac = C - A
    ap = P - A
    d1 = np.sum(ab * ap, axis=-1)
    d2 = np.sum(ac * ap, axis=-1)
    bp = P - B
    d3 = np.sum(ab * bp, axis=-1)
    d4 = np.sum(ac * bp, axis=-1)
    cp = P - C
    d5 = np.sum(ab * cp, axis=-1)
    d6 = np.sum(ac * cp, axis=-1)
    va = d3 * d6 - d5 * d4
    vb = d5 * d2 - d1 * d6
    vc = d1 * d4 - d3 * d2

    shape = np.broadcast_shapes(d1.shape)
    # Barycentric weights (v, w) of the closest point: Q = A + v ab + w ac.
    v = np.zeros(shape)
    w = np.zeros(shape)
    done = np.zeros(shape, dtype=bool)

    def assign(mask, vv, ww):
        nonlocal done
        m = mask & ~done
        v[m] = np.broadcast_to(vv, shape)[m]
        w[m] = np.broadcast_to(ww, shape)[m]
        done |= m

    with np.errstate(divide="ignore", invalid="ignore"):
        assign((d1 <= 0) & (d2 <= 0), 0.0, 0.0)                       # vertex A
        assign((d3 >= 0) & (d4 <= d3), 1.0, 0.0)                      # vertex B
        assign((vc <= 0) & (d1 >= 0) & (d3 <= 0), d1 / (d1 - d3), 0.0)  # edge AB
        assign((d6 >= 0) & (d5 <= d6), 0.0, 1.0)                      # vertex C
        assign((vb <= 0) & (d2 >= 0) & (d6 <= 0), 0.0, d2 / (d2 - d6))  # edge AC
        e = (d4 - d3) + (d5 - d6)
        wbc = (d4 - d3) / e
        assign((va <= 0) & ((d4 - d3) >= 0) & ((d5 - d6) >= 0), 1.0 - wbc, wbc)  # edge BC
        denom = 1.0 / (va + vb + vc)
        assign(np.ones(shape, dtype=bool), vb * denom, vc * denom)   # face interior
    Q = A + v[..., None] * ab + w[..., None] * ac
    d = P - Q
    return np.sum(d * d, axis=-1)


def winding_number(P, A, B, C):
    """Generalized winding number of the closed triangle soup at points P (M,)."""
    a = A[None, :, :] - P[:, None, :]
    b = B[None, :, :] - P[:, None, :]
    c = C[None, :, :] - P[:, None, :]
    la = np.linalg.norm(a, axis=-1)
    lb = np.linalg.norm(b, axis=-1)
    lc = np.linalg.norm(c, axis=-1)
    det = np.sum(a * np.cross(b, c), axis=-1)
    den = la * lb * lc + np.sum(a * b, axis=-1) * lc + np.sum(b * c, axis=-1) * la \
        + np.sum(c * a, axis=-1) * lb
    omega = 2.0 * np.arctan2(det, den)
    return np.sum(omega, axis=1) / (4.0 * np.pi)


def signed_distance(P, V, T, chunk=2048):
    """Signed distance of points P (M,3) to the closed mesh (V, T)."""
    V = np.asarray(V, dtype=np.float64)
    T = np.asarray(T, dtype=np.int64)
    A, B, C = V[T[:, 0]], V[T[:, 1]], V[T[:, 2]]
    out = np.empty(len(P))
    tchunk = max(1, (1 << 21) // max(1, len(T)))
    step = min(chunk, tchunk)
    for s in range(0, len(P), step):
        p = P[s:s + step]
        d = np.sqrt(np.min(closest_point_sqdist(p, A, B, C), axis=1))
        wn = winding_number(p, A, B, C)
        out[s:s + step] = np.where(np.abs(wn) > 0.5, -d, d)
    return out


def create_signed_distance(V, T, I, J, K, boundary=0.5):
    """grid3.py:320-339 with the libigl call replaced by ``signed_distance``."""
    if boundary < 0:
        raise ValueError()
    V = np.asarray(V, dtype=np.float64)
    min_coord = (V.min(axis=0) - boundary).ravel()
    max_coord = (V.max(axis=0) + boundary).ravel()
    grid = Grid(min_coord, max_coord, I, J, K)
    grid.values = signed_distance(nodes_array(grid), V, T)
    return grid
