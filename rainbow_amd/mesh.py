"""Triangle meshes: generators, vertex/triangle ordering, OBJ input.

Generators keep the vertex and triangle tables of
rainbow/geometry/surface_mesh.py (create_box :286-329, profile_sweep :85-171,
create_sphere :211-218) because the triangle order fixes the contact order.
``axis_sort`` (surface_mesh.py:56-82) and ``direction_of_most_variance``
(rainbow/math/functions.py:27-48; the reference returns a *row* of numpy's
eigenvector matrix) decide the final ordering, as in api.create_mesh
(api.py:159-173).  Icospheres, torus knots and the OBJ reader feed the
non-convex narrow-phase stress scene.
"""
import numpy as np

from . import math3 as Q


class Mesh:
    def __init__(self, V, T):
        self.V = np.array(V, dtype=np.float64, copy=True)
        self.T = np.array(T, copy=True)


def translate(mesh, r):
    mesh.V[:, :] = mesh.V[:, :] + r
    return mesh


def rotate(mesh, q):
    mesh.V = Q.rotate_array(q, mesh.V)
    return mesh


def scale(mesh, a, b, c):
    mesh.V = mesh.V * np.array([a, b, c], dtype=np.float64)
    return mesh


def aabb(mesh):
    return np.amin(mesh.V, axis=0), np.amax(mesh.V, axis=0)


def scale_to_unit(mesh):
    lo, hi = aabb(mesh)
    center = (lo + hi) / 2.0
    translate(mesh, -center)
    s = 1.0 / np.max(hi - lo)
    scale(mesh, s, s, s)
    translate(mesh, center)


def direction_of_most_variance(P):
    """functions.py:27-48: row ``argmax(eigenvalues)`` of numpy's eigenvector matrix."""
    P = np.asarray(P, dtype=np.float64)
    mean = np.mean(P.T, axis=1)
    Cm = P - mean
    cov = np.cov(Cm.T)
    values, vectors = np.linalg.eig(cov)
    return vectors[np.argmax(values)]


def axis_sort(mesh, axis):
    """surface_mesh.py:56-82: stable sort of vertices by projection on ``axis``,
    remap triangles, stable sort of triangles by the mean of their indices."""
    n = Q.unit3(axis)
    d = np.array([np.dot(mesh.V[i, :], n) for i in range(len(mesh.V))])
    new2old = np.argsort(d, kind="stable")
    mesh.V = mesh.V[new2old]
    old2new = np.empty_like(new2old)
    old2new[new2old] = np.arange(len(new2old))
    mesh.T = old2new[mesh.T]
    tmean = np.mean(mesh.T, axis=1)
    mesh.T = mesh.T[np.argsort(tmean, kind="stable")]


def create_box(width, height, depth):
    """surface_mesh.py:286-329 (vertex and triangle tables)."""
    V = np.zeros((8, 3), dtype=np.float64)
    V[0] = np.array([-width, -height, -depth]) * 0.5
    V[1] = np.array([width, -height, -depth]) * 0.5
    V[3] = np.array([width, height, -depth]) * 0.5
    V[2] = np.array([-width, height, -depth]) * 0.5
    V[4] = np.array([-width, -height, depth]) * 0.5
    V[5] = np.array([width, -height, depth]) * 0.5
    V[7] = np.array([width, height, depth]) * 0.5
    V[6] = np.array([-width, height, depth]) * 0.5
    T = np.array([(0, 2, 3), (0, 3, 1), (4, 5, 7), (4, 7, 6), (1, 3, 7), (1, 7, 5),
                  (0, 4, 6), (0, 6, 2), (0, 1, 5), (0, 5, 4), (3, 2, 6), (3, 6, 7)], dtype=np.int64)
    return V, T


def profile_sweep(profile, slices):
    """surface_mesh.py:85-171: sweep a profile around the y axis."""
    N = len(profile)
    J = slices
    if N <= 2:
        raise RuntimeError("profile_sweep(): Profile must have at least 3 points")
    if J <= 2:
        raise RuntimeError("profile_sweep(): Sweep must have at least 3 slices to be a proper volume.")
    K = (N - 2) * J + 2
    bottom, top = K - 1, K - 2
    H = N - 2
    F = 2 * J * (N - 2)
    V = np.zeros((K, 3), dtype=np.float64)
    T = np.zeros((F, 3), dtype=np.int32)
    dtheta = 2.0 * np.pi / J
    v = 0
    for j in range(J):
        R = Q.Ru(j * dtheta, np.array([0.0, 1.0, 0.0]))
        for i in range(H):
            V[v, :] = Q.rotate(R, profile[i + 1])
            v += 1
    V[v, :] = profile[N - 1]
    v += 1
    V[v, :] = profile[0]
    f = 0
    for j in range(J):
        left, right = j, (j + 1) % J
        T[f, :] = (bottom, H * right, H * left)
        f += 1
    for i in range(H - 1):
        for j in range(J):
            left, right = j, (j + 1) % J
            vi, vj = left * H + i, right * H + i
            vk, vm = right * H + i + 1, left * H + i + 1
            T[f, :] = (vi, vj, vk)
            f += 1
            T[f, :] = (vi, vk, vm)
            f += 1
    for j in range(J):
        left, right = j, (j + 1) % J
        T[f, :] = (top, left * H + N - 3, right * H + N - 3)
        f += 1
    return V, T


def create_sphere(radius, slices=12, segments=12):
    """surface_mesh.py:211-218."""
    profile = []
    dtheta = np.pi / (segments - 1)
    for i in range(segments):
        R = Q.Ru(dtheta * i, np.array([0.0, 0.0, 1.0]))
        profile.append(Q.rotate(R, np.array([0.0, -radius, 0.0])))
    return profile_sweep(profile, slices)


def create_icosphere(radius, frequency):
    """Geodesic sphere: icosahedron faces subdivided ``frequency`` times per
    edge, vertices projected to the sphere (10 f^2 + 2 vertices)."""
    t = (1.0 + np.sqrt(5.0)) / 2.0
    base = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
                     [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    f = int(frequency)
    verts = {}
    V = []
    T = []

    def vid(key, p):
        if key not in verts:
            verts[key] = len(V)
            V.append(p)
        return verts[key]

    def key_of(a, b, c, i, j):
        # barycentric lattice point (i, j, f-i-j) on face (a, b, c), keyed canonically
        w = {a: f - i - j, b: i, c: j}
        return tuple(sorted((k, v) for k, v in w.items() if v > 0))

    for (a, b, c) in faces:
        A, B, Cc = base[a], base[b], base[c]
        idx = {}
        for i in range(f + 1):
            for j in range(f + 1 - i):
                p = (A * (f - i - j) + B * i + Cc * j) / f
                idx[(i, j)] = vid(key_of(a, b, c, i, j), p)
        for i in range(f):
            for j in range(f - i):
                T.append((idx[(i, j)], idx[(i + 1, j)], idx[(i, j + 1)]))
                if i + j < f - 1:
                    T.append((idx[(i + 1, j)], idx[(i + 1, j + 1)], idx[(i, j + 1)]))
    V = np.array(V)
    V = radius * V / np.linalg.norm(V, axis=1)[:, None]
    return V, np.array(T, dtype=np.int64)


def create_torus_knot(p=2, q=3, R=1.0, r=0.4, tube=0.15, segments=200, sides=25):
    """(p, q) torus knot swept by a circular tube (segments x sides vertices)."""
    s = np.arange(segments) * 2.0 * np.pi / segments

    def curve(u):
        rr = R + r * np.cos(q * u)
        return np.stack([rr * np.cos(p * u), r * np.sin(q * u), rr * np.sin(p * u)], axis=-1)
    c = curve(s)
    dc = curve(s + 1e-4) - curve(s - 1e-4)
    tan = dc / np.linalg.norm(dc, axis=1)[:, None]
    ref = np.array([0.0, 1.0, 0.0])
    nrm = np.cross(tan, ref)
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    bin_ = np.cross(tan, nrm)
    a = np.arange(sides) * 2.0 * np.pi / sides
    V = (c[:, None, :] + tube * (np.cos(a)[None, :, None] * nrm[:, None, :]
                                 + np.sin(a)[None, :, None] * bin_[:, None, :])).reshape(-1, 3)
    T = []
    for i in range(segments):
        i2 = (i + 1) % segments
        for j in range(sides):
            j2 = (j + 1) % sides
            a0, a1, b0, b1 = i * sides + j, i * sides + j2, i2 * sides + j, i2 * sides + j2
            T.append((a0, b1, b0))          # outward normals
            T.append((a0, a1, b1))
    return V, np.array(T, dtype=np.int64)


def create_noisy_icosphere(rng, frequency, amplitude=0.2):
    """Unit icosphere whose vertices are pushed radially by a factor
    1 + amplitude * U(-1, 1), one draw per vertex from ``rng``."""
    V, T = create_icosphere(1.0, frequency)
    V = V * (1.0 + amplitude * rng.uniform(-1.0, 1.0, size=len(V)))[:, None]
    return V, T


def create_funnel(wide=0.5, narrow=0.17, wall=0.05, height=1.0 / 3.0, rings=10, sides=32):
    """Closed funnel: a conical shell around the y axis, open at both ends.

    The outer surface narrows linearly from radius ``wide`` at y = -height/2
    to ``narrow`` at y = +height/2; the inner surface lies ``wall`` closer to
    the axis at every ring; annular rims close the shell at both ends.
    Faces are wound with outward normals (the inner surface faces the axis),
    so the solid is the shell.  The wide end is at -y, as in the mesh that
    scenes/create_funnel.py:84-110 turns upside down with Rx(pi).
    ``rings`` x ``sides`` vertices per surface.
    """
    k = np.arange(rings) / (rings - 1.0)
    y = -0.5 * height + height * k
    ro = wide + (narrow - wide) * k
    ri = ro - wall
    th = np.arange(sides) * 2.0 * np.pi / sides
    c, s_ = np.cos(th), np.sin(th)

    def ring(rad):
        return np.stack([rad[:, None] * c[None, :], np.broadcast_to(y[:, None], (rings, sides)),
                         rad[:, None] * s_[None, :]], axis=-1).reshape(-1, 3)
    V = np.concatenate([ring(ro), ring(ri)])
    o = lambda i, j: i * sides + j % sides
    n = lambda i, j: rings * sides + i * sides + j % sides
    T = []
    for i in range(rings - 1):
        for j in range(sides):
            T.append((o(i, j), o(i + 1, j), o(i, j + 1)))
            T.append((o(i + 1, j), o(i + 1, j + 1), o(i, j + 1)))
            T.append((n(i, j), n(i, j + 1), n(i + 1, j)))
            T.append((n(i + 1, j), n(i, j + 1), n(i + 1, j + 1)))
    b, t = 0, rings - 1
    for j in range(sides):
        T.append((o(b, j), o(b, j + 1), n(b, j)))          # wide rim, facing -y
        T.append((n(b, j), o(b, j + 1), n(b, j + 1)))
        T.append((o(t, j), n(t, j), o(t, j + 1)))          # narrow rim, facing +y
        T.append((n(t, j), n(t, j + 1), o(t, j + 1)))
    return V, np.array(T, dtype=np.int64)


def read_obj(path):
    """Vertices and triangles of a Wavefront OBJ file (polygons fanned)."""
    V, T = [], []
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "v":
                V.append([float(x) for x in parts[1:4]])
            elif parts[0] == "f":
                idx = [int(p.split("/")[0]) for p in parts[1:]]
                idx = [i - 1 if i > 0 else len(V) + i for i in idx]
                for k in range(1, len(idx) - 1):
                    T.append((idx[0], idx[k], idx[k + 1]))
    return np.array(V, dtype=np.float64), np.array(T, dtype=np.int64)
