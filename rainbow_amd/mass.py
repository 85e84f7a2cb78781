"""Mass properties of closed triangle meshes and the body frame.

The reference integrates with Mirtich's projected face integrals
(rainbow/simulators/prox_rigid_bodies/mass.py:138-283).  Here the same volume
integrals are taken over the signed tetrahedra (origin, a, b, c) of the
triangles: volume det[a b c]/6, first moment vol (a+b+c)/4 and second moments
vol/20 (a a^T + b b^T + c c^T + s s^T), s = a+b+c; the inertia tensor about
the origin is tr(C) I - C.  Tiny values are clamped to zero as the reference
does (mass.py:267-282), then the body frame follows xform_model_2_body_space
(mass.py:285-392): parallel-axis shift to the centre of mass, eigenvectors of
the inertia tensor ordered by decreasing moment, a sign flip for a proper
rotation.
"""
import numpy as np

from . import math3 as Q

TOO_SMALL = 1.0e-10


class MassProperties:
    def __init__(self):
        self.mass = 0.0
        self.x = self.y = self.z = 0.0
        self.Ixx = self.Iyy = self.Izz = 0.0
        self.Ixy = self.Ixz = self.Iyz = 0.0


def compute_mass_properties(V, T, density):
    V = np.asarray(V, dtype=np.float64)
    T = np.asarray(T, dtype=np.int64)
    a, b, c = V[T[:, 0]], V[T[:, 1]], V[T[:, 2]]
    vol = np.einsum("ij,ij->i", a, np.cross(b, c)) / 6.0
    s = a + b + c
    total = vol.sum()
    if not total > 0.0:
        raise RuntimeError("compute_mass_properties(): internal error, bad mesh encountered")
    com = (vol[:, None] * s).sum(axis=0) / 4.0 / total
    Cv = (np.einsum("i,ij,ik->jk", vol, a, a) + np.einsum("i,ij,ik->jk", vol, b, b)
          + np.einsum("i,ij,ik->jk", vol, c, c) + np.einsum("i,ij,ik->jk", vol, s, s)) / 20.0
    I = np.trace(Cv) * np.eye(3) - Cv
    p = MassProperties()
    p.mass = density * total
    p.x, p.y, p.z = com
    p.Ixx, p.Iyy, p.Izz = density * I[0, 0], density * I[1, 1], density * I[2, 2]
    p.Ixy, p.Iyz, p.Ixz = density * I[0, 1], density * I[1, 2], density * I[0, 2]
    p.mass = p.mass if p.mass > TOO_SMALL else 0.0
    for k in ("Ixx", "Iyy", "Izz"):
        setattr(p, k, getattr(p, k) if getattr(p, k) > TOO_SMALL else 0.0)
    for k in ("Ixy", "Iyz", "Ixz", "x", "y", "z"):
        setattr(p, k, getattr(p, k) if abs(getattr(p, k)) > TOO_SMALL else 0.0)
    return p


def xform_model_2_body_space(prop):
    """Returns (r_bf2mf, q_bf2mf, mass, I_body): body -> model transform and the
    principal moments in decreasing order."""
    m = prop.mass
    x, y, z = -prop.x, -prop.y, -prop.z
    Ixx = prop.Ixx - m * (y * y + z * z)
    Iyy = prop.Iyy - m * (x * x + z * z)
    Izz = prop.Izz - m * (x * x + y * y)
    Ixy = prop.Ixy + m * (x * y)
    Ixz = prop.Ixz + m * (x * z)
    Iyz = prop.Iyz + m * (y * z)
    Ixy = Ixy if abs(Ixy) > TOO_SMALL else 0.0
    Ixz = Ixz if abs(Ixz) > TOO_SMALL else 0.0
    Iyz = Iyz if abs(Iyz) > TOO_SMALL else 0.0
    A = np.array([[Ixx, Ixy, Ixz], [Ixy, Iyy, Iyz], [Ixz, Iyz, Izz]])
    d, R = np.linalg.eig(A)
    d = np.real(d)
    R = np.real(R)
    order = [0, 0, 0]
    for p in ([0, 1, 2], [0, 2, 1], [1, 0, 2], [2, 0, 1], [1, 2, 0], [2, 1, 0]):
        if d[p[0]] >= d[p[1]] >= d[p[2]]:
            order = p
            break
    R = R[:, order]
    if np.linalg.det(R) < 0.0:
        R[:, 0] = -R[:, 0]
    q = Q.from_matrix(R)
    return np.array([prop.x, prop.y, prop.z]), q, m, d[order]
