"""Mass properties of closed triangle meshes and the body frame.

The integrals are Mirtich's ("Fast and Accurate Computation of Polyhedral
Mass Properties", J. Graphics Tools 1(2), 1996): each triangle is projected
onto the coordinate plane most nearly parallel to it, the projection
integrals are taken over its three edges, lifted back to face integrals
through the face plane n.x + w = 0, and summed into volume integrals by the
divergence theorem.  The reference computes the same integrals
(rainbow/simulators/prox_rigid_bodies/mass.py:58-283).

Bodies get their frame from the eigenvectors of the inertia tensor, and for a
body with tied principal moments (cubes, spheres) those eigenvectors depend
on the last bits of the tensor.  So every rounded operation here happens in
the reference's order: the edge terms and their sums in edge order, each
face's contribution in triangle order (np.cumsum accumulates left to right),
and each face normal, its length and plane offset from the same numpy calls
on the same 3-vectors (np.cross, np.linalg.norm, np.dot).  The mass, centre
of mass, inertia tensor and body frame are then bitwise the reference's.

Tiny values are clamped to zero as the reference does (mass.py:267-282),
then the body frame follows xform_model_2_body_space (mass.py:285-392):
parallel-axis shift to the centre of mass, eigenvectors of the inertia tensor
ordered by decreasing moment, a sign flip for a proper rotation.
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


def _face_planes(V, T):
    """Unit normal n, offset w (n . x + w = 0) and dominant axis of every
    triangle, from the same numpy calls on 3-vectors as the reference
    (mass.py:176-183)."""
    M = len(T)
    normals = np.empty((M, 3))
    offsets = np.empty(M)
    for t in range(M):
        p0, p1, p2 = V[T[t, 0], :], V[T[t, 1], :], V[T[t, 2], :]
        c = np.cross(p1 - p0, p2 - p0, axis=0)
        n = c / np.linalg.norm(c)
        normals[t] = n
        offsets[t] = -np.dot(n, (p0 + p1 + p2) / 3.0)
    m = np.fabs(normals)
    # vector3.py:127-133: strict comparisons, the later axis wins ties
    dom = np.where((m[:, 0] > m[:, 1]) & (m[:, 0] > m[:, 2]), 0, np.where(m[:, 1] > m[:, 2], 1, 2))
    return normals, offsets, dom


def _projection_integrals(A, B):
    """Integrals over the projected triangle of 1, a, a^2, a^3, b, b^2, b^3,
    ab, a^2 b, a b^2 (Mirtich section 4.3); A, B: (3, M) projected vertex
    coordinates, vertex order as in T.  Edge terms are summed in edge order."""
    M = A.shape[1]
    P1, Pa, Paa, Paaa, Pb, Pbb, Pbbb, Pab, Paab, Pabb = (np.zeros(M) for _ in range(10))
    for e in range(3):
        a0, b0 = A[e], B[e]
        a1, b1 = A[(e + 1) % 3], B[(e + 1) % 3]
        da, db = a1 - a0, b1 - b0
        a0_2 = a0 * a0
        a0_3 = a0_2 * a0
        a0_4 = a0_3 * a0
        b0_2 = b0 * b0
        b0_3 = b0_2 * b0
        b0_4 = b0_3 * b0
        a1_2 = a1 * a1
        a1_3 = a1_2 * a1
        b1_2 = b1 * b1
        b1_3 = b1_2 * b1
        C1 = a1 + a0
        Ca = a1 * C1 + a0_2
        Caa = a1 * Ca + a0_3
        Caaa = a1 * Caa + a0_4
        Cb = b1 * (b1 + b0) + b0_2
        Cbb = b1 * Cb + b0_3
        Cbbb = b1 * Cbb + b0_4
        Cab = 3.0 * a1_2 + 2.0 * a1 * a0 + a0_2
        Kab = a1_2 + 2.0 * a1 * a0 + 3.0 * a0_2
        Caab = a0 * Cab + 4.0 * a1_3
        Kaab = a1 * Kab + 4.0 * a0_3
        Cabb = 4.0 * b1_3 + 3.0 * b1_2 * b0 + 2.0 * b1 * b0_2 + b0_3
        Kabb = b1_3 + 2.0 * b1_2 * b0 + 3.0 * b1 * b0_2 + 4.0 * b0_3
        P1 += db * C1
        Pa += db * Ca
        Paa += db * Caa
        Paaa += db * Caaa
        Pb += da * Cb
        Pbb += da * Cbb
        Pbbb += da * Cbbb
        Pab += db * (b1 * Cab + b0 * Kab)
        Paab += db * (b1 * Caab + b0 * Kaab)
        Pabb += da * (a1 * Cabb + a0 * Kabb)
    return (P1 / 2.0, Pa / 6.0, Paa / 12.0, Paaa / 20.0, Pb / -6.0, Pbb / -12.0, Pbbb / -20.0, Pab / 24.0,
            Paab / 60.0, Pabb / -60.0)


def compute_mass_properties(V, T, density):
    """Mass, centre of mass and inertia tensor about the model origin of the
    closed mesh (V, T) with uniform ``density`` (mass.py:186-283)."""
    V = np.asarray(V, dtype=np.float64)
    T = np.asarray(T, dtype=np.int64)
    M = len(T)
    n, w, Z = _face_planes(V, T)
    X = (Z + 1) % 3
    Y = (X + 1) % 3
    idx = np.arange(M)
    nA, nB, nC = n[idx, X], n[idx, Y], n[idx, Z]
    k1 = 1.0 / nC
    k2 = k1 * k1
    k3 = k2 * k1
    k4 = k3 * k1
    tri = V[T]                                   # (M, 3 vertices, 3 coordinates)
    A = tri[idx, :, X].T if M else np.zeros((3, 0))
    B = tri[idx, :, Y].T if M else np.zeros((3, 0))
    P1, Pa, Paa, Paaa, Pb, Pbb, Pbbb, Pab, Paab, Pabb = _projection_integrals(A, B)
    # face integrals through the plane (Mirtich section 4.2)
    Fa = k1 * Pa
    Fb = k1 * Pb
    Fc = -k2 * (nA * Pa + nB * Pb + w * P1)
    Faa = k1 * Paa
    Fbb = k1 * Pbb
    Fcc = k3 * (nA * nA * Paa + 2.0 * nA * nB * Pab + nB * nB * Pbb + w * (2.0 * (nA * Pa + nB * Pb) + w * P1))
    Faaa = k1 * Paaa
    Fbbb = k1 * Pbbb
    Fccc = -k4 * (nA * nA * nA * Paaa + 3.0 * nA * nA * nB * Paab + 3.0 * nA * nB * nB * Pabb
                  + nB * nB * nB * Pbbb + 3.0 * w * (nA * nA * Paa + 2.0 * nA * nB * Pab + nB * nB * Pbb)
                  + w * w * (3.0 * (nA * Pa + nB * Pb) + w * P1))
    Faab = k1 * Paab
    Fbbc = -k2 * (nA * Pabb + nB * Pbbb + w * Pbb)
    Fcca = k3 * (nA * nA * Paaa + 2.0 * nA * nB * Paab + nB * nB * Pabb + w * (2.0 * (nA * Paa + nB * Pab) + w * Pa))

    def total(x):
        # running sum in triangle order from 0.0, as the reference's loop adds them
        return float(np.cumsum(x)[-1]) if len(x) else 0.0

    # volume: the x-component of the face integral of x (the reference's T0)
    vol = total(np.where(X == 0, n[:, 0] * Fa, np.where(Y == 0, n[:, 0] * Fb, n[:, 0] * Fc)))
    # each axis receives one term per face: from the face's a-, b- or c-axis
    first, second, product = np.zeros(3), np.zeros(3), np.zeros(3)
    for c in range(3):
        first[c] = total(np.where(X == c, nA * Faa, np.where(Y == c, nB * Fbb, nC * Fcc)))
        second[c] = total(np.where(X == c, nA * Faaa, np.where(Y == c, nB * Fbbb, nC * Fccc)))
        product[c] = total(np.where(X == c, nA * Faab, np.where(Y == c, nB * Fbbc, nC * Fcca)))
    first /= 2.0
    second /= 3.0
    product /= 2.0
    if not vol > 0.0:
        raise RuntimeError("compute_mass_properties(): internal error, bad mesh encountered")
    p = MassProperties()
    p.mass = density * vol
    p.x, p.y, p.z = first[0] / vol, first[1] / vol, first[2] / vol
    p.Ixx = density * (second[1] + second[2])
    p.Iyy = density * (second[2] + second[0])
    p.Izz = density * (second[0] + second[1])
    p.Ixy = -density * product[0]
    p.Iyz = -density * product[1]
    p.Ixz = -density * product[2]
    p.mass = p.mass if p.mass > TOO_SMALL else 0.0
    for k in ("Ixx", "Iyy", "Izz"):
        setattr(p, k, getattr(p, k) if getattr(p, k) > TOO_SMALL else 0.0)
    for k in ("Ixy", "Iyz", "Ixz", "x", "y", "z"):
        setattr(p, k, getattr(p, k) if abs(getattr(p, k)) > TOO_SMALL else 0.0)
    return p


def xform_model_2_body_space(prop):
    """Returns (r_bf2mf, q_bf2mf, mass, I_body): body -> model transform and the
    principal moments in decreasing order (mass.py:285-392)."""
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
    A = np.array([[Ixx, Ixy, Ixz], [Ixy, Iyy, Iyz], [Ixz, Iyz, Izz]], dtype=np.float64)
    d, R = np.linalg.eig(A)
    order = [0, 0, 0]
    for p in ([0, 1, 2], [0, 2, 1], [1, 0, 2], [2, 0, 1], [1, 2, 0], [2, 1, 0]):
        if d[p[0]] >= d[p[1]] >= d[p[2]]:
            order = p
            break
    R = R[:, order]
    if np.linalg.det(R) < 0.0:
        R[:, 0] = -R[:, 0]
    q = Q.from_matrix(R)
    return np.array([-x, -y, -z]), q, m, d[order]
