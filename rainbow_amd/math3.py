"""Quaternion and vector helpers used by the host-side scene API.

Same conventions as rainbow/math/quaternion.py and vector3.py: q = [s, x, y, z],
rotation r -> (q (0, r) q*)[1:].
"""
from math import cos, sin, sqrt

import numpy as np


def identity():
    return np.array([1.0, 0.0, 0.0, 0.0])


def unit(q):
    q = np.asarray(q, dtype=np.float64)
    n = np.linalg.norm(q)
    return q / n if n > 0 else identity()


def conjugate(q):
    return np.array([q[0], -q[1], -q[2], -q[3]], dtype=np.float64)


def _prod(a0, a1, a2, a3, b0, b1, b2, b3):
    # [a0 b0 - A.B, a0 B + b0 A + A x B] on Python floats
    return (a0 * b0 - (a1 * b1 + a2 * b2 + a3 * b3),
            a0 * b1 + a1 * b0 + (a2 * b3 - a3 * b2),
            a0 * b2 + a2 * b0 + (a3 * b1 - a1 * b3),
            a0 * b3 + a3 * b0 + (a1 * b2 - a2 * b1))


def prod(a, b):
    return np.array(_prod(float(a[0]), float(a[1]), float(a[2]), float(a[3]),
                          float(b[0]), float(b[1]), float(b[2]), float(b[3])), dtype=np.float64)


def rotate(q, r):
    s, x, y, z = (float(c) for c in q)
    t = _prod(*_prod(s, x, y, z, 0.0, float(r[0]), float(r[1]), float(r[2])), s, -x, -y, -z)
    return np.array(t[1:], dtype=np.float64)


def rotate_array(q, V):
    """Rotate each row of V (quaternion sandwich written out per component)."""
    V = np.asarray(V, dtype=np.float64)
    s, x, y, z = q
    vx, vy, vz = V[:, 0], V[:, 1], V[:, 2]
    i0 = s * 0.0 - x * vx - y * vy - z * vz
    i1 = s * vx + x * 0.0 + y * vz - z * vy
    i2 = s * vy - x * vz + y * 0.0 + z * vx
    i3 = s * vz + x * vy - y * vx + z * 0.0
    bs, bx, by, bz = s, -x, -y, -z
    return np.stack([i0 * bx + i1 * bs + i2 * bz - i3 * by,
                     i0 * by - i1 * bz + i2 * bs + i3 * bx,
                     i0 * bz + i1 * by - i2 * bx + i3 * bs], axis=1)


def Ru(radians, axis):
    c, s = cos(radians / 2.0), sin(radians / 2.0)
    n = np.asarray(axis, dtype=np.float64) / np.linalg.norm(axis)
    return np.array([c, s * n[0], s * n[1], s * n[2]])


def Rx(radians):
    return np.array([cos(radians / 2.0), sin(radians / 2.0), 0.0, 0.0])


def Ry(radians):
    return np.array([cos(radians / 2.0), 0.0, sin(radians / 2.0), 0.0])


def Rz(radians):
    return np.array([cos(radians / 2.0), 0.0, 0.0, sin(radians / 2.0)])


def to_matrix(q):
    s, x, y, z = unit(q)
    return np.array([
        [1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - s * z), 2.0 * (s * y + x * z)],
        [2.0 * (x * y + s * z), 1.0 - 2.0 * (x * x + z * z), 2.0 * (z * y - s * x)],
        [2.0 * (-s * y + x * z), 2.0 * (z * y + s * x), 1.0 - 2.0 * (y * y + x * x)]])


def from_matrix(M):
    """Rotation matrix to quaternion (largest-component branch selection)."""
    M = np.asarray(M, dtype=np.float64)
    tr = M[0, 0] + M[1, 1] + M[2, 2]
    if tr < 0.0:
        i = int(np.argmax([M[0, 0], M[1, 1], M[2, 2]]))
        if M[1, 1] > M[0, 0]:
            i = 1
        if M[2, 2] > M[i, i]:
            i = 2
        if i == 0:
            r = sqrt((M[0, 0] - (M[1, 1] + M[2, 2])) + 1.0)
            qx = 0.5 * r
            r = 0.5 / r
            return np.array([(M[2, 1] - M[1, 2]) * r, qx, (M[0, 1] + M[1, 0]) * r, (M[2, 0] + M[0, 2]) * r])
        if i == 1:
            r = sqrt((M[1, 1] - (M[2, 2] + M[0, 0])) + 1.0)
            qy = 0.5 * r
            r = 0.5 / r
            return np.array([(M[0, 2] - M[2, 0]) * r, (M[0, 1] + M[1, 0]) * r, qy, (M[1, 2] + M[2, 1]) * r])
        r = sqrt((M[2, 2] - (M[0, 0] + M[1, 1])) + 1.0)
        qz = 0.5 * r
        r = 0.5 / r
        return np.array([(M[1, 0] - M[0, 1]) * r, (M[2, 0] + M[0, 2]) * r, (M[1, 2] + M[2, 1]) * r, qz])
    r = sqrt(tr + 1.0)
    qs = 0.5 * r
    r = 0.5 / r
    return np.array([qs, (M[2, 1] - M[1, 2]) * r, (M[0, 2] - M[2, 0]) * r, (M[1, 0] - M[0, 1]) * r])


def unit3(a):
    a = np.asarray(a, dtype=np.float64)
    return a / np.linalg.norm(a)


def make_orthonormal_vectors(n):
    """vector3.py:75-113: (t, b, n) with t guessed from the coordinate axis of
    the smallest |n_i| (a later axis wins a tie), b = unit(n x t), t = b x n."""
    n = np.asarray(n, dtype=np.float64)
    n = n / np.linalg.norm(n)
    ax = np.fabs(n)
    if ax[0] <= ax[1] and ax[0] <= ax[2]:
        t = np.array([1.0, 0.0, 0.0])
    if ax[1] <= ax[0] and ax[1] <= ax[2]:
        t = np.array([0.0, 1.0, 0.0])
    if ax[2] <= ax[0] and ax[2] <= ax[1]:
        t = np.array([0.0, 0.0, 1.0])
    b = np.cross(n, t)
    b = b / np.linalg.norm(b)
    t = np.cross(b, n)
    return t, b, n
