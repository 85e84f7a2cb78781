"""Quaternion helpers, restated from rainbow/math/quaternion.py and vector3.py.

Convention: q = [s, x, y, z].  Pinned by tests/test_oracle_math.py against the
reference unit-test expectations (test_math_quaternion.py rotate / to_matrix
cases) and against rotation-matrix closed forms.
"""
import numpy as np


def unit(q):
    """quaternion.py:76-80 (identity if the norm is zero)."""
    q = np.asarray(q, dtype=np.float64)
    n = np.linalg.norm(q)
    if n > 0:
        return q / n
    return np.array([1.0, 0.0, 0.0, 0.0])


def conjugate(q):
    """quaternion.py:182-183."""
    return np.array([q[0], -q[1], -q[2], -q[3]], dtype=np.float64)


def prod(a, b):
    """quaternion.py:186-193: [a0 b0 - A.B, a0 B + b0 A + A x B].

    The dot product is summed left to right and the cross product written out
    (np.dot on 3-vectors goes through BLAS, whose rounding order is not fixed)."""
    s = a[0] * b[0] - (a[1] * b[1] + a[2] * b[2] + a[3] * b[3])
    x = a[0] * b[1] + a[1] * b[0] + (a[2] * b[3] - a[3] * b[2])
    y = a[0] * b[2] + a[2] * b[0] + (a[3] * b[1] - a[1] * b[3])
    z = a[0] * b[3] + a[3] * b[0] + (a[1] * b[2] - a[2] * b[1])
    return np.array([s, x, y, z], dtype=np.float64)


def rotate(q, r):
    """quaternion.py:196-198: (q (0,r) q*)[1:]."""
    qr = np.array([0.0, r[0], r[1], r[2]], dtype=np.float64)
    return prod(prod(q, qr), conjugate(q))[1:]


def _prod_qa_array(qa, B):
    """quaternion.py:201-221 (single quaternion times an array)."""
    return np.stack([
        qa[0] * B[:, 0] - qa[1] * B[:, 1] - qa[2] * B[:, 2] - qa[3] * B[:, 3],
        qa[0] * B[:, 1] + qa[1] * B[:, 0] + qa[2] * B[:, 3] - qa[3] * B[:, 2],
        qa[0] * B[:, 2] - qa[1] * B[:, 3] + qa[2] * B[:, 0] + qa[3] * B[:, 1],
        qa[0] * B[:, 3] + qa[1] * B[:, 2] - qa[2] * B[:, 1] + qa[3] * B[:, 0],
    ], axis=1)


def _prod_array_qb(A, qb):
    """quaternion.py:222-235 (array times a single quaternion)."""
    return np.stack([
        A[:, 0] * qb[0] - A[:, 1] * qb[1] - A[:, 2] * qb[2] - A[:, 3] * qb[3],
        A[:, 0] * qb[1] + A[:, 1] * qb[0] + A[:, 2] * qb[3] - A[:, 3] * qb[2],
        A[:, 0] * qb[2] - A[:, 1] * qb[3] + A[:, 2] * qb[0] + A[:, 3] * qb[1],
        A[:, 0] * qb[3] + A[:, 1] * qb[2] - A[:, 2] * qb[1] + A[:, 3] * qb[0],
    ], axis=1)


def rotate_array(q, V):
    """quaternion.py:244-250: rotate every row of V by q."""
    V = np.asarray(V, dtype=np.float64)
    qr = np.concatenate([np.zeros((len(V), 1)), V], axis=1)
    inner = _prod_qa_array(q, qr)
    outer = _prod_array_qb(inner, conjugate(q))
    return outer[:, 1:]


def to_matrix(q):
    """quaternion.py:108-129 (normalises first)."""
    q = unit(q)
    s, x, y, z = q
    R = np.zeros((3, 3))
    R[0, 0] = 1.0 - 2.0 * (y * y + z * z)
    R[1, 1] = 1.0 - 2.0 * (x * x + z * z)
    R[2, 2] = 1.0 - 2.0 * (y * y + x * x)
    R[1, 0] = 2.0 * (x * y + s * z)
    R[0, 1] = 2.0 * (x * y - s * z)
    R[2, 0] = 2.0 * (-s * y + x * z)
    R[0, 2] = 2.0 * (s * y + x * z)
    R[2, 1] = 2.0 * (z * y + s * x)
    R[1, 2] = 2.0 * (z * y - s * x)
    return R


def update_inertia_tensor(R, Bv):
    """mass.py:395-452: R diag(Bv) R^T written out element by element."""
    W = np.zeros((3, 3))
    W[0, 0] = R[0, 0] * R[0, 0] * Bv[0] + R[0, 1] * R[0, 1] * Bv[1] + R[0, 2] * R[0, 2] * Bv[2]
    W[1, 1] = R[1, 0] * R[1, 0] * Bv[0] + R[1, 1] * R[1, 1] * Bv[1] + R[1, 2] * R[1, 2] * Bv[2]
    W[2, 2] = R[2, 0] * R[2, 0] * Bv[0] + R[2, 1] * R[2, 1] * Bv[1] + R[2, 2] * R[2, 2] * Bv[2]
    W[1, 0] = W[0, 1] = R[1, 0] * R[0, 0] * Bv[0] + R[1, 1] * R[0, 1] * Bv[1] + R[1, 2] * R[0, 2] * Bv[2]
    W[2, 0] = W[0, 2] = R[2, 0] * R[0, 0] * Bv[0] + R[2, 1] * R[0, 1] * Bv[1] + R[2, 2] * R[0, 2] * Bv[2]
    W[2, 1] = W[1, 2] = R[1, 0] * R[2, 0] * Bv[0] + R[1, 1] * R[2, 1] * Bv[1] + R[1, 2] * R[2, 2] * Bv[2]
    return W


def make_orthonormal_vectors(n):
    """vector3.py:75-113: returns (t, b, n).

    The tangent guess is the axis with the smallest |n_i|; the three ``if``
    statements are evaluated in order, so on ties the later axis wins.
    """
    n = np.array(n, dtype=np.float64)
    n = n / np.linalg.norm(n)
    ax, ay, az = np.fabs(n)
    t = None
    if ax <= ay and ax <= az:
        t = np.array([1.0, 0.0, 0.0])
    if ay <= ax and ay <= az:
        t = np.array([0.0, 1.0, 0.0])
    if az <= ax and az <= ay:
        t = np.array([0.0, 0.0, 1.0])
    b = np.cross(n, t)
    b = b / np.linalg.norm(b)
    t = np.cross(b, n)
    return t, b, n
