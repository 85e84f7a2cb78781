"""Oracle quaternion/frame helpers and SDF construction."""
import numpy as np
import pytest

from oracle import quat as Q
from oracle import sdf as S
from oracle.grid import Grid, nodes_array
from rainbow_amd import mesh as MESH


def axis_angle_matrix(axis, ang):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K


def quat_from_axis_angle(axis, ang):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    return np.array([np.cos(ang / 2), *(np.sin(ang / 2) * a)])


def test_rotate_matches_rodrigues():
    rng = np.random.default_rng(1)
    for _ in range(50):
        axis, ang = rng.normal(size=3), rng.uniform(-np.pi, np.pi)
        q = quat_from_axis_angle(axis, ang)
        R = axis_angle_matrix(axis, ang)
        V = rng.normal(size=(7, 3))
        assert np.allclose(Q.rotate_array(q, V), V @ R.T, atol=1e-13)
        assert np.allclose(Q.rotate(q, V[0]), R @ V[0], atol=1e-13)
        assert np.allclose(Q.to_matrix(q), R, atol=1e-13)
        assert np.allclose(Q.to_matrix(3.0 * q), R, atol=1e-13)   # to_matrix normalises


def test_reference_quaternion_cases():
    # test_math_quaternion.py rotate cases: Rx(90) maps j to k, Ry(90) maps k to i
    qx = np.array([np.cos(np.pi / 4), np.sin(np.pi / 4), 0, 0])
    assert np.allclose(Q.rotate(qx, [0, 1, 0]), [0, 0, 1], atol=1e-15)
    qy = np.array([np.cos(np.pi / 4), 0, np.sin(np.pi / 4), 0])
    assert np.allclose(Q.rotate(qy, [0, 0, 1]), [1, 0, 0], atol=1e-15)
    assert np.allclose(Q.unit([0, 0, 0, 0]), [1, 0, 0, 0])


def test_orthonormal_frame_rule():
    # vector3.py:75-113: orthonormal; tangent guess = axis of smallest |n_i|, later axis wins ties
    rng = np.random.default_rng(2)
    for _ in range(100):
        n = rng.normal(size=3)
        t, b, nn = Q.make_orthonormal_vectors(n)
        M = np.stack([t, b, nn])
        assert np.allclose(M @ M.T, np.eye(3), atol=1e-12)
        assert np.allclose(np.cross(t, b), nn, atol=1e-12)
    t, b, n = Q.make_orthonormal_vectors([0, 1, 0])   # |nx| = |nz| = 0: z wins
    assert np.allclose(b, np.cross([0, 1, 0], [0, 0, 1]))
    t, b, n = Q.make_orthonormal_vectors([0, 0, 1])   # x and y tie: y wins
    assert np.allclose(b, np.cross([0, 0, 1], [0, 1, 0]))


def test_sdf_box_is_analytic():
    # the box mesh is the box: SDF = analytic box distance (in/out)
    V, T = MESH.create_box(1.0, 2.0, 0.5)
    g = S.create_signed_distance(V, T, 9, 9, 9, boundary=0.4)
    P = nodes_array(g)
    h = np.array([0.5, 1.0, 0.25])
    q = np.abs(P) - h
    exact = np.linalg.norm(np.maximum(q, 0), axis=1) + np.minimum(np.max(q, axis=1), 0)
    assert np.allclose(g.values, exact, atol=1e-12)


def test_sdf_sphere_bruteforce():
    V, T = MESH.create_sphere(0.5, 8, 8)
    rng = np.random.default_rng(3)
    P = rng.uniform(-0.8, 0.8, size=(40, 3))
    d = S.signed_distance(P, V, T)
    # brute force: dense barycentric sampling of every triangle
    u, v = np.meshgrid(np.linspace(0, 1, 41), np.linspace(0, 1, 41))
    m = (u + v) <= 1
    u, v = u[m], v[m]
    A, B, C = V[T[:, 0]], V[T[:, 1]], V[T[:, 2]]
    pts = (A[:, None] + u[None, :, None] * (B - A)[:, None] + v[None, :, None] * (C - A)[:, None]).reshape(-1, 3)
    bf = np.min(np.linalg.norm(P[:, None] - pts[None], axis=-1), axis=1)
    assert np.all(np.abs(np.abs(d) - bf) < 0.01)
    assert np.all(np.abs(d) <= bf + 1e-12)          # exact distance never exceeds a sampled one
    inside = np.linalg.norm(P, axis=1) < 0.45
    outside = np.linalg.norm(P, axis=1) > 0.51
    assert np.all(d[inside] < 0) and np.all(d[outside] > 0)
