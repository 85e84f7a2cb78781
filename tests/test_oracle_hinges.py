"""Oracle hinge problem (oracle/hinges.py) and the multi-problem solve driver.

Every check here comes from outside the restated formulas: finite differences
of the anchor separation, rotations that a hinge must leave free or must
stop, placed joint errors, numpy.linalg.solve of the dense system, a
pendulum's anchor at rest after a step, and hand-computed driver sequences.
"""
import types

import numpy as np
import pytest

import oracle.hinges as OH
import oracle.problem as OP
import oracle.scene as OS
import oracle.stepper as OST
import rainbow_amd.api as API
from rainbow_amd import math3 as Q
from rainbow_amd import scenes


def rot(axis, angle):
    """Rodrigues rotation matrix."""
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = OH.star(a)
    return np.eye(3) + np.sin(angle) * K + (1.0 - np.cos(angle)) * K @ K


def rand_q(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def fake_engine(rng, n=2):
    shape = types.SimpleNamespace(r=np.zeros(3), q=Q.identity())
    bodies = {}
    for k in range(n):
        b = types.SimpleNamespace(name="b%d" % k, idx=k, r=rng.normal(size=3), q=rand_q(rng), shape=shape)
        bodies[b.name] = b
    return types.SimpleNamespace(hinges={}, bodies=bodies)


def test_set_hinge_sockets_reproduce_the_joint_frame():
    rng = np.random.default_rng(3)
    eng = fake_engine(rng)
    p, c = eng.bodies["b0"], eng.bodies["b1"]
    o, s = rng.normal(size=3), rng.normal(size=3)
    API.create_hinge(eng, "h")
    with pytest.raises(RuntimeError):
        API.create_hinge(eng, "h")
    API.set_hinge(eng, "h", "b0", "b1", o, s, mode="world")
    h = eng.hinges["h"]
    su = s / np.linalg.norm(s)
    assert np.allclose(Q.rotate(p.q, h.arm_p) + p.r, o) and np.allclose(Q.rotate(c.q, h.arm_c) + c.r, o)
    assert np.allclose(Q.rotate(p.q, h.axis_p), su) and np.allclose(Q.rotate(c.q, h.axis_c), su)
    # body mode: origin and axis in the parent's body frame
    API.set_hinge(eng, "h", "b0", "b1", o, s, mode="body")
    ow, sw = p.r + rot_q(p.q) @ o, rot_q(p.q) @ su
    assert np.allclose(Q.rotate(c.q, h.arm_c) + c.r, ow) and np.allclose(Q.rotate(c.q, h.axis_c), sw)
    # model mode: first undo the parent shape's model-to-body transform
    p.shape = types.SimpleNamespace(r=np.array([0.3, -0.2, 0.1]), q=rand_q(rng))
    API.set_hinge(eng, "h", "b0", "b1", o, s, mode="model")
    ob = rot_q(p.shape.q).T @ (o - p.shape.r)
    sb = rot_q(p.shape.q).T @ su
    assert np.allclose(Q.rotate(p.q, h.arm_p) + p.r, p.r + rot_q(p.q) @ ob)
    assert np.allclose(Q.rotate(p.q, h.axis_p), rot_q(p.q) @ sb)
    with pytest.raises(ValueError):
        API.set_hinge(eng, "h", "b0", "b1", o, s, mode="joint")
    with pytest.raises(RuntimeError):
        API.set_hinge(eng, "nope", "b0", "b1", o, s)
    with pytest.raises(RuntimeError):
        API.set_hinge(eng, "h", "b0", "nope", o, s)


def rot_q(q):
    return np.asarray(Q.to_matrix(q))


def two_body_scene(rng, pre=False, fixed_parent=False):
    """A parent and a child with random states and a hinge at a random point."""
    eng = fake_engine(rng)
    API.create_hinge(eng, "h")
    API.set_hinge(eng, "h", "b0", "b1", rng.normal(size=3), rng.normal(size=3))
    sc = types.SimpleNamespace(N=2, r=np.array([b.r for b in eng.bodies.values()]),
                               q=np.array([b.q for b in eng.bodies.values()]),
                               v=rng.normal(size=(2, 3)), w=rng.normal(size=(2, 3)),
                               params=dict(OS.PARAM_DEFAULTS, use_pre_stabilization=pre))
    H = OH.from_engine(eng, {"b0": 0, "b1": 1})
    Winv = np.zeros((2, 6, 6))
    for k in range(2):
        if fixed_parent and k == 0:
            continue
        R = rot_q(sc.q[k])
        Winv[k, :3, :3] = np.eye(3) / (1.0 + k)
        Winv[k, 3:, 3:] = R @ np.diag([1.0, 0.5, 0.25]) @ R.T
    return sc, H, Winv


def jrows(P, sc, v, w):
    u = np.concatenate([v, w], axis=1)
    return P.JP[0] @ u[0] + P.JC[0] @ u[1]


def test_jacobian_rows_are_the_anchor_separation_rate():
    rng = np.random.default_rng(5)
    for _ in range(5):
        sc, H, Winv = two_body_scene(rng)
        P = OH.assemble(sc, H, Winv, np.zeros((2, 6)), 0.01)
        arms = [rot_q(sc.q[0]) @ H.arm_p[0], rot_q(sc.q[1]) @ H.arm_c[0]]

        def sep(t):
            a = [sc.r[k] + sc.v[k] * t + rot(sc.w[k], np.linalg.norm(sc.w[k]) * t) @ arms[k] for k in range(2)]
            return a[0] - a[1]
        h = 1e-6
        fd = (sep(h) - sep(-h)) / (2 * h)
        Ju = jrows(P, sc, sc.v, sc.w)
        assert np.allclose(Ju[:3], fd, atol=1e-7)
        # a common spin and a relative spin about the joint axis are free
        n_p = rot_q(sc.q[0]) @ H.axis_p[0]
        w_same = np.array([sc.w[0], sc.w[0]])
        assert np.allclose(jrows(P, sc, np.zeros((2, 3)), w_same)[3:], 0.0, atol=1e-12)
        w_axis = np.array([sc.w[0], sc.w[0] + 0.7 * n_p])
        assert np.allclose(jrows(P, sc, np.zeros((2, 3)), w_axis)[3:], 0.0, atol=1e-12)
        # a relative spin about t_p (b_p) shows only in row 3 (4), with the child's sign
        t_p, b_p, _ = Q.make_orthonormal_vectors(n_p)
        r3 = jrows(P, sc, np.zeros((2, 3)), np.array([np.zeros(3), t_p]))[3:]
        r4 = jrows(P, sc, np.zeros((2, 3)), np.array([np.zeros(3), b_p]))[3:]
        assert np.allclose(r3, [-1.0, 0.0], atol=1e-12) and np.allclose(r4, [0.0, -1.0], atol=1e-12)


def test_error_vector_of_placed_joints():
    rng = np.random.default_rng(7)
    sc, H, Winv = two_body_scene(rng, pre=True)
    dt = 0.02
    rate = sc.params["gap_reduction"] / dt
    # put the child so that both sockets coincide (same orientation as the parent)
    sc.q[1] = sc.q[0].copy()
    H.arm_c[0] = H.arm_p[0].copy()
    H.axis_c[0] = H.axis_p[0].copy()
    sc.r[1] = sc.r[0].copy()
    P = OH.assemble(sc, H, Winv, np.zeros((2, 6)), dt)
    assert np.allclose(P.g[0], 0.0, atol=1e-14)
    d = np.array([0.01, -0.02, 0.03])
    sc.r[1] = sc.r[0] + d
    P = OH.assemble(sc, H, Winv, np.zeros((2, 6)), dt)
    assert np.allclose(P.g[0, :3], rate * d, atol=1e-12) and np.allclose(P.g[0, 3:], 0.0, atol=1e-14)
    sc.r[1] = sc.r[0].copy()
    n_p = rot_q(sc.q[0]) @ H.axis_p[0]
    t_p, b_p, _ = Q.make_orthonormal_vectors(n_p)
    th = 0.05
    for axis, row in ((t_p, 3), (b_p, 4)):
        # turn the child about the axis through its socket: anchors still meet
        Rt = rot(axis, th)
        q_turn = Q.from_matrix(Rt)
        sc.q[1] = Q.prod(q_turn, sc.q[0])
        anchor = sc.r[0] + rot_q(sc.q[0]) @ H.arm_p[0]
        sc.r[1] = anchor - rot_q(sc.q[1]) @ H.arm_c[0]
        P = OH.assemble(sc, H, Winv, np.zeros((2, 6)), dt)
        want = np.zeros(5)
        want[row] = rate * np.sin(th)
        assert np.allclose(P.g[0], want, atol=1e-12), (row, P.g[0], want)


def chain_problem(rng, nb=4, fixed0=True):
    """Hinges (k, k + 1) on nb bodies with random states."""
    eng = fake_engine(rng, nb)
    for k in range(nb - 1):
        API.create_hinge(eng, "h%d" % k)
        API.set_hinge(eng, "h%d" % k, "b%d" % k, "b%d" % (k + 1), rng.normal(size=3), rng.normal(size=3))
    sc = types.SimpleNamespace(N=nb, r=np.array([b.r for b in eng.bodies.values()]),
                               q=np.array([b.q for b in eng.bodies.values()]),
                               v=rng.normal(size=(nb, 3)), w=rng.normal(size=(nb, 3)),
                               params=dict(OS.PARAM_DEFAULTS))
    H = OH.from_engine(eng, {"b%d" % k: k for k in range(nb)})
    Winv = np.zeros((nb, 6, 6))
    for k in range(nb):
        if fixed0 and k == 0:
            continue
        R = rot_q(sc.q[k])
        Winv[k, :3, :3] = np.eye(3) / (1.0 + 0.5 * k)
        Winv[k, 3:, 3:] = R @ np.diag(rng.uniform(0.3, 2.0, 3)) @ R.T
    du = rng.normal(size=(nb, 6))
    return sc, OH.assemble(sc, H, Winv, du, 0.01), Winv, du


def dense(P, Winv):
    K, N = P.K, P.N
    J = np.zeros((5 * K, 6 * N))
    for k in range(K):
        J[5 * k:5 * k + 5, 6 * P.parent[k]:6 * P.parent[k] + 6] += P.JP[k]
        J[5 * k:5 * k + 5, 6 * P.child[k]:6 * P.child[k] + 6] += P.JC[k]
    W = np.zeros((6 * N, 6 * N))
    for b in range(N):
        W[6 * b:6 * b + 6, 6 * b:6 * b + 6] = Winv[b]
    return J, W


def test_assembly_against_dense_matrices():
    rng = np.random.default_rng(11)
    sc, P, Winv, du = chain_problem(rng)
    J, W = dense(P, Winv)
    u = np.concatenate([sc.v, sc.w], axis=1).ravel()
    b = J @ u + J @ (0.01 * du.ravel())
    assert np.allclose(P.b.ravel(), b, atol=1e-12)
    A = J @ W @ J.T
    assert np.allclose(P.diag.ravel(), np.diag(A), atol=1e-12)
    x = rng.normal(size=(P.K, 5))
    assert np.allclose(OH.body_w(P, x).ravel(), W @ J.T @ x.ravel(), atol=1e-12)


def test_converged_solve_matches_dense_solve():
    rng = np.random.default_rng(13)
    sc, P, Winv, du = chain_problem(rng)
    J, W = dense(P, Winv)
    A = J @ W @ J.T
    x_ref = np.linalg.solve(A, -P.b.ravel())
    # too_small_merit_change = 0: no r-factor growth (with the default 0.01 the
    # driver stalls on this problem, as the reference's would: r shrinks after
    # rejects until x stops moving)
    prm = dict(sc.params, max_iterations=5000, absolute_tolerance=0.0, relative_tolerance=0.0,
               too_small_merit_change=0.0)
    res = OP.solve_multi([OH.Solvable(P)], prm)
    assert res["exitcode"] == 0
    assert np.allclose(P.sol.ravel(), x_ref, rtol=1e-7, atol=1e-7 * np.abs(x_ref).max())
    # the constraint velocity of the solved system vanishes
    v_new = A @ P.sol.ravel() + P.b.ravel()
    assert np.abs(v_new).max() < 1e-6 * max(1.0, np.abs(P.b).max())


def test_one_sweep_is_serial_gauss_seidel():
    rng = np.random.default_rng(17)
    sc, P, Winv, du = chain_problem(rng)
    J, W = dense(P, Winv)
    A = J @ W @ J.T
    x = rng.normal(size=5 * P.K)
    P.x = x.reshape(P.K, 5).copy()
    OH.sweep(P)
    # block Gauss-Seidel with a block Jacobi step inside each hinge
    y = x.copy()
    r = P.r.ravel()
    b = P.b.ravel()
    for k in range(P.K):
        blk = slice(5 * k, 5 * k + 5)
        z = y[blk] - r[blk] * (A[blk] @ y + b[blk])
        y[blk] = z
    assert np.allclose(P.x.ravel(), y, atol=1e-12)


def test_pendulum_anchor_at_rest_after_a_step(oracle_bank32):
    eng = scenes.cube_hinge_chain(oracle_bank32, n=2, edge=1.0, backend=None,
                                  params=dict(max_iterations=3000, too_small_merit_change=0.0))
    sc = OS.from_engine(eng)
    sc.w[1] = np.array([0.3, -0.2, 0.5])
    q_before = sc.q[1].copy()
    out = OST.step(sc, 0.01)
    assert out["contacts"].K == 0 and out["log"]["exitcode"] == 0
    P = out["hinges"]
    arm = rot_q(q_before) @ P.arm_c[0]     # the rows hold the arm at the start of the step
    # the child's point at the joint moves with the fixed parent (rest); only
    # the spin about the joint axis (world z) survives
    v_anchor = sc.v[1] + np.cross(sc.w[1], arm)
    assert np.abs(v_anchor).max() < 1e-8
    assert np.abs(sc.w[1, :2]).max() < 1e-8 and abs(sc.w[1, 2]) > 1e-3


def test_hinge_only_step_logs_the_hinge_solve(oracle_bank32):
    eng = scenes.cube_hinge_chain(oracle_bank32, n=3, edge=1.0, backend=None,
                                  params=dict(max_iterations=50, absolute_tolerance=1e-6))
    sc = OS.from_engine(eng)
    out = OST.step(sc, 0.01)
    lg = out["log"]
    assert out["contacts"].K == 0
    assert lg["exitcode"] in (0, 1, 2) and np.isfinite(lg["residuals"][0])
    assert "natural_residual" not in lg


class Scripted:
    """A problem whose sweep sets x to the next scripted value."""

    def __init__(self, xs):
        self.xs = [np.array([float(v)]) for v in xs]
        self.k = 0
        self.x = np.zeros(1)
        self.sol = np.zeros(1)
        self.r = np.ones(1)

    def sweep(self):
        self.x = self.xs[min(self.k, len(self.xs) - 1)].copy()
        self.k += 1


def test_solve_multi_scripted_sequences():
    prm = dict(OS.PARAM_DEFAULTS, max_iterations=4, absolute_tolerance=0.0, relative_tolerance=0.0,
               too_small_merit_change=0.0)
    # problem A: merits 4, 1 (accept), then a jump -> reject (keeps 1); B: merits 2, 3 (reject), ...
    a = Scripted([4.0, 3.0, 10.0, 3.5])
    b = Scripted([2.0, 5.0, 2.5, 2.5])
    res = OP.solve_multi([a, b], prm)
    # it 0: A 4 (acc), B 2 (acc) -> 4
    # it 1: A |3-4| = 1 (acc), B |5-2| = 3 > 2 (rej, x back to 2) -> max(1, 2) = 2
    # it 2: A |10-3| = 7 > 1 (rej) -> 1;  B |2.5-2| = 0.5 (acc) -> 0.5; max = 1
    # it 3: A |3.5-3| = 0.5 (acc);  B |2.5-2.5| = 0 (acc) -> 0.5
    assert np.allclose(res["residuals"], [4.0, 2.0, 1.0, 0.5])
    assert res["iterations"] == 4 and res["exitcode"] == 0 and res["rejects"] == 1
    assert b.r[0] == pytest.approx(0.7) and a.r[0] == pytest.approx(0.7)
    # absolute stop on the global merit
    a = Scripted([4.0, 3.0, 10.0, 3.5])
    b = Scripted([2.0, 5.0, 2.5, 2.5])
    res = OP.solve_multi([a, b], dict(prm, absolute_tolerance=1.5))
    assert res["iterations"] == 2 and res["exitcode"] == 1


def test_solve_multi_with_one_problem_is_solve(oracle_bank32):
    eng = scenes.box_stack(oracle_bank32, backend=None, nx=2, ny=2, nz=1,
                           params=dict(max_iterations=30, absolute_tolerance=1e-7, relative_tolerance=1e-5))
    sc = OS.from_engine(eng)
    Winv, _, du = OP.body_mass_blocks(sc)
    from oracle import collision
    c = collision.detect(sc, {})
    assert c.K > 0
    P1 = OP.assemble(sc, c, Winv, du, 0.01)
    P2 = OP.assemble(sc, c, Winv, du, 0.01)
    groups, _ = OST.sweep_groups(sc, c)
    r1 = OP.solve(P1, groups, sc.params)
    r2 = OP.solve_multi([OP.ContactSolvable(P2, groups, sc.params)], sc.params)
    assert r1["iterations"] == r2["iterations"] and r1["exitcode"] == r2["exitcode"]
    assert r1["rejects"] == r2["rejects"]
    assert np.array_equal(r1["residuals"], r2["residuals"]) and np.array_equal(P1.sol, P2.sol)
