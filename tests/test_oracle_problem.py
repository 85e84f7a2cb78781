"""Oracle contact problem: prox operators, Jacobian, sweeps, r-factor driver."""
import numpy as np
import pytest
from scipy.optimize import brentq

import oracle.scene as OS
from oracle import collision as CD
from oracle import coloring as COL
from oracle import problem as PB
from oracle import quat as Q
from oracle.scene import PARAM_DEFAULTS
from rainbow_amd import scenes


def test_prox_sphere_closed_forms():
    # inside -> identity, outside -> radial scaling to mu x_n, x_n <= 0 -> origin
    assert np.allclose(PB.prox_sphere(0.1, 0.0, 0.0, 0.5, 1.0), [[0.1, 0, 0]])
    out = PB.prox_sphere(3.0, 4.0, 0.0, 0.5, 2.0)[0]
    assert np.linalg.norm(out) == pytest.approx(1.0) and np.allclose(out / np.linalg.norm(out), [0.6, 0.8, 0])
    assert np.allclose(PB.prox_sphere(3.0, 4.0, 0.0, 0.5, 0.0), 0.0)
    assert np.allclose(PB.prox_sphere(3.0, 4.0, 0.0, 0.5, -1.0), 0.0)


def _ellipsoid_root(z, a):
    """Independent projection onto {sum (y_i/a_i)^2 <= 1}: root of
    sum a_i^2 z_i^2 / (a_i^2 + t)^2 = 1 solved with brentq."""
    z, a = np.asarray(z, float), np.asarray(a, float)
    if np.sum((z / a) ** 2) <= 1:
        return z
    f = lambda t: np.sum(a ** 2 * z ** 2 / (a ** 2 + t) ** 2) - 1
    hi = 1.0
    while f(hi) > 0:
        hi *= 2
    t = brentq(f, 0, hi, xtol=1e-15, rtol=1e-15)
    return a ** 2 * z / (a ** 2 + t)


def test_prox_ellipsoid_against_root_solve():
    prm = dict(PARAM_DEFAULTS)
    rng = np.random.default_rng(7)
    for _ in range(200):
        mu = rng.uniform(0.1, 1.0, 3)
        xn = rng.uniform(0.1, 2.0)
        z = rng.normal(scale=2.0, size=3)
        got = PB.prox_ellipsoid(z, mu, xn, prm)
        assert np.allclose(got, _ellipsoid_root(z, mu * xn), atol=1e-7)
    # isotropic reduces to the sphere; axis-aligned z clamps; x_n <= 0 -> 0
    z = np.array([2.0, -1.0, 0.5])
    assert np.allclose(PB.prox_ellipsoid(z, [0.5] * 3, 1.0, prm), PB.prox_sphere(*z, 0.5, 1.0)[0])
    assert np.allclose(PB.prox_ellipsoid([3.0, 0, 0], [0.5, 0.2, 0.1], 2.0, prm), [1.0, 0, 0], atol=1e-8)
    assert np.allclose(PB.prox_ellipsoid(z, [0.5, 0.2, 0.1], 0.0, prm), 0.0)
    assert np.allclose(PB.prox_ellipsoid([0.01, 0.01, 0.0], [0.5, 0.2, 0.1], 1.0, prm), [0.01, 0.01, 0.0])


def scene_with_contacts(bank, seed=1):
    from tests.test_oracle_collision import random_scene
    sc = OS.from_engine(random_scene(bank, 6, seed=seed, spread=0.7))
    # sink every body so there are ground contacts
    sc.r[1:, 1] -= 0.25
    return sc


def test_jacobian_equals_literal_assembly(oracle_bank32):
    sc = scene_with_contacts(oracle_bank32)
    c = CD.detect(sc)
    assert c.K > 0
    Winv, f_ext, du = PB.body_mass_blocks(sc)
    P = PB.assemble(sc, c, Winv, du, 0.01)
    N = sc.N
    # contacts.py:118-185 written out: dense J (4K x 6N)
    J = np.zeros((4 * c.K, 6 * N))
    for k in range(c.K):
        rA, rB = c.p[k] - sc.r[c.ia[k]], c.p[k] - sc.r[c.ib[k]]
        vs, vt, vn = Q.make_orthonormal_vectors(c.n[k])
        JA_v = -np.array([vn, vs, vt, np.zeros(3)])
        JA_w = -np.array([np.cross(rA, vn), np.cross(rA, vs), np.cross(rA, vt), vn])
        JB_v = np.array([vn, vs, vt, np.zeros(3)])
        JB_w = np.array([np.cross(rB, vn), np.cross(rB, vs), np.cross(rB, vt), vn])
        for i in range(4):
            J[4 * k + i, 6 * c.ia[k]:6 * c.ia[k] + 3] += JA_v[i]
            J[4 * k + i, 6 * c.ia[k] + 3:6 * c.ia[k] + 6] += JA_w[i]
            J[4 * k + i, 6 * c.ib[k]:6 * c.ib[k] + 3] += JB_v[i]
            J[4 * k + i, 6 * c.ib[k] + 3:6 * c.ib[k] + 6] += JB_w[i]
    W = np.zeros((6 * N, 6 * N))
    for b in range(N):
        W[6 * b:6 * b + 6, 6 * b:6 * b + 6] = Winv[b]
    u = np.concatenate([sc.v, sc.w], axis=1).ravel()
    b_ref = J @ u + J @ (0.01 * du.ravel())
    A = J @ W @ J.T
    diag = np.diag(A).copy()
    diag[diag == 0] = 1
    assert np.allclose(P.b.ravel(), b_ref, atol=1e-12)
    assert np.allclose(P.diag.ravel(), diag, rtol=1e-12, atol=1e-14)
    # one GS sweep == serial z = x - r (A x + b) updates on the dense form
    prm = dict(sc.params, friction_model="isotropic")
    x = np.zeros(4 * c.K)
    r = P.r.ravel()
    for k in range(c.K):
        blk = slice(4 * k, 4 * k + 4)
        z = x[blk] - r[blk] * (A[blk] @ x + b_ref[blk])
        xn = max(0.0, z[0])
        x[blk] = np.concatenate([[xn], PB.prox_sphere(z[1], z[2], z[3], P.mu[k], xn)[0]])
    PB.sweep_gs(P, COL.serial_groups(c.K), prm)
    assert np.allclose(P.x.ravel(), x, atol=1e-12)


def test_single_contact_converges_to_closed_form(oracle_bank32):
    # a box sunk into the ground with one contact: at convergence x_n = max(0, -b_n / A_nn)
    sc = scene_with_contacts(oracle_bank32)
    c = CD.detect(sc)
    c1 = c.take(np.array([0]))
    Winv, f_ext, du = PB.body_mass_blocks(sc)
    P = PB.assemble(sc, c1, Winv, du, 0.01)
    prm = dict(sc.params, max_iterations=500, absolute_tolerance=0.0, relative_tolerance=0.0)
    PB.solve(P, COL.serial_groups(1), prm)
    A = np.einsum("ij,jl->il", P.JA[0], P.WJA[0]) + np.einsum("ij,jl->il", P.JB[0], P.WJB[0])
    xn = max(0.0, -P.b[0, 0] / A[0, 0]) if np.allclose(P.sol[0, 1:], 0) else P.sol[0, 0]
    assert P.sol[0, 0] >= 0
    assert np.linalg.norm(P.sol[0, 1:]) <= P.mu[0] * P.sol[0, 0] + 1e-12
    if np.allclose(P.sol[0, 1:], 0):
        assert P.sol[0, 0] == pytest.approx(xn, rel=1e-9, abs=1e-12)


def test_solve_invariants_and_driver(oracle_bank32):
    sc = scene_with_contacts(oracle_bank32, seed=2)
    c = CD.detect(sc)
    Winv, f_ext, du = PB.body_mass_blocks(sc)
    P = PB.assemble(sc, c, Winv, du, 0.01)
    prm = dict(sc.params, max_iterations=60, absolute_tolerance=0.0, relative_tolerance=0.0)
    groups = COL.groups_from_colors(COL.color(c.ia, c.ib, sc.fixed, 3, 0))
    res = PB.solve(P, groups, prm)
    assert res["iterations"] == 60 and res["exitcode"] == 0
    assert np.all(P.sol[:, 0] >= 0)
    assert np.all(np.linalg.norm(P.sol[:, 1:], axis=1) <= P.mu * P.sol[:, 0] + 1e-12)
    # the logged merit is the max |x - sol| of accepted sweeps (last merit on rejects)
    r = res["residuals"]
    assert np.all(np.isfinite(r)) and np.all(r >= 0)
    # a large initial r-factor forces rejections (rollback + r *= 0.7)
    P2 = PB.assemble(sc, c, Winv, du, 0.01)
    P2.r = P2.r * 100
    res2 = PB.solve(P2, groups, prm)
    assert res2["rejects"] > 0


def test_absolute_and_relative_stopping():
    # gauss_seidel_solver.py:78-97 with a one-contact problem built by hand
    P = PB.Problem()
    P.K, P.N = 1, 2
    P.ia, P.ib = np.array([0]), np.array([1])
    P.JA = np.zeros((1, 4, 6)); P.JB = np.zeros((1, 4, 6))
    P.JB[0, 0, 1] = 1.0
    W = np.eye(6)
    P.WJA = np.zeros((1, 6, 4)); P.WJB = np.einsum("ij,kj->ik", W, P.JB[0])[None]
    P.b = np.array([[-1.0, 0, 0, 0]])
    P.r = np.array([[1.0, 1, 1, 1]])
    P.mu = np.array([0.5]); P.mu3 = np.array([[0.5] * 3])
    P.x = np.zeros((1, 4)); P.sol = np.zeros((1, 4))
    prm = dict(PARAM_DEFAULTS, max_iterations=10, absolute_tolerance=1e-3, relative_tolerance=0.0)
    res = PB.solve(P, [np.array([0])], prm)
    # sweep 0: x_n = 1 (merit 1); sweep 1: x_n stays 1 (merit 0 < 1e-3) -> exit code 1 at iteration 1
    assert res["exitcode"] == 1 and res["iterations"] == 1
    assert P.sol[0, 0] == pytest.approx(1.0)


# ------------------------------------------------------------ natural-map residual
def one_contact(b, mu, r_scale=1.0):
    """One contact between a fixed body 0 and a free body 1 (mass 1, inertia
    2 I, arm 0) with normal +y: A = J W J^T = diag(1, 1, 1, 1/2)."""
    n = np.array([0.0, 1.0, 0.0])
    s, t, _ = Q.make_orthonormal_vectors(n)
    P = PB.Problem()
    P.K, P.N = 1, 2
    P.ia, P.ib = np.array([0]), np.array([1])
    P.JA = np.zeros((1, 4, 6))
    P.JB = np.zeros((1, 4, 6))
    P.JB[0, 0, :3], P.JB[0, 1, :3], P.JB[0, 2, :3], P.JB[0, 3, 3:] = n, s, t, n
    W = np.diag([1.0, 1.0, 1.0, 0.5, 0.5, 0.5])
    P.WJA = np.zeros((1, 6, 4))
    P.WJB = (W @ P.JB[0].T)[None]
    P.b = np.array([b], dtype=float)
    P.r0 = np.array([[1.0, 1.0, 1.0, 2.0]])
    P.r = P.r0 * r_scale          # the solve's current r-factors (not used by the residual)
    P.mu = np.array([mu])
    P.mu3 = np.array([[mu] * 3])
    return P


def test_natural_residual_zero_at_closed_form_fixed_point():
    # x_n = -b_n / A_nn = 1.5 with zero friction rows is a fixed point of the prox map
    P = one_contact([-1.5, 0.0, 0.0, 0.0], 0.2, r_scale=3.0)
    assert PB.natural_residual(P, PARAM_DEFAULTS, x=np.array([[1.5, 0.0, 0.0, 0.0]])) == 0.0


def test_natural_residual_hand_value():
    # x = (2, 0.9, 0, 0), b = (-1.5, 0.5, 0, 0), mu = 0.2, r0 = 1/diag(A):
    # z = x - r0 (A x + b) = (1.5, -0.5, 0, 0); prox: y_n = 1.5, friction ball of
    # radius 0.3 -> y_s = -0.3; |x - y|_inf = 1.2 (row s); / max(1, |x|_inf = 2) = 0.6.
    # (with the solve's r = 3 r0 instead: 0.75; without the cone: 0.7; unscaled: 1.2)
    P = one_contact([-1.5, 0.5, 0.0, 0.0], 0.2, r_scale=3.0)
    x = np.array([[2.0, 0.9, 0.0, 0.0]])
    assert PB.natural_residual(P, PARAM_DEFAULTS, x=x) == pytest.approx(0.6, abs=1e-15)
    # the torsion row has A = 1/2, so r0 = 2: z_tau = 0.4 - 2 (0.2 + 0.1) = -0.2, inside the cone
    P = one_contact([-1.5, 0.0, 0.0, 0.1], 0.5)
    x = np.array([[2.0, 0.0, 0.0, 0.4]])
    assert PB.natural_residual(P, PARAM_DEFAULTS, x=x) == pytest.approx(0.3, abs=1e-15)


def test_natural_residual_equals_dense_prox_map():
    # two contacts among three free bodies with random Jacobian blocks and SPD
    # inverse masses: the residual equals the prox fixed-point map on the dense
    # A = J W J^T, and vanishes at the converged Gauss-Seidel solution
    rng = np.random.default_rng(11)
    P = PB.Problem()
    P.K, P.N = 2, 3
    P.ia, P.ib = np.array([0, 1]), np.array([1, 2])
    P.JA, P.JB = rng.normal(size=(2, 4, 6)), rng.normal(size=(2, 4, 6))
    Ws = []
    for _ in range(3):
        M = rng.normal(size=(6, 6))
        Ws.append(M @ M.T + 6 * np.eye(6))
    P.WJA = np.einsum("kij,klj->kil", np.array(Ws)[P.ia], P.JA)
    P.WJB = np.einsum("kij,klj->kil", np.array(Ws)[P.ib], P.JB)
    J = np.zeros((8, 18))
    for k in range(2):
        J[4 * k:4 * k + 4, 6 * P.ia[k]:6 * P.ia[k] + 6] = P.JA[k]
        J[4 * k:4 * k + 4, 6 * P.ib[k]:6 * P.ib[k] + 6] = P.JB[k]
    Wd = np.zeros((18, 18))
    for i in range(3):
        Wd[6 * i:6 * i + 6, 6 * i:6 * i + 6] = Ws[i]
    A = J @ Wd @ J.T
    P.b = rng.normal(size=(2, 4))
    P.b[:, 0] = -np.abs(P.b[:, 0]) - 1.0
    P.r0 = (1.0 / np.diag(A)).reshape(2, 4)
    P.r = P.r0.copy()
    P.mu = np.array([0.3, 0.7])
    P.mu3 = np.repeat(P.mu[:, None], 3, axis=1)

    def dense(x):
        z = x.ravel() - P.r0.ravel() * (A @ x.ravel() + P.b.ravel())
        y = np.empty(8)
        for k in range(2):
            yn = max(0.0, z[4 * k])
            zf = z[4 * k + 1:4 * k + 4]
            nf = np.linalg.norm(zf)
            y[4 * k] = yn
            y[4 * k + 1:4 * k + 4] = zf if nf <= P.mu[k] * yn else zf * (P.mu[k] * yn / nf)
        return np.abs(x.ravel() - y).max() / max(1.0, np.abs(x).max())

    for _ in range(5):
        x = rng.normal(size=(2, 4))
        x[:, 0] = np.abs(x[:, 0])
        assert PB.natural_residual(P, PARAM_DEFAULTS, x=x) == pytest.approx(dense(x), rel=1e-12, abs=1e-15)
    P.x, P.sol = np.zeros((2, 4)), np.zeros((2, 4))
    prm = dict(PARAM_DEFAULTS, max_iterations=2000, absolute_tolerance=0.0, relative_tolerance=0.0,
               too_small_merit_change=0.0)
    PB.solve(P, COL.serial_groups(2), prm)
    assert PB.natural_residual(P, prm) < 1e-10 and dense(P.sol) < 1e-10


def test_faithful_sweep_equals_blocked_sweep(oracle_bank32):
    # contacts.py:66-107 on the reference's scipy.sparse structures (a full
    # J.dot(w) and WJT.tocsr() per contact) against the blocked sweep, serial
    # and coloured orders, three sweeps each: equal up to summation order
    from oracle import faithful
    sc = scene_with_contacts(oracle_bank32, seed=3)
    c = CD.detect(sc)
    assert c.K > 4
    Winv, f_ext, du = PB.body_mass_blocks(sc)
    prm = dict(sc.params)
    for groups in (COL.serial_groups(c.K), COL.groups_from_colors(COL.color(c.ia, c.ib, sc.fixed, 5, 0))):
        P = PB.assemble(sc, c, Winv, du, 0.01)
        Jm, W, WJT = faithful.sparse_problem(P, Winv)
        assert np.allclose(Jm.toarray() @ W.toarray(), WJT.toarray().T, atol=1e-14)
        x = np.zeros(4 * c.K)
        order = np.concatenate(groups)
        for _ in range(3):
            PB.sweep_gs(P, groups, prm)
            assert faithful.sweep(Jm, WJT, x, P.r.ravel(), P.b.ravel(), P.mu, order) == c.K
            assert np.allclose(x, P.x.ravel(), rtol=1e-10, atol=1e-12)
