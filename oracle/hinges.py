"""Hinge-joint problem, restated in FP64 (test infrastructure: the checker of
the GPU hinge solve; nothing shipped imports it).

* ``from_engine``   Engine.hinges in creation order (hinge.idx) with the
                    sockets set by api.set_hinge (api.py:86-156, types.py:329-353):
                    arms and joint axes in the parent's and the child's body
                    frames.
* ``assemble``      Hinges.initialize (constraints/hinges.py:28-48) with
                    _compute_jacobian_matrix (:103-185) and _compute_error_vector
                    (:187-218).  Five rows per hinge: rows 0-2 the anchor
                    velocity difference (parent [I | -[arm_p]x], child
                    [-I | [arm_c]x]), rows 3-4 the relative angular velocity
                    along t_p and b_p, the two directions normal to the
                    parent's joint axis (make_orthonormal_vectors).  J and
                    W J^T are kept as per-hinge dense blocks (5 x 6 per body),
                    which is the reference's csr matrix; sums over a J row run
                    in the csr column order (the body of lower index first).
* ``sweep``         Hinges.sweep (:59-101): w = W J^T x recomputed at the start,
                    then serial Gauss-Seidel over the hinges with the identity
                    prox, x_b <- x_b - r_b (J w + b)_b, w += W J^T_b delta.
* ``body_w``        W J^T x per body (finalize_solution, :50-57).

Pinned by tests/test_oracle_hinges.py: J u against the finite-difference
derivative of the anchor separation and against rotations about the joint
axis (rows 3-4 vanish); the error vector by placed configurations (anchors
apart by d -> rate d, axes tilted about t_p -> a b_p row only); the converged
solve against numpy.linalg.solve of the dense J W J^T x = -b built from the
same blocks; a step's post-velocities against J u_new = 0 (a pendulum's
anchor stays at rest); api.set_hinge's sockets against the joint frame they
came from in world, body and model mode.
"""
import numpy as np

from . import quat as Q


class HingeSet:
    pass


def from_engine(engine, body_index):
    """Hinge records of ``engine`` (Engine.hinges, creation order); body_index
    maps a RigidBody name to its scene index."""
    hs = sorted(getattr(engine, "hinges", {}).values(), key=lambda h: h.idx)
    H = HingeSet()
    H.K = len(hs)
    H.parent = np.array([body_index[h.parent.name] for h in hs], dtype=np.int64)
    H.child = np.array([body_index[h.child.name] for h in hs], dtype=np.int64)
    H.arm_p = np.array([h.arm_p for h in hs], dtype=np.float64).reshape(-1, 3)
    H.arm_c = np.array([h.arm_c for h in hs], dtype=np.float64).reshape(-1, 3)
    H.axis_p = np.array([h.axis_p for h in hs], dtype=np.float64).reshape(-1, 3)
    H.axis_c = np.array([h.axis_c for h in hs], dtype=np.float64).reshape(-1, 3)
    return H


def star(a):
    return np.array([[0.0, -a[2], a[1]], [a[2], 0.0, -a[0]], [-a[1], a[0], 0.0]])


def _rowdot(P, k, JP, JC, vp, vc):
    """One J row times a body-space vector, summed term by term in the csr
    column order (the body of lower index first)."""
    terms = [(JP, vp), (JC, vc)] if P.parent[k] < P.child[k] else [(JC, vc), (JP, vp)]
    s = 0.0
    for J, v in terms:
        for j in range(6):
            s += float(J[j]) * float(v[j])
    return s


def _wjt(Winv, J):
    """W J^T for one body: (6, 5), entry (e, i) = sum_j W[e, j] J[i, j] with
    the inverse-mass block's nonzeros (1/m on the diagonal, the inverse
    inertia's row)."""
    out = np.zeros((6, 5))
    for i in range(5):
        for e in range(3):
            out[e, i] = Winv[e, e] * J[i, e]
        for e in range(3, 6):
            out[e, i] = (Winv[e, 3] * J[i, 3] + Winv[e, 4] * J[i, 4]) + Winv[e, 5] * J[i, 5]
    return out


def assemble(scene, H, Winv, du_ext, dt):
    """Hinges.initialize: J, W J^T, g, b = J u + J (dt du_ext) + g, diag, r = 1/diag, x = sol = 0."""
    K = H.K
    prm = scene.params
    P = HingeSet()
    for name, val in vars(H).items():
        setattr(P, name, val)
    P.N = scene.N
    P.JP = np.zeros((K, 5, 6))
    P.JC = np.zeros((K, 5, 6))
    g = np.zeros((K, 5))
    for k in range(K):
        p, c = H.parent[k], H.child[k]
        arm_p = Q.rotate(scene.q[p], H.arm_p[k])
        arm_c = Q.rotate(scene.q[c], H.arm_c[k])
        n_p = Q.rotate(scene.q[p], H.axis_p[k])
        t_p, b_p, n_p = Q.make_orthonormal_vectors(n_p)
        P.JP[k, 0:3, 0:3] = np.eye(3)
        P.JP[k, 0:3, 3:6] = -star(arm_p)
        P.JP[k, 3, 3:6] = t_p
        P.JP[k, 4, 3:6] = b_p
        P.JC[k, 0:3, 0:3] = -np.eye(3)
        P.JC[k, 0:3, 3:6] = star(arm_c)
        P.JC[k, 3, 3:6] = -t_p
        P.JC[k, 4, 3:6] = -b_p
        if prm["use_pre_stabilization"]:
            r_p = arm_p + scene.r[p]
            r_c = arm_c + scene.r[c]
            n_c = Q.rotate(scene.q[c], H.axis_c[k])
            u = np.cross(n_p, n_c)
            g[k, 0:3] = r_c - r_p
            g[k, 3] = t_p.dot(u)
            g[k, 4] = b_p.dot(u)
    if prm["use_pre_stabilization"]:
        g = g * (prm["gap_reduction"] / dt)
    P.g = g
    P.WP = np.array([_wjt(Winv[H.parent[k]], P.JP[k]) for k in range(K)]).reshape(K, 6, 5)
    P.WC = np.array([_wjt(Winv[H.child[k]], P.JC[k]) for k in range(K)]).reshape(K, 6, 5)
    u = np.concatenate([scene.v, scene.w], axis=1)
    dtu = dt * du_ext
    P.b = np.zeros((K, 5))
    P.diag = np.zeros((K, 5))
    for k in range(K):
        p, c = H.parent[k], H.child[k]
        for i in range(5):
            ju = _rowdot(P, k, P.JP[k, i], P.JC[k, i], u[p], u[c])
            jd = _rowdot(P, k, P.JP[k, i], P.JC[k, i], dtu[p], dtu[c])
            P.b[k, i] = (ju + jd) + g[k, i]
            dg = _rowdot(P, k, P.JP[k, i], P.JC[k, i], P.WP[k, :, i], P.WC[k, :, i])
            P.diag[k, i] = 1.0 if dg == 0 else dg
    P.r = 1.0 / P.diag
    P.x = np.zeros((K, 5))
    P.sol = np.zeros((K, 5))
    P.rejects = 0
    return P


def body_w(P, x):
    """W J^T x per body (N, 6): one running sum per body entry over the
    hinges touching the body, in hinge order, rows 0..4 within a hinge."""
    w = np.zeros((P.N, 6))
    for k in range(P.K):
        for W, b in ((P.WP[k], P.parent[k]), (P.WC[k], P.child[k])):
            for e in range(6):
                s = w[b, e]
                for i in range(5):
                    s += W[e, i] * x[k, i]
                w[b, e] = s
    return w


def sweep(P):
    """Hinges.sweep (hinges.py:59-101)."""
    w = body_w(P, P.x)
    for k in range(P.K):
        p, c = P.parent[k], P.child[k]
        old = P.x[k].copy()
        new = old.copy()
        for i in range(5):
            jw = _rowdot(P, k, P.JP[k, i], P.JC[k, i], w[p], w[c])
            new[i] = old[i] - P.r[k, i] * (jw + P.b[k, i])
        P.x[k] = new
        delta = new - old
        for W, b in ((P.WP[k], p), (P.WC[k], c)):
            for e in range(6):
                t = 0.0
                for i in range(5):
                    t += W[e, i] * delta[i]
                w[b, e] += t


class Solvable:
    """The hinge problem as seen by the multi-problem solve driver."""

    def __init__(self, P):
        self.P = P

    def sweep(self):
        sweep(self.P)

    def __getattr__(self, name):
        return getattr(self.__dict__["P"], name)

    def __setattr__(self, name, value):
        if name == "P":
            self.__dict__["P"] = value
        else:
            setattr(self.__dict__["P"], name, value)
