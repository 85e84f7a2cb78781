"""Contact problem, prox operators and the r-factor solve, restated in FP64.

* ``assemble``           Contacts.initialize (constraints/contacts.py:31-55) with
                         _compute_jacobian_matrix (:110-186), _compute_error_vector
                         (:189-220), _get_friction_coefficient_vector (:238-256),
                         _get_restitution_vector (:259-287).  J and W J^T are kept
                         as per-contact dense blocks (4x6 for each of the two
                         bodies), which is the same matrix as the reference's csr.
* ``prox_sphere``        prox_operators.py:14-34.
* ``prox_ellipsoid``     proximal_contact/proximal_operators.py:42-139.
* ``sweep_gs``           Contacts.sweep (contacts.py:66-107): serial Gauss-Seidel
                         over the contacts in a given order, with w = W J^T x
                         recomputed at the start of the sweep.  ``order`` groups
                         contacts into colours; contacts inside one colour share no
                         free body, so updating them together is the same
                         arithmetic as updating them one after another.
* ``sweep_jacobi``       proximal_contact/jacobi_solver.py:6-57 (x_n from z of the
                         old iterate; the friction radius uses the old x_n, or the
                         new one in the hybrid scheme).
* ``solve``              gauss_seidel_solver.py:14-101 (single problem): merit
                         ||x - sol||_inf, reject -> r *= nu_reduce and x = sol,
                         accept -> sol = x and r *= nu_increase when the merit
                         dropped by less than too_small_merit_change; absolute and
                         relative stopping tests; exit codes 0/1/2.
* ``solve_multi``        the same driver over several problems (contacts, then hinges),
                         stopping on the largest of their merits.
* ``natural_residual``   ||x - prox(x - r (A x + b))||_inf / max(1, ||x||_inf).

Pinned by tests/test_oracle_problem.py: prox operators by closed forms (inside
point -> identity, outside -> radial scaling, x_n <= 0 -> origin, isotropic
ellipsoid == sphere, axis-aligned z -> clamp) and by an independent root solve
of sum a_i^2 z_i^2 / (a_i^2 + t)^2 = 1; the sweep by the dense-matrix form
z = x - r (J W J^T x + b) on random problems and by the single-contact closed
form x_n = max(0, -b_n / A_nn); the block Jacobian by the reference's scalar
assembly loop (contacts.py:130-185) restated literally in the test; the solve
driver by hand-computed accept/reject sequences; solve_multi by the same
sequences (one problem: identical to ``solve``; two scripted problems: the
residual is the larger contribution, a rejected problem's last merit).
"""
import numpy as np

from . import quat as Q


class Problem:
    pass


def body_mass_blocks(scene):
    """Per-body inverse-mass blocks and external-force velocity change.

    state_storage.py:108-185: W_b = blockdiag(1/m I3, R diag(1/I) R^T), zero for
    fixed bodies (ValueError for a free body without mass); f_ext = (sum F_i,
    -w x (I_w w) + sum T_i); delta_u_ext = W f_ext.
    """
    N = scene.N
    Winv = np.zeros((N, 6, 6))
    f_ext = np.zeros((N, 6))
    for k in range(N):
        if scene.fixed[k]:
            continue
        if scene.mass[k] == 0:
            raise ValueError("Body {} has no mass.".format(k))
        R = Q.to_matrix(scene.q[k])
        I_wcs = Q.update_inertia_tensor(R, scene.inertia[k])
        w = scene.w[k]
        v = scene.v[k]
        F = np.zeros(3)
        T = -np.cross(w, np.dot(I_wcs, w))
        for f in scene.forces[k]:
            if f[0] == "gravity":
                F = F + (-scene.mass[k] * f[1] * f[2])
            else:
                F = F + (-v * f[1])
                T = T + (-w * f[2])
        f_ext[k, :3] = F
        f_ext[k, 3:] = T
        m = 1.0 / scene.mass[k]
        Winv[k, :3, :3] = np.diag([m, m, m])
        Winv[k, 3:, 3:] = Q.update_inertia_tensor(R, 1.0 / scene.inertia[k])
    du_ext = np.einsum("nij,nj->ni", Winv, f_ext)
    return Winv, f_ext, du_ext


def assemble(scene, contacts, Winv, du_ext, dt):
    """Contacts.initialize; returns a Problem with per-contact blocks."""
    K = contacts.K
    prm = scene.params
    P = Problem()
    P.K = K
    P.ia, P.ib = contacts.ia.copy(), contacts.ib.copy()
    JA = np.zeros((K, 4, 6))
    JB = np.zeros((K, 4, 6))
    frames = np.zeros((K, 3, 3))
    for k in range(K):
        rA = contacts.p[k] - scene.r[P.ia[k]]
        rB = contacts.p[k] - scene.r[P.ib[k]]
        vs, vt, vn = Q.make_orthonormal_vectors(contacts.n[k])
        frames[k] = (vs, vt, vn)
        JA[k, :, :3] = -np.array([vn, vs, vt, np.zeros(3)])
        JA[k, :, 3:] = -np.array([np.cross(rA, vn), np.cross(rA, vs), np.cross(rA, vt), vn])
        JB[k, :, :3] = np.array([vn, vs, vt, np.zeros(3)])
        JB[k, :, 3:] = np.array([np.cross(rB, vn), np.cross(rB, vs), np.cross(rB, vt), vn])
    P.JA, P.JB, P.frames = JA, JB, frames
    # W J^T blocks (6x4 per body side).
    P.WJA = np.einsum("kij,klj->kil", Winv[P.ia], JA)
    P.WJB = np.einsum("kij,klj->kil", Winv[P.ib], JB)
    u = np.concatenate([scene.v, scene.w], axis=1)
    v = np.einsum("kij,kj->ki", JA, u[P.ia]) + np.einsum("kij,kj->ki", JB, u[P.ib])
    # pre-stabilisation gap vector, contacts.py:189-220
    g = np.zeros((K, 4))
    if prm["use_pre_stabilization"]:
        rate = prm["gap_reduction"] / dt
        upper = -prm["max_gap_value"] / dt
        lower = -prm["min_gap_value"]
        for k in range(K):
            if contacts.gap[k] < lower and v[k, 0] <= 0.0:
                g[k, 0] = max(upper, rate * contacts.gap[k])
    e = np.zeros((K, 4))
    if prm["use_bounce"]:
        e[:, 0] = scene.eps[scene.material[P.ia], scene.material[P.ib]]
    mu3 = scene.mu[scene.material[P.ia], scene.material[P.ib]]   # (K, 3)
    P.mu = mu3[:, 0].copy()           # contacts.py:255: isotropic mu[0]
    P.mu3 = mu3.copy()
    dtu = dt * du_ext
    P.b = np.multiply(1 + e, v) + (np.einsum("kij,kj->ki", JA, dtu[P.ia])
                                   + np.einsum("kij,kj->ki", JB, dtu[P.ib])) + g
    diag = np.einsum("kij,kji->ki", JA, P.WJA) + np.einsum("kij,kji->ki", JB, P.WJB)
    diag[diag == 0] = 1
    P.diag = diag
    P.r = 1.0 / diag
    P.r0 = P.r.copy()
    P.x = np.zeros((K, 4))
    P.sol = np.zeros((K, 4))
    P.N = scene.N
    P.gap = contacts.gap.copy()
    return P


def prox_sphere(z_s, z_t, z_tau, mu, x_n):
    """prox_operators.py:14-34 (vectorised over arrays)."""
    z_s, z_t, z_tau, mu, x_n = np.broadcast_arrays(*(np.atleast_1d(np.asarray(a, dtype=np.float64))
                                                      for a in (z_s, z_t, z_tau, mu, x_n)))
    radius = mu * x_n
    sq = z_s * z_s + z_t * z_t + z_tau * z_tau
    out = np.stack([z_s, z_t, z_tau], axis=-1).astype(np.float64).copy()
    with np.errstate(divide="ignore", invalid="ignore"):
        scale = radius / np.sqrt(sq)
    outside = sq > radius * radius
    out[outside] = out[outside] * scale[outside][:, None]
    out[x_n <= 0.0] = 0.0
    return out


def prox_ellipsoid(z, mu3, x_n, prm):
    """proximal_operators.py:42-139 for one contact; z = (z_s, z_t, z_tau)."""
    z_s, z_t, z_tau = (float(v) for v in z)
    mu_s, mu_t, mu_tau = (float(v) for v in mu3)
    if x_n <= 0.0:
        return np.zeros(3)
    if mu_s == 0 and mu_t == 0 and mu_tau == 0:
        return np.zeros(3)
    if mu_s == mu_t and mu_s == mu_tau:
        return prox_sphere([z_s], [z_t], [z_tau], [mu_s], [x_n])[0]
    a, b, c = mu_s * x_n, mu_t * x_n, mu_tau * x_n
    scale = 1 / max(1, a, b, c, abs(z_s), abs(z_t), abs(z_tau))
    sa, sb, sc, sx, sy, sz = scale * a, scale * b, scale * c, scale * z_s, scale * z_t, scale * z_tau
    saa, sbb, scc, sxx, syy, szz = sa * sa, sb * sb, sc * sc, sx * sx, sy * sy, sz * sz
    f0 = (sxx / saa) + (syy / sbb) + (szz / scc) - 1
    if f0 < prm["ellipsoid_tolerance"]:
        return np.array([z_s, z_t, z_tau])
    t0 = 0.0
    t1 = max(sa, sb, sc) * np.linalg.norm([sx, sy, sz])

    def gfun(t):
        return ((saa * sxx) / ((saa + t) ** 2) + (sbb * syy) / ((sbb + t) ** 2)
                + (scc * szz) / ((scc + t) ** 2) - 1)
    g1 = gfun(t1)
    while g1 > 0:
        t1 *= prm["ellipsoid_expansion"]
        g1 = gfun(t1)
    tk = (t0 + t1) * 0.5
    for _ in range(int(prm["ellipsoid_max_iterations"])):
        if abs(t1 - t0) < prm["ellipsoid_tolerance"]:
            break
        gk = gfun(tk)
        if abs(gk) < prm["ellipsoid_tolerance"]:
            break
        if gk > 0:
            t0 = tk
        else:
            t1 = tk
        tk = (t0 + t1) * 0.5
    tk /= scale * scale
    return np.array([(a * a * z_s) / (a * a + tk), (b * b * z_t) / (b * b + tk),
                     (c * c * z_tau) / (c * c + tk)])


def _friction(P, idx, z, xn, prm):
    if prm["friction_model"] == "anisotropic":
        return np.array([prox_ellipsoid(z[i], P.mu3[k], xn[i], prm) for i, k in enumerate(idx)]
                        ).reshape(-1, 3)
    return prox_sphere(z[:, 0], z[:, 1], z[:, 2], P.mu[idx], xn)


def body_w(P, x):
    """w = W J^T x, per body (N, 6)."""
    w = np.zeros((P.N, 6))
    np.add.at(w, P.ia, np.einsum("kil,kl->ki", P.WJA, x))
    np.add.at(w, P.ib, np.einsum("kil,kl->ki", P.WJB, x))
    return w


def sweep_gs(P, groups, prm, r_eff=None):
    """One Gauss-Seidel sweep in the order given by ``groups`` (list of index
    arrays, each a set of contacts sharing no free body)."""
    r = P.r if r_eff is None else r_eff
    w = body_w(P, P.x)
    for idx in groups:
        Jw = np.einsum("kij,kj->ki", P.JA[idx], w[P.ia[idx]]) + \
            np.einsum("kij,kj->ki", P.JB[idx], w[P.ib[idx]])
        x_old = P.x[idx].copy()
        z = x_old - np.multiply(r[idx], Jw + P.b[idx])
        xn = np.maximum(0.0, z[:, 0])
        xf = _friction(P, idx, z[:, 1:], xn, prm)
        x_new = np.concatenate([xn[:, None], xf], axis=1)
        P.x[idx] = x_new
        delta = x_new - x_old
        np.add.at(w, P.ia[idx], np.einsum("kil,kl->ki", P.WJA[idx], delta))
        np.add.at(w, P.ib[idx], np.einsum("kil,kl->ki", P.WJB[idx], delta))


def sweep_jacobi(P, prm, hybrid=False):
    """jacobi_solver.py:6-57: every contact from the same w = W J^T x."""
    w = body_w(P, P.x)
    Jw = np.einsum("kij,kj->ki", P.JA, w[P.ia]) + np.einsum("kij,kj->ki", P.JB, w[P.ib])
    z = P.x - np.multiply(P.r, Jw + P.b)
    xn = np.maximum(0.0, z[:, 0])
    xn_f = xn if hybrid else P.x[:, 0]
    xf = _friction(P, np.arange(P.K), z[:, 1:], xn_f, prm)
    P.x = np.concatenate([xn[:, None], xf], axis=1)


def solve(P, groups, prm, scheme=None):
    """gauss_seidel_solver.py:14-101 for the single contact problem.

    Returns dict(residuals, iterations, exitcode, rejects).
    """
    scheme = scheme or prm["proximal_solver"]
    nu_reduce, nu_increase = prm["nu_reduce"], prm["nu_increase"]
    too_small = prm["too_small_merit_change"]
    max_it = int(prm["max_iterations"])
    residuals = np.full(max_it, np.inf)
    last_merit = np.inf
    global_last_merit = np.inf
    rejects = 0
    for it in range(max_it):
        if scheme in ("gauss_seidel", "parallel_gauss_seidel"):
            sweep_gs(P, groups, prm)
        elif scheme == "parallel_jacobi":
            sweep_jacobi(P, prm, hybrid=False)
        elif scheme in ("parallel_jacobi_hybrid", "parallel_jacboi_hybrid"):
            sweep_jacobi(P, prm, hybrid=True)
        else:
            raise NotImplementedError(scheme)
        merit = np.linalg.norm((P.x - P.sol).ravel(), np.inf) if P.K else 0.0
        if merit > last_merit:
            P.r = P.r * nu_reduce
            P.x = P.sol.copy()
            global_merit = max(0.0, last_merit)
            rejects += 1
        else:
            P.sol = P.x.copy()
            if last_merit - merit < too_small:
                P.r = P.r * nu_increase
            last_merit = merit
            global_merit = max(0.0, merit)
        residuals[it] = global_merit
        if global_merit < prm["absolute_tolerance"]:
            return dict(residuals=residuals, iterations=it, exitcode=1, rejects=rejects)
        if np.abs(global_merit - global_last_merit) < prm["relative_tolerance"] * global_last_merit:
            return dict(residuals=residuals, iterations=it, exitcode=2, rejects=rejects)
        global_last_merit = global_merit
    return dict(residuals=residuals, iterations=max_it, exitcode=0, rejects=rejects)


class ContactSolvable:
    """The contact problem as seen by the multi-problem solve driver."""

    def __init__(self, P, groups, prm, scheme=None):
        self.__dict__.update(P=P, groups=groups, prm=prm, scheme=scheme or prm["proximal_solver"])

    def sweep(self):
        P, scheme = self.__dict__["P"], self.__dict__["scheme"]
        if scheme in ("gauss_seidel", "parallel_gauss_seidel"):
            sweep_gs(P, self.__dict__["groups"], self.__dict__["prm"])
        elif scheme == "parallel_jacobi":
            sweep_jacobi(P, self.__dict__["prm"], hybrid=False)
        elif scheme in ("parallel_jacobi_hybrid", "parallel_jacboi_hybrid"):
            sweep_jacobi(P, self.__dict__["prm"], hybrid=True)
        else:
            raise NotImplementedError(scheme)

    def __getattr__(self, name):
        return getattr(self.__dict__["P"], name)

    def __setattr__(self, name, value):
        setattr(self.__dict__["P"], name, value)


def solve_multi(problems, prm):
    """gauss_seidel_solver.py:14-101 over several problems (steppers.py:68-83:
    the contact problem first, then the hinges).  Each problem keeps its own
    merit ||x - sol||_inf, accept/reject and r-factors; the residual log and
    the stopping tests use the largest contribution, where a rejected problem
    contributes its last accepted merit.  ``rejects`` counts the first
    problem's rejected sweeps.
    """
    nu_reduce, nu_increase = prm["nu_reduce"], prm["nu_increase"]
    too_small = prm["too_small_merit_change"]
    max_it = int(prm["max_iterations"])
    residuals = np.full(max_it, np.inf)
    last = [np.inf for _ in problems]
    global_last_merit = np.inf
    rejects = 0
    for it in range(max_it):
        global_merit = 0.0
        for i, p in enumerate(problems):
            p.sweep()
            merit = float(np.max(np.abs(p.x - p.sol))) if p.x.size else 0.0
            if merit > last[i]:
                p.r = p.r * nu_reduce
                p.x = p.sol.copy()
                global_merit = max(global_merit, last[i])
                rejects += 1 if i == 0 else 0
            else:
                p.sol = p.x.copy()
                if last[i] - merit < too_small:
                    p.r = p.r * nu_increase
                last[i] = merit
                global_merit = max(global_merit, merit)
        residuals[it] = global_merit
        if global_merit < prm["absolute_tolerance"]:
            return dict(residuals=residuals, iterations=it, exitcode=1, rejects=rejects)
        if np.abs(global_merit - global_last_merit) < prm["relative_tolerance"] * global_last_merit:
            return dict(residuals=residuals, iterations=it, exitcode=2, rejects=rejects)
        global_last_merit = global_merit
    return dict(residuals=residuals, iterations=max_it, exitcode=0, rejects=rejects)


def natural_residual(P, prm, x=None):
    """||x - prox(x - r0 (A x + b))||_inf / max(1, ||x||_inf) for the impulses x
    (default: the accepted solution), with the initial r-factors r0 = 1/diag."""
    x = P.sol if x is None else x
    if P.K == 0:
        return 0.0
    w = body_w(P, x)
    Jw = np.einsum("kij,kj->ki", P.JA, w[P.ia]) + np.einsum("kij,kj->ki", P.JB, w[P.ib])
    z = x - np.multiply(P.r0, Jw + P.b)
    xn = np.maximum(0.0, z[:, 0])
    xf = _friction(P, np.arange(P.K), z[:, 1:], xn, prm)
    y = np.concatenate([xn[:, None], xf], axis=1)
    return float(np.max(np.abs(x - y)) / max(1.0, float(np.max(np.abs(x)))))
