"""One semi-implicit time step, restated from prox_rigid_bodies/steppers.py:47-106.

copy_from_engine (state_storage.py:43-53) -> delta_u = dt W f_ext ->
collision detection -> contact problem + prox solve -> delta_u += W J^T sol ->
velocity_update (state_storage.py:239-248) -> position_update (:251-302:
r += v dt; q <- unit(q + 0.5 dt (0, w) q), free bodies only) -> energies
(:188-237) and the largest penetration (contacts.py:223-236) for the log.

With use_post_stabilization the step ends with ``post_stabilize``
(steppers.py:15-36, 90-91; constraints/post_stabilization.py), pinned by
tests/test_oracle_stepper.py: a sphere sunk into the ground is lifted by
exactly its deepest penetration without turning, and contacts shallower than
min_gap_value leave the scene untouched.

With hinges (Engine.hinges) the contact problem and the hinge problem
(oracle/hinges.py) are solved by one driver, problem.solve_multi
(steppers.py:68-83), and delta_u adds both problems' W J^T sol.
"""
import numpy as np

from . import coloring, collision, hinges, problem
from . import quat as Q


def sweep_groups(scene, contacts):
    prm = scene.params
    scheme = prm["proximal_solver"]
    if scheme == "gauss_seidel":
        return coloring.serial_groups(contacts.K), None
    C = coloring.color(contacts.ia, contacts.ib, scene.fixed, int(prm["coloring_seed"]),
                       int(scene.step_index))
    return coloring.groups_from_colors(C), C


def energies(scene):
    """state_storage.py:188-237 (gravity direction from all engine gravity forces)."""
    kin, pot = 0.0, 0.0
    G = np.zeros(3)
    for g, up in scene.gravity:
        G = G + up * g
    for k in range(scene.N):
        if scene.fixed[k]:
            continue
        R = Q.to_matrix(scene.q[k])
        I_wcs = Q.update_inertia_tensor(R, scene.inertia[k])
        if scene.gravity:
            up = G / np.linalg.norm(G)
            pot += scene.mass[k] * np.linalg.norm(G) * up.dot(scene.r[k])
        w = scene.w[k]
        v = scene.v[k]
        kin += 0.5 * (scene.mass[k] * v.dot(v) + np.dot(w, np.dot(I_wcs, w)))
    return kin, pot


def integrate(scene, delta_u, dt):
    u = np.concatenate([scene.v, scene.w], axis=1) + delta_u
    scene.v = u[:, :3].copy()
    scene.w = u[:, 3:].copy()
    for k in range(scene.N):
        if scene.fixed[k]:
            continue
        scene.r[k] = scene.r[k] + scene.v[k] * dt
        q = scene.q[k] + Q.prod(np.array([0.0, *scene.w[k]]), scene.q[k]) * dt * 0.5
        scene.q[k] = Q.unit(q)


def post_stabilize(scene, P, groups, prm):
    """steppers.py:15-36 with constraints/post_stabilization.py:21-104.

    The contact problem's J, W J^T and diag are kept; the right-hand side is
    the gap g of contacts with g < -min_gap_value on the normal row (zero
    elsewhere); r = 1/diag; the sweep projects x_n onto R+ and sets the
    friction rows to zero (here: mu = 0, whose projections are the origin);
    the solve driver is the contact solve's (gauss_seidel_solver.solve, same
    sweep order); delta_r = W J^T sol is applied by position_update with
    dt = 1 (state_storage.py:251-302).  Returns the solve record, or None
    when no contact is deep enough (the reference returns before solving).
    """
    g = np.where(P.gap < -prm["min_gap_value"], P.gap, 0.0)
    if not g.any():
        return None
    Q2 = problem.Problem()
    for name, val in vars(P).items():
        setattr(Q2, name, val)
    Q2.b = np.zeros((P.K, 4))
    Q2.b[:, 0] = g
    Q2.r = 1.0 / P.diag
    Q2.r0 = Q2.r.copy()
    Q2.x = np.zeros((P.K, 4))
    Q2.sol = np.zeros((P.K, 4))
    Q2.mu = np.zeros_like(P.mu)
    Q2.mu3 = np.zeros_like(P.mu3)
    res = problem.solve(Q2, groups, prm)
    dr = problem.body_w(Q2, Q2.sol)
    for k in range(scene.N):
        if scene.fixed[k]:
            continue
        scene.r[k] = scene.r[k] + dr[k, :3] * 1.0
        q = scene.q[k] + Q.prod(np.array([0.0, *dr[k, 3:]]), scene.q[k]) * 1.0 * 0.5
        scene.q[k] = Q.unit(q)
    res["problem"] = Q2
    return res


def step(scene, dt, margins=None, contacts=None):
    """Advance ``scene`` in place by dt.  Returns a dict with the contacts, the
    solved problem, the colouring and the per-step log.  ``margins`` (dict)
    receives the decision margins of the candidate triangles, keyed by
    (SDF owner, triangle owner, triangle).  ``contacts`` (a collision.Contacts
    in the reference's order) replaces collision detection: the rest of the
    step runs on that contact set."""
    prm = scene.params
    Winv, f_ext, du_ext = problem.body_mass_blocks(scene)
    delta_u = dt * du_ext
    stats = {}
    if contacts is None:
        contacts = collision.detect(scene, stats, margins)
    else:
        stats["number_of_overlaps"] = None
    out = dict(contacts=contacts, problem=None, colors=None, log={})
    log = out["log"]
    log["number_of_overlaps"] = stats["number_of_overlaps"]
    H = scene.hinges if scene.hinges is not None and scene.hinges.K > 0 else None
    if contacts.K > 0 and H is None:
        P = problem.assemble(scene, contacts, Winv, du_ext, dt)
        groups, C = sweep_groups(scene, contacts)
        res = problem.solve(P, groups, prm)
        delta_u = delta_u + problem.body_w(P, P.sol)
        out.update(problem=P, colors=C, groups=groups)
        log.update(residuals=res["residuals"], iterations=res["iterations"],
                   exitcode=res["exitcode"], rejects=res["rejects"])
        log["natural_residual"] = problem.natural_residual(P, prm)
    elif H is not None:
        # steppers.py:68-83: contacts (when there are any), then hinges, one solve
        probs = []
        if contacts.K > 0:
            P = problem.assemble(scene, contacts, Winv, du_ext, dt)
            groups, C = sweep_groups(scene, contacts)
            probs.append(problem.ContactSolvable(P, groups, prm))
            out.update(problem=P, colors=C, groups=groups)
        HP = hinges.assemble(scene, H, Winv, du_ext, dt)
        probs.append(hinges.Solvable(HP))
        res = problem.solve_multi(probs, prm)
        if contacts.K > 0:
            delta_u = delta_u + problem.body_w(P, P.sol)
        delta_u = delta_u + hinges.body_w(HP, HP.sol)
        out["hinges"] = HP
        log.update(residuals=res["residuals"], iterations=res["iterations"],
                   exitcode=res["exitcode"], rejects=res["rejects"])
        if contacts.K > 0:
            log["natural_residual"] = problem.natural_residual(P, prm)
    integrate(scene, delta_u, dt)
    if contacts.K > 0 and prm.get("use_post_stabilization"):
        post = post_stabilize(scene, out["problem"], out["groups"], prm)
        if post is not None:
            log.update(post_stabilization_residuals=post["residuals"],
                       post_stabilization_iterations=post["iterations"],
                       post_stabilization_exitcode=post["exitcode"])
            out["post"] = post
    scene.step_index += 1
    kin, pot = energies(scene)
    log.update(dt=dt, contact_points=contacts.K, kinetic_energy=kin, potential_energy=pot,
               max_penetration=-min(0.0, float(contacts.gap.min()) if contacts.K else 0.0))
    out["delta_u"] = delta_u
    return out


def simulate(scene, T):
    """api.py:676-693: substeps dt = min(T_left, time_step)."""
    if T <= 0:
        raise ValueError("Time must be positive")
    logs = []
    T_left = T
    while T_left:
        dt = min(T_left, scene.params["time_step"])
        logs.append(step(scene, dt)["log"])
        T_left -= dt
    return logs
