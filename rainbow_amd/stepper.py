"""GPU time stepper: the drop-in replacement of SemiImplicitStepper.

The reference plugs its stepper into ``engine.stepper``
(prox_rigid_bodies/api.py:42-53); ``api.simulate`` calls
``stepper.step(dt, engine, profiling_on)`` and ``api.get_log`` reads
``stepper.log`` (api.py:676-705).  ``HipStepper`` has the same two members.
The whole step of SemiImplicitStepper.step (steppers.py:47-106) runs in
librbh on the device; the host only marshals body state in and out and turns
the device log into the reference's profile dict.

Body state lives on the device between the sub-steps of one ``run`` /
``simulate`` call; ``engine.bodies`` is updated when the call returns.
``engine.contact_points`` holds the contacts of the last step as a lazy
sequence of ContactPoint objects (built on access).
"""
import numpy as np

from . import math3 as Q
from . import native
from .types import ContactPoint

_CONTEXTS = {}


def shared_context(device=0, precision="fp32"):
    """One librbh context per (device, precision) for the process."""
    key = (int(device), precision)
    if key not in _CONTEXTS or _CONTEXTS[key].closed:
        _CONTEXTS[key] = native.Context(device, precision)
    return _CONTEXTS[key]


def params_struct(p):
    """types.Parameters -> rbh_params."""
    s = native.Params()
    for name in ("time_step", "envelope", "absolute_tolerance", "relative_tolerance", "nu_reduce", "nu_increase",
                 "too_small_merit_change", "contact_optimization_tolerance", "gap_reduction", "min_gap_value",
                 "max_gap_value", "ellipsoid_tolerance", "ellipsoid_expansion"):
        setattr(s, name, float(getattr(p, name)))
    s.max_iterations = int(p.max_iterations)
    s.contact_optimization_max_iterations = int(p.contact_optimization_max_iterations)
    s.ellipsoid_max_iterations = int(p.ellipsoid_max_iterations)
    s.use_bounce = int(bool(p.use_bounce))
    s.use_pre_stabilization = int(bool(p.use_pre_stabilization))
    s.use_post_stabilization = int(bool(getattr(p, "use_post_stabilization", False)))
    scheme = getattr(p, "proximal_solver", "parallel_gauss_seidel")
    if scheme not in native.SCHEMES:
        raise NotImplementedError(
            "Unknown solver scheme: {}, you can use one of 'gauss_seidel', 'parallel_gauss_seidel', "
            "'parallel_jacobi', 'parallel_jacboi_hybrid'.".format(scheme))
    s.proximal_solver = native.SCHEMES[scheme]
    s.friction_model = native.FRICTION[getattr(p, "friction_model", "isotropic")]
    s.coloring_seed = int(getattr(p, "coloring_seed", 0)) & ((1 << 64) - 1)
    return s


def params_key(p):
    return tuple(getattr(p, k, None) for k in (
        "time_step", "max_iterations", "use_bounce", "use_pre_stabilization", "use_post_stabilization",
        "gap_reduction", "min_gap_value", "max_gap_value", "absolute_tolerance", "relative_tolerance",
        "ellipsoid_max_iterations", "ellipsoid_expansion", "ellipsoid_tolerance", "nu_reduce", "nu_increase",
        "too_small_merit_change", "contact_optimization_max_iterations", "contact_optimization_tolerance",
        "envelope", "proximal_solver", "friction_model", "coloring_seed"))


def material_table(engine, bodies):
    """Materials in first-use order and the (mu0, mu1, mu2, epsilon) table of
    SurfacesInteractionLibrary.get_interaction (types.py:42-53)."""
    mats = ["default"]
    for b in bodies:
        if b.material not in mats:
            mats.append(b.material)
    M = len(mats)
    tab = np.zeros((M, M, 4))
    for i, a in enumerate(mats):
        for j, c in enumerate(mats):
            beh = engine.surfaces_interactions.get_interaction(a, c)
            mu = np.asarray(beh.mu, dtype=np.float64).ravel()
            tab[i, j, :3] = mu[:3]
            tab[i, j, 3] = beh.epsilon
    return mats, tab


def check_masses(bodies):
    """state_storage.py:169-171: the step refuses a free body without mass."""
    for b in bodies:
        if not b.is_fixed and b.mass == 0:
            raise ValueError("Body {} has no mass.".format(b.name))


def engine_desc(engine, ctx):
    """rbh_scene_desc of an engine (state_storage.py:108-185 static inputs)."""
    bodies = list(engine.bodies.values())
    if not bodies:
        raise ValueError("the engine has no bodies")
    check_masses(bodies)
    shape_ids = {}
    for b in bodies:
        if b.shape is None:
            raise RuntimeError("body {} has no shape".format(b.name))
        sh = b.shape
        if sh.name not in shape_ids:
            g = sh.grid
            shape_ids[sh.name] = ctx.shape_upload(sh.mesh.V, sh.mesh.T, np.asarray(g.values).ravel(), g.I, g.J,
                                                  g.K, g.min_coord, g.max_coord)
    mats, tab = material_table(engine, bodies)
    n = len(bodies)
    grav = np.zeros((n, 3))
    damp = np.zeros((n, 2))
    for i, b in enumerate(bodies):
        for f in b.forces:
            if f.force_type == "Gravity":
                grav[i] += f.g * np.asarray(f.up, dtype=np.float64)
            else:
                damp[i] += (f.alpha, f.beta)
    desc, keep = ctx.make_desc(
        shape=[shape_ids[b.shape.name] for b in bodies], fixed=[int(b.is_fixed) for b in bodies],
        material=[mats.index(b.material) for b in bodies], mass=[b.mass for b in bodies],
        inertia=np.array([b.inertia for b in bodies], dtype=np.float64), gravity=grav, damping=damp,
        interactions=tab, params=params_struct(engine.params))
    return desc, keep


def gather_state(bodies):
    r = np.array([b.r for b in bodies], dtype=np.float64).reshape(-1, 3)
    q = np.array([b.q for b in bodies], dtype=np.float64).reshape(-1, 4)
    v = np.array([b.v for b in bodies], dtype=np.float64).reshape(-1, 3)
    w = np.array([b.w for b in bodies], dtype=np.float64).reshape(-1, 3)
    return r, q, v, w


def energies(engine, bodies, r, q, v, w):
    """state_storage.py:188-237 (gravity direction from every engine Gravity force)."""
    G = np.zeros(3)
    has_g = False
    for f in engine.forces.values():
        if f.force_type == "Gravity":
            G = G + f.up * f.g
            has_g = True
    kin = pot = 0.0
    for k, b in enumerate(bodies):
        if b.is_fixed:
            continue
        R = Q.to_matrix(q[k])
        Iw = R @ np.diag(b.inertia) @ R.T
        if has_g:
            pot += b.mass * np.linalg.norm(G) * (G / np.linalg.norm(G)).dot(r[k])
        kin += 0.5 * (b.mass * v[k].dot(v[k]) + w[k].dot(Iw @ w[k]))
    return kin, pot


class ContactList:
    """Read-only sequence of ContactPoint built on access from device arrays."""

    def __init__(self, bodies, arrays):
        self._bodies = bodies
        self.arrays = arrays

    def __len__(self):
        return len(self.arrays["ia"])

    def __getitem__(self, k):
        if isinstance(k, slice):
            return [self[i] for i in range(*k.indices(len(self)))]
        if k < 0:
            k += len(self)
        a = self.arrays
        cp = ContactPoint(self._bodies[a["ia"][k]], self._bodies[a["ib"][k]], a["p"][k].copy(), a["n"][k].copy(),
                          float(a["gap"][k]))
        cp.impulse = a["impulse"][k].copy()
        return cp

    def __iter__(self):
        for k in range(len(self)):
            yield self[k]

    def __bool__(self):
        return len(self) > 0


def upload_hinges(ctx, sid, engine):
    """Engine.hinges (creation order) -> rbh_scene_set_hinges."""
    index = {b.name: k for k, b in enumerate(engine.bodies.values())}
    hs = sorted(engine.hinges.values(), key=lambda h: h.idx)
    for h in hs:
        if h.parent is None or h.child is None:
            raise RuntimeError("hinge %s has no parent or child: call set_hinge() first" % h.name)
    ctx.set_hinges(sid, [index[h.parent.name] for h in hs], [index[h.child.name] for h in hs],
                   [h.arm_p for h in hs], [h.arm_c for h in hs], [h.axis_p for h in hs], [h.axis_c for h in hs])


class HipStepper:
    """Drop-in for steppers.SemiImplicitStepper running on one MI355X."""

    def __init__(self):
        self.log = []
        self.ctx = None
        self.scene = None
        self._key = None
        self._bodies = None

    # -- device scene management
    def _prepare(self, engine):
        check_masses(engine.bodies.values())
        prec = getattr(engine.params, "precision", "fp32")
        ctx = shared_context(getattr(engine.params, "device", 0), prec)
        key = (id(engine), getattr(engine, "_version", 0), params_key(engine.params), prec)
        if self.ctx is not ctx or key != self._key:
            if self.scene is not None and self.ctx is not None:
                self.ctx.scene_destroy(self.scene)
                self.scene = None
            self.ctx = ctx
            desc, keep = engine_desc(engine, ctx)
            self.scene = ctx.scene_create(desc)
            del keep
            if engine.hinges:
                upload_hinges(ctx, self.scene, engine)
            self._key = key
            self._bodies = list(engine.bodies.values())
        self.ctx.set_step_index(self.scene, getattr(engine, "step_index", 0))
        self.ctx.set_state(self.scene, *gather_state(self._bodies))

    def _finish(self, engine):
        r, q, v, w = self.ctx.get_state(self.scene, len(self._bodies))
        for k, b in enumerate(self._bodies):
            if b.is_fixed:
                continue
            b.r, b.q, b.v, b.w = r[k].copy(), q[k].copy(), v[k].copy(), w[k].copy()
        engine.contact_points = ContactList(self._bodies, self.ctx.get_contacts(self.scene))
        return r, q, v, w

    def _profile(self, engine, dt):
        lg = self.ctx.get_log(self.scene, engine.params.max_iterations)
        t = self.ctx.get_timings()
        r, q, v, w = self.ctx.get_state(self.scene, len(self._bodies))
        kin, pot = energies(engine, self._bodies, r, q, v, w)
        cd = t["update_bvh"] + t["narrow_phase"] + t["contact_determination"] + t["contact_point_reduction"]
        # collision_detection.py:193-201: the reference times three parts of the
        # contact determination.  One kernel does all three here (each lane moves
        # its triangle into the SDF's frame, runs Frank-Wolfe and writes the
        # point), so its whole time is contact_optimization and the other two
        # parts, which have no kernel of their own, are 0.
        d = dict(update_bvh=t["update_bvh"], narrow_phase=t["narrow_phase"],
                 number_of_overlaps=int(lg["number_of_overlaps"]), model_space_update=0.0,
                 contact_optimization=t["contact_determination"], contact_point_generation=0.0,
                 contact_determination=t["contact_determination"],
                 contact_point_reduction=t["contact_point_reduction"], collision_detection_time=cd)
        if lg["contact_points"] > 0 or engine.hinges:
            d.update(residuals=lg["residuals"].copy(), iterations=int(lg["iterations"]),
                     exitcode=int(lg["exitcode"]), solver_time=t["coloring"] + t["solver_time"])
        if lg["post_exitcode"] >= 0:
            # steppers.py:33-34 with gauss_seidel_solver.py:29-34, 98-101: the
            # post-stabilisation solve's keys carry its prefix
            d.update(post_stabilization_residuals=lg["post_residuals"].copy(),
                     post_stabilization_iterations=int(lg["post_iterations"]),
                     post_stabilization_exitcode=int(lg["post_exitcode"]),
                     post_stabilization_solver_time=t["post_stabilization_solver_time"])
        d.update(stepper_time=t["stepper_time"], dt=dt, contact_points=int(lg["contact_points"]),
                 kinetic_energy=kin, potential_energy=pot, max_penetration=float(lg["max_penetration"]),
                 natural_residual=float(lg["natural_residual"]), rejects=int(lg["rejects"]),
                 colors=int(lg["colors"]))
        return d

    # -- stepper interface
    def step(self, dt, engine, profiling_on):
        """steppers.py:47-106."""
        self.run(engine, [dt], profiling_on)

    def run(self, engine, dts, profiling_on=False):
        """Take the steps ``dts`` in one go, state staying on the device."""
        if not dts:
            return
        self._prepare(engine)
        if profiling_on:
            for dt in dts:
                self.ctx.step([self.scene], 1, dt, native.STEP_TIMING | native.STEP_COUNT_OVERLAPS)
                self.log.append(self._profile(engine, dt))
        else:
            i = 0
            while i < len(dts):
                j = i
                while j < len(dts) and dts[j] == dts[i]:
                    j += 1
                self.ctx.step([self.scene], j - i, dts[i], 0)
                i = j
        engine.step_index = getattr(engine, "step_index", 0) + len(dts)
        self._finish(engine)
