"""Plain-array scene description consumed by the oracle.

The oracle works on arrays, not on the reference's object model: one record per
body (the fields StateStorage reads, state_storage.py:43-185 and
types.py:111-152), one per shape (mesh in body frame + SDF grid,
types.py:82-108), the material interaction table (types.py:20-79) and the
Parameters fields (types.py:356-418).  ``from_engine`` reads any object with
the reference's Engine attribute layout (``bodies`` dict in insertion order,
``shapes``, ``forces``, ``surfaces_interactions``, ``params``).
"""
import copy

import numpy as np

from .grid import Grid

PARAM_DEFAULTS = dict(
    time_step=0.01, max_iterations=200, use_bounce=False, use_pre_stabilization=False,
    use_post_stabilization=False, gap_reduction=0.5, min_gap_value=0.001, max_gap_value=0.01,
    absolute_tolerance=0.001, relative_tolerance=0.0001, ellipsoid_max_iterations=100,
    ellipsoid_expansion=1.5, ellipsoid_tolerance=10e-10, nu_reduce=0.7, nu_increase=1.3,
    too_small_merit_change=0.01, contact_optimization_max_iterations=8,
    contact_optimization_tolerance=0.0, bvh_chunk_size=255, K=3, envelope=0.1, resolution=64,
    # knobs of this build (not in the reference's Parameters)
    proximal_solver="parallel_gauss_seidel", friction_model="isotropic", coloring_seed=0,
)


class ShapeArrays:
    def __init__(self, V, T, grid):
        self.V = np.asarray(V, dtype=np.float64)
        self.T = np.asarray(T, dtype=np.int64)
        self.grid = grid


class Scene:
    """Body state + static data of one scene, in body insertion order."""

    def __init__(self):
        self.r = np.zeros((0, 3))
        self.q = np.zeros((0, 4))
        self.v = np.zeros((0, 3))
        self.w = np.zeros((0, 3))
        self.mass = np.zeros(0)
        self.inertia = np.zeros((0, 3))
        self.fixed = np.zeros(0, dtype=bool)
        self.shape = np.zeros(0, dtype=np.int64)
        self.material = np.zeros(0, dtype=np.int64)
        self.forces = []          # per body: list of ('gravity', g, up) / ('damping', alpha, beta)
        self.gravity = []         # every Gravity force of the engine: (g, up); energy uses all of them
        self.shapes = []
        self.mu = np.ones((1, 1, 3))
        self.eps = np.zeros((1, 1))
        self.params = dict(PARAM_DEFAULTS)
        self.step_index = 0       # number of steps taken (keys the colouring RNG)
        self.hinges = None        # oracle.hinges.HingeSet (Engine.hinges) or None

    @property
    def N(self):
        return len(self.r)

    def copy(self):
        return copy.deepcopy(self)


def from_engine(engine):
    """Build a Scene from an Engine-like object (reference attribute layout)."""
    sc = Scene()
    bodies = list(engine.bodies.values())
    shape_names = list(engine.shapes.keys())
    shape_index = {n: i for i, n in enumerate(shape_names)}
    for n in shape_names:
        sh = engine.shapes[n]
        g = sh.grid
        grid = Grid(g.min_coord, g.max_coord, g.I, g.J, g.K, values=np.asarray(g.values).ravel())
        sc.shapes.append(ShapeArrays(sh.mesh.V, sh.mesh.T, grid))
    mats = ["default"]
    for b in bodies:
        if b.material not in mats:
            mats.append(b.material)
    for key in engine.surfaces_interactions.storage:
        for m in key:
            if m not in mats:
                mats.append(m)
    M = len(mats)
    sc.mu = np.zeros((M, M, 3))
    sc.eps = np.zeros((M, M))
    for i, a in enumerate(mats):
        for j, b in enumerate(mats):
            beh = engine.surfaces_interactions.get_interaction(a, b)
            sc.mu[i, j] = np.asarray(beh.mu, dtype=np.float64)[:3]
            sc.eps[i, j] = beh.epsilon
    sc.r = np.array([b.r for b in bodies], dtype=np.float64).reshape(-1, 3)
    sc.q = np.array([b.q for b in bodies], dtype=np.float64).reshape(-1, 4)
    sc.v = np.array([b.v for b in bodies], dtype=np.float64).reshape(-1, 3)
    sc.w = np.array([b.w for b in bodies], dtype=np.float64).reshape(-1, 3)
    sc.mass = np.array([b.mass for b in bodies], dtype=np.float64)
    sc.inertia = np.array([b.inertia for b in bodies], dtype=np.float64).reshape(-1, 3)
    sc.fixed = np.array([b.is_fixed for b in bodies], dtype=bool)
    sc.shape = np.array([shape_index[b.shape.name] for b in bodies], dtype=np.int64)
    sc.material = np.array([mats.index(b.material) for b in bodies], dtype=np.int64)
    for b in bodies:
        fl = []
        for f in b.forces:
            if f.force_type == "Gravity":
                fl.append(("gravity", float(f.g), np.asarray(f.up, dtype=np.float64)))
            else:
                fl.append(("damping", float(f.alpha), float(f.beta)))
        sc.forces.append(fl)
    for f in engine.forces.values():
        if f.force_type == "Gravity":
            sc.gravity.append((float(f.g), np.asarray(f.up, dtype=np.float64)))
    from . import hinges
    sc.hinges = hinges.from_engine(engine, {b.name: k for k, b in enumerate(bodies)})
    p = engine.params
    sc.params = {k: getattr(p, k, v) for k, v in PARAM_DEFAULTS.items()}
    sc.step_index = int(getattr(engine, "step_index", 0))
    return sc
