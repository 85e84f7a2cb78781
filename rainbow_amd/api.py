"""Scene-building and simulation API of the rigid-body simulator.

Same functions, argument order, defaults and exceptions as
rainbow/simulators/prox_rigid_bodies/api.py (:21-705), so a script written
against the reference runs unchanged.  Differences:

* ``create_engine(backend="hip")`` installs the GPU stepper (stepper.HipStepper);
  ``backend=None`` leaves ``engine.stepper`` empty, like ``Engine()``.
* ``create_shape`` builds the signed-distance grid on the GPU
  (librbh's rbh_sdf_build replaces igl.signed_distance, api.py:234-241) unless
  ``engine.sdf_builder`` supplies another builder with the signature of
  grid3.create_signed_distance(V, T, I, J, K, boundary).
* ``connect_shape`` builds no BVH: the GPU broad phase computes the same
  candidate-triangle set from per-triangle boxes every step.
* ``simulate`` hands all sub-steps to a HIP stepper in one call, keeping state
  on the device between them.
"""
import datetime
import random

import numpy as np

from . import math3 as Q
from . import mesh as MESH
from . import mass as MASS
from .types import (ContactPoint, Damping, Engine, Gravity, Grid, Hinge, Parameters, RigidBody, Shape,  # noqa: F401
                    SurfacesInteraction)


def generate_unique_name(name: str) -> str:
    """api.py:21-39."""
    unique_name = name + "_" + str(random.random()) + "_" + str(datetime.datetime.now())
    for ic in [" ", ".", ":", "-"]:
        unique_name = unique_name.replace(ic, "_")
    return unique_name


def create_engine(backend="hip") -> Engine:
    """api.py:42-53."""
    engine = Engine()
    if backend == "hip":
        from .stepper import HipStepper
        engine.stepper = HipStepper()
    elif backend is not None:
        raise ValueError("unknown backend " + str(backend))
    return engine


def _touch(engine):
    engine._version = getattr(engine, "_version", 0) + 1


def create_rigid_body(engine, body_name: str) -> None:
    if body_name in engine.bodies:
        raise RuntimeError("create_rigid_body() rigid body already exist with that name")
    body = RigidBody(body_name)
    body.idx = len(engine.bodies)
    engine.bodies[body_name] = body
    _touch(engine)


def create_hinge(engine, hinge_name: str) -> None:
    if hinge_name in engine.hinges:
        raise RuntimeError("create_hinge() hinge already exist with that name")
    hinge = Hinge(hinge_name)
    hinge.idx = len(engine.hinges)
    engine.hinges[hinge_name] = hinge
    _touch(engine)


def set_hinge(engine, hinge_name: str, parent_name: str, child_name: str, origin: np.ndarray, axis: np.ndarray,
              mode: str = "world") -> None:
    """api.py:86-156: the joint frame (origin, axis as its local z) in world,
    parent-body or parent-model coordinates, stored as sockets in both bodies'
    frames (types.py:329-353: arm = socket origin, axis = socket z)."""
    if hinge_name in engine.hinges:
        hinge = engine.hinges[hinge_name]
    else:
        raise RuntimeError("set_hinge() no such rigid body exist with that name")
    if parent_name in engine.bodies:
        parent = engine.bodies[parent_name]
    else:
        raise RuntimeError("set_hinge() no such rigid body exist with that name")
    if child_name in engine.bodies:
        child = engine.bodies[child_name]
    else:
        raise RuntimeError("set_hinge() no such rigid body exist with that name")
    origin = np.asarray(origin, dtype=np.float64)
    axis = np.asarray(axis, dtype=np.float64)
    if mode == "world":
        o_world, s_world = origin.copy(), axis.copy()
    elif mode == "body":
        o_world = Q.rotate(parent.q, origin) + parent.r
        s_world = Q.rotate(parent.q, axis)
    elif mode == "model":
        o_body = Q.rotate(Q.conjugate(parent.shape.q), origin - parent.shape.r)
        s_body = Q.rotate(Q.conjugate(parent.shape.q), axis)
        o_world = Q.rotate(parent.q, o_body) + parent.r
        s_world = Q.rotate(parent.q, s_body)
    else:
        raise ValueError(f"set_hinge() no such mode {mode} is supported")
    t, b, n = Q.make_orthonormal_vectors(s_world)
    q_world = Q.from_matrix(np.column_stack([t, b, n]))
    z = np.array([0.0, 0.0, 1.0])
    q_ps = Q.prod(Q.conjugate(parent.q), q_world)
    q_cs = Q.prod(Q.conjugate(child.q), q_world)
    hinge.parent, hinge.child = parent, child
    hinge.socket_p = (Q.rotate(Q.conjugate(parent.q), o_world - parent.r), q_ps)
    hinge.socket_c = (Q.rotate(Q.conjugate(child.q), o_world - child.r), q_cs)
    hinge.arm_p = hinge.socket_p[0].copy()
    hinge.arm_c = hinge.socket_c[0].copy()
    hinge.axis_p = Q.rotate(q_ps, z)
    hinge.axis_c = Q.rotate(q_cs, z)
    _touch(engine)


def create_mesh(V: np.ndarray, T: np.ndarray) -> MESH.Mesh:
    """api.py:159-173: spatial re-ordering along the direction of most variance."""
    axis = MESH.direction_of_most_variance(V)
    mesh = MESH.Mesh(V, T)
    MESH.axis_sort(mesh, axis)
    return mesh


def _gpu_signed_distance(engine, V, T, I, J, K, boundary):
    from .stepper import shared_context
    if boundary < 0:
        raise ValueError()
    min_coord = (V.min(axis=0) - boundary).ravel()
    max_coord = (V.max(axis=0) + boundary).ravel()
    grid = Grid(min_coord, max_coord, I, J, K)
    ctx = shared_context(getattr(engine.params, "device", 0), "fp64")
    grid.values = ctx.sdf_build(V, T, min_coord, max_coord, I, J, K)
    return grid


def create_shape(engine, shape_name: str, mesh: MESH.Mesh, transform_to_body_frame: bool = True) -> None:
    """api.py:176-242."""
    if shape_name in engine.shapes:
        raise RuntimeError("create_shape(): shape with that name already exist: " + shape_name)
    shape = Shape(shape_name)
    shape.mesh = mesh
    prop = MASS.compute_mass_properties(shape.mesh.V, shape.mesh.T, 1.0)
    (shape.r, shape.q, shape.mass, shape.inertia) = MASS.xform_model_2_body_space(prop)
    if transform_to_body_frame:
        MESH.translate(shape.mesh, -shape.r)
        MESH.rotate(shape.mesh, Q.conjugate(shape.q))
    max_length = (shape.mesh.V.max(axis=0) - shape.mesh.V.min(axis=0)).max()
    boundary = max(max_length * 0.1, engine.params.envelope * 2)
    res = engine.params.resolution
    builder = getattr(engine, "sdf_builder", None)
    if builder is not None:
        shape.grid = builder(shape.mesh.V, shape.mesh.T, res, res, res, boundary)
    else:
        shape.grid = _gpu_signed_distance(engine, shape.mesh.V, shape.mesh.T, res, res, res, boundary)
    engine.shapes[shape_name] = shape
    _touch(engine)


def connect_shape(engine, body_name: str, shape_name: str) -> None:
    """api.py:245-281."""
    if body_name in engine.bodies:
        body = engine.bodies[body_name]
    else:
        raise RuntimeError("connect() no such rigid body exist with that name")
    if shape_name in engine.shapes:
        shape = engine.shapes[shape_name]
    else:
        raise RuntimeError("connect() no such shape exist with that name")
    body.shape = shape
    _touch(engine)


def create_gravity_force(engine, force_name: str, g: float, up: np.ndarray) -> None:
    if force_name in engine.forces:
        raise RuntimeError("create_gravity(): Force already exist with that name")
    if g <= 0.0:
        raise RuntimeError("create_gravity(): Illegal value for gravitational acceleration")
    gravity = Gravity(force_name)
    gravity.up = Q.unit3(up)
    gravity.g = g
    engine.forces[force_name] = gravity
    _touch(engine)


def create_damping_force(engine, force_name: str, alpha: float, beta: float) -> None:
    if force_name in engine.forces:
        raise RuntimeError("create_damping(): Force already exist with that name")
    if alpha <= 0:
        raise RuntimeError("create_damping(): Illegal value for alpha")
    if beta <= 0:
        raise RuntimeError("create_damping(): Illegal value for beta")
    damping = Damping(force_name)
    damping.alpha = alpha
    damping.beta = beta
    engine.forces[force_name] = damping
    _touch(engine)


def connect_force(engine, body_name: str, force_name: str) -> None:
    if body_name in engine.bodies:
        body = engine.bodies[body_name]
    else:
        raise RuntimeError("connect_force() no such rigid body exist with that name")
    if force_name in engine.forces:
        force = engine.forces[force_name]
    else:
        raise RuntimeError("connect_force() no such force exist with that name")
    if force in body.forces:
        raise RuntimeError("connect_force() force was already connected to body")
    body.forces.append(force)
    _touch(engine)


def set_position(engine, body_name: str, r: np.ndarray, use_model_frame=False) -> None:
    """api.py:358-415."""
    if body_name in engine.bodies:
        body = engine.bodies[body_name]
    else:
        raise RuntimeError("set_position() no such rigid body exist with that name")
    if use_model_frame:
        q_bf2wcs = np.copy(body.q)
        r_bf2mf = np.copy(body.shape.r)
        q_bf2mf = np.copy(body.shape.q)
        q_mf2wcs = Q.prod(q_bf2wcs, Q.conjugate(q_bf2mf))
        body.r = Q.rotate(q_mf2wcs, r_bf2mf) + np.copy(r)
    else:
        body.r = np.copy(r)


def get_position(engine, body_name: str) -> np.ndarray:
    if body_name in engine.bodies:
        return np.copy(engine.bodies[body_name].r)
    raise RuntimeError("get_position() no such rigid body exist with that name")


def set_orientation(engine, body_name: str, q: np.ndarray, use_model_frame=False) -> None:
    """api.py:433-489."""
    if body_name in engine.bodies:
        body = engine.bodies[body_name]
    else:
        raise RuntimeError("set_position() no such rigid body exist with that name")
    if use_model_frame:
        r_bf2wcs = np.copy(body.r)
        q_bf2wcs = np.copy(body.q)
        r_bf2mf = np.copy(body.shape.r)
        q_bf2mf = np.copy(body.shape.q)
        q_mf2wcs = Q.prod(q_bf2wcs, Q.conjugate(q_bf2mf))
        r_mf2wcs = r_bf2wcs - Q.rotate(q_mf2wcs, r_bf2mf)
        q_mf2wcs = Q.unit(q)
        body.q = Q.prod(q_mf2wcs, q_bf2mf)
        body.r = Q.rotate(q_mf2wcs, r_bf2mf) + r_mf2wcs
    else:
        body.q = Q.unit(q)


def get_orientation(engine, body_name: str) -> np.ndarray:
    if body_name in engine.bodies:
        return np.copy(engine.bodies[body_name].q)
    raise RuntimeError("get_orientation() no such rigid body exist with that name")


def set_velocity(engine, body_name: str, v: np.ndarray) -> None:
    if body_name in engine.bodies:
        engine.bodies[body_name].v = np.copy(v)
    else:
        raise RuntimeError("set_velocity() no such rigid body exist with that name")


def get_velocity(engine, body_name: str) -> np.ndarray:
    if body_name in engine.bodies:
        return np.copy(engine.bodies[body_name].v)
    raise RuntimeError("get_velocity() no such rigid body exist with that name")


def set_spin(engine, body_name: str, w: np.ndarray) -> None:
    if body_name in engine.bodies:
        engine.bodies[body_name].w = np.copy(w)
    else:
        raise RuntimeError("set_spin() no such rigid body exist with that name")


def get_spin(engine, body_name: str) -> np.ndarray:
    if body_name in engine.bodies:
        return np.copy(engine.bodies[body_name].w)
    raise RuntimeError("get_spin() no such rigid body exist with that name")


def set_mass_properties(engine, body_name: str, density: float) -> None:
    if body_name in engine.bodies:
        body = engine.bodies[body_name]
    else:
        raise RuntimeError("set_mass_properties() no such rigid body exist with that name")
    if body.shape is None:
        raise RuntimeError("set_mass_properties() rigid body did not have a shape")
    body.mass = body.shape.mass * density
    body.inertia = body.shape.inertia * density
    _touch(engine)


def set_body_type(engine, body_name: str, body_type: str) -> None:
    if body_name not in engine.bodies:
        raise RuntimeError("set_body_type() no such rigid body exist with that name")
    body = engine.bodies[body_name]
    if body_type in ["fixed", "Fixed", "FIXED"]:
        body.is_fixed = True
    elif body_type in ["free", "Free", "FREE"]:
        body.is_fixed = False
    else:
        raise RuntimeError("set_body_type(): Unsupported body type found")
    _touch(engine)


def is_fixed_body(engine, body_name: str) -> bool:
    if body_name not in engine.bodies:
        raise RuntimeError("is_fixed_body() no such rigid body exist with that name")
    return engine.bodies[body_name].is_fixed


def set_body_material(engine, body_name: str, material_name: str) -> None:
    if body_name not in engine.bodies:
        raise RuntimeError("set_body_material() no such rigid body exist with that name")
    body = engine.bodies[body_name]
    if not engine.surfaces_interactions.exist_material(material_name):
        raise RuntimeError("set_body_material() no such material exist")
    body.material = material_name
    _touch(engine)


def create_surfaces_interaction(engine, A: str, B: str, epsilon: float, mu) -> None:
    """api.py:646-673."""
    if engine.surfaces_interactions.exist_interaction(A, B):
        raise RuntimeError("create_surfaces_interaction() behaviour already exist")
    if epsilon < 0.0:
        raise RuntimeError("create_surfaces_interaction() illegal epsilon value")
    if mu[0] < 0.0:
        raise RuntimeError("create_surfaces_interaction() illegal mu_x value")
    if mu[1] < 0.0:
        raise RuntimeError("create_surfaces_interaction() illegal mu_y value")
    if mu[2] < 0.0:
        raise RuntimeError("create_surfaces_interaction() illegal mu_z value")
    key = tuple(sorted([A, B]))
    behaviour = SurfacesInteraction()
    behaviour.epsilon = epsilon
    behaviour.mu = mu
    engine.surfaces_interactions.storage[key] = behaviour
    _touch(engine)


def substeps(T, time_step):
    """The dt sequence of api.py:689-693 (while T_left: dt = min(T_left, time_step))."""
    out = []
    T_left = T
    while T_left:
        dt = min(T_left, time_step)
        out.append(dt)
        T_left -= dt
    return out


def simulate(engine: Engine, T: float, profiling_on: bool = False) -> None:
    """api.py:676-693."""
    if T <= 0:
        raise ValueError("Time must be positive")
    if engine.stepper is None:
        raise ValueError("A time-stepper must be initialized")
    run = getattr(engine.stepper, "run", None)
    if run is not None:
        run(engine, substeps(T, engine.params.time_step), profiling_on)
        return
    for dt in substeps(T, engine.params.time_step):
        engine.stepper.step(dt, engine, profiling_on)


def get_log(engine: Engine):
    """api.py:696-705."""
    return engine.stepper.log
