"""Scene data model of the rigid-body simulator.

Mirrors rainbow/simulators/prox_rigid_bodies/types.py (names, fields and
defaults) so code written against the reference reads the same here:
SurfacesInteraction(Library) (:20-79), Shape (:82-108), RigidBody (:111-152),
Gravity / Damping (:189-256), ContactPoint (:259-301), Hinge (:304-353),
Parameters (:356-418) and Engine (:421-445).

Parameters adds the knobs of this build after the reference's fields:
``proximal_solver`` (the scheme names of proximal_contact/prox_solvers.py:6-37;
"parallel_gauss_seidel" sweeps contacts in random-colour order on the GPU,
"gauss_seidel" keeps the reference's serial contact order), ``friction_model``
("isotropic" = prox_sphere with mu[0]; "anisotropic" = prox_ellipsoid with the
three mu values), ``coloring_seed``, ``precision`` ("fp32" product path,
"fp64" for the FP64 build of the same kernels) and ``device``.
"""
from typing import Optional

import numpy as np

from . import math3 as Q


class Grid:
    """Regular node grid holding a signed-distance field (rainbow/geometry/grid3.py:5-45)."""

    def __init__(self, min_coord, max_coord, I, J, K):
        min_coord = np.asarray(min_coord, dtype=np.float64).ravel()
        max_coord = np.asarray(max_coord, dtype=np.float64).ravel()
        if np.any(min_coord > max_coord) or I < 2 or J < 2 or K < 2:
            raise ValueError()
        self.min_coord = min_coord
        self.max_coord = max_coord
        self.I, self.J, self.K = int(I), int(J), int(K)
        self.spacing = np.divide(max_coord - min_coord, np.array([I - 1, J - 1, K - 1], dtype=np.float64))
        self.values = None


class SurfacesInteraction:
    def __init__(self):
        self.mu = np.ones(3)
        self.epsilon = 0.0


class SurfacesInteractionLibrary:
    def __init__(self):
        self.storage = {("default", "default"): SurfacesInteraction()}

    def get_interaction(self, A, B):
        key = (A, B) if A < B else (B, A)
        return self.storage.get(key, self.storage[("default", "default")])

    def exist_interaction(self, A, B):
        key = (A, B) if A < B else (B, A)
        return key in self.storage

    def exist_material(self, name):
        return any(name in key for key in self.storage)


class Shape:
    def __init__(self, name):
        self.name = name
        self.mesh = None
        self.grid: Optional[Grid] = None
        self.mass = 0.0
        self.inertia = np.zeros(3)
        self.r = np.zeros(3)
        self.q = Q.identity()


class RigidBody:
    def __init__(self, name):
        self.name = name
        self.idx: Optional[int] = None
        self.q = Q.identity()
        self.r = np.zeros(3)
        self.v = np.zeros(3)
        self.w = np.zeros(3)
        self.mass = 0.0
        self.inertia = np.zeros(3)
        self.shape: Optional[Shape] = None
        self.is_fixed = False
        self.forces = []
        self.material = "default"
        self.bvh = None


class ForceCalculator:
    def __init__(self, force_type, name):
        self.force_type = force_type
        self.name = name


class Gravity(ForceCalculator):
    def __init__(self, name):
        super().__init__("Gravity", name)
        self.g = 9.81
        self.up = np.array([0.0, 1.0, 0.0])

    def compute(self, body, r, q, v, w):
        return -body.mass * self.g * self.up, np.zeros(3)


class Damping(ForceCalculator):
    def __init__(self, name):
        super().__init__("Damping", name)
        self.alpha = 0.001
        self.beta = 0.001

    def compute(self, body, r, q, v, w):
        return -v * self.alpha, -w * self.beta


class ContactPoint:
    def __init__(self, bodyA, bodyB, position=None, normal=None, gap=0.0):
        position = np.zeros(3) if position is None else position
        normal = np.array([0.0, 0.0, 1.0]) if normal is None else normal
        if abs(1.0 - np.linalg.norm(normal)) > 0.1:
            raise RuntimeError("ContactPoint.init() was called with non-unit size normal")
        self.bodyA = bodyA
        self.bodyB = bodyB
        self.p = position
        self.n = normal
        self.g = gap


class Hinge:
    def __init__(self, name):
        self.name = name
        self.idx = None
        self.parent = None
        self.child = None
        self.socket_p = None
        self.socket_c = None
        self.arm_p = None
        self.arm_c = None
        self.axis_p = None
        self.axis_c = None


class Parameters:
    def __init__(self):
        self.time_step = 0.01
        self.max_iterations = 200
        self.use_bounce = False
        self.use_pre_stabilization = False
        self.use_post_stabilization = False
        self.gap_reduction = 0.5
        self.min_gap_value = 0.001
        self.max_gap_value = 0.01
        self.absolute_tolerance = 0.001
        self.relative_tolerance = 0.0001
        self.ellipsoid_max_iterations = 100
        self.ellipsoid_expansion = 1.5
        self.ellipsoid_tolerance = 10e-10
        self.nu_reduce = 0.7
        self.nu_increase = 1.3
        self.too_small_merit_change = 0.01
        self.contact_optimization_max_iterations = 8
        self.contact_optimization_tolerance = 0
        self.bvh_chunk_size = 255
        self.K = 3
        self.envelope = 0.1
        self.resolution = 64
        # knobs of this build
        self.proximal_solver = "parallel_gauss_seidel"
        self.friction_model = "isotropic"
        self.coloring_seed = 0
        self.precision = "fp32"
        self.device = 0


class Engine:
    def __init__(self):
        self.simulator_type = "rigid_body"
        self.bodies = dict()
        self.forces = dict()
        self.shapes = dict()
        self.hinges = dict()
        self.contact_points = []
        self.surfaces_interactions = SurfacesInteractionLibrary()
        self.params = Parameters()
        self.stepper = None
        # number of steps taken (keys the colouring); builder for shape SDFs
        self.step_index = 0
        self.sdf_builder = None
        self._version = 0
