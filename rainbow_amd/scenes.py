"""Scene generators for the benchmark and test workloads.

Built only with the public API (api.py), following the reference's scene
helpers: ``create_ground`` is scenes/create_ground.py:13-48 (a 100 x 1 x 100
fixed box whose top face is the plane y = 0), bodies are
surface_mesh.create_box / create_sphere shapes placed with
``set_position(..., use_model_frame=True)``.

Workloads (seeds from numpy.random.default_rng):

* ``box_stack``      64 unit boxes in a 4x4x4 lattice with 0.05 gaps above the
                     ground, +-0.01 jitter (seed 1); mu 0.5, dt 0.01,
                     100 iterations, 32^3 SDFs.
* ``sphere_rain``    n spheres r = 0.5 (16 x 16) uniform in
                     [-9.5, 9.5] x [1, 40] x [-9.5, 9.5], centre distance > 1,
                     v ~ N(0, 0.5^2) (seed 2); mu 0.3, dt 0.005, 50 iterations.
* ``mixed_drop``     one scene of the batched workload: N log-uniform in
                     [8, 512], half boxes half spheres, dropped from heights
                     U[1, 10]; mu ~ U[0.1, 0.8]; dt 0.01, 30 iterations.
* ``serving_request`` a sphere-rain style scene with N log-uniform in [8, 1024].
* ``funnel_stress``  n bodies (half noisy icospheres of geodesic frequency 22,
                     10 f^2 + 2 = 4,842 vertices, radial noise 1 + 0.2 U(-1, 1)
                     per vertex; half (2, 3) torus knots of 200 x 25 = 5,000
                     vertices), each scaled to a unit bounding box, dropped in
                     a lattice with random orientations (seed 5) into a fixed
                     funnel placed as in scenes/create_funnel.py:83-110
                     (scaled to 2R x H x 2R, at y = 1.5 H, turned by Rx(pi));
                     64^3 SDFs, mu 0.4, dt 0.005, 50 iterations.

* ``cube_hinge_chain`` the reference's hinge scene: 5 cubes of edge 10 in a
                     row joined by 4 hinges about z, the first cube fixed.

``sdf_builder`` (optional) replaces the GPU SDF construction, e.g. for
CPU-only tests.
"""
import numpy as np

from . import api as API
from . import mesh as MESH
from . import math3 as Q
from .xml_io import read_xml, write_xml   # noqa: F401  (scenes/write_xml.py, and its reader)


class ShapeBank:
    """Shapes shared by many engines (the SDF of each is built once)."""

    def __init__(self, resolution=32, envelope=0.1, sdf_builder=None, device=0):
        self.resolution = resolution
        self.envelope = envelope
        self.sdf_builder = sdf_builder
        self.device = device
        self.shapes = {}

    def get(self, name, make_VT):
        if name not in self.shapes:
            eng = API.create_engine(backend=None)
            eng.params.resolution = self.resolution
            eng.params.envelope = self.envelope
            eng.params.device = self.device
            eng.sdf_builder = self.sdf_builder
            V, T = make_VT()
            API.create_shape(eng, name, API.create_mesh(V, T))
            self.shapes[name] = eng.shapes[name]
        return self.shapes[name]


def new_engine(bank, params=None, backend="hip"):
    eng = API.create_engine(backend=backend)
    eng.params.resolution = bank.resolution
    eng.params.envelope = bank.envelope
    eng.sdf_builder = bank.sdf_builder
    for k, v in (params or {}).items():
        setattr(eng.params, k, v)
    return eng


def add_shape(eng, shape):
    if shape.name not in eng.shapes:
        eng.shapes[shape.name] = shape
        API._touch(eng)


def add_body(eng, name, shape, r, q=None, v=None, w=None, fixed=False, density=1.0, forces=("gravity",)):
    add_shape(eng, shape)
    API.create_rigid_body(eng, name)
    API.connect_shape(eng, name, shape.name)
    API.set_mass_properties(eng, name, density)
    if q is not None:
        API.set_orientation(eng, name, q, use_model_frame=True)
    API.set_position(eng, name, np.asarray(r, dtype=np.float64), use_model_frame=True)
    if fixed:
        API.set_body_type(eng, name, "fixed")
    if v is not None:
        API.set_velocity(eng, name, np.asarray(v, dtype=np.float64))
    if w is not None:
        API.set_spin(eng, name, np.asarray(w, dtype=np.float64))
    for f in forces:
        if f in eng.forces and not fixed:
            API.connect_force(eng, name, f)


def add_ground(eng, bank):
    """scenes/create_ground.py:13-48 with r = 0, q = identity."""
    shape = bank.get("ground_shape", lambda: MESH.create_box(100.0, 1.0, 100.0))
    add_body(eng, "ground_body", shape, np.array([0.0, -0.5, 0.0]), q=Q.identity(), fixed=True, forces=())


def set_friction(eng, mu, epsilon=0.0):
    beh = eng.surfaces_interactions.storage[("default", "default")]
    beh.mu = np.array([mu, mu, mu], dtype=np.float64)
    beh.epsilon = epsilon


def box_stack(bank=None, seed=1, backend="hip", params=None, nx=4, ny=4, nz=4):
    bank = bank or ShapeBank(resolution=32)
    prm = dict(time_step=0.01, max_iterations=100, absolute_tolerance=0.0, relative_tolerance=0.0)
    prm.update(params or {})
    eng = new_engine(bank, prm, backend)
    set_friction(eng, 0.5)
    API.create_gravity_force(eng, "gravity", 9.81, np.array([0.0, 1.0, 0.0]))
    add_ground(eng, bank)
    box = bank.get("box_shape", lambda: MESH.create_box(1.0, 1.0, 1.0))
    rng = np.random.default_rng(seed)
    for i in range(nx):
        for j in range(nz):
            for k in range(ny):
                c = np.array([(i - 1.5) * 1.05, 0.55 + k * 1.05, (j - 1.5) * 1.05]) + rng.uniform(-0.01, 0.01, 3)
                add_body(eng, "box_%d_%d_%d" % (i, j, k), box, c, q=Q.identity())
    return eng


def cube_hinge_chain(bank=None, n=5, edge=10.0, backend="hip", params=None, ground=False):
    """scenes/create_cube_hinge_chain.py:13-87: n cubes turned by Rz(pi/4) in
    a row along x (corner gaps of 0.1 half-diagonal), the first one fixed,
    neighbours joined by hinges at the midpoint of their centres about the
    world z axis; gravity along -y.  ``ground`` adds the ground plane one
    half-diagonal below the chain."""
    bank = bank or ShapeBank(resolution=32)
    prm = dict(time_step=0.01, max_iterations=100, absolute_tolerance=0.0, relative_tolerance=0.0)
    prm.update(params or {})
    eng = new_engine(bank, prm, backend)
    set_friction(eng, 0.5)
    API.create_gravity_force(eng, "gravity", 9.81, np.array([0.0, 1.0, 0.0]))
    diameter = np.sqrt(2.0 * edge * edge)
    radius = diameter / 2.0
    margin = 0.1 * radius
    length = n * diameter + (n - 1) * margin
    if ground:
        add_ground(eng, bank)
    cube = bank.get("cube_%g" % edge, lambda: MESH.create_box(edge, edge, edge))
    names = []
    y = 2.0 * radius if ground else 0.0
    for k in range(n):
        name = "link_%d" % k
        x = k * (diameter + margin) + radius - length / 2.0
        add_body(eng, name, cube, np.array([x, y, 0.0]), q=Q.Rz(np.pi / 4), fixed=(k == 0))
        names.append(name)
    for k in range(n - 1):
        hname = names[k] + "_" + names[k + 1]
        API.create_hinge(eng, hname)
        o = (API.get_position(eng, names[k]) + API.get_position(eng, names[k + 1])) / 2.0
        API.set_hinge(eng, hname, names[k], names[k + 1], o, np.array([0.0, 0.0, 1.0]), mode="world")
    return eng


def _place(rng, n, lo, hi, radii, max_tries=2000000):
    """Rejection sampling of n centres in the box [lo, hi] with |ci - cj| > ri + rj."""
    cell = 2.0 * float(np.max(radii))
    grid = {}
    out = []
    tries = 0
    for i in range(n):
        while True:
            tries += 1
            if tries > max_tries:
                raise RuntimeError("could not place bodies without overlap")
            p = rng.uniform(lo, hi)
            key = tuple(np.floor(p / cell).astype(int))
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        for j in grid.get((key[0] + dx, key[1] + dy, key[2] + dz), ()):
                            if np.linalg.norm(out[j] - p) <= radii[i] + radii[j]:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                grid.setdefault(key, []).append(i)
                out.append(p)
                break
    return np.array(out)


def sphere_rain(n=4096, bank=None, seed=2, backend="hip", params=None, extent=9.5, ylo=1.0, yhi=40.0):
    bank = bank or ShapeBank(resolution=32)
    prm = dict(time_step=0.005, max_iterations=50, absolute_tolerance=0.0, relative_tolerance=0.0,
               coloring_seed=seed)
    prm.update(params or {})
    eng = new_engine(bank, prm, backend)
    set_friction(eng, 0.3)
    API.create_gravity_force(eng, "gravity", 9.81, np.array([0.0, 1.0, 0.0]))
    add_ground(eng, bank)
    sph = bank.get("sphere_shape", lambda: MESH.create_sphere(0.5, 16, 16))
    rng = np.random.default_rng(seed)
    P = _place(rng, n, np.array([-extent, ylo, -extent]), np.array([extent, yhi, extent]), np.full(n, 0.5))
    V = rng.normal(0.0, 0.5, size=(n, 3))
    for i in range(n):
        add_body(eng, "sphere_%d" % i, sph, P[i], q=Q.identity(), v=V[i])
    return eng


def mixed_drop(seed, bank=None, backend="hip", params=None, nmin=8, nmax=512):
    bank = bank or ShapeBank(resolution=32)
    rng = np.random.default_rng(seed)
    N = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
    mu = float(rng.uniform(0.1, 0.8))
    prm = dict(time_step=0.01, max_iterations=30, absolute_tolerance=0.0, relative_tolerance=0.0,
               coloring_seed=seed)
    prm.update(params or {})
    eng = new_engine(bank, prm, backend)
    set_friction(eng, mu)
    API.create_gravity_force(eng, "gravity", 9.81, np.array([0.0, 1.0, 0.0]))
    add_ground(eng, bank)
    box = bank.get("box_shape", lambda: MESH.create_box(1.0, 1.0, 1.0))
    sph = bank.get("sphere_shape", lambda: MESH.create_sphere(0.5, 16, 16))
    kinds = rng.uniform(size=N) < 0.5
    radii = np.where(kinds, np.sqrt(3.0) / 2.0, 0.5)
    side = 1.5 * np.ceil(np.sqrt(N))
    P = _place(rng, N, np.array([-side / 2, 1.0, -side / 2]), np.array([side / 2, 10.0, side / 2]), radii)
    for i in range(N):
        add_body(eng, "body_%d" % i, box if kinds[i] else sph, P[i], q=Q.identity())
    return eng


def serving_request(seed, bank=None, backend="hip", params=None, nmin=8, nmax=1024):
    rng = np.random.default_rng(seed)
    N = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
    extent = max(2.0, 0.5 * np.cbrt(N / 0.15 / 40.0) * 19.0 / 2.0)
    extent = min(extent, 9.5)
    return sphere_rain(N, bank=bank, seed=seed, backend=backend, params=params, extent=extent,
                       ylo=1.0, yhi=1.0 + max(4.0, N * 0.52 / (0.15 * (2 * extent) ** 2)))


def _unit_shape(bank, name, make_VT):
    """A shape whose mesh is scaled to a unit bounding box (surface_mesh.scale_to_unit)."""
    def make():
        V, T = make_VT()
        m = MESH.Mesh(V.copy(), T)
        MESH.scale_to_unit(m)
        return m.V, m.T
    return bank.get(name, make)


def funnel_shape(bank, height, radius):
    """The funnel mesh scaled to 2 radius x height x 2 radius (create_funnel.py:84-95)."""
    def make():
        m = MESH.Mesh(*MESH.create_funnel())
        MESH.scale_to_unit(m)
        lo, hi = MESH.aabb(m)
        ext = hi - lo
        MESH.scale(m, 2.0 * radius / ext[0], height / ext[1], 2.0 * radius / ext[2])
        return m.V, m.T
    return bank.get("funnel_%g_%g" % (height, radius), make)


def funnel_inner_radius(y, height, radius):
    """Inner radius of the placed funnel at world height y (height <= y <= 2 height)."""
    s = (2.0 * height - y) / height              # 0 at the wide top rim, 1 at the spout
    wide, narrow, wall = 0.5, 0.17, 0.05
    return 2.0 * radius * ((wide - wall) + (narrow - wide) * s)


def funnel_stress(seed=5, n=256, bank=None, backend="hip", params=None, frequency=22, height=8.0, radius=8.0,
                  spacing=1.5, clearance=1.0):
    """Noisy icospheres and torus knots poured into a funnel (see the module docstring).

    Bodies fill the funnel from the spout up: lattice sites ``spacing`` apart
    (more than twice the largest body radius, so no two bodies touch at the
    start), layer by layer from 0.9 above the spout, each layer's sites in
    (x, z) order, keeping the sites whose axis distance is at most the
    funnel's inner radius at that height (its top-rim radius above the rim)
    minus ``clearance``.  Even bodies are icospheres, odd ones knots.
    Random draws from default_rng(seed): the vertex noise of each icosphere
    (in body order), then one orientation quaternion per body.
    """
    bank = bank or ShapeBank(resolution=64)
    prm = dict(time_step=0.005, max_iterations=50, absolute_tolerance=0.0, relative_tolerance=0.0,
               coloring_seed=seed)
    prm.update(params or {})
    eng = new_engine(bank, prm, backend)
    set_friction(eng, 0.4)
    API.create_gravity_force(eng, "gravity", 9.81, np.array([0.0, 1.0, 0.0]))
    add_ground(eng, bank)
    rng = np.random.default_rng(seed)
    shapes = []
    for i in range(n):
        if i % 2 == 0:
            V, T = MESH.create_noisy_icosphere(rng, frequency)
            shapes.append(_unit_shape(bank, "noisy_ico_f%d_s%d_%d" % (frequency, seed, i), lambda V=V, T=T: (V, T)))
        else:
            shapes.append(_unit_shape(bank, "torus_knot_23", MESH.create_torus_knot))
    sites = []
    half = int(np.ceil(radius / spacing))
    k = 0
    while len(sites) < n:
        y = height + 0.9 + k * spacing
        rho = funnel_inner_radius(min(y, 2.0 * height), height, radius) - clearance
        for i in range(-half, half + 1):
            for j in range(-half, half + 1):
                x, z = i * spacing, j * spacing
                if np.hypot(x, z) <= rho and len(sites) < n:
                    sites.append((x, y, z))
        k += 1
    for b, r in enumerate(sites):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        add_body(eng, "body_%d" % b, shapes[b], np.array(r), q=q)
    add_body(eng, "funnel", funnel_shape(bank, height, radius), np.array([0.0, 1.5 * height, 0.0]),
             q=Q.Rx(np.pi), fixed=True, forces=())
    return eng
