"""The body-pair reach bound of the GPU broad phase, checked with the oracle.

librbh drops a body pair before comparing its triangles when the bodies'
origins are farther apart than reach_A + reach_B + envelope + max(hd_A, hd_B)
(reach: the largest vertex distance from the body-frame origin; hd: the SDF
cell diagonal).  The argument: a contact needs phi_B(x) <= envelope at a point
x of a triangle of A inside B's grid (collision_detection.py:172-176), and the
trilinear interpolant of exact signed distances is at least the signed
distance minus the cell diagonal, so such pairs give no contact.  Here the
oracle's collision detection (all pairs, as the reference) must find no
contact for pairs placed just beyond that distance, in random directions and
orientations, and must find contacts once the pair moves closer."""
import numpy as np
import pytest

import oracle.collision as OC
import oracle.scene as OS
import rainbow_amd.api as API
from oracle.sdf import create_signed_distance
from rainbow_amd import math3 as Q
from rainbow_amd import mesh as MESH
from rainbow_amd import scenes


@pytest.fixture(scope="module")
def bank():
    return scenes.ShapeBank(resolution=16, sdf_builder=create_signed_distance)


def reach_and_diag(shape):
    V = np.asarray(shape.mesh.V, dtype=np.float64)
    g = shape.grid
    h = (np.asarray(g.max_coord) - np.asarray(g.min_coord)) / (np.array([g.I, g.J, g.K]) - 1)
    return float(np.sqrt((V * V).sum(axis=1)).max()), float(np.sqrt((h * h).sum()))


def pair_scene(bank, kind, dist, direction, qa, qb):
    eng = scenes.new_engine(bank, dict(time_step=0.01), backend=None)
    scenes.set_friction(eng, 0.5)
    make = (lambda: MESH.create_sphere(0.5, 16, 16)) if kind == "sphere" else (lambda: MESH.create_box(1.0, 1.0, 1.0))
    sh = bank.get(kind + "_shape", make)
    scenes.add_body(eng, "a", sh, np.zeros(3), q=qa, forces=())
    scenes.add_body(eng, "b", sh, dist * direction, q=qb, forces=())
    return eng


def random_unit(rng, n):
    x = rng.normal(size=n)
    return x / np.linalg.norm(x)


@pytest.mark.parametrize("kind", ["sphere", "box"])
def test_pairs_beyond_reach_give_no_contact(bank, kind):
    rng = np.random.default_rng(11)
    probe = pair_scene(bank, kind, 5.0, np.array([1.0, 0.0, 0.0]), Q.identity(), Q.identity())
    bodies = list(probe.bodies.values())
    reach, hd = reach_and_diag(bodies[0].shape)
    env = probe.params.envelope
    limit = 2 * reach + env + hd
    closest = []
    for trial in range(8):
        u = random_unit(rng, 3)
        qa, qb = random_unit(rng, 4), random_unit(rng, 4)
        far = OC.detect(OS.from_engine(pair_scene(bank, kind, limit * (1.0 + 1e-9), u, qa, qb)), {})
        assert far.K == 0, (trial, far.K)
        # the bound is not vacuous: moving the pair closer from the limit, contacts appear
        eng = pair_scene(bank, kind, limit, u, qa, qb)
        for d in np.arange(limit, 0.0, -0.01):
            API.set_position(eng, "b", d * u, use_model_frame=True)
            if OC.detect(OS.from_engine(eng), {}).K:
                closest.append(limit - d)
                break
    assert len(closest) == 8
    print(kind, "limit", limit, "first contact within", max(closest), "of it")


def test_contact_triangles_lie_within_reach(bank):
    """The per-triangle form of the bound: every contact the oracle finds in a
    pile of boxes and spheres comes from a triangle whose box (grown by the
    envelope) is within reach + envelope + cell diagonal of the SDF owner's
    origin, so skipping triangles beyond it loses no contact."""
    rng = np.random.default_rng(5)
    eng = scenes.new_engine(bank, dict(time_step=0.01), backend=None)
    scenes.set_friction(eng, 0.5)
    shapes = [bank.get("sphere_shape", lambda: MESH.create_sphere(0.5, 16, 16)),
              bank.get("box_shape", lambda: MESH.create_box(1.0, 1.0, 1.0))]
    for i in range(14):
        scenes.add_body(eng, "b%d" % i, shapes[i % 2], rng.uniform(-1.0, 1.0, 3), q=random_unit(rng, 4), forces=())
    sc = OS.from_engine(eng)
    c = OC.detect(sc, {})
    assert c.K > 20
    bodies = list(eng.bodies.values())
    env = eng.params.envelope
    for a, b, t in zip(c.ia, c.ib, c.tri):
        reach, hd = reach_and_diag(bodies[a].shape)
        lim = reach + hd + env
        V = np.asarray(bodies[b].shape.mesh.V, dtype=np.float64)
        T = np.asarray(bodies[b].shape.mesh.T)
        R = Q.to_matrix(sc.q[b])
        P = (R @ V[T[t]].T).T + sc.r[b]
        lo, hi = P.min(axis=0) - env, P.max(axis=0) + env
        gap = np.maximum(np.maximum(lo - sc.r[a], sc.r[a] - hi), 0.0)
        assert np.sqrt((gap * gap).sum()) <= lim, (a, b, t)
