"""GPU: the funnel stress workload (noisy icospheres and torus knots in a funnel).

* A small funnel scene (icospheres of frequency 4, light knots, 32^3 SDFs)
  in the FP64 build: contact sets, colourings, impulses, states and residual
  histories equal to the oracle's, as for the box and sphere scenes.
* The full workload (256 bodies, 2.5 M triangles, 64^3 SDFs) in the FP32
  product build: impulses inside the friction cone with x_n >= 0 and finite
  states every step; and, for bodies resting on the funnel after 60 steps,
  the contact set of one step equal to the FP64 oracle's collision
  detection run on that body and the funnel alone (from the same state),
  up to candidates within EPS_MARGIN of a decision boundary.
"""
import types

import numpy as np
import pytest

import oracle.collision as OC
import oracle.scene as OS
import oracle.stepper as OST
import rainbow_amd.api as API
from rainbow_amd import mesh as MESH
from rainbow_amd import scenes

from test_gpu_parity import EPS_MARGIN, assert_fp64_parity, gpu_step, resync

pytestmark = pytest.mark.gpu


def mini_funnel(bank, precision, seed=11, height=4.0, radius=4.0):
    """Two noisy icospheres and two light knots placed on the inner wall of a
    small funnel (centres 0.45 from the wall), one icosphere in the spout."""
    eng = scenes.new_engine(bank, dict(time_step=0.01, max_iterations=40, precision=precision,
                                       absolute_tolerance=0.0, relative_tolerance=0.0, coloring_seed=seed))
    scenes.set_friction(eng, 0.4)
    API.create_gravity_force(eng, "gravity", 9.81, np.array([0.0, 1.0, 0.0]))
    scenes.add_ground(eng, bank)
    rng = np.random.default_rng(seed)
    meshes = [MESH.create_noisy_icosphere(rng, 4) for _ in range(3)]
    ico = [scenes._unit_shape(bank, "mini_ico_%d_%d" % (seed, i), lambda m=m: m) for i, m in enumerate(meshes)]
    knot = scenes._unit_shape(bank, "mini_knot", lambda: MESH.create_torus_knot(segments=48, sides=8))
    slope = 2.0 * radius * 0.33 / height                # d(inner radius)/dy of the placed funnel
    cos_a = 1.0 / np.sqrt(1.0 + slope * slope)
    for k in range(4):
        th = k * np.pi / 2.0
        y = 1.4 * height + 0.5 * (k % 2)
        rho = scenes.funnel_inner_radius(y, height, radius) - 0.45 / cos_a
        q = rng.normal(size=4)
        scenes.add_body(eng, "b%d" % k, ico[k // 2] if k % 2 == 0 else knot,
                        np.array([rho * np.cos(th), y, rho * np.sin(th)]), q=q / np.linalg.norm(q))
    scenes.add_body(eng, "spout", ico[2], np.array([0.0, height + 0.4, 0.0]), q=np.array([1.0, 0.0, 0.0, 0.0]))
    scenes.add_body(eng, "funnel", scenes.funnel_shape(bank, height, radius),
                    np.array([0.0, 1.5 * height, 0.0]), q=np.array([0.0, 1.0, 0.0, 0.0]), fixed=True, forces=())
    return eng


def test_mini_funnel_fp64():
    bank = scenes.ShapeBank(resolution=32)
    eng = mini_funnel(bank, "fp64")
    sc = OS.from_engine(eng)
    total = exact = 0
    funnel = len(eng.bodies) - 1
    on_funnel = 0
    for _ in range(3):
        m = {}
        out = OST.step(sc, 0.01, m)
        c, state, lg = gpu_step(eng, 0.01)
        exact += assert_fp64_parity(c, state, lg, out, sc, m)
        total += len(c["ia"])
        on_funnel += int(np.sum((c["ia"] == funnel) | (c["ib"] == funnel)))
        resync(eng, sc)
    assert total > 0 and on_funnel > 0 and exact >= 2


def cone_ok(imp, mu):
    xn, f = imp[:, 0], imp[:, 1:4]
    return bool(np.all(xn >= 0.0)) and bool(np.all(np.linalg.norm(f, axis=1) <= mu * xn * (1 + 1e-5) + 1e-7))


def sub_scene(eng, names):
    """An engine-like view holding only the named bodies (and their shapes)."""
    bodies = {n: eng.bodies[n] for n in names}
    shapes = {}
    for b in bodies.values():
        shapes[b.shape.name] = b.shape
    return types.SimpleNamespace(bodies=bodies, shapes=shapes, surfaces_interactions=eng.surfaces_interactions,
                                 forces=eng.forces, params=eng.params, step_index=0)


def test_funnel_stress_fp32():
    bank = scenes.ShapeBank(resolution=64)
    eng = scenes.funnel_stress(bank=bank)
    names = list(eng.bodies.keys())
    funnel = names.index("funnel")
    assert len(names) == 258 and sum(len(b.shape.mesh.T) for b in eng.bodies.values()) > 2_500_000
    dt = eng.params.time_step
    seen = 0
    for step in range(60):
        c, state, lg = gpu_step(eng, dt)
        assert all(np.all(np.isfinite(a)) for a in state), step
        if len(c["ia"]):
            assert cone_ok(c["impulse"], 0.4), step
            assert np.isfinite(lg["natural_residual"])
        seen += len(c["ia"])
    assert seen > 0
    # one more step; its contacts come from the state the engine holds now
    before = {n: (b.r.copy(), b.q.copy()) for n, b in eng.bodies.items()}
    c, state, lg = gpu_step(eng, dt)
    on = (c["ia"] == funnel) | (c["ib"] == funnel)
    other = np.where(c["ia"] == funnel, c["ib"], c["ia"])[on]
    assert on.any(), "no body rests on the funnel after 60 steps"
    picked = [int(b) for b in np.unique(other)[:2]]
    for b in picked:
        sub_names = [names[b], "funnel"]
        view = sub_scene(eng, sub_names)
        for n in sub_names:
            view.bodies[n] = types.SimpleNamespace(**{**vars(eng.bodies[n]), "r": before[n][0], "q": before[n][1]})
        sc = OS.from_engine(view)
        margins = {}
        oc = OC.detect(sc, {}, margins)
        to_sub = {b: 0, funnel: 1}
        sel = on & ((c["ia"] == b) | (c["ib"] == b))
        kg = {(to_sub[int(x)], to_sub[int(y)], int(t)): i
              for i, (x, y, t) in enumerate(zip(c["ia"], c["ib"], c["tri"])) if sel[i]}
        ko = {(int(x), int(y), int(t)): i for i, (x, y, t) in enumerate(zip(oc.ia, oc.ib, oc.tri))}
        assert ko, names[b]
        diff = set(kg) ^ set(ko)
        bad = sorted((k, margins.get(k)) for k in diff if margins.get(k, np.inf) >= EPS_MARGIN)
        assert not bad, (names[b], bad[:5], len(kg), len(ko))
        for k in set(kg) & set(ko):
            if margins.get(k, np.inf) < EPS_MARGIN:
                continue
            i, j = kg[k], ko[k]
            assert np.allclose(c["p"][i], oc.p[j], atol=2e-4), (k, c["p"][i], oc.p[j])
            assert np.allclose(c["n"][i], oc.n[j], atol=2e-3), (k, c["n"][i], oc.n[j])
