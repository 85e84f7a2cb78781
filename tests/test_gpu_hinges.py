"""GPU: hinge joints (constraints/hinges.py) solved with the contacts.

* FP64 build against the oracle, step by step: a swinging chain of cubes
  (hinges only: the hinge solve's own stopping tests and residual log), two
  boxes resting on the ground joined by a hinge along their common top edge
  (contacts and hinges in one solve: the stopping tests on the larger merit),
  with and without pre-stabilisation, under the serial and the coloured
  Gauss-Seidel schemes and both solve layouts.
* FP32 build: the reference's hinge scene (five cubes of edge 10) swings for
  100 steps with its joint anchors staying together; a batch of scenes with
  and without hinges steps bitwise like each scene alone.
"""
import numpy as np
import pytest

import oracle.scene as OS
import oracle.stepper as OST
import rainbow_amd.api as API
from rainbow_amd import math3 as Q
from rainbow_amd import mesh as MESH
from rainbow_amd import native, scenes
from rainbow_amd.stepper import engine_desc, gather_state, upload_hinges

from test_gpu_parity import assert_fp64_parity, gpu_step, resync

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bank():
    return scenes.ShapeBank(resolution=32)


def hinged_pair(bank, params):
    """Two unit boxes on the ground, 0.02 apart, joined by a hinge about z
    through the middle of their facing top edges."""
    prm = dict(time_step=0.01, max_iterations=40, absolute_tolerance=0.0, relative_tolerance=0.0)
    prm.update(params)
    eng = scenes.new_engine(bank, prm)
    scenes.set_friction(eng, 0.5)
    API.create_gravity_force(eng, "gravity", 9.81, np.array([0.0, 1.0, 0.0]))
    scenes.add_ground(eng, bank)
    box = bank.get("box_shape", lambda: MESH.create_box(1.0, 1.0, 1.0))
    scenes.add_body(eng, "a", box, np.array([-0.51, 0.52, 0.0]), q=Q.identity())
    scenes.add_body(eng, "b", box, np.array([0.51, 0.52, 0.0]), q=Q.Ry(0.05))
    API.create_hinge(eng, "ab")
    API.set_hinge(eng, "ab", "a", "b", np.array([0.0, 1.02, 0.0]), np.array([0.0, 0.0, 1.0]))
    return eng


def check_step(eng, sc, dt):
    out = OST.step(sc, dt)
    c, state, lg = gpu_step(eng, dt)
    oc = out["contacts"]
    if oc.K:
        assert_fp64_parity(c, state, lg, out, sc)
    else:
        assert len(c["ia"]) == 0
        r, q, v, w = state
        assert np.abs(r - sc.r).max() <= 1e-10 and np.abs(v - sc.v).max() <= 1e-9
        assert np.abs(q - sc.q).max() <= 1e-10 and np.abs(w - sc.w).max() <= 1e-8
        assert lg["iterations"] == out["log"]["iterations"] and lg["exitcode"] == out["log"]["exitcode"]
        assert lg["rejects"] == out["log"]["rejects"]
        ro, rg = out["log"]["residuals"], lg["residuals"]
        fin = np.isfinite(ro)
        assert np.array_equal(fin, np.isfinite(rg))
        assert np.allclose(rg[fin], ro[fin], rtol=1e-8, atol=1e-14)
    resync(eng, sc)
    return oc.K, out


@pytest.mark.parametrize("tol", [(0.0, 0.0), (1e-5, 1e-3)])
def test_chain_fp64(bank, tol):
    eng = scenes.cube_hinge_chain(bank, n=4, edge=1.0,
                                  params=dict(precision="fp64", max_iterations=60, absolute_tolerance=tol[0],
                                              relative_tolerance=tol[1]))
    sc = OS.from_engine(eng)
    stops = set()
    for _ in range(5):
        k, out = check_step(eng, sc, 0.01)
        assert k == 0
        stops.add(out["log"]["exitcode"])
    if tol[0] > 0:
        assert stops & {1, 2}, stops


@pytest.mark.parametrize("scheme,flags,pre,post", [
    ("parallel_gauss_seidel", native.STEP_FORCE_MULTI, False, False),
    ("parallel_gauss_seidel", native.STEP_FORCE_SINGLE, True, False),
    ("gauss_seidel", 0, False, False),
    ("parallel_gauss_seidel", native.STEP_FORCE_SINGLE, False, True),   # post-stabilisation: contacts only
    ("parallel_jacobi", 0, False, False),          # Jacobi contacts beside the serial hinge sweep
    ("parallel_jacobi_hybrid", 0, True, False),
])
def test_hinged_pair_on_ground_fp64(bank, scheme, flags, pre, post):
    eng = hinged_pair(bank, dict(precision="fp64", proximal_solver=scheme, use_pre_stabilization=pre,
                                 use_post_stabilization=post, absolute_tolerance=1e-7, relative_tolerance=1e-4))
    if post:
        for b in eng.bodies.values():
            if not b.is_fixed:
                b.r = b.r - np.array([0.0, 0.03, 0.0])    # sunk below min_gap_value: the post solve runs
    sc = OS.from_engine(eng)
    seen = 0
    for _ in range(4):
        out = OST.step(sc, 0.01)
        st = eng.stepper
        st._prepare(eng)
        st.ctx.step([st.scene], 1, 0.01, flags)
        c = st.ctx.get_contacts(st.scene)
        r, q, v, w = st.ctx.get_state(st.scene, len(eng.bodies))
        lg = st.ctx.get_log(st.scene, eng.params.max_iterations)
        for k, b in enumerate(eng.bodies.values()):
            b.r, b.q, b.v, b.w = r[k].copy(), q[k].copy(), v[k].copy(), w[k].copy()
        eng.step_index += 1
        assert out["contacts"].K > 0
        assert_fp64_parity(c, (r, q, v, w), lg, out, sc)
        if post and "post" in out:
            assert lg["post_iterations"] == out["log"]["post_stabilization_iterations"]
            assert lg["post_exitcode"] == out["log"]["post_stabilization_exitcode"]
        seen += 1
        resync(eng, sc)
    assert seen == 4


def anchor_gap(eng):
    gaps = []
    for h in eng.hinges.values():
        p, c = h.parent, h.child
        gaps.append(np.linalg.norm((Q.rotate(p.q, h.arm_p) + p.r) - (Q.rotate(c.q, h.arm_c) + c.r)))
    return max(gaps)


def test_reference_chain_fp32(bank):
    """100 steps of the reference's chain: the FP64 oracle's joints drift apart
    by 0.0067 (velocity-level constraints, no pre-stabilisation) while the tip
    falls by 4.95; the FP32 product path follows it."""
    eng = scenes.cube_hinge_chain(bank, n=5, edge=10.0, params=dict(precision="fp32", max_iterations=100))
    sc = OS.from_engine(eng)
    start = {n: b.r.copy() for n, b in eng.bodies.items()}
    for _ in range(10):
        API.simulate(eng, 0.1)
        assert all(np.all(np.isfinite(b.r)) for b in eng.bodies.values())
    for _ in range(100):
        OST.step(sc, 0.01)
    assert anchor_gap(eng) < 0.02
    tip = eng.bodies["link_4"]
    assert tip.r[1] < start["link_4"][1] - 4.0      # the free end has swung down
    got = np.array([b.r for b in eng.bodies.values()])
    assert np.abs(got - sc.r).max() < 0.05, np.abs(got - sc.r).max()


def test_hinge_only_profile_keys(bank):
    """steppers.py:78-83 and gauss_seidel_solver.py:75-101: a step with hinges
    and no contacts still logs the solve (residuals, iterations, exitcode)."""
    eng = scenes.cube_hinge_chain(bank, n=3, edge=1.0, params=dict(precision="fp32", max_iterations=30))
    API.simulate(eng, 0.02, profiling_on=True)
    lg = API.get_log(eng)
    assert len(lg) == 2
    for d in lg:
        assert d["contact_points"] == 0
        assert {"residuals", "iterations", "exitcode"} <= set(d)
        assert len(d["residuals"]) == 30 and np.isfinite(d["residuals"][0])


def test_set_hinges_rejects_bad_input(bank):
    eng = scenes.cube_hinge_chain(bank, n=3, edge=1.0, params=dict(precision="fp32"))
    ctx = native.Context(0, "fp32")
    try:
        desc, keep = engine_desc(eng, ctx)
        sid = ctx.scene_create(desc)
        z = np.zeros((1, 3))
        with pytest.raises(native.RbhError):
            ctx.set_hinges(sid, [1], [1], z, z, z, z)          # parent and child are one body
        with pytest.raises(native.RbhError):
            ctx.set_hinges(sid, [0], [7], z, z, z, z)          # no body 7 in a 3-body scene
        n = 1025
        zz = np.zeros((n, 3))
        with pytest.raises(native.RbhError):
            ctx.set_hinges(sid, [0] * n, [1] * n, zz, zz, zz, zz)   # more hinges than a scene holds
        ctx.set_hinges(sid, [0], [1], z, z, z, z)                # a valid one is accepted
        ctx.set_hinges(sid, [], [], zz[:0], zz[:0], zz[:0], zz[:0])   # and removed again
    finally:
        ctx.close()


def test_graph_replay_with_hinges(bank):
    """hipGraph replays of a batch with hinged scenes equal eager steps,
    replay after replay, with eager work and synchronisations between."""
    def run(graph):
        ctx = native.Context(0, "fp32")
        try:
            engines = [hinged_pair(bank, dict(precision="fp32")),
                       scenes.cube_hinge_chain(bank, n=4, edge=1.0, params=dict(precision="fp32"))]
            ids = []
            for e in engines:
                desc, keep = engine_desc(e, ctx)
                sid = ctx.scene_create(desc)
                ctx.set_state(sid, *gather_state(list(e.bodies.values())))
                upload_hinges(ctx, sid, e)
                ids.append(sid)
            snaps = []
            for _ in range(5):
                ctx.step(ids, 1, 0.01, native.STEP_GRAPH if graph else 0)
                ctx.sync()
                snaps.append([list(ctx.get_state(sid, len(e.bodies))) for sid, e in zip(ids, engines)])
            return snaps
        finally:
            ctx.close()

    a, b = run(False), run(True)
    for sa, sb in zip(a, b):
        for x, y in zip(sa, sb):
            for u, v in zip(x, y):
                assert np.array_equal(u, v)


def test_batch_with_hinges_bitwise(bank):
    def make():
        return [hinged_pair(bank, dict(precision="fp32")),
                scenes.cube_hinge_chain(bank, n=3, edge=1.0, params=dict(precision="fp32")),
                scenes.mixed_drop(3300, bank, nmax=48, params=dict(precision="fp32"))]

    def run(single):
        ctx = native.Context(0, "fp32")
        try:
            engines = make()
            ids = []
            for e in engines:
                desc, keep = engine_desc(e, ctx)
                sid = ctx.scene_create(desc)
                ctx.set_state(sid, *gather_state(list(e.bodies.values())))
                if e.hinges:
                    upload_hinges(ctx, sid, e)
                ids.append(sid)
            snaps = []
            for _ in range(6):
                if single:
                    for sid in ids:
                        ctx.step([sid], 1, 0.01)
                else:
                    ctx.step(ids, 1, 0.01)
                snaps.append([list(ctx.get_state(sid, len(e.bodies))) for sid, e in zip(ids, engines)])
            return snaps
        finally:
            ctx.close()

    a, b = run(False), run(True)
    for sa, sb in zip(a, b):
        for x, y in zip(sa, sb):
            for u, v in zip(x, y):
                assert np.array_equal(u, v)
