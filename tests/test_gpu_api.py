"""The drop-in API on the GPU against the oracle: api.simulate over unequal
substeps with profiling on (api.py:676-705, steppers.py:47-106), FP64 build,
including the options no other GPU test runs -- damping (types.py:239-256,
state_storage.py:108-156) and restitution (use_bounce, contacts.py:259-287)
-- and post-stabilisation with its prefixed log keys (steppers.py:15-36,
gauss_seidel_solver.py:29-34, 98-101).  Every substep's log (dt, contact
count, solver iterations, exit code and residual history, energies, largest
penetration, the post-stabilisation solve's record) and the end state must
equal the oracle's within the FP64 parity bounds."""
import numpy as np
import pytest

import oracle.scene as OS
import oracle.stepper as OST
import rainbow_amd.api as API
from rainbow_amd import scenes

pytestmark = pytest.mark.gpu


def sunk_boxes(damping, bounce, post):
    """Four unit boxes 0.35 into the ground (the smoke scene), moving down at 1 m/s."""
    bank = scenes.ShapeBank(resolution=32)
    eng = scenes.box_stack(bank, nx=2, ny=2, nz=1, params=dict(
        precision="fp64", max_iterations=20, proximal_solver="parallel_gauss_seidel", use_bounce=bounce,
        use_post_stabilization=post))
    if bounce:
        scenes.set_friction(eng, 0.5, epsilon=0.5)
    if damping:
        API.create_damping_force(eng, "damping", 0.3, 0.2)
    for b in eng.bodies.values():
        if b.is_fixed:
            continue
        b.r = b.r - np.array([0.0, 0.4, 0.0])
        API.set_velocity(eng, b.name, np.array([0.1, -1.0, 0.05]))
        API.set_spin(eng, b.name, np.array([0.0, 0.5, 0.2]))
        if damping:
            API.connect_force(eng, b.name, "damping")
    return eng


def close(a, b, tol):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    fin = np.isfinite(b)
    assert np.array_equal(np.isfinite(a), fin)
    assert np.allclose(a[fin], b[fin], rtol=tol, atol=1e-13), np.abs(a[fin] - b[fin]).max()


@pytest.mark.parametrize("damping,bounce,post", [(False, False, False), (True, True, False), (True, False, True)])
def test_simulate_fp64_matches_oracle(damping, bounce, post):
    eng = sunk_boxes(damping, bounce, post)
    sc = OS.from_engine(eng)
    T = 0.025                                  # substeps 0.01, 0.01, 0.005 (api.py:689-693)
    logs_o = OST.simulate(sc, T)
    API.simulate(eng, T, profiling_on=True)
    logs = API.get_log(eng)
    assert len(logs) == len(logs_o) == 3
    post_runs = 0
    for lg, lo in zip(logs, logs_o):
        assert lg["dt"] == lo["dt"]
        assert lg["contact_points"] == lo["contact_points"] > 0
        assert lg["iterations"] == lo["iterations"] and lg["exitcode"] == lo["exitcode"]
        close(lg["residuals"], lo["residuals"], 1e-8)
        close(lg["kinetic_energy"], lo["kinetic_energy"], 1e-9)
        close(lg["potential_energy"], lo["potential_energy"], 1e-9)
        assert abs(lg["max_penetration"] - lo["max_penetration"]) <= 1e-12
        assert ("post_stabilization_iterations" in lg) == ("post_stabilization_iterations" in lo)
        if "post_stabilization_iterations" in lo:
            post_runs += 1
            assert lg["post_stabilization_iterations"] == lo["post_stabilization_iterations"]
            assert lg["post_stabilization_exitcode"] == lo["post_stabilization_exitcode"]
            close(lg["post_stabilization_residuals"], lo["post_stabilization_residuals"], 1e-8)
            assert lg["post_stabilization_solver_time"] > 0.0
        # the collision-detection stages add up (collision_detection.py:293-318)
        parts = (lg["update_bvh"] + lg["narrow_phase"] + lg["contact_determination"]
                 + lg["contact_point_reduction"])
        assert lg["contact_point_reduction"] > 0.0 and abs(lg["collision_detection_time"] - parts) < 1e-9
    assert (post_runs >= 1) if post else post_runs == 0
    for k, b in enumerate(eng.bodies.values()):
        assert np.abs(b.r - sc.r[k]).max() <= 1e-10 and np.abs(b.q - sc.q[k]).max() <= 1e-10
        assert np.abs(b.v - sc.v[k]).max() <= 1e-9 and np.abs(b.w - sc.w[k]).max() <= 1e-8


def test_scene_churn_keeps_memory_bounded():
    """A server creates and destroys one scene per request: the interaction
    tables, residual records and arenas of destroyed scenes are reused, so
    device memory does not grow with the number of requests served."""
    from rainbow_amd import native
    from rainbow_amd.stepper import engine_desc, gather_state
    bank = scenes.ShapeBank(resolution=32)
    engs = [scenes.mixed_drop(3000 + s, bank, nmax=24) for s in range(4)]
    ctx = native.Context(0, "fp32")
    try:
        descs = [engine_desc(e, ctx) for e in engs]
        for (d, _), e in zip(descs, engs):      # one round: workspaces and tables at their sizes
            sid = ctx.scene_create(d)
            ctx.set_state(sid, *gather_state(list(e.bodies.values())))
            ctx.step([sid], 2, 0.0, 0)
            ctx.scene_destroy(sid)
        m0 = ctx.memory_usage()
        for _ in range(50):
            for (d, _), e in zip(descs, engs):
                sid = ctx.scene_create(d)
                ctx.set_state(sid, *gather_state(list(e.bodies.values())))
                ctx.step([sid], 1, 0.0, 0)
                ctx.scene_destroy(sid)
        assert ctx.memory_usage() == m0
    finally:
        ctx.close()
