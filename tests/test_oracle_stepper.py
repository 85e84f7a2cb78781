"""Oracle time step: closed forms for contact-free motion and a resting stack."""
import numpy as np
import pytest

import oracle.scene as OS
import oracle.stepper as OST
from rainbow_amd import scenes, mesh as MESH
import rainbow_amd.api as API


def single_body(bank, r, v=None, w=None, q=None, gravity=True):
    eng = scenes.new_engine(bank, backend=None)
    if gravity:
        API.create_gravity_force(eng, "gravity", 9.81, np.array([0.0, 1.0, 0.0]))
    box = bank.get("box_shape", lambda: MESH.create_box(1.0, 1.0, 1.0))
    scenes.add_body(eng, "b", box, np.asarray(r, float), q=q, v=v, w=w)
    return OS.from_engine(eng)


def test_free_fall_closed_form(oracle_bank32):
    sc = single_body(oracle_bank32, [0, 50, 0], v=[1.0, 2.0, -0.5])
    r0, v0 = sc.r[0].copy(), sc.v[0].copy()
    dt, n = 0.01, 20
    for _ in range(n):
        out = OST.step(sc, dt)
        assert out["contacts"].K == 0
    g = np.array([0, -9.81, 0])
    assert np.allclose(sc.v[0], v0 + n * dt * g, atol=1e-12)
    # semi-implicit Euler: r_n = r0 + dt sum_{k=1..n} v_k
    assert np.allclose(sc.r[0], r0 + dt * (n * v0 + dt * g * n * (n + 1) / 2), atol=1e-10)


def test_torque_free_principal_spin(oracle_bank32):
    # spin about a principal axis of a symmetric body: w constant, gyroscopic term zero
    sc = single_body(oracle_bank32, [0, 50, 0], w=[0.0, 3.0, 0.0], gravity=False)
    for _ in range(10):
        OST.step(sc, 0.01)
    assert np.allclose(sc.w[0], [0, 3.0, 0], atol=1e-12)
    ang = 2 * np.arctan2(np.linalg.norm(sc.q[0, 1:]), sc.q[0, 0])
    assert ang == pytest.approx(10 * 0.01 * 3.0, rel=2e-3)


def test_resting_box_impulse_balances_gravity(oracle_bank32):
    # a box resting on the ground (0.05 gap, inside the envelope): the normal
    # impulses cancel m g dt and the box barely moves
    eng = scenes.new_engine(oracle_bank32, dict(max_iterations=200, absolute_tolerance=0.0, relative_tolerance=0.0),
                            backend=None)
    scenes.set_friction(eng, 0.5)
    API.create_gravity_force(eng, "gravity", 9.81, np.array([0.0, 1.0, 0.0]))
    scenes.add_ground(eng, oracle_bank32)
    box = oracle_bank32.get("box_shape", lambda: MESH.create_box(1.0, 1.0, 1.0))
    scenes.add_body(eng, "box", box, np.array([0.0, 0.55, 0.0]), q=np.array([1.0, 0, 0, 0]))
    sc = OS.from_engine(eng)
    out = OST.step(sc, 0.01)
    P = out["problem"]
    assert P is not None and P.K >= 1
    # 200 sweeps with r-factor adaptation leave a small residual
    assert np.sum(P.sol[:, 0]) == pytest.approx(1.0 * 9.81 * 0.01, rel=1e-4)
    assert abs(sc.v[1, 1]) < 1e-4
    assert out["log"]["exitcode"] == 0 and out["log"]["iterations"] == 200


def test_simulate_substeps(oracle_bank32):
    sc = single_body(oracle_bank32, [0, 50, 0])
    logs = OST.simulate(sc, 0.025)
    assert [l["dt"] for l in logs] == pytest.approx([0.01, 0.01, 0.005])
    with pytest.raises(ValueError):
        OST.simulate(sc, 0.0)
