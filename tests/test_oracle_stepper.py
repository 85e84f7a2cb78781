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


def sunk_sphere(bank, depth, params=None):
    eng = scenes.new_engine(bank, params, backend=None)
    API.create_gravity_force(eng, "gravity", 9.81, np.array([0.0, 1.0, 0.0]))
    scenes.add_ground(eng, bank)
    sph = bank.get("sphere_shape", lambda: MESH.create_sphere(0.5, 16, 16))
    scenes.add_body(eng, "s", sph, np.array([0.0, 0.5 - depth, 0.0]), q=np.array([1.0, 0.0, 0.0, 0.0]))
    return OS.from_engine(eng)


def test_post_stabilization_lifts_by_the_deepest_penetration(oracle_bank32):
    """A sphere sunk into the ground: its deepest contact is the bottom pole,
    whose arm passes through the centre of mass along the ground's normal, and
    every other contact is shallower.  The post-stabilisation problem's
    solution is therefore a pure lift by exactly that penetration, no turn."""
    import oracle.collision as OC
    import oracle.problem as OP
    sc = sunk_sphere(oracle_bank32, 0.02, params=dict(max_iterations=200, absolute_tolerance=0.0,
                                                      relative_tolerance=0.0))
    contacts = OC.detect(sc)
    assert contacts.K > 1
    Winv, f_ext, du_ext = OP.body_mass_blocks(sc)
    P = OP.assemble(sc, contacts, Winv, du_ext, 0.01)
    groups, _ = OST.sweep_groups(sc, contacts)
    r0, q0 = sc.r.copy(), sc.q.copy()
    deepest = -float(contacts.gap.min())
    assert deepest == pytest.approx(0.02, abs=1e-9)
    res = OST.post_stabilize(sc, P, groups, sc.params)
    assert res is not None
    assert sc.r[1] - r0[1] == pytest.approx([0.0, deepest, 0.0], abs=1e-9)
    assert np.allclose(sc.q[1], q0[1], atol=1e-9)
    assert np.array_equal(sc.r[0], r0[0])          # the fixed ground does not move
    # friction rows stay zero, normal impulses non-negative
    assert np.all(res["problem"].sol[:, 1:] == 0.0) and np.all(res["problem"].sol[:, 0] >= 0.0)


def test_post_stabilization_skips_shallow_contacts(oracle_bank32):
    """No contact deeper than min_gap_value: nothing is solved or moved."""
    import oracle.collision as OC
    import oracle.problem as OP
    sc = sunk_sphere(oracle_bank32, 0.0005)
    contacts = OC.detect(sc)
    assert contacts.K > 0 and contacts.gap.min() > -sc.params["min_gap_value"]
    Winv, f_ext, du_ext = OP.body_mass_blocks(sc)
    P = OP.assemble(sc, contacts, Winv, du_ext, 0.01)
    groups, _ = OST.sweep_groups(sc, contacts)
    r0 = sc.r.copy()
    assert OST.post_stabilize(sc, P, groups, sc.params) is None
    assert np.array_equal(sc.r, r0)


def test_energies_closed_forms():
    # state_storage.py:188-237: KE = sum over free bodies of 1/2 (m v.v + w^T I_w w)
    # with I_w = R diag(I) R^T, PE = sum of m g (up . r) for the summed gravity
    # vector; fixed bodies count for nothing.  The rotational term is checked in
    # body coordinates, 1/2 sum_i I_i (R^T w)_i^2, and for a spin about a
    # principal axis, 1/2 I_1 |w|^2.
    from oracle import quat as Q
    rng = np.random.default_rng(5)
    sc = OS.Scene()
    sc.r = np.array([[0.0, -3.0, 0.0], [1.0, 2.5, -2.0], [0.5, 7.0, 3.0]])
    qa = rng.normal(size=4)
    qa /= np.linalg.norm(qa)
    sc.q = np.array([[1.0, 0.0, 0.0, 0.0], qa, [np.cos(0.2), 0.0, np.sin(0.2), 0.0]])
    sc.mass = np.array([50.0, 2.0, 3.0])
    sc.inertia = np.array([[1.0, 1.0, 1.0], [0.9, 0.5, 0.2], [0.3, 0.3, 0.3]])
    sc.fixed = np.array([True, False, False])
    Ra = Q.to_matrix(qa)
    sc.v = np.array([[5.0, 5.0, 5.0], [0.3, -1.2, 0.7], [0.0, 0.0, 0.0]])
    w_body = np.array([0.4, -2.0, 1.1])
    sc.w = np.array([[9.0, 9.0, 9.0], Ra @ w_body, [0.0, 0.0, 0.0]])
    sc.gravity = [(9.81, np.array([0.0, 1.0, 0.0])), (1.0, np.array([1.0, 0.0, 0.0]))]
    kin, pot = OST.energies(sc)
    ke = 0.5 * 2.0 * (0.3 ** 2 + 1.2 ** 2 + 0.7 ** 2) + 0.5 * np.sum(sc.inertia[1] * w_body ** 2)
    G = np.array([1.0, 9.81, 0.0])
    pe = 2.0 * G.dot(sc.r[1]) + 3.0 * G.dot(sc.r[2])
    assert kin == pytest.approx(ke, rel=1e-13) and pot == pytest.approx(pe, rel=1e-13)
    # spin about the first principal axis only
    sc.v[1] = 0.0
    sc.w[1] = Ra[:, 0] * 3.0
    kin, _ = OST.energies(sc)
    assert kin == pytest.approx(0.5 * 0.9 * 9.0, rel=1e-13)
    # no gravity force: no potential energy
    sc.gravity = []
    assert OST.energies(sc)[1] == 0.0
