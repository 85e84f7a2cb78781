"""Host-side API (the reference's api.py surface), meshes and mass properties."""
import numpy as np
import pytest

import rainbow_amd.api as API
from rainbow_amd import mass as MASS
from rainbow_amd import math3 as Q
from rainbow_amd import mesh as MESH


def test_engine_without_stepper_cannot_simulate():
    eng = API.Engine()
    with pytest.raises(ValueError):
        API.simulate(eng, 0.1)
    eng = API.create_engine(backend=None)
    with pytest.raises(ValueError):
        API.simulate(eng, 1.0)
    with pytest.raises(ValueError):
        API.simulate(API.create_engine(backend=None), -1.0)


def test_api_errors_and_names():
    eng = API.create_engine(backend=None)
    API.create_rigid_body(eng, "a")
    with pytest.raises(RuntimeError):
        API.create_rigid_body(eng, "a")
    with pytest.raises(RuntimeError):
        API.set_position(eng, "nope", np.zeros(3))
    with pytest.raises(RuntimeError):
        API.create_gravity_force(eng, "g", -1.0, np.array([0, 1, 0.0]))
    with pytest.raises(RuntimeError):
        API.create_damping_force(eng, "d", 0.0, 1.0)
    with pytest.raises(RuntimeError):
        API.create_surfaces_interaction(eng, "default", "default", 0.0, [0.5, 0.5, 0.5])
    with pytest.raises(RuntimeError):
        API.create_surfaces_interaction(eng, "a", "b", -0.1, [0.5, 0.5, 0.5])
    with pytest.raises(RuntimeError):
        API.create_surfaces_interaction(eng, "a", "b", 0.1, [0.5, -0.5, 0.5])
    API.create_surfaces_interaction(eng, "b", "a", 0.2, np.array([0.1, 0.2, 0.3]))
    assert eng.surfaces_interactions.get_interaction("a", "b").epsilon == 0.2
    with pytest.raises(RuntimeError):
        API.set_body_material(eng, "a", "unknown")
    API.set_body_material(eng, "a", "b")
    with pytest.raises(RuntimeError):
        API.set_body_type(eng, "a", "floating")
    API.set_body_type(eng, "a", "fixed")
    assert API.is_fixed_body(eng, "a")
    with pytest.raises(RuntimeError):
        API.set_mass_properties(eng, "a", 1.0)      # no shape
    n = API.generate_unique_name("x y.z:w-v")
    assert all(c not in n for c in " .:-")


def test_box_mass_properties():
    # unit_tests/test_mass.py:33-130: a 4 x 5 x 6 box has mass 120 and body
    # moments exactly 610, 520, 410 (largest first) under translations
    # (assertEqual) and within 1e-7 (assertAlmostEqual) under rotations
    w, h, d = 4.0, 5.0, 6.0
    I = 120.0 / 12.0 * np.array([h * h + d * d, w * w + d * d, w * w + h * h])
    assert I.tolist() == [610.0, 520.0, 410.0]
    for xf in ([0.0, 1.0, 1.0], [2.0, 1.0, 1.0], [2.0, 10.0, -1.0], Q.Rx(0.3 * np.pi), Q.Ry(0.3 * np.pi),
               Q.Rz(0.3 * np.pi)):
        V, T = MESH.create_box(w, h, d)
        V = V + np.array(xf) if len(xf) == 3 else Q.rotate_array(xf, V)
        mesh = API.create_mesh(V, T)
        p = MASS.compute_mass_properties(mesh.V, mesh.T, 1.0)
        r, q, m, Ib = MASS.xform_model_2_body_space(p)
        if len(xf) == 3:
            assert m == 120.0 and Ib.tolist() == I.tolist()
        else:
            assert np.abs(Ib - I).max() < 1e-7 and abs(m - 120.0) < 1e-9
        R = Q.to_matrix(q)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-12) and np.linalg.det(R) > 0


def test_mesh_tables():
    # surface_mesh.py:286-329 (box) and 85-171 (profile sweep counts)
    V, T = MESH.create_box(1.0, 1.0, 1.0)
    assert V.shape == (8, 3) and T.shape == (12, 3)
    assert T[0].tolist() == [0, 2, 3] and T[11].tolist() == [3, 6, 7]
    V, T = MESH.create_sphere(0.5, 16, 16)
    assert len(V) == (16 - 2) * 16 + 2 and len(T) == 2 * 16 * (16 - 2)
    assert np.allclose(np.linalg.norm(V, axis=1), 0.5)
    # closed: every edge is shared by exactly two triangles
    edges = {}
    for t in T:
        for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
            edges[tuple(sorted((a, b)))] = edges.get(tuple(sorted((a, b))), 0) + 1
    assert set(edges.values()) == {2}
    p = MASS.compute_mass_properties(V, T, 1.0)
    assert 0.5 < p.mass < 4 / 3 * np.pi * 0.125


def test_direction_of_most_variance_row_convention():
    # functions.py:40-48 returns vectors[argmax(values)], a row of the eigenvector
    # matrix (test_math_functions.py:102-115)
    rng = np.random.default_rng(0)
    P = rng.normal(size=(100, 3)) * np.array([3.0, 1.0, 0.5])
    mean = P.mean(axis=0)
    vals, vecs = np.linalg.eig(np.cov((P - mean).T))
    assert np.array_equal(MESH.direction_of_most_variance(P), vecs[np.argmax(vals)])


def test_axis_sort_permutes_consistently():
    V, T = MESH.create_sphere(0.5, 8, 8)
    m = API.create_mesh(V, T)
    tri_old = {tuple(sorted(map(tuple, np.round(V[t], 12)))) for t in T}
    tri_new = {tuple(sorted(map(tuple, np.round(m.V[t], 12)))) for t in m.T}
    assert tri_old == tri_new
    axis = MESH.direction_of_most_variance(V)
    d = m.V @ (axis / np.linalg.norm(axis))
    assert np.all(np.diff(d) >= -1e-15)
    assert np.all(np.diff(m.T.mean(axis=1)) >= 0)


def test_substeps_match_reference_loop():
    assert API.substeps(0.025, 0.01) == pytest.approx([0.01, 0.01, 0.005])
    assert len(API.substeps(1.0, 0.005)) >= 200


def test_smoke_scene_has_contacts():
    """The scene smoke() steps on the GPU gives the oracle contacts to compare."""
    import __graft_entry__ as G
    eng, ref, out = G.smoke_scene()
    assert out["contacts"].K > 0


def test_massless_free_body_error():
    # state_storage.py:169-171: the step refuses a free body without mass, by name
    # (raised before any device work)
    from oracle.sdf import create_signed_distance
    from rainbow_amd import scenes
    bank = scenes.ShapeBank(resolution=8, sdf_builder=create_signed_distance)
    eng = scenes.box_stack(bank, nx=1, ny=2, nz=1)
    name = list(eng.bodies)[-1]
    eng.bodies[name].mass = 0.0
    with pytest.raises(ValueError, match="^Body {} has no mass.$".format(name)):
        eng.stepper.step(0.01, eng, False)


def test_host_energies_equal_oracle(oracle_bank32):
    # the log's kinetic/potential energies (state_storage.py:188-237) from the
    # product's host code against the oracle's on random states
    import oracle.scene as OS
    import oracle.stepper as OST
    from rainbow_amd import scenes
    from rainbow_amd.stepper import energies, gather_state
    eng = scenes.box_stack(oracle_bank32, nx=2, ny=2, nz=2, backend=None)
    rng = np.random.default_rng(2)
    bodies = list(eng.bodies.values())
    for b in bodies:
        if not b.is_fixed:
            q = rng.normal(size=4)
            b.q, b.v, b.w = q / np.linalg.norm(q), rng.normal(size=3), rng.normal(size=3)
    kin, pot = energies(eng, bodies, *gather_state(bodies))
    ko, po = OST.energies(OS.from_engine(eng))
    assert kin == pytest.approx(ko, rel=1e-12) and pot == pytest.approx(po, rel=1e-12)
