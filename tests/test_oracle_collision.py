"""Oracle narrow phase: candidate set, Frank-Wolfe, contact reduction."""
import numpy as np
import pytest

import oracle.scene as OS
from oracle import collision as CD
from oracle import quat as Q
from rainbow_amd import scenes


def random_scene(bank, n=6, seed=0, spread=1.6):
    eng = scenes.new_engine(bank, backend=None)
    scenes.set_friction(eng, 0.4)
    import rainbow_amd.api as API
    API.create_gravity_force(eng, "gravity", 9.81, np.array([0.0, 1.0, 0.0]))
    scenes.add_ground(eng, bank)
    box = bank.get("box_shape", lambda: __import__("rainbow_amd.mesh", fromlist=["x"]).create_box(1.0, 1.0, 1.0))
    sph = bank.get("sphere_shape", lambda: __import__("rainbow_amd.mesh", fromlist=["x"]).create_sphere(0.5, 16, 16))
    rng = np.random.default_rng(seed)
    for i in range(n):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        r = np.array([rng.uniform(-spread, spread), rng.uniform(0.2, 1.2), rng.uniform(-spread, spread)])
        scenes.add_body(eng, "b%d" % i, box if i % 2 else sph, r, q=q)
    return eng


def test_candidates_equal_pairwise_loop(oracle_bank32):
    sc = OS.from_engine(random_scene(oracle_bank32, 5, seed=4, spread=0.8))
    env = sc.params["envelope"]
    boxes = [CD.triangle_aabbs(CD.world_vertices(sc, b), sc.shapes[sc.shape[b]].T, env) for b in range(sc.N)]
    nonempty = 0
    for a in range(sc.N):
        for b in range(a + 1, sc.N):
            ca, cb, cnt = CD.candidates(boxes[a], boxes[b])
            sa, sb, c2 = set(), set(), 0
            (la, ha), (lb, hb) = boxes[a], boxes[b]
            for i in range(len(la)):
                for j in range(len(lb)):
                    if not (np.any(ha[i] < lb[j]) or np.any(hb[j] < la[i])):
                        sa.add(i); sb.add(j); c2 += 1
            assert set(ca.tolist()) == sa and set(cb.tolist()) == sb and cnt == c2
            assert list(ca) == sorted(ca) and list(cb) == sorted(cb)
            nonempty += cnt > 0
    assert nonempty > 0


def test_frank_wolfe_vectorised_equals_scalar(oracle_bank32):
    sc = OS.from_engine(random_scene(oracle_bank32, 6, seed=5, spread=0.7))
    checked = 0
    for a in range(sc.N):
        for b in range(sc.N):
            if a == b:
                continue
            tris = np.arange(len(sc.shapes[sc.shape[a]].T))[:80]
            pw, nw, gap, tt = CD.frank_wolfe_contacts(sc, a, b, tris)
            got = {int(t): (p, n, g) for p, n, g, t in zip(pw, nw, gap, tt)}
            for t in tris:
                ref = CD.fw_scalar(sc, a, b, int(t))
                assert (ref is None) == (int(t) not in got)
                if ref is not None:
                    p, n, g = got[int(t)]
                    assert np.allclose(p, ref[0], atol=1e-13) and np.allclose(n, ref[1], atol=1e-13)
                    assert g == pytest.approx(ref[2], abs=1e-13)
                    checked += 1
    assert checked > 0


def test_box_on_ground_vertex_contacts(oracle_bank32):
    # a unit box resting 0.05 above the ground: the first FW step (alpha = 1)
    # lands on a vertex and the lowest corner never loses the argmin, so every
    # contact sits on a bottom corner with gap 0.05 and normal +y.  Which
    # corners appear depends on argmin ties between equal-height corners.
    eng = scenes.new_engine(oracle_bank32, backend=None)
    scenes.add_ground(eng, oracle_bank32)
    box = oracle_bank32.get("box_shape", lambda: __import__("rainbow_amd.mesh", fromlist=["x"]).create_box(1.0, 1.0, 1.0))
    scenes.add_body(eng, "box", box, np.array([0.0, 0.55, 0.0]), q=np.array([1.0, 0, 0, 0]))
    sc = OS.from_engine(eng)
    c = CD.detect(sc)
    assert 1 <= c.K <= 4
    assert np.all(c.ia == 0) and np.all(c.ib == 1)      # SDF owner = ground, triangle owner = box
    corners = {tuple(np.round(p, 9)) for p in c.p}
    assert len(corners) == c.K
    assert corners <= {(x, 0.05, z) for x in (-0.5, 0.5) for z in (-0.5, 0.5)}
    assert np.allclose(c.n, [0, 1, 0], atol=1e-9)
    assert np.allclose(c.gap, 0.05, atol=1e-9)


def test_reduction_equals_quadratic_loop():
    rng = np.random.default_rng(6)
    c = CD.Contacts()
    K = 60
    c.ia = rng.integers(0, 3, K)
    c.ib = (c.ia + 1 + rng.integers(0, 2, K)) % 4
    base = rng.normal(size=(8, 3))
    c.p = base[rng.integers(0, 8, K)]
    c.n = np.tile([0, 1.0, 0], (K, 1))
    c.gap = np.zeros(K)
    c.tri = np.arange(K)
    r = CD.reduce_contacts(c)
    keep = []
    for k in range(K):
        if not any({c.ia[k], c.ib[k]} == {c.ia[j], c.ib[j]} and np.all(c.p[k] == c.p[j]) for j in keep):
            keep.append(k)
    assert list(r.tri) == keep


# ------------------------------------------------------------ Frank-Wolfe hand cases
# Body 1 owns a tabulated field on a 9^3 grid over [-4, 4]^3 (spacing 1); body
# 0 owns one triangle whose corners are interior grid nodes.  At a node that
# is the lower corner of its enclosing cell the local coordinates are 0, so
# the gradient is the forward difference of the node values (whatever the
# local coordinates are divided by).  The base field is linear, phi = y - c;
# a few nodes are moved to set the gradient at chosen corners.  Each case has
# an endpoint worked out by hand (collision_detection.py:134-165) and names
# the plausible misreading that would give another endpoint.

def fw_case(P, base, edits):
    from oracle.grid import Grid, nodes_array
    g = Grid([-4.0, -4.0, -4.0], [4.0, 4.0, 4.0], 9, 9, 9)
    nodes = nodes_array(g)
    vals = nodes[:, 1] - base
    for (i, j, k), v in edits.items():
        vals[(i + 4) + 9 * ((j + 4) + 9 * (k + 4))] = v
    g.values = vals
    sc = OS.Scene()
    V = np.array(P, dtype=float)
    sc.shapes = [OS.ShapeArrays(V, np.array([[0, 1, 2]]), None), OS.ShapeArrays(V, np.array([[0, 1, 2]]), g)]
    sc.shape = np.array([0, 1])
    sc.r = np.zeros((2, 3))
    sc.q = np.tile([1.0, 0.0, 0.0, 0.0], (2, 1))
    sc.fixed = np.array([False, True])
    return sc


def fw_endpoint(sc):
    p, n, gap, tri = CD.frank_wolfe_contacts(sc, 0, 1, np.array([0]))
    s = CD.fw_scalar(sc, 0, 1, 0)
    assert len(p) == 1 and s is not None
    assert np.array_equal(p[0], s[0]) and np.array_equal(n[0], s[1]) and gap[0] == s[2]
    return p[0], n[0], gap[0]


def test_fw_step_before_break_and_first_index_ties():
    # phi = y - 1 except at p2 = (-1, 1, 0), whose gradient is (1, 0, 0): all
    # three corner gradients have norm 1, so the start is corner 0 (first index);
    # its gradient (0, 1, 0) gives objectives s.g = (3, 1, 1), a tie won by
    # corner 1 (first index); the step (alpha = 1) moves x onto corner 1 before
    # the break test (1 > 0) ends the loop.  Break before the step: (0, 3, 0);
    # last index on the objective tie: (-1, 1, 0); last index on the start tie:
    # start at p2, objectives (0, 1, -1), endpoint (-1, 1, 0).
    sc = fw_case([(0, 3, 0), (1, 1, 0), (-1, 1, 0)], 1.0, {(0, 1, 0): 1.0, (-1, 2, 0): 0.0})
    p, n, gap = fw_endpoint(sc)
    assert p.tolist() == [1.0, 1.0, 0.0] and n.tolist() == [0.0, 1.0, 0.0] and gap == 0.0


def test_fw_start_at_smallest_gradient_norm():
    # phi = y - 2; at p2 = (-2, 2, 0) the gradient is (0.2, 0.4, 0) (norm 0.45,
    # the others 1): the start is p2, objectives (1.2, 0.8, 0.4) pick p2 itself
    # and 0.4 > 0 ends the loop: the contact is at p2 with gap 0.  Starting at
    # corner 0 instead: objectives (3, 1, 2) pick p1 = (2, 1, 0), gap -1.
    sc = fw_case([(0, 3, 0), (2, 1, 0), (-2, 2, 0)], 2.0, {(-1, 2, 0): 0.2, (-2, 3, 0): 0.4})
    p, n, gap = fw_endpoint(sc)
    assert p.tolist() == [-2.0, 2.0, 0.0] and gap == 0.0
    assert np.allclose(n, np.array([0.2, 0.4, 0.0]) / np.sqrt(0.2))


def test_fw_objective_is_s_dot_gradient():
    # phi = y - 1; gradient (0, 0.5, 0) at p2 = (-2, 2, 0) (the start: smallest
    # norm) and (1, 1, 0) at p1 = (2, 1, 0).  Objectives s.g at p2 are
    # (1.5, 0.5, 1.0): x steps onto p1 and 0.5 > 0 ends the loop, so the
    # contact is at p1 with the normal (1, 1, 0)/sqrt(2).  With the objective
    # (s - x).g the minimum is never positive, the loop goes on, and the
    # gradient at p1 drags x towards p2 (to (-2/3, 5/3, 0) at the next step).
    sc = fw_case([(0, 3, 0), (2, 1, 0), (-2, 2, 0)], 1.0, {(-2, 3, 0): 1.5, (3, 1, 0): 1.0})
    p, n, gap = fw_endpoint(sc)
    assert p.tolist() == [2.0, 1.0, 0.0] and gap == 0.0
    assert np.allclose(n, [np.sqrt(0.5), np.sqrt(0.5), 0.0], atol=1e-15)
