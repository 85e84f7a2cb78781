"""Oracle SDF lookups (grid3.py) against the reference's own test cases
(unit_tests/test_geometry_grid3.py) and closed forms."""
import numpy as np
import pytest

from oracle import grid as G


def make(I=4, f=None, lo=-1.0, hi=1.0):
    g = G.Grid(np.full(3, lo), np.full(3, hi), I, I, I)
    if f is not None:
        G.eval_on_grid(g, f)
    return g


def test_enclosing_cell_reference_cases():
    # test_geometry_grid3.py:56-91
    A = make(4)
    cases = [((0, 0, 0), (1, 1, 1)), ((-0.5,) * 3, (0, 0, 0)), ((0.5,) * 3, (2, 2, 2)), ((-1,) * 3, (0, 0, 0)),
             ((1,) * 3, (2, 2, 2)), ((-2,) * 3, (0, 0, 0)), ((2,) * 3, (2, 2, 2))]
    for p, idx in cases:
        assert A.get_enclosing_cell_idx(np.array(p, dtype=float)) == idx


@pytest.mark.parametrize("f,grad", [
    (lambda p: p[0], (1, 0, 0)), (lambda p: p[1], (0, 1, 0)), (lambda p: p[2], (0, 0, 1)),
    (lambda p: 2 * p[2], (0, 0, 2)), (lambda p: 2 * p[2] + 3, (0, 0, 2)), (lambda p: 2 * p[0] + 3 * p[1], (2, 3, 0)),
    (lambda p: 2 * p[0] + 3 * p[1] - 4 * p[2], (2, 3, -4))])
def test_linear_fields_reference_cases(f, grad):
    # test_geometry_grid3.py:108-191, 255-508: value and gradient of linear fields
    A = make(5, f)
    for p in [(0, 0, 0), (-0.5,) * 3, (0.5,) * 3, (-1,) * 3, (1,) * 3, (-2,) * 3, (2,) * 3, (0.3, -0.7, 0.1)]:
        p = np.array(p, dtype=float)
        assert G.get_value(A, p) == pytest.approx(f(p), abs=1e-12)
        assert np.allclose(G.get_gradient(A, p), grad, atol=1e-12)


def test_is_inside_reference_cases():
    # test_geometry_grid3.py:193-239
    A = make(4)
    for b, p, exp in [(0, -1, True), (0, 0, True), (0, 1, True), (0, -1.1, False), (0, 1.1, False),
                      (0.5, -1, False), (0.5, 0, True), (0.5, 1, False), (0.5, -1.1, False), (0.5, 1.1, False)]:
        assert G.is_inside(A, np.full(3, float(p)), b) is exp
    with pytest.raises(ValueError):
        G.is_inside(A, np.zeros(3), -1)


def _trilinear_closed_form(c, s, t, u):
    """Value and (d/ds, d/dt, d/du) of the trilinear polynomial of cell corners c[i,j,k]."""
    val = 0.0
    ds = dt = du = 0.0
    for i in (0, 1):
        for j in (0, 1):
            for k in (0, 1):
                ws = s if i else 1 - s
                wt = t if j else 1 - t
                wu = u if k else 1 - u
                val += c[i, j, k] * ws * wt * wu
                ds += c[i, j, k] * (1 if i else -1) * wt * wu
                dt += c[i, j, k] * ws * (1 if j else -1) * wu
                du += c[i, j, k] * ws * wt * (1 if k else -1)
    return val, np.array([ds, dt, du])


def test_value_and_gradient_closed_form_nonlinear():
    # any tabulated field: value = trilinear interpolant; gradient = its
    # derivative at local coordinates (p - cell_origin) / min_coord (grid3.py:209-211)
    rng = np.random.default_rng(0)
    g = G.Grid(np.array([-1.3, -0.7, -2.1]), np.array([0.9, 1.6, 0.4]), 6, 5, 7)
    g.values = rng.normal(size=6 * 5 * 7)
    for _ in range(200):
        p = rng.uniform(g.min_coord - 0.3, g.max_coord + 0.3)
        i, j, k = g.get_enclosing_cell_idx(p)
        c = np.empty((2, 2, 2))
        for a in (0, 1):
            for b in (0, 1):
                for e in (0, 1):
                    c[a, b, e] = g.get_node_value(i + a, j + b, k + e)
        org = g.min_coord + g.spacing * np.array([i, j, k])
        val, _ = _trilinear_closed_form(c, *((p - org) / g.spacing))
        assert G.get_value(g, p) == pytest.approx(val, rel=1e-12, abs=1e-12)
        loc = (p - org) / g.min_coord
        _, dloc = _trilinear_closed_form(c, *loc)
        assert np.allclose(G.get_gradient(g, p), dloc / g.spacing, rtol=1e-10, atol=1e-12)
        # the vectorised forms agree with the scalar ones
        assert G.get_value_v(g, p[None])[0] == pytest.approx(G.get_value(g, p), abs=1e-14)
        assert np.allclose(G.get_gradient_v(g, p[None])[0], G.get_gradient(g, p), atol=1e-14)
        assert bool(G.is_inside_v(g, p[None], 0.5)[0]) == G.is_inside(g, p, 0.5)


def test_gradient_uses_min_coord_not_spacing():
    # a plant: dividing the local coordinate by spacing (the value path's
    # convention) changes the gradient of a non-linear field
    g = G.Grid(np.array([-2.0, -2.0, -2.0]), np.array([2.0, 2.0, 2.0]), 5, 5, 5)
    G.eval_on_grid(g, lambda p: p[0] * p[1])
    p = np.array([0.4, 0.3, 0.2])
    got = G.get_gradient(g, p)
    # at local coordinates s,t ~ (p - origin)/min_coord (small, negative) the
    # trilinear x-derivative is ~ y at the cell's lower edge, not at p
    assert abs(got[0] - 0.3) > 1e-3
    assert got[0] == pytest.approx(0.0 + (0.3 - 0.0) / -2.0 * 1.0, abs=1e-12)
