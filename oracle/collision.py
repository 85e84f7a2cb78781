"""Collision detection restated from prox_rigid_bodies/collision_detection.py.

* ``world_vertices`` / ``triangle_aabbs``: _update_bvh (:10-44) and the leaf
  volumes of refit_bvh (kdop_bvh.py:146-167 with kdop.py:53-75, K=3 so the slab
  directions are the x, y, z axes): per-triangle min/max of the world vertices
  grown by ``envelope``.
* ``candidates``: _narrow_phase (:47-84).  The BVH tandem traversal returns every
  leaf pair whose volumes overlap (kdop.py:21-38: closed intervals, ``upper <
  lower`` rejects); internal volumes are exact unions of their children, so the
  result equals the brute-force set of overlapping triangle-AABB pairs.  The
  candidate triangles of A are ``unique(results[:, 0])`` (:234-236), sorted.
* ``frank_wolfe_contacts``: _compute_contacts (:87-201), vectorised over the
  candidate triangles of one (triangle owner, SDF owner) direction.
* ``reduce_contacts``: _contact_reduction (:255-290), keep the first of
  contacts with equal position within one unordered body pair.
* ``detect``: run_collision_detection (:293-318) with the contact order of
  _contact_determination (:204-252): body pairs in ``combinations`` order,
  A's triangles against B's SDF, then B's triangles against A's SDF.

Pinned by tests/test_oracle_collision.py: the candidate set against an explicit
per-pair loop over all triangle pairs; Frank-Wolfe against a scalar
line-by-line restatement (oracle.collision.fw_scalar) and against hand-worked
cases on the analytic box SDF (box resting on the ground: the FW iterate is a
vertex after iteration 0 and the loop stops at once); contact reduction against
an O(K^2) loop.
"""
import numpy as np

from . import grid as G
from . import quat as Q


def world_vertices(scene, b):
    sh = scene.shapes[scene.shape[b]]
    return Q.rotate_array(scene.q[b], sh.V) + scene.r[b]


def triangle_aabbs(Vw, T, envelope):
    P = Vw[T]                                  # (nt, 3, 3)
    lo = P.min(axis=1) - envelope
    hi = P.max(axis=1) + envelope
    return lo, hi


def _overlap_matrix(loA, hiA, loB, hiB):
    """kdop.py:21-38 for all pairs: overlap unless upper < lower on some axis."""
    sep = (hiA[:, None, :] < loB[None, :, :]) | (hiB[None, :, :] < loA[:, None, :])
    return ~np.any(sep, axis=-1)


def _slack(loA, hiA, loB, hiB):
    """Overlap slack of boxes: min over axes of min(hiA - loB, hiB - loA);
    >= 0 exactly when kdop.overlap reports an overlap."""
    return np.min(np.minimum(hiA[:, None, :] - loB[None, :, :], hiB[None, :, :] - loA[:, None, :]), axis=-1)


def candidates(aabbA, aabbB, chunk=4096, slack=None):
    """Candidate triangles of A and of B and the number of overlapping pairs.

    ``slack`` (optional dict) receives, per triangle of A ("A") and of B
    ("B"), the largest overlap slack against the other body's triangles (the
    decision margin of the candidate test; root-volume slack for triangles
    outside the other root)."""
    loA, hiA = aabbA
    loB, hiB = aabbB
    rootA = (loA.min(axis=0), hiA.max(axis=0))
    rootB = (loB.min(axis=0), hiB.max(axis=0))
    if slack is not None:
        slack["A"] = _slack(loA, hiA, rootB[0][None], rootB[1][None])[:, 0]
        slack["B"] = _slack(loB, hiB, rootA[0][None], rootA[1][None])[:, 0]
    if np.any(rootA[1] < rootB[0]) or np.any(rootB[1] < rootA[0]):
        return np.zeros(0, np.int64), np.zeros(0, np.int64), 0
    # Only triangles overlapping the other body's root volume can overlap a leaf.
    ia = np.nonzero(~(np.any(hiA < rootB[0], axis=1) | np.any(rootB[1] < loA, axis=1)))[0]
    ib = np.nonzero(~(np.any(hiB < rootA[0], axis=1) | np.any(rootA[1] < loB, axis=1)))[0]
    hitA = np.zeros(len(ia), dtype=bool)
    hitB = np.zeros(len(ib), dtype=bool)
    count = 0
    bestB = np.full(len(ib), -np.inf)
    for s in range(0, len(ia), chunk):
        m = _overlap_matrix(loA[ia[s:s + chunk]], hiA[ia[s:s + chunk]], loB[ib], hiB[ib])
        count += int(m.sum())
        hitA[s:s + chunk] |= m.any(axis=1)
        hitB |= m.any(axis=0)
        if slack is not None and len(ib):
            sl = _slack(loA[ia[s:s + chunk]], hiA[ia[s:s + chunk]], loB[ib], hiB[ib])
            slack["A"][ia[s:s + chunk]] = sl.max(axis=1)
            bestB = np.maximum(bestB, sl.max(axis=0))
    if slack is not None and len(ib):
        slack["B"][ib] = bestB
    return ia[hitA], ib[hitB], count


def _argmin_first(vals):
    """np.argmin over the last axis (first index on ties)."""
    return np.argmin(vals, axis=-1)


def _gap2(v, scale):
    """Gap between the smallest and second smallest of 3 values per row, in
    units of ``scale`` (a fixed magnitude, not the values themselves: where the
    SDF is flat the values are rounding noise and any gap between them is
    meaningless)."""
    srt = np.sort(v, axis=1)
    return (srt[:, 1] - srt[:, 0]) / np.maximum(scale, 1e-300)


def _cell_margin(grid, P):
    """Distance (in cells) of the lookup points P to the nearest cell face that
    changes grid3.get_enclosing_cell_idx (the floor decision); the
    min_coord-scaled gradient jumps across those faces."""
    f = (P - grid.min_coord) / grid.spacing
    hi = np.array([grid.I - 2, grid.J - 2, grid.K - 2], dtype=float)
    d = np.abs(f - np.round(f))
    # faces outside the clipped index range do not change the cell
    d = np.where((f < 0.5) | (f > hi + 0.5), np.inf, d)
    return d.min(axis=1)


def frank_wolfe_contacts(scene, a, b, tris, margins=None):
    """Contacts of triangles ``tris`` of body ``a`` against the SDF of body ``b``.

    Returns (p_world, n_world, gap, tri_index) for the triangles that produce a
    contact, in the order of ``tris``.  The ContactPoint built from each has
    bodyA = b (SDF owner) and bodyB = a (triangle owner), collision_detection.py:187.
    """
    prm = scene.params
    shA = scene.shapes[scene.shape[a]]
    shB = scene.shapes[scene.shape[b]]
    grid = shB.grid
    if len(tris) == 0:
        z = np.zeros((0, 3))
        return z, z, np.zeros(0), np.zeros(0, np.int64)
    V_w = Q.rotate_array(scene.q[a], shA.V) + scene.r[a]
    V_b = Q.rotate_array(Q.conjugate(scene.q[b]), V_w - scene.r[b])
    P = V_b[shA.T[tris]]                                   # (C, 3, 3)
    C = len(tris)
    gn = np.stack([np.linalg.norm(G.get_gradient_v(grid, P[:, v, :]), axis=1) for v in range(3)], axis=1)
    x = P[np.arange(C), _argmin_first(gn)]
    margin = None
    if margins is not None:
        # gradient norms of an SDF are ~1 where they carry information
        margin = _gap2(gn, np.maximum(gn.max(axis=1), 1.0))
        for v in range(3):
            margin = np.minimum(margin, _cell_margin(grid, P[:, v, :]))
    active = np.ones(C, dtype=bool)
    tol = prm["contact_optimization_tolerance"]
    for i in range(int(prm["contact_optimization_max_iterations"])):
        idx = np.nonzero(active)[0]
        if len(idx) == 0:
            break
        g = G.get_gradient_v(grid, x[idx])
        if margin is not None:
            margin[idx] = np.minimum(margin[idx], _cell_margin(grid, x[idx]))
        Pi = P[idx]
        obj = np.stack([Pi[:, v, 0] * g[:, 0] + Pi[:, v, 1] * g[:, 1] + Pi[:, v, 2] * g[:, 2]
                        for v in range(3)], axis=1)
        vtx = _argmin_first(obj)
        s_i = Pi[np.arange(len(idx)), vtx]
        alpha = 2 / (i + 2)
        x[idx] = x[idx] + alpha * (s_i - x[idx])
        best = obj[np.arange(len(idx)), vtx]
        if margin is not None:
            # objectives s . grad: scale |s| (|grad| ~ 1 for an SDF)
            scale = np.maximum(np.abs(Pi).max(axis=(1, 2)), 1e-6)
            margin[idx] = np.minimum(margin[idx], np.minimum(_gap2(obj, scale), np.abs(best - tol) / scale))
        active[idx[best > tol]] = False
    inside = G.is_inside_v(grid, x, 0.5)
    phi = np.full(C, np.inf)
    phi[inside] = G.get_value_v(grid, x[inside])
    keep = inside & (phi <= prm["envelope"])
    if margin is not None:
        h = float(np.min(grid.spacing))
        ins = np.minimum(np.min(np.abs(x - (grid.max_coord - 0.5)), axis=1),
                         np.min(np.abs(x - (grid.min_coord + 0.5)), axis=1)) / h
        margin = np.minimum(margin, ins)
        phv = np.abs(G.get_value_v(grid, x) - prm["envelope"]) / h
        margin = np.minimum(margin, phv)
        margin = np.minimum(margin, _cell_margin(grid, x))
        for t, m in zip(tris, margin):
            margins[(int(b), int(a), int(t))] = min(margins.get((int(b), int(a), int(t)), np.inf), float(m))
    idx = np.nonzero(keep)[0]
    n_b = G.get_gradient_v(grid, x[idx]) if len(idx) else np.zeros((0, 3))
    qb = scene.q[b]
    n_w = np.array([Q.rotate(qb, nb) for nb in n_b]).reshape(-1, 3)
    nrm = np.linalg.norm(n_w, axis=1)
    ok = nrm > 0
    idx, n_w, nrm = idx[ok], n_w[ok], nrm[ok]
    p_w = np.array([Q.rotate(qb, x[k]) + scene.r[b] for k in idx]).reshape(-1, 3)
    return p_w, n_w / nrm[:, None], phi[idx], np.asarray(tris)[idx]


def fw_scalar(scene, a, b, t):
    """Line-by-line scalar restatement of collision_detection.py:134-188 for
    one triangle; returns (p_w, n_w, gap) or None."""
    prm = scene.params
    shA = scene.shapes[scene.shape[a]]
    grid = scene.shapes[scene.shape[b]].grid
    V_w = Q.rotate_array(scene.q[a], shA.V) + scene.r[a]
    V_b = Q.rotate_array(Q.conjugate(scene.q[b]), V_w - scene.r[b])
    p_bs = V_b[shA.T[t], :]
    gradients = np.linalg.norm([G.get_gradient(grid, p) for p in p_bs], axis=1)
    x_i = p_bs[np.argmin(gradients)]
    for i in range(int(prm["contact_optimization_max_iterations"])):
        g = G.get_gradient(grid, x_i)
        objectives = [np.dot(s_i, g) for s_i in p_bs]
        vertex = np.argmin(objectives)
        s_i = p_bs[vertex]
        alpha = 2 / (i + 2)
        x_i = x_i + alpha * (s_i - x_i)
        if objectives[vertex] > prm["contact_optimization_tolerance"]:
            break
    if G.is_inside(grid, x_i):
        phi = G.get_value(grid, x_i)
        if phi <= prm["envelope"]:
            n_w = Q.rotate(scene.q[b], G.get_gradient(grid, x_i))
            if np.linalg.norm(n_w) > 0:
                p_w = Q.rotate(scene.q[b], x_i) + scene.r[b]
                return p_w, n_w / np.linalg.norm(n_w), phi
    return None


class Contacts:
    """Contact list in the reference's order (SoA)."""

    def __init__(self):
        self.ia = np.zeros(0, np.int64)    # ContactPoint.bodyA: SDF owner
        self.ib = np.zeros(0, np.int64)    # ContactPoint.bodyB: triangle owner
        self.p = np.zeros((0, 3))
        self.n = np.zeros((0, 3))
        self.gap = np.zeros(0)
        self.tri = np.zeros(0, np.int64)   # provenance: triangle of bodyB

    @property
    def K(self):
        return len(self.ia)

    def take(self, idx):
        c = Contacts()
        c.ia, c.ib, c.p, c.n, c.gap, c.tri = (self.ia[idx], self.ib[idx], self.p[idx], self.n[idx],
                                              self.gap[idx], self.tri[idx])
        return c


def reduce_contacts(c):
    """collision_detection.py:255-290: keep the first contact of each
    (unordered body pair, position) class."""
    keep = []
    seen = {}
    for k in range(c.K):
        key = (min(c.ia[k], c.ib[k]), max(c.ia[k], c.ib[k]))
        lst = seen.setdefault(key, [])
        pk = c.p[k]
        if any(np.all(pk == c.p[j]) for j in lst):
            continue
        lst.append(k)
        keep.append(k)
    return c.take(np.array(keep, dtype=np.int64))


def detect(scene, stats=None, margins=None):
    """run_collision_detection: returns the reduced Contacts.

    ``stats`` (dict) receives ``number_of_overlaps`` (overlapping triangle pairs)
    and ``candidates`` (candidate triangles over both directions).
    """
    env = scene.params["envelope"]
    N = scene.N
    aabbs = []
    for b in range(N):
        Vw = world_vertices(scene, b)
        aabbs.append(triangle_aabbs(Vw, scene.shapes[scene.shape[b]].T, env))
    ia, ib, p, n, gap, tri = [], [], [], [], [], []
    n_overlaps = 0
    n_cand = 0
    # body pairs in combinations order; pairs whose root volumes are disjoint
    # have no overlapping triangle pair and are skipped (found vectorised)
    rlo = np.array([lo.min(axis=0) for lo, _ in aabbs]).reshape(-1, 3)
    rhi = np.array([hi.max(axis=0) for _, hi in aabbs]).reshape(-1, 3)
    pairs = []
    for a in range(N):
        ok = ~np.any((rhi[a] < rlo[a + 1:]) | (rhi[a + 1:] < rlo[a]), axis=1)
        if margins is not None:
            # keep near-touching roots too, so candidate margins are recorded
            sl = np.min(np.minimum(rhi[a] - rlo[a + 1:], rhi[a + 1:] - rlo[a]), axis=1)
            ok |= np.abs(sl) < 1e-4
        pairs.extend((a, a + 1 + j) for j in np.nonzero(ok)[0])
    for a, b in pairs:
        if True:
            slack = {} if margins is not None else None
            ca, cb, cnt = candidates(aabbs[a], aabbs[b], slack=slack)
            if margins is not None:
                for (owner, sdf, key) in ((a, b, "A"), (b, a, "B")):
                    sl = slack[key]
                    near = np.nonzero(np.abs(sl) < 1e-4)[0]
                    for t in near:
                        k = (int(sdf), int(owner), int(t))
                        margins[k] = min(margins.get(k, np.inf), float(abs(sl[t])))
            if cnt == 0:
                continue
            n_overlaps += cnt
            n_cand += len(ca) + len(cb)
            for (owner, sdf, tr) in ((a, b, ca), (b, a, cb)):
                pw, nw, g, t = frank_wolfe_contacts(scene, owner, sdf, tr, margins)
                ia.append(np.full(len(t), sdf))
                ib.append(np.full(len(t), owner))
                p.append(pw)
                n.append(nw)
                gap.append(g)
                tri.append(t)
    c = Contacts()
    if ia:
        c.ia = np.concatenate(ia).astype(np.int64)
        c.ib = np.concatenate(ib).astype(np.int64)
        c.p = np.concatenate(p).reshape(-1, 3)
        c.n = np.concatenate(n).reshape(-1, 3)
        c.gap = np.concatenate(gap)
        c.tri = np.concatenate(tri).astype(np.int64)
    if stats is not None:
        stats["number_of_overlaps"] = n_overlaps
        stats["candidates"] = n_cand
        stats["raw_contacts"] = c.K
    if margins is not None and c.K:
        # contact reduction compares positions exactly: a raw contact close
        # to another contact of the same body pair sits near that decision
        # boundary (recorded up to a relative 1e-5, the reach of FP32 rounding)
        for k in range(c.K):
            same = np.nonzero((np.minimum(c.ia, c.ib) == min(c.ia[k], c.ib[k]))
                              & (np.maximum(c.ia, c.ib) == max(c.ia[k], c.ib[k])))[0]
            d = np.abs(c.p[same] - c.p[k]).max(axis=1) / max(1.0, np.abs(c.p[k]).max())
            d[same == k] = np.inf
            if len(d) and d.min() < 1e-5:
                key = (int(c.ia[k]), int(c.ib[k]), int(c.tri[k]))
                margins[key] = min(margins.get(key, np.inf), float(d.min()))
    return reduce_contacts(c)
