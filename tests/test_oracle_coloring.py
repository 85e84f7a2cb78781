"""Oracle random colouring (blocking.py:62-111, deterministic form)."""
import numpy as np

from oracle import coloring as C


def path_contacts(n):
    # contact k joins bodies k and k+1: consecutive contacts share a body
    ia = np.arange(n)
    ib = np.arange(1, n + 1)
    return ia, ib, np.zeros(n + 1, dtype=bool)


def test_reference_path_graph():
    # test_random_graph_coloring.py:20-32 expects {0,2,4}, {1,3} for the path
    # 0-1-2-3-4 (its comparison is computed but not asserted: palettes can run
    # empty and take a fresh colour).  Every draw sequence gives a proper
    # colouring, and a two-colouring is always this one.
    ia, ib, fixed = path_contacts(5)
    two = 0
    for seed in range(40):
        col = C.color(ia, ib, fixed, seed, 0)
        assert C.is_proper(col, ia, ib, fixed)
        groups = sorted(sorted(g.tolist()) for g in C.groups_from_colors(col))
        if len(groups) == 2:
            assert groups == [[0, 2, 4], [1, 3]]
            two += 1
    assert two > 0


def literal_blocking(ia, ib, fixed, seed, step, max_iter=200):
    """blocking.py:62-111 line by line (sets, per-island maximum, draws from C.draw)."""
    nb = C.neighbours(ia, ib, fixed)
    isl = C.islands(ia, ib, fixed)
    K = len(ia)
    pal = {v: set(range(len(nb[v]))) for v in range(K)}
    U = set(range(K))
    Cc = {}
    it = 0
    while U and it < max_iter:
        for v in sorted(U):
            if pal[v]:
                lst = sorted(pal[v])
                Cc[v] = lst[C.draw(seed, step, v, it, len(lst))]
        I = set()
        for v in U:
            ncol = {Cc.get(u, 0) for u in nb[v] if u in U}
            if Cc.get(v, 0) not in ncol:
                I.add(v)
                for u in nb[v]:
                    if u in U:
                        pal[u].discard(Cc.get(v, 0))
        U -= I
        for v in range(K):
            Cc.setdefault(v, 0)
        mx = {}
        for v in range(K):
            mx[isl[v]] = max(mx.get(isl[v], -1), Cc[v])
        for v in sorted(U):
            if not pal[v]:
                mx[isl[v]] += 1
                pal[v].add(mx[isl[v]])
        it += 1
    if U:
        mx = {}
        for v in range(K):
            mx[isl[v]] = max(mx.get(isl[v], -1), Cc[v])
        for v in sorted(U):
            mx[isl[v]] += 1
            Cc[v] = mx[isl[v]]
    return np.array([Cc[v] for v in range(K)])


def random_contacts(rng, nbody, K, nfixed=1):
    ia = rng.integers(0, nbody, K)
    ib = (ia + rng.integers(1, nbody, K)) % nbody
    fixed = np.zeros(nbody, dtype=bool)
    fixed[:nfixed] = True
    return ia, ib, fixed


def test_proper_complete_and_equal_to_literal():
    rng = np.random.default_rng(8)
    for trial in range(30):
        ia, ib, fixed = random_contacts(rng, rng.integers(3, 12), rng.integers(1, 60))
        col = C.color(ia, ib, fixed, trial, trial * 7)
        assert C.is_proper(col, ia, ib, fixed)
        assert np.array_equal(col, literal_blocking(ia, ib, fixed, trial, trial * 7))


def test_cliques_and_islands():
    # 12 contacts on one free body + the ground: a 12-clique, 12 colours;
    # a second independent body starts its own island (colours restart at 0)
    ia = np.array([0] * 12 + [0] * 5)
    ib = np.array([1] * 12 + [2] * 5)
    fixed = np.array([True, False, False])
    col = C.color(ia, ib, fixed, 1, 0)
    assert len(set(col[:12].tolist())) == 12 and len(set(col[12:].tolist())) == 5
    assert C.is_proper(col, ia, ib, fixed)
    assert col[12:].min() == 0 and col[:12].min() == 0


def test_draw_is_uniform():
    n = 7
    cnt = np.bincount([C.draw(5, 3, v, 2, n) for v in range(70000)], minlength=n)
    assert np.all(np.abs(cnt / 10000.0 - 1) < 0.05)


def test_stuck_rounds_shortcut_equals_full_run():
    """Stopping at a round that can only repeat gives the colouring of all 200 rounds."""
    rng = np.random.default_rng(11)
    cases = [(np.array([0] * 61), np.array([1] * 61), np.array([True, False])),     # a 61-clique
             (np.array([0, 0]), np.array([1, 1]), np.array([True, False]))]           # a stuck pair
    for _ in range(6):
        nbod = int(rng.integers(3, 9))
        K = int(rng.integers(5, 40))
        ia = rng.integers(0, nbod, K)
        ib = (ia + rng.integers(1, nbod, K)) % nbod
        fixed = rng.uniform(size=nbod) < 0.3
        cases.append((ia, ib, fixed))
    for ia, ib, fixed in cases:
        for step in (0, 5):
            fast = C.color(ia, ib, fixed, 3, step)
            full = C.color(ia, ib, fixed, 3, step, shortcut=False)
            assert np.array_equal(fast, full)
            assert C.is_proper(fast, ia, ib, fixed)
