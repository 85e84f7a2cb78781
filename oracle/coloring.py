"""Random graph colouring of the contact conflict graph (deterministic form).

Restates the Vivace-style colouring of proximal_contact/blocking.py:62-111 in
the form this build defines for the GPU, so that a colouring is a function of
(contacts, seed, step) and the oracle reproduces it exactly:

* Conflict graph (blocking.py:6-42, rigid bodies): two contacts are adjacent if
  they share a body -- here only a *free* body, because fixed bodies have a zero
  inverse mass block and never carry velocity updates.
* Islands: connected components of the free bodies linked by contacts.  Each
  island is coloured independently; contacts whose bodies are all fixed form
  islands of their own.
* Palettes (blocking.py:70-73): node v starts with {0, ..., deg(v)-1}.
* Round r (blocking.py:78-98): every uncoloured node with a non-empty palette
  draws a tentative colour uniformly from its palette; the draw is element
  floor(hi32(h) * |P| / 2^32) of the palette in increasing order, with
  h = mix(mix(mix(mix(seed) ^ step) ^ v) ^ r) and ``mix`` the splitmix64
  finaliser.  A node keeps its colour if no uncoloured neighbour drew the same
  colour; kept colours are removed from the palettes of uncoloured neighbours.
  Nodes left with an empty palette then receive max+1, max+2, ... in increasing
  node order, where max is the largest colour value present in the island
  (kept or tentative; 0 for nodes that never drew).
* After 200 rounds, remaining nodes get fresh colours max+1, ... in node order
  (blocking.py:100-104).  A round with no winner in which every uncoloured node
  holds a one-colour palette (two adjacent nodes left with the same single
  colour, say) repeats forever, so the rounds stop there; the result is that
  of running all 200.
* Sweep order: colours in increasing value; a colour's contacts are mutually
  independent, so their order inside the colour does not change the result.

Deterministic choices where blocking.py leaves the order to Python's set
iteration or to numpy's global RNG: the counter-based draw above, fresh colours
in increasing node order, per-island maxima, colours swept in increasing value.

Pinned by tests/test_oracle_coloring.py: the reference's path-graph case
(test_random_graph_coloring.py:20-32, groups {0,2,4} and {1,3}), properness and
completeness invariants on random graphs, and agreement with a literal
set-based restatement of blocking.py's loop driven by the same draws.
"""
import numpy as np

M64 = (1 << 64) - 1
MAX_ROUNDS = 200


def mix64(x):
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def draw(seed, step, v, rnd, n):
    h = mix64(mix64(mix64(mix64(seed & M64) ^ (step & M64)) ^ (v & M64)) ^ (rnd & M64))
    return ((h >> 32) * n) >> 32


def islands(ia, ib, fixed):
    """Island label per contact: smallest free body index of its component,
    or -1 - k for a contact with no free body."""
    N = len(fixed)
    parent = list(range(N))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    for a, b in zip(ia, ib):
        if not fixed[a] and not fixed[b]:
            ra, rb = find(a), find(b)
            if ra != rb:
                parent[max(ra, rb)] = min(ra, rb)
    lab = np.empty(len(ia), dtype=np.int64)
    for k, (a, b) in enumerate(zip(ia, ib)):
        if not fixed[a]:
            lab[k] = find(a)
        elif not fixed[b]:
            lab[k] = find(b)
        else:
            lab[k] = -1 - k
    return lab


def neighbours(ia, ib, fixed):
    """Adjacency lists of the free-body conflict graph."""
    by_body = {}
    for k, (a, b) in enumerate(zip(ia, ib)):
        for x in {int(a), int(b)}:
            if not fixed[x]:
                by_body.setdefault(x, []).append(k)
    nb = [set() for _ in range(len(ia))]
    for lst in by_body.values():
        for k in lst:
            nb[k].update(lst)
    for k in range(len(ia)):
        nb[k].discard(k)
    return [sorted(s) for s in nb]


def color(ia, ib, fixed, seed, step, max_rounds=MAX_ROUNDS, shortcut=True):
    """Colour per contact (int array).  shortcut=False runs every round up to
    max_rounds (the literal loop); the result is the same."""
    K = len(ia)
    if K == 0:
        return np.zeros(0, dtype=np.int64)
    nb = neighbours(ia, ib, fixed)
    isl = islands(ia, ib, fixed)
    pal = [set(range(len(nb[v]))) for v in range(K)]
    C = np.zeros(K, dtype=np.int64)
    U = np.ones(K, dtype=bool)
    rnd = 0
    while U.any() and rnd < max_rounds:
        for v in np.nonzero(U)[0]:
            if pal[v]:
                srt = sorted(pal[v])
                C[v] = srt[draw(seed, step, int(v), rnd, len(srt))]
        acc = []
        for v in np.nonzero(U)[0]:
            if all(not U[u] or C[u] != C[v] for u in nb[v]):
                acc.append(v)
        for v in acc:
            for u in nb[v]:
                if U[u]:
                    pal[u].discard(int(C[v]))
        U[acc] = False
        mx = {}
        for v in range(K):
            mx[isl[v]] = max(mx.get(isl[v], -1), int(C[v]))
        filled = False
        for v in np.nonzero(U)[0]:
            if not pal[v]:
                mx[isl[v]] += 1
                pal[v] = {mx[isl[v]]}
                filled = True
        rnd += 1
        if shortcut and not acc and not filled and all(len(pal[v]) == 1 for v in np.nonzero(U)[0]):
            # No winner and every palette a single colour: every later round
            # draws the same colours and keeps none, so the remaining rounds
            # change nothing; go to the leftovers directly.
            break
    if U.any():
        mx = {}
        for v in range(K):
            mx[isl[v]] = max(mx.get(isl[v], -1), int(C[v]))
        for v in np.nonzero(U)[0]:
            mx[isl[v]] += 1
            C[v] = mx[isl[v]]
    return C


def groups_from_colors(C):
    """Sweep groups: colours in increasing value, members in index order."""
    return [np.nonzero(C == c)[0] for c in np.unique(C)]


def serial_groups(K):
    """The reference's serial order (contacts.py:73): one contact per group."""
    return [np.array([k]) for k in range(K)]


def is_proper(C, ia, ib, fixed):
    nb = neighbours(ia, ib, fixed)
    return all(C[u] != C[v] for v in range(len(C)) for u in nb[v])
