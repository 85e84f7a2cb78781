"""The reference's contact sweep on its own data structures (test and baseline
infrastructure only; nothing in the product imports it).

constraints/contacts.py:31-55 and :66-107 restated with scipy.sparse the way
the reference holds its matrices: J a csr matrix of 4K x 6N rows, W a bsr
matrix of 3x3 blocks (state_storage.py:158-185), WJT = W.dot(J.T) (bsr), and
for every contact of a sweep a full product J.dot(w) and a conversion
WJT.tocsr() (contacts.py:89, 107).  A sweep therefore costs O(K nnz(J)),
quadratic in the contact count.  The module serves two purposes:

* the "faithful" CPU baseline: the reference's sweep timed on one host core;
* a check of the oracle's blocked sweep (problem.sweep_gs), which must give
  the same iterate up to summation order (tests/test_oracle_problem.py).
"""
import time

import numpy as np
import scipy.sparse as sparse

from . import problem as PR


def sparse_problem(P, Winv):
    """J (csr, 4K x 6N), W (bsr 3x3 blocks, 6N x 6N) and WJT = W J^T (bsr) of
    an assembled oracle Problem; Winv: (N, 6, 6) per-body inverse masses."""
    K, N = P.K, P.N
    rows, cols, vals = [], [], []
    for k in range(K):
        for i in range(4):
            for side, J, body in ((0, P.JA, P.ia[k]), (1, P.JB, P.ib[k])):
                for j in range(6):
                    rows.append(4 * k + i)
                    cols.append(6 * body + j)
                    vals.append(J[k, i, j])
    Jm = sparse.csr_matrix((np.asarray(vals), (np.asarray(rows), np.asarray(cols))), shape=(4 * K, 6 * N))
    blocks = []
    for k in range(N):
        blocks.append(Winv[k, :3, :3])
        blocks.append(Winv[k, 3:, 3:])
    W = sparse.block_diag(blocks).tobsr(blocksize=(3, 3))
    WJT = W.dot(Jm.T)
    return Jm, W, WJT


def sweep(Jm, WJT, x, r, b, mu, order, deadline=None):
    """One Gauss-Seidel sweep of contacts.py:66-107 over the contacts in
    ``order`` (flat x, r, b of length 4K; mu of length K), updating x in
    place.  With ``deadline`` (a time.perf_counter() value) the sweep stops
    after the first contact that ends past it.  Returns the number of contacts
    updated."""
    w = WJT.dot(x)
    done = 0
    for k in order:
        block = range(4 * k, 4 * k + 4)
        x_b = x[block]
        delta = x_b.copy()
        z_b = x_b - np.multiply(r[block], Jm.dot(w)[block] + b[block])
        x_b[0] = np.max([0.0, z_b[0]])
        x_b[1:] = PR.prox_sphere([z_b[1]], [z_b[2]], [z_b[3]], [mu[k]], [x_b[0]])[0]
        x[block] = x_b
        np.subtract(x_b, delta, delta)
        w += WJT.tocsr()[:, block].dot(delta)
        done += 1
        if deadline is not None and time.perf_counter() > deadline:
            break
    return done


def time_sweep(P, Winv, order, seconds):
    """Contact-iterations per second of the reference's sweep on this problem:
    sweeps from x = 0 with r = 1/diag until ``seconds`` have passed.  Returns
    (contact_iters_per_s, contact_iters, seconds_spent, seconds_to_build)."""
    t0 = time.perf_counter()
    Jm, W, WJT = sparse_problem(P, Winv)
    t_build = time.perf_counter() - t0
    x = np.zeros(4 * P.K)
    r, b = P.r0.ravel().copy(), P.b.ravel().copy()
    done = 0
    t1 = time.perf_counter()
    deadline = t1 + seconds
    while time.perf_counter() < deadline and P.K:
        done += sweep(Jm, WJT, x, r, b, P.mu, order, deadline)
    t = time.perf_counter() - t1
    return done / t if t > 0 else 0.0, done, t, t_build
