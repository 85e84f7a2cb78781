"""Admission and batching logic of the serving loop, on the CPU with a fake device context."""
import numpy as np

from rainbow_amd import serving


class FakeCtx:
    """Records batches and device memory the way librbh would account them."""

    def __init__(self):
        self.next = 0
        self.live = {}
        self.batches = []
        self.used = 0
        self.peak = 0

    def scene_create(self, desc):
        sid = self.next
        self.next += 1
        self.live[sid] = dict(fp=desc["fp"], n=desc["n"], steps=0)
        self.used += desc["fp"]
        self.peak = max(self.peak, self.used)
        return sid

    def set_state(self, sid, r, q, v, w):
        self.live[sid]["r"] = r.copy()

    def set_step_index(self, sid, k):
        self.live[sid]["step_index"] = k

    def step(self, ids, n, dt, flags):
        self.batches.append(list(ids))
        for i in ids:
            self.live[i]["steps"] += n
            self.live[i]["r"] = self.live[i]["r"] + 1.0

    def sync(self):
        pass

    def get_state(self, sid, n):
        r = self.live[sid]["r"]
        return r, np.zeros((n, 4)), np.zeros((n, 3)), np.zeros((n, 3))

    def scene_destroy(self, sid):
        self.used -= self.live.pop(sid)["fp"]


class FakeEngine:
    def __init__(self, n):
        self.bodies = {i: None for i in range(n)}


def make_requests(fps, steps, arrivals):
    reqs = []
    for i, (fp, st, t) in enumerate(zip(fps, steps, arrivals)):
        r = serving.Request(i, FakeEngine(3), st, t)
        r._desc = dict(fp=fp, n=3)
        r._state0 = (np.zeros((3, 3)), np.zeros((3, 4)), np.zeros((3, 3)), np.zeros((3, 3)))
        r.footprint = fp
        reqs.append(r)
    return reqs


def test_budget_fifo_and_completion():
    ctx = FakeCtx()
    srv = serving.Server(ctx=ctx, budget=100, graph=False)
    reqs = make_requests([40, 40, 40, 30, 90], [5, 3, 4, 2, 1], [0.0, 0.0, 0.0, 0.0, 0.0])
    rep = srv.run(reqs, clock=iter(np.arange(0, 1e6, 1e-3)).__next__)
    assert ctx.peak <= 100
    assert rep["scene_steps"] == 15 and rep["requests"] == 5
    for r in reqs:
        assert r.finished >= r.admitted >= r.arrival
        assert np.all(r.state[0] == r.n_steps)          # stepped exactly n_steps times
    # first in, first admitted: request 2 (third) waits for a slot, 3 is admitted after 2
    order = sorted(reqs, key=lambda r: r.admitted)
    assert [r.rid for r in order] == sorted(r.rid for r in reqs)
    assert all(len(b) > 0 for b in ctx.batches)
    assert ctx.used == 0 and not ctx.live


def test_arrivals_respected():
    ctx = FakeCtx()
    srv = serving.Server(ctx=ctx, budget=1000, graph=False)
    reqs = make_requests([10, 10], [2, 2], [0.0, 0.5])
    t = iter(np.arange(0, 1e6, 0.1))
    rep = srv.run(reqs, clock=t.__next__)
    assert reqs[1].admitted >= 0.5
    assert rep["p99_wait_ms"] >= 0.0


def test_poisson_arrivals_rate():
    a = serving.poisson_arrivals(200.0, 20000, seed=4)
    assert np.all(np.diff(a) > 0)
    assert abs(len(a) / a[-1] - 200.0) < 5.0
