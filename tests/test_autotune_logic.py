"""The autotuner's trial loop, on the CPU with a fake device context.

The fake context charges each step a cost that depends on the tuning and
moves every scene's state and step index, so the tests see whether the
tuner picks the cheapest candidate, restores states and step indices before
every trial and at the end, and puts the caller's tuning back when a trial
fails.
"""
import numpy as np
import pytest

from rainbow_amd import autotune


class FakeCtx:
    def __init__(self, n_scenes, cost):
        self.cost = cost              # tuning dict -> simulated seconds per step
        self.tune = dict(autotune.DEFAULT)
        self.state = {s: np.full((2, 3), float(s)) for s in range(n_scenes)}
        self.index = {s: 7 for s in range(n_scenes)}
        self.clock = 0.0
        self.trial_starts = []        # (tuning, states, indices) seen by each timed step call
        self.fail_on = None

    def get_tuning(self):
        return dict(self.tune)

    def set_tuning(self, **kw):
        for k in kw:
            if k not in autotune.DEFAULT:
                raise KeyError(k)
        self.tune.update(kw)

    def get_state(self, sid, n):
        r = self.state[sid].copy()
        return r, np.zeros((n, 4)), np.zeros((n, 3)), np.zeros((n, 3))

    def set_state(self, sid, r, q, v, w):
        self.state[sid] = r.copy()

    def get_log(self, sid, m):
        return {"step_index": self.index[sid]}

    def set_step_index(self, sid, k):
        self.index[sid] = k

    def step(self, ids, n, dt, flags):
        if self.fail_on is not None and self.tune["unit_size"] == self.fail_on:
            raise RuntimeError("launch failed")
        if n > 1:
            self.trial_starts.append((dict(self.tune), {s: self.state[s].copy() for s in ids},
                                      {s: self.index[s] for s in ids}))
        for s in ids:
            self.state[s] = self.state[s] + n
            self.index[s] += n
        self.clock += n * self.cost(self.tune)

    def sync(self):
        pass


@pytest.fixture
def fake_time(monkeypatch):
    box = {}
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: box["ctx"].clock)
    return box


def cost_of(t):
    # waves 8 with one lane per scene and units of 128 is cheapest
    c = 1.0
    c += 0.0 if t["waves_per_group"] == 8 else 0.5
    c += 0.0 if t["lanes_per_scene"] == 1 else 0.25
    c += 0.0 if t["unit_size"] == 128 else 0.125
    return c * 1e-3


def test_picks_fastest_and_restores(fake_time):
    ctx = FakeCtx(3, cost_of)
    fake_time["ctx"] = ctx
    before = {s: ctx.state[s].copy() for s in ctx.state}
    res = autotune.autotune(ctx, [0, 1, 2], [2, 2, 2], steps=3)
    assert res["best"]["waves_per_group"] == 8
    assert res["best"]["lanes_per_scene"] == 1
    assert res["best"]["unit_size"] == 128
    assert ctx.tune == res["best"]
    assert abs(res["ms_per_step"] - 1.0) < 1e-9
    assert len(res["trials"]) == len(autotune.batch_candidates())
    # every timed trial starts from the saved states and indices
    for tune, st, idx in ctx.trial_starts:
        for s in st:
            assert np.array_equal(st[s], before[s])
            assert idx[s] == 7
    for s in ctx.state:
        assert np.array_equal(ctx.state[s], before[s])
        assert ctx.index[s] == 7


def test_lone_scene_uses_cooperative_candidates(fake_time):
    ctx = FakeCtx(1, lambda t: (1.0 + 0.1 * abs(t["lanes_per_group"] - 2)) * 1e-3)
    fake_time["ctx"] = ctx
    res = autotune.autotune(ctx, [0], [2], steps=2)
    assert len(res["trials"]) == len(autotune.single_candidates())
    assert res["best"]["lanes_per_group"] == 2


def test_small_gains_keep_the_baseline(fake_time):
    # a candidate 2 % faster than the first one is noise under the default margin
    ctx = FakeCtx(2, lambda t: (1.0 - 0.02 * (t["unit_size"] == 128)) * 1e-3)
    fake_time["ctx"] = ctx
    res = autotune.autotune(ctx, [0, 1], [2, 2], steps=2)
    assert res["best"] == res["trials"][0][0]
    assert ctx.tune == res["trials"][0][0]
    res = autotune.autotune(ctx, [0, 1], [2, 2], steps=2, min_gain=0.01)
    assert res["best"]["unit_size"] == 128


def test_failed_trial_restores_caller_tuning(fake_time):
    ctx = FakeCtx(2, cost_of)
    fake_time["ctx"] = ctx
    ctx.tune["unit_size"] = 256
    start = dict(ctx.tune)
    before = {s: ctx.state[s].copy() for s in ctx.state}
    ctx.fail_on = 128
    with pytest.raises(RuntimeError):
        autotune.autotune(ctx, [0, 1], [2, 2], steps=2)
    assert ctx.tune == start
    for s in ctx.state:
        assert np.array_equal(ctx.state[s], before[s])
        assert ctx.index[s] == 7


def test_rejects_bad_arguments():
    ctx = FakeCtx(1, cost_of)
    with pytest.raises(ValueError):
        autotune.autotune(ctx, [0], [2, 2])
    with pytest.raises(ValueError):
        autotune.autotune(ctx, [], [])
    with pytest.raises(ValueError):
        autotune.autotune(ctx, [0], [2], steps=0)
