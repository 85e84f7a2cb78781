"""GPU: launch shapes and the serving loop change no result.

* Every tuning of the step (unit size, sweep lanes per scene, resident LDS
  capacities small enough to force the velocity-only and global layouts,
  the half LDS layout of two solve workgroups per CU, per-body colour
  bitmaps too small for some or all colouring units) gives bitwise the same
  contacts, colours, impulses and states.
* The autotuner, which steps the batch under each candidate, leaves every
  scene where it found it: the steps after it are bitwise the untuned ones.
* A request served by continuous batching, among other requests arriving
  over time, ends in bitwise the same state as the same request run alone.
"""
import numpy as np
import pytest

from rainbow_amd import autotune, native, scenes, serving
from rainbow_amd.stepper import engine_desc, gather_state

pytestmark = pytest.mark.gpu

TUNINGS = [
    dict(),                                                   # defaults: resident groups
    dict(unit_size=32, lanes_per_scene=2),
    dict(unit_size=128, lanes_per_scene=8, lanes_per_group=2),
    dict(waves_per_group=8),                                  # 512-thread solve groups
    dict(waves_per_group=8, lanes_per_scene=3),
    dict(unit_size=256, lanes_per_scene=1),
    dict(resident_contacts=64, resident_bodies=64),           # groups too big: velocity-only layout
    dict(color_body_bits=1),                                  # small units coloured through body lists
    dict(unit_size=256, color_body_bits=24),                  # bitmaps for the units with few bodies only
    dict(solve_groups_per_cu=2),                              # half LDS layout: two solve groups per CU
    dict(solve_groups_per_cu=2, waves_per_group=8, lanes_per_scene=1),
]


@pytest.fixture(scope="module")
def bank():
    return scenes.ShapeBank(resolution=32)


def run_batch(bank, tuning, seeds, steps, single=False, tune_first=False):
    ctx = native.Context(0, "fp32")
    try:
        if tuning:
            ctx.set_tuning(**tuning)
        engines = [scenes.mixed_drop(s, bank, nmax=96, params=dict(precision="fp32")) for s in seeds]
        for e in engines:
            for b in e.bodies.values():
                if not b.is_fixed:
                    b.r = b.r - np.array([0.0, 0.8 * b.r[1] - 0.3, 0.0])   # start near the ground: contacts at once
        ids = []
        for e in engines:
            desc, keep = engine_desc(e, ctx)
            sid = ctx.scene_create(desc)
            ctx.set_state(sid, *gather_state(list(e.bodies.values())))
            ids.append(sid)
        if tune_first:
            cands = autotune.batch_candidates()[::3] if len(ids) > 1 else autotune.single_candidates()[::4]
            res = autotune.autotune(ctx, ids, [len(e.bodies) for e in engines], dt=0.01, steps=2,
                                    candidates=cands, graph=False)
            assert ctx.get_tuning() == res["best"]
        out = []
        for _ in range(steps):
            if single:
                for sid in ids:
                    ctx.step([sid], 1, 0.01)
            else:
                ctx.step(ids, 1, 0.01)
            snap = []
            for sid, e in zip(ids, engines):
                c = ctx.get_contacts(sid)
                snap.append([c["tri"], c["color"], c["impulse"]] + list(ctx.get_state(sid, len(e.bodies))))
            out.append(snap)
        total = sum(ctx.scene_totals(s)["contacts"] for s in ids)
        return out, total
    finally:
        ctx.close()


def test_tunings_bitwise(bank):
    ref, total = run_batch(bank, TUNINGS[0], (3200, 3201, 3202), 4)
    assert total > 0
    for t in TUNINGS[1:]:
        got, _ = run_batch(bank, t, (3200, 3201, 3202), 4)
        for a, b in zip(ref, got):
            for sa, sb in zip(a, b):
                for x, y in zip(sa, sb):
                    assert np.array_equal(x, y), t


def test_single_scene_schedules_bitwise(bank):
    """A lone scene (cooperative grid) equals the same scene inside a batch."""
    ref, _ = run_batch(bank, {}, (3203,), 4, single=True)
    got, _ = run_batch(bank, {}, (3203, 3204), 4)
    for a, b in zip(ref, got):
        for x, y in zip(a[0], b[0]):
            assert np.array_equal(x, y)


def test_autotune_leaves_results_unchanged(bank):
    """Trials restore every scene: stepping after the tuner equals stepping without it."""
    for seeds in ((3205, 3206, 3207), (3208,)):
        ref, _ = run_batch(bank, {}, seeds, 3, single=len(seeds) == 1)
        got, _ = run_batch(bank, {}, seeds, 3, single=len(seeds) == 1, tune_first=True)
        for a, b in zip(ref, got):
            for sa, sb in zip(a, b):
                for x, y in zip(sa, sb):
                    assert np.array_equal(x, y), seeds


def test_tuning_rejects_bad_values():
    ctx = native.Context(0, "fp32")
    try:
        with pytest.raises(Exception):
            ctx.set_tuning(resident_contacts=100)          # not a multiple of 64
        with pytest.raises(Exception):
            ctx.set_tuning(resident_contacts=1 << 16, resident_bodies=1 << 16)   # beyond the LDS layout
        with pytest.raises(Exception):
            ctx.set_tuning(waves_per_group=16)             # only 4 or 8
    finally:
        ctx.close()


def test_serving_snapshot_of_a_running_request(bank, tmp_path):
    """A request snapshot between ticks (from on_tick) and read back finishes
    its remaining steps bitwise like the request left running."""
    from rainbow_amd import xml_io
    ctx = native.Context(0, "fp32")
    try:
        srv = serving.Server(ctx=ctx, budget=1 << 30)
        a = scenes.serving_request(4400, bank, nmax=40, params=dict(precision="fp32"))
        b = scenes.serving_request(4401, bank, nmax=40, params=dict(precision="fp32"))
        reqs = [srv.prepare(serving.Request(0, a, 14, 0.0)), srv.prepare(serving.Request(1, b, 9, 0.0))]
        fn = str(tmp_path / "mid.xml")
        taken = {}

        def on_tick(tick, n, dt):
            if tick == 5:
                taken["n"] = srv.checkpoint(reqs[0], fn)

        srv.run(reqs, on_tick=on_tick)
        assert taken["n"] == 5
        rest = srv.prepare(serving.Request(2, xml_io.read_xml(fn), 14 - 5, 0.0))
        srv.run([rest])
        for x, y in zip(rest.state, reqs[0].state):
            assert np.array_equal(x, y)
    finally:
        ctx.close()


def test_large_unit_colouring_paths_bitwise(bank):
    """A pile of 1024 spheres (one island of more than 2048 contacts) is
    coloured the same with its colouring state in LDS and in global memory
    (color_lds_contacts = 64), and steps to the same states."""
    ctx = native.Context(0, "fp32")
    try:
        eng = scenes.serving_request(4321, bank, nmin=1024, nmax=1024, params=dict(precision="fp32"))
        desc, keep = engine_desc(eng, ctx)
        sid = ctx.scene_create(desc)
        ctx.set_state(sid, *gather_state(list(eng.bodies.values())))
        ctx.step([sid], 300, 0.0, 0)
        ctx.sync()
        start = ctx.get_state(sid, len(eng.bodies))
        runs = []
        for cap in (0, 64):
            ctx.set_tuning(color_lds_contacts=cap)
            ctx.set_state(sid, *start)
            ctx.set_step_index(sid, 300)
            snaps = []
            for _ in range(3):
                ctx.step([sid], 1, 0.0, 0)
                c = ctx.get_contacts(sid)
                snaps.append([c["color"], c["impulse"]] + list(ctx.get_state(sid, len(eng.bodies))))
            runs.append(snaps)
        assert len(runs[0][-1][0]) > 2048, len(runs[0][-1][0])
        for a, b in zip(*runs):
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
    finally:
        ctx.close()


def test_half_layout_global_groups_bitwise(bank):
    """A batch scene of more bodies than the half LDS layout holds (800
    spheres; about 700 fit) solves from the per-contact items in global
    memory under solve_groups_per_cu = 2, and from the velocity-only layout
    under 1: both give bitwise the same colours, impulses and states."""
    ctx = native.Context(0, "fp32")
    try:
        engs = [scenes.serving_request(4400, bank, nmin=800, nmax=800, params=dict(precision="fp32")),
                scenes.mixed_drop(3210, bank, nmax=96, params=dict(precision="fp32"))]
        sids = []
        for e in engs:
            desc, keep = engine_desc(e, ctx)
            sid = ctx.scene_create(desc)
            ctx.set_state(sid, *gather_state(list(e.bodies.values())))
            sids.append(sid)
        ctx.step(sids, 300, 0.0, 0)
        ctx.sync()
        start = [ctx.get_state(sid, len(e.bodies)) for sid, e in zip(sids, engs)]
        runs = []
        for groups in (1, 2):
            ctx.set_tuning(solve_groups_per_cu=groups, lanes_per_scene=1)
            for sid, st in zip(sids, start):
                ctx.set_state(sid, *st)
                ctx.set_step_index(sid, 300)
            snaps = []
            for _ in range(3):
                ctx.step(sids, 1, 0.0, 0)
                for sid, e in zip(sids, engs):
                    c = ctx.get_contacts(sid)
                    snaps.append([c["color"], c["impulse"]] + list(ctx.get_state(sid, len(e.bodies))))
            runs.append(snaps)
        assert len(runs[0][0][0]) > 500, len(runs[0][0][0])
        for a, b in zip(*runs):
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
    finally:
        ctx.close()


def test_serving_alone_equals_mixed(bank):
    ctx = native.Context(0, "fp32")
    try:
        srv = serving.Server(ctx=ctx, budget=1 << 30)
        arrivals = serving.poisson_arrivals(400.0, 6, seed=7)
        reqs = []
        for i in range(6):
            eng = scenes.serving_request(4100 + i, bank, nmax=40, params=dict(precision="fp32"))
            reqs.append(srv.prepare(serving.Request(i, eng, 10 + 4 * i, float(arrivals[i]))))
        rep = srv.run(reqs)
        assert rep["scene_steps"] == sum(r.n_steps for r in reqs)
        assert rep["ticks"] >= max(r.n_steps for r in reqs)
        for r in reqs:
            alone = srv.run_alone(r)
            for x, y in zip(alone, r.state):
                assert np.array_equal(x, y), r.rid
    finally:
        ctx.close()


def test_serving_hinges_and_checkpoint(bank, tmp_path):
    """A hinged request is served like any other (bitwise its run alone); a
    finished request checkpointed to a scene file and read back continues
    bitwise like the same request run for all the steps at once."""
    from rainbow_amd import xml_io
    ctx = native.Context(0, "fp32")
    try:
        srv = serving.Server(ctx=ctx, budget=1 << 30)
        chain = scenes.cube_hinge_chain(bank, n=4, edge=1.0, params=dict(precision="fp32"))
        drop = scenes.serving_request(4200, bank, nmax=40, params=dict(precision="fp32"))
        reqs = [srv.prepare(serving.Request(0, chain, 12, 0.0)), srv.prepare(serving.Request(1, drop, 10, 0.0))]
        srv.run(reqs)
        for r in reqs:
            for x, y in zip(srv.run_alone(r), r.state):
                assert np.array_equal(x, y), r.rid
        fn = str(tmp_path / "drop.xml")
        srv.checkpoint(reqs[1], fn)
        resumed = srv.prepare(serving.Request(2, xml_io.read_xml(fn), 8, 0.0))
        srv.run([resumed])
        whole = srv.prepare(serving.Request(3, drop, 18, 0.0))
        for x, y in zip(srv.run_alone(whole), resumed.state):
            assert np.array_equal(x, y)
    finally:
        ctx.close()
