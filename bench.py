#!/usr/bin/env python3
"""Throughput of the prox rigid-body time step on one MI355X.

Workload: 4096 spheres (create_sphere(0.5, 16, 16), 32^3 SDF) dropped from
[-9.5, 9.5] x [1, 40] x [-9.5, 9.5] onto a fixed ground box, mu = 0.3,
dt = 0.005, 50 coloured parallel Gauss-Seidel sweeps per step (tolerances 0),
FP32 state, FP64 merit/residual.  The run is 200 steps long; the timed window
is its last --steps steps (never starting before step --warmup + 1): the
steps before the window run untimed, so under any --steps/--warmup the window
covers the run's contact regime (the spheres dropped from low heights have
landed and piled up by then).

metric: contact-iterations per second over the whole step (collision
detection + assembly + colouring + solve + integration), inputs resident in
HBM.  A contact-iteration is one projection of one contact's 4-row block in
one sweep, rejected sweeps included (gauss_seidel_solver.py:38-71).

roofline: the solve kernel (the largest single kernel of the step) against
HBM peak, with 288 algorithmic bytes per contact-iteration; its duration is
the mean over the timed steps of HIP events recorded on the library's stream.
traffic: HBM bytes per solve launch from rocprofv3 FETCH_SIZE/WRITE_SIZE runs
of this same command, read from the committed summary under profiles/.  The
timed window is bracketed by roctxProfilerResume/Pause, so a profile taken
with rocprofv3 --selected-regions covers exactly the timed steps.
roofline.latency: the solve is bound by its dependent colour phases (each
sweep visits the colours of its longest sweep lane one after another); one
step after the timed region gives colours x sweeps and the microseconds per
phase.

cpu_baseline: the FP64 NumPy restatement of the step, single-threaded, on one
step of the same scene from the GPU's state at the end of the timed window
("port", O(K) per sweep); beside it ("faithful") the reference's own sweep on
its scipy.sparse structures (a full J.dot(w) and WJT.tocsr() per contact,
contacts.py:89, 107) on the same contact problem for a bounded time on one
core.  The host's CPU model and core count are recorded.

batched (on by default, --no-batched to skip): 1024 independent scenes of
8..512 boxes and spheres (log-uniform counts, mu ~ U[0.1, 0.8], dt = 0.01,
30 GS sweeps) advanced together by one hipGraph replay per step, timed after
the drops have landed (steps 156..175); scene-steps/s and per-step p50/p99
beside the plain path (one scene per launch) on a sample of the same scenes
at the same steps, whose end states must equal the batched ones; the FP64
oracle on the host's cores (one scene per process) as the CPU baseline.

serving (on by default, --no-serving to skip): Poisson arrivals of
sphere-drop requests (8..1024 bodies, 50..500 steps) admitted against an HBM
budget and advanced by continuous batching, 64 requests at 10, 50 and 200
requests/s, 256 requests all waiting at t = 0 (saturation), and the 64
requests waiting under a quarter of their footprint (admission queues);
launch shapes from the autotuner; p50/p99 per-step latency, completed
scene-steps/s, admission wait, the plain path (one request per launch) on a
sample of the requests over their whole lifetimes, and the alone-vs-mixed
difference of the checked requests.

stress (on by default, --no-stress to skip): 128 noisy icospheres and 128
torus knots (about 10^4 triangles each, 64^3 SDFs) poured into a fixed funnel,
300 steps from rest; contact-iters/s over the run, and per-stage times of the
last 50 steps.

hinges (on by default, --no-hinges to skip): the reference's chain of five
hinged cubes alone and as a batch of 256 scenes, and a chain of 256 cubes;
100 sweeps per step; hinge-iterations/s beside the FP64 oracle's.

--copy-calibration BYTES runs only a plain device-to-device copy of BYTES
bytes (the reference kernel for scaling the profiler's byte counters).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BYTES_PER_CONTACT_ITER = 288
HBM_PEAK_GBS = 8000.0
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_solve_multi_spheres4096_pmc.txt")
WORKLOAD = "sphere_rain_4096"
RUN_STEPS = 200      # length of the 4096-sphere run; the timed window is its last steps


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def copy_calibration(nbytes):
    import torch
    n = nbytes // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log("copied %d bytes 3 times" % (4 * n))


def read_traffic():
    """HBM bytes per solve launch from the committed rocprofv3 summary."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def host_info():
    """CPU model and core counts of this host."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "os_cpu_count": os.cpu_count(), "usable_cores": usable}


class ProfilerRegion:
    """roctxProfilerResume(0) / roctxProfilerPause(0) of the ROCm marker
    library: under `rocprofv3 --selected-regions` only what runs between them
    is traced or counted.  Without the profiler (or the library) both are
    no-ops."""

    def __init__(self):
        import ctypes
        self.lib = None
        for name in ("librocprofiler-sdk-roctx.so.1", "/opt/rocm/lib/librocprofiler-sdk-roctx.so.1"):
            try:
                self.lib = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if self.lib is not None:
            for fn in ("roctxProfilerResume", "roctxProfilerPause"):
                getattr(self.lib, fn).argtypes = [ctypes.c_uint64]
                getattr(self.lib, fn).restype = ctypes.c_int

    def resume(self):
        if self.lib is not None:
            self.lib.roctxProfilerResume(0)

    def pause(self):
        if self.lib is not None:
            self.lib.roctxProfilerPause(0)


def cpu_baseline(eng, state, steps_done, faithful_seconds=10.0):
    """One oracle step of the scene from ``state`` (the GPU's state after
    ``steps_done`` steps), and the reference's own sweep on that step's
    contact problem for ``faithful_seconds``; both on one thread."""
    from threadpoolctl import threadpool_limits
    from oracle import faithful, problem as OP, scene as OS, stepper as OST
    r, q, v, w = state
    for k, b in enumerate(eng.bodies.values()):
        if not b.is_fixed:
            b.r, b.q, b.v, b.w = r[k].copy(), q[k].copy(), v[k].copy(), w[k].copy()
    sc = OS.from_engine(eng)
    sc.step_index = steps_done
    Winv = OP.body_mass_blocks(sc.copy())[0]
    dt = eng.params.time_step
    with threadpool_limits(1):
        t0 = time.perf_counter()
        out = OST.step(sc, dt)
        t = time.perf_counter() - t0
        lg = out["log"]
        K = int(lg["contact_points"])
        sweeps = int(lg.get("iterations", 0)) + (1 if lg.get("exitcode", 0) else 0) if K else 0
        fa = None
        if K:
            order = np.concatenate(out["groups"])
            rate, done, tf, tbuild = faithful.time_sweep(out["problem"], Winv, order, faithful_seconds)
            fa = {"value": rate, "unit": "contact-iters/s", "cores": 1, "kind": "port",
                  "sample": "the reference's sweep (contacts.py:66-107: scipy.sparse J and bsr W J^T, a full "
                            "J.dot(w) and WJT.tocsr() per contact) on the same step's contact problem (K=%d) in "
                            "colour order from x = 0, %d contact updates in %.1f s, 1 thread; a whole step's %d "
                            "sweeps would take %.0f s" % (K, done, tf, sweeps, K * sweeps / rate if rate else 0),
                  "seconds": tf, "contact_updates": done, "matrix_build_s": tbuild}
    return dict(value=K * sweeps / t if t > 0 else 0.0, unit="contact-iters/s", cores=1, kind="port",
                sample="one full time step (collision detection, assembly, colouring, %d GS sweeps) of the "
                       "same 4096-sphere scene from the GPU state after step %d, FP64 NumPy restatement "
                       "(O(K) vectorised sweep), 1 thread: K=%d contacts, %.2f s" % (sweeps, steps_done, K, t),
                seconds=t, contacts=K, sweeps=sweeps, natural_residual=float(lg.get("natural_residual", 0.0)),
                faithful=fa, host=host_info())


def batched_leg(n_scenes, steps, settle, plain_sample, plain_steps, cpu=True):
    """1024 mixed box/sphere scenes stepped as one batch vs one scene per
    launch, both timed after the drops have landed: the batch first advances
    `settle` steps untimed (1.5 s at dt = 0.01; the highest drop, 10 m, lands
    after 1.4 s), then `plain_steps` steps whose end states the plain path
    must reproduce, then the timed `steps` steps.  The plain path steps the
    sample scenes from their states at step `settle`, one scene per launch."""
    import numpy as np
    from rainbow_amd import autotune, native, scenes
    from rainbow_amd.stepper import engine_desc, gather_state, shared_context
    t0 = time.perf_counter()
    ctx = shared_context(0, "fp32")
    bank = scenes.ShapeBank(resolution=32)
    engs = [scenes.mixed_drop(3000 + s, bank) for s in range(n_scenes)]
    sids, init = [], []
    for e in engs:
        desc, keep = engine_desc(e, ctx)
        sid = ctx.scene_create(desc)
        st0 = gather_state(list(e.bodies.values()))
        ctx.set_state(sid, *st0)
        sids.append(sid)
        init.append((desc, keep, st0, len(e.bodies)))
    n_bodies = sum(len(e.bodies) for e in engs)
    log("batched: %d scenes, %d bodies built in %.1f s" % (n_scenes, n_bodies, time.perf_counter() - t0))
    sample = list(range(0, n_scenes, max(1, n_scenes // plain_sample)))[:plain_sample]
    # untimed: the drops land
    ctx.step(sids, settle, 0.0, native.STEP_GRAPH)
    ctx.sync()
    s_settle = {i: ctx.get_state(sids[i], init[i][3]) for i in sample}
    ctx.step(sids, plain_steps, 0.0, native.STEP_GRAPH)
    ctx.sync()
    s_ref = {i: ctx.get_state(sids[i], init[i][3]) for i in sample}
    at = settle + plain_steps
    log("batched: %d steps untimed (landed)" % at)
    # launch shapes picked on this batch (the trials restore every scene's state)
    tuned = autotune.autotune(ctx, sids, [x[3] for x in init], steps=3, log=log)
    for sid in sids:
        ctx.scene_totals(sid, clear=True)
    t1 = time.perf_counter()
    ctx.step(sids, steps, 0.0, native.STEP_GRAPH)
    ctx.sync()
    tb = time.perf_counter() - t1
    tot = [ctx.scene_totals(sid) for sid in sids]
    lat = []
    for _ in range(10):
        t2 = time.perf_counter()
        ctx.step(sids, 1, 0.0, native.STEP_GRAPH)
        ctx.sync()
        lat.append(time.perf_counter() - t2)
    citers = sum(t["contact_iters"] for t in tot)
    contacts = sum(t["contacts"] for t in tot)
    res = np.array([ctx.get_log(sid, 30)["natural_residual"] for sid in sids])
    # one more batch step launched kernel by kernel with stage events: the batched
    # solve kernel's duration for its roofline entry (a graph replay has no stage marks)
    for sid in sids:
        ctx.scene_totals(sid, clear=True)
    ctx.step(sids, 1, 0.0, native.STEP_TIMING)
    stage = ctx.get_timings()
    solve_ci = sum(ctx.scene_totals(sid)["contact_iters"] for sid in sids)
    solve_ms = 1e3 * stage["solver_time"]
    solve_gbs = BYTES_PER_CONTACT_ITER * solve_ci / (solve_ms * 1e-3) / 1e9 if solve_ms > 0 else 0.0
    end_states = {i: ctx.get_state(sids[i], init[i][3]) for i in range(n_scenes)} if cpu else None
    # plain path: the sample scenes from their states at step `settle`, one scene per launch
    dup = []
    for i in sample:
        desc, keep, st0, n = init[i]
        sid = ctx.scene_create(desc)
        ctx.set_state(sid, *s_settle[i])
        ctx.set_step_index(sid, settle)
        ctx.step([sid], 1, 0.0, 0)           # first launch of each scene (workspace sizing), untimed
        ctx.set_state(sid, *s_settle[i])
        ctx.set_step_index(sid, settle)
        dup.append(sid)
    ctx.sync()
    t3 = time.perf_counter()
    for _ in range(plain_steps):
        for sid in dup:
            ctx.step([sid], 1, 0.0, 0)
    ctx.sync()
    tp = time.perf_counter() - t3
    diff = 0.0
    for i, sid in zip(sample, dup):
        got = ctx.get_state(sid, init[i][3])
        diff = max(diff, max(float(np.abs(a - b).max()) for a, b in zip(got, s_ref[i])))
    for sid in dup:
        ctx.scene_destroy(sid)
    for sid in sids:
        ctx.scene_destroy(sid)
    ctx.set_tuning(**autotune.DEFAULT)
    batched = n_scenes * steps / tb
    plain = len(dup) * plain_steps / tp
    out = {"workload": "mixed_drop_%d_scenes" % n_scenes, "bodies": n_bodies, "steps": steps,
           "steps_timed": "%d..%d" % (at + 1, at + steps),
           "scene_steps_per_s": batched, "ms_per_step": 1e3 * tb / steps,
           "p50_step_ms": 1e3 * float(np.percentile(lat, 50)), "p99_step_ms": 1e3 * float(np.percentile(lat, 99)),
           "contact_iters_per_s": citers / tb if tb > 0 else 0.0,
           "contacts_per_scene_step": contacts / (n_scenes * steps),
           "roofline": {"kernel": "k_solve", "bound": "hbm", "achieved": solve_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": solve_gbs / HBM_PEAK_GBS, "bytes_per_launch": BYTES_PER_CONTACT_ITER * solve_ci,
                        "launch_ms": solve_ms,
                        "note": "one eager batch step after the timed ones; 288 B per contact-iteration"},
           "stages_ms": {k: round(1e3 * v, 4) for k, v in stage.items()},
           "natural_residual_last_step": {"median": float(np.median(res)), "max": float(res.max())},
           "plain": {"scene_steps_per_s": plain, "sample_scenes": len(dup), "steps": plain_steps,
                     "steps_timed": "%d..%d" % (settle + 1, at),
                     "ms_per_scene_step": 1e3 * tp / (len(dup) * plain_steps)},
           "speedup_vs_plain": batched / plain if plain > 0 else None,
           "max_abs_state_diff_vs_plain": diff,
           "tuning": tuning_report(tuned)}
    if cpu:
        log("batched: cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline_scenes(engs, end_states, at + steps + 11, "mixed_drop")
    return out


def batched_trace_leg(n_scenes, settle=155, steps=5):
    """The batched leg's scenes stepped kernel by kernel, for a profiler's kernel
    trace: `settle` steps (the drops land), then `steps` steps with one sweep
    lane per scene (the shape the autotuner picks on this batch), one launch
    sequence and one synchronisation each.  No hipGraph: a graph replay of this
    batch under rocprofv3's kernel trace ended the process."""
    from rainbow_amd import native, scenes
    from rainbow_amd.stepper import engine_desc, gather_state, shared_context
    ctx = shared_context(0, "fp32")
    bank = scenes.ShapeBank(resolution=32)
    engs = [scenes.mixed_drop(3000 + s, bank) for s in range(n_scenes)]
    sids = []
    for e in engs:
        desc, keep = engine_desc(e, ctx)
        sid = ctx.scene_create(desc)
        ctx.set_state(sid, *gather_state(list(e.bodies.values())))
        sids.append(sid)
    log("batched trace: %d scenes built; %d steps to land" % (n_scenes, settle))
    ctx.step(sids, settle, 0.0, 0)
    ctx.sync()
    ctx.set_tuning(lanes_per_scene=1)
    ms = []
    for k in range(steps):
        t0 = time.perf_counter()
        ctx.step(sids, 1, 0.0, 0)
        ctx.sync()
        ms.append(1e3 * (time.perf_counter() - t0))
        log("batched trace: step %d" % (settle + k + 1))
    for sid in sids:
        ctx.scene_destroy(sid)
    from rainbow_amd import autotune
    ctx.set_tuning(**autotune.DEFAULT)
    return {"workload": "mixed_drop_%d_scenes" % n_scenes, "steps_traced": "%d..%d" % (settle + 1, settle + steps),
            "ms_per_step_eager": ms}


def _oracle_scene_steps(job):
    """Worker: one oracle step of each (scene, dt) in job on one thread; seconds per scene."""
    from threadpoolctl import threadpool_limits
    from oracle import stepper as OST
    out = []
    with threadpool_limits(1):
        for sc, dt in job:
            t0 = time.perf_counter()
            OST.step(sc, dt)
            out.append(time.perf_counter() - t0)
    return out


def cpu_baseline_scenes(engs, states, step_index, what, seconds=15.0):
    """The FP64 oracle's step (O(K) vectorised sweep) of independent scenes
    from the GPU's states (step_index: the steps each state has taken, one
    number or one per scene): on one core over the first scenes for about
    `seconds`, then one scene per process on every usable core over about
    `seconds` of work per core."""
    import multiprocessing as mp
    from oracle import scene as OS
    hi = host_info()

    def oracle_scene(i):
        e = engs[i]
        r, q, v, w = states[i]
        for k, b in enumerate(e.bodies.values()):
            if not b.is_fixed:
                b.r, b.q, b.v, b.w = r[k].copy(), q[k].copy(), v[k].copy(), w[k].copy()
        sc = OS.from_engine(e)
        sc.step_index = step_index[i] if isinstance(step_index, (list, tuple)) else step_index
        return sc, e.params.time_step

    # one core: scenes in order until `seconds` pass
    one = []
    t0 = time.perf_counter()
    for i in range(len(engs)):
        one += _oracle_scene_steps([oracle_scene(i)])
        if time.perf_counter() - t0 > seconds:
            break
    per_scene = float(np.mean(one))
    # all cores: enough scenes for about `seconds` per core, one process per core (fork: the workers
    # only run NumPy and leave with os._exit, so they never touch the GPU)
    cores = hi["usable_cores"]
    n_all = min(len(engs), max(cores, int(cores * seconds / max(per_scene, 1e-3))))
    jobs = [oracle_scene(i) for i in range(n_all)]
    chunks = [jobs[k::cores] for k in range(cores) if jobs[k::cores]]
    t1 = time.perf_counter()
    with mp.get_context("fork").Pool(len(chunks)) as pool:
        res = pool.map(_oracle_scene_steps, chunks)
    wall = time.perf_counter() - t1
    done = sum(len(r) for r in res)
    return {"value": done / wall, "unit": "scene-steps/s", "cores": len(chunks), "kind": "port",
            "sample": "one FP64 oracle step (NumPy, O(K) sweep) of each of %d %s scenes from the GPU states %s, "
                      "one scene per process on %d cores, %.1f s wall" % (
                          done, what, "at the end of their runs" if isinstance(step_index, (list, tuple))
                          else "after step %d" % step_index, len(chunks), wall),
            "single_core": {"value": 1.0 / per_scene, "unit": "scene-steps/s", "cores": 1,
                            "sample": "%d scenes in order, %.1f s" % (len(one), sum(one))},
            "host": hi}


def tuning_report(res):
    keep = lambda t: {k: v for k, v in t.items() if v}
    return {"chosen": keep(res["best"]), "ms_per_step": round(res["ms_per_step"], 4),
            "trials": [[keep(t), round(ms, 4)] for t, ms in res["trials"]]}


def serving_leg(n_req, rates, budget_gib, alone_checks, plain_sample=8, sat_requests=256, cpu=True):
    """Poisson arrivals served by continuous batching, one run per arrival rate
    on the same n_req requests, and a saturation run (every one of
    sat_requests requests waiting at t = 0; the first n_req are the same
    requests); the n_req requests again all waiting under a budget of a
    quarter of their footprint, so admission queues; the plain path (one
    request per launch, no graph) on a sample, each over its whole lifetime.
    Launch shapes come from the autotuner, run on the 32 largest requests
    after 300 steps (when their bodies have piled up)."""
    import numpy as np
    from rainbow_amd import autotune, native, scenes, serving
    t0 = time.perf_counter()
    bank = scenes.ShapeBank(resolution=32)
    rng = np.random.default_rng(4)
    srv = serving.Server(budget=budget_gib << 30)
    reqs = []
    for i in range(max(n_req, sat_requests)):
        eng = scenes.serving_request(4000 + i, bank)
        reqs.append(srv.prepare(serving.Request(i, eng, int(rng.integers(50, 501)), 0.0)))
    log("serving: %d requests (%d bodies, %d scene-steps) prepared in %.1f s" % (
        len(reqs), sum(len(r.engine.bodies) for r in reqs), sum(r.n_steps for r in reqs), time.perf_counter() - t0))
    ctx = srv.ctx
    # launch shapes for the served batches
    tsids = []
    big = sorted(reqs, key=lambda r: -len(r.engine.bodies))[:32]
    for r in big:
        sid = ctx.scene_create(r._desc)
        ctx.set_state(sid, *r._state0)
        tsids.append(sid)
    ctx.step(tsids, 300, 0.0, native.STEP_GRAPH)
    ctx.sync()
    tuned = autotune.autotune(ctx, tsids, [len(r.engine.bodies) for r in big], steps=3, log=log)
    for sid in tsids:
        ctx.scene_destroy(sid)
    runs = []
    for rate in list(rates) + ["saturation"]:
        sub = reqs[:n_req] if rate != "saturation" else reqs[:sat_requests]
        arr = serving.poisson_arrivals(rate, len(sub), seed=4) if rate != "saturation" else np.zeros(len(sub))
        for r, a in zip(sub, arr):
            r.arrival = float(a)
        rep = srv.run(sub)
        rep["rate_per_s"] = rate
        runs.append(rep)
        log("serving at %s req/s (%d requests): %.0f scene-steps/s, p50 %.2f ms, p99 %.2f ms, wait p99 %.1f ms" % (
            rate, len(sub), rep["scene_steps_per_s"], rep["p50_step_ms"], rep["p99_step_ms"], rep["p99_wait_ms"]))
    diffs = []
    for r in reqs[:alone_checks]:
        alone = srv.run_alone(r)
        diffs.append(max(float(np.abs(a - b).max()) for a, b in zip(alone, r.state)))
    # a budget that binds: the first n_req requests all waiting at t = 0 under a quarter of their footprint
    sub = reqs[:n_req]
    for r in sub:
        r.arrival = 0.0
    full_budget = srv.budget
    srv.budget = max(max(r.footprint for r in sub), sum(r.footprint for r in sub) // 4)
    rep = srv.run(sub)
    rep["rate_per_s"] = "saturation, budget 1/4 of the requests' footprint"
    srv.budget = full_budget
    bound = rep
    log("serving under a binding budget (%d of %d bytes): %.0f scene-steps/s, max %d active, wait p99 %.1f ms" % (
        bound["budget_bytes"], sum(r.footprint for r in sub), bound["scene_steps_per_s"], bound["max_active"],
        bound["p99_wait_ms"]))
    # plain path: a sample of the requests, each alone over its whole lifetime (one scene per launch, no
    # graph, one step per call as a serving loop without batching would take them)
    sample = reqs[:n_req][::max(1, n_req // plain_sample)][:plain_sample]
    tp, psteps = 0.0, 0
    for r in sample:
        sid = ctx.scene_create(r._desc)
        ctx.set_state(sid, *r._state0)
        ctx.step([sid], 1, 0.0, 0)               # workspace sizing, untimed
        ctx.set_state(sid, *r._state0)
        ctx.set_step_index(sid, 0)
        ctx.sync()
        t1 = time.perf_counter()
        for _ in range(r.n_steps):
            ctx.step([sid], 1, 0.0, 0)
        ctx.sync()
        tp += time.perf_counter() - t1
        psteps += r.n_steps
        got = ctx.get_state(sid, len(r.engine.bodies))
        diffs.append(max(float(np.abs(a - b).max()) for a, b in zip(got, r.state)))
        ctx.scene_destroy(sid)
    ctx.set_tuning(**autotune.DEFAULT)
    plain = psteps / tp
    sat = runs[-1]
    cb = None
    if cpu:
        log("serving: cpu baseline ...")
        done = [r for r in reqs[:sat_requests] if r.state is not None]
        cb = cpu_baseline_scenes([r.engine for r in done], [r.state for r in done], [r.n_steps for r in done],
                                 "served-request")
    return {"workload": "sphere_drop_requests", "requests": n_req, "saturation_requests": sat_requests,
            "budget_bytes": budget_gib << 30, "tuning": tuning_report(tuned), "runs": runs,
            "binding_budget": bound,
            "plain": {"scene_steps_per_s": plain, "sample_requests": len(sample), "scene_steps": psteps,
                      "note": "each sampled request alone over its whole lifetime, one step per launch",
                      "ms_per_scene_step": 1e3 * tp / psteps},
            "speedup_vs_plain_at_saturation": sat["scene_steps_per_s"] / plain if plain > 0 else None,
            "max_abs_diff_alone_vs_mixed": max(diffs) if diffs else None, "alone_checked": len(diffs),
            "cpu_baseline": cb}


def stress_leg(steps):
    """Noisy icospheres and torus knots poured into a funnel (64^3 SDFs), from rest."""
    from rainbow_amd import native, scenes
    t0 = time.perf_counter()
    eng = scenes.funnel_stress()
    st = eng.stepper
    st._prepare(eng)
    ctx, sid = st.ctx, st.scene
    dt = eng.params.time_step
    tris = sum(len(b.shape.mesh.T) for b in eng.bodies.values())
    log("stress: %d bodies, %d triangles, built in %.1f s" % (len(eng.bodies), tris, time.perf_counter() - t0))
    ctx.scene_totals(sid, clear=True)
    t1 = time.perf_counter()
    ctx.step([sid], steps - 50, dt, 0)
    ctx.sync()
    part = ctx.scene_totals(sid, clear=True)
    t2 = time.perf_counter()
    ctx.step([sid], 50, dt, native.STEP_TIMING)      # the last 50 steps with stage timings
    ctx.sync()
    t3 = time.perf_counter()
    last = ctx.scene_totals(sid)
    tm = ctx.get_timings()
    lg = ctx.get_log(sid, eng.params.max_iterations)
    tot = {k: part[k] + last[k] for k in part}
    # the narrow phase's roofline over the last 50 steps: Frank-Wolfe's algorithmic bytes
    # (48 B per candidate triangle, 32 B per trilinear cell evaluation) over its stage time
    fw_bytes = (48 * last["candidates"] + 32 * last["fw_cell_evals"]) / 50
    fw_ms = 1e3 * tm["contact_determination"]
    fw_gbs = fw_bytes / (fw_ms * 1e-3) / 1e9 if fw_ms > 0 else 0.0
    ctx.scene_destroy(sid)
    st.scene = None
    return {"workload": "funnel_stress_256", "bodies": len(eng.bodies), "triangles": tris, "sdf_resolution": 64,
            "steps": steps, "contact_iters_per_s": tot["contact_iters"] / (t3 - t1),
            "ms_per_step": 1e3 * (t3 - t1) / steps,
            "last_50": {"ms_per_step": 1e3 * (t3 - t2) / 50, "contacts_per_step": last["contacts"] / 50,
                        "candidates_per_step": last["candidates"] / 50,
                        "contact_iters_per_s": last["contact_iters"] / (t3 - t2),
                        "stages_ms": {k: round(1e3 * v, 4) for k, v in tm.items()},
                        "narrow_phase": {"kernel": "k_frank_wolfe", "bound": "hbm", "bytes_per_step": fw_bytes,
                                         "launch_ms": fw_ms, "achieved_GBs": fw_gbs,
                                         "frac": fw_gbs / HBM_PEAK_GBS,
                                         "cell_evals_per_step": last["fw_cell_evals"] / 50}},
            "last_step": {"contacts": lg["contact_points"], "colors": lg["colors"], "sweeps": lg["sweeps"],
                          "natural_residual": lg["natural_residual"]}}


def hinges_leg(steps, graph=True):
    """Hinge chains (constraints/hinges.py): the reference's chain of five cubes
    of edge 10 alone and as a batch of 256 copies, and one chain of 256 unit
    cubes; 100 sweeps per step (tolerances 0), FP32.  A hinge-iteration is one
    hinge's 5-row update in one sweep.  CPU: the FP64 oracle on one step of
    the same scenes (10 sweeps for the long chain)."""
    import oracle.scene as OS
    import oracle.stepper as OST
    from rainbow_amd import native, scenes
    from rainbow_amd.stepper import engine_desc, gather_state, upload_hinges
    bank = scenes.ShapeBank(resolution=32)
    prm = dict(precision="fp32", max_iterations=100, absolute_tolerance=0.0, relative_tolerance=0.0)
    out = {}
    for name, n, edge, copies in (("reference_chain_5", 5, 10.0, 1), ("reference_chain_5_x256", 5, 10.0, 256),
                                  ("chain_256", 256, 1.0, 1)):
        ctx = native.Context(0, "fp32")
        try:
            engines = [scenes.cube_hinge_chain(bank, n=n, edge=edge, params=prm) for _ in range(copies)]
            ids = []
            for e in engines:
                desc, keep = engine_desc(e, ctx)
                sid = ctx.scene_create(desc)
                ctx.set_state(sid, *gather_state(list(e.bodies.values())))
                upload_hinges(ctx, sid, e)
                ids.append(sid)
            flags = native.STEP_GRAPH if (copies > 1 and graph) else 0
            ctx.step(ids, 10, 0.01, flags)
            ctx.sync()
            t0 = time.perf_counter()
            ctx.step(ids, steps, 0.01, flags)
            ctx.sync()
            sec = time.perf_counter() - t0
            ctx.step(ids[:1], 1, 0.01, native.STEP_TIMING)
            tm = ctx.get_timings()
            lg = ctx.get_log(ids[0], 100)
            hi = copies * (n - 1) * lg["sweeps"] * steps
            out[name] = {"scenes": copies, "hinges_per_scene": n - 1, "sweeps": lg["sweeps"],
                         "contacts": lg["contact_points"], "ms_per_step": 1e3 * sec / steps,
                         "hinge_iters_per_s": hi / sec, "solver_ms_one_scene": round(1e3 * tm["solver_time"], 4)}
            if copies == 1:
                sweeps = 100 if n <= 8 else 10
                sc = OS.from_engine(scenes.cube_hinge_chain(
                    scenes.ShapeBank(resolution=32), n=n, edge=edge, backend=None,
                    params=dict(max_iterations=sweeps, absolute_tolerance=0.0, relative_tolerance=0.0)))
                t0 = time.perf_counter()
                OST.step(sc, 0.01)
                cs = time.perf_counter() - t0
                out[name]["cpu_baseline"] = {"value": (n - 1) * sweeps / cs, "unit": "hinge-iters/s", "cores": 1,
                                             "kind": "port", "sample": "one oracle step with %d sweeps" % sweeps,
                                             "seconds": cs}
        finally:
            ctx.close()
        log("hinges %s: %.3g hinge-iters/s, %.3f ms/step" % (name, out[name]["hinge_iters_per_s"],
                                                             out[name]["ms_per_step"]))
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=180)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--bodies", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-scenes leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--batched-scenes", type=int, default=1024)
    ap.add_argument("--serving-requests", type=int, default=64)
    ap.add_argument("--serving-rates", default="10,50,200",
                    help="arrival rates in requests/s (a saturation run follows)")
    ap.add_argument("--no-stress", action="store_true", help="skip the funnel stress leg")
    ap.add_argument("--stress-steps", type=int, default=300)
    ap.add_argument("--no-hinges", action="store_true", help="skip the hinge-chain leg")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the hinge leg's batched steps one kernel at a time instead of replaying a hipGraph")
    ap.add_argument("--only", choices=("batched", "batched-trace", "serving", "stress", "hinges"), default=None,
                    help="run one of the extra legs alone (for profiling) and print its JSON object")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the step of one scene (or of one batch of scenes) is measured on "
              "a single MI355X; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("RANK", "0")) != 0:
        return
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.only == "batched":
        print(json.dumps(batched_leg(args.batched_scenes, 20, 150, 64, 5, cpu=not args.no_cpu)), flush=True)
        return
    if args.only == "batched-trace":
        print(json.dumps(batched_trace_leg(args.batched_scenes)), flush=True)
        return
    rates = [float(x) for x in args.serving_rates.split(",") if x]
    if args.only == "serving":
        print(json.dumps(serving_leg(args.serving_requests, rates, 64, 3, cpu=not args.no_cpu)), flush=True)
        return
    if args.only == "stress":
        print(json.dumps(stress_leg(args.stress_steps)), flush=True)
        return
    if args.only == "hinges":
        print(json.dumps(hinges_leg(50, graph=not args.no_graph)), flush=True)
        return

    from rainbow_amd import native, scenes
    t0 = time.perf_counter()
    bank = scenes.ShapeBank(resolution=32)
    eng = scenes.sphere_rain(args.bodies, bank)
    st = eng.stepper
    st._prepare(eng)
    ctx, sid = st.ctx, st.scene
    dt = eng.params.time_step
    log("scene built: %d bodies in %.1f s" % (len(eng.bodies), time.perf_counter() - t0))

    # untimed: the warm-up steps and the run up to the timed window, which is
    # the last args.steps steps of the RUN_STEPS-step run (or, for a longer
    # window, the args.steps steps after the warm-up)
    start = max(args.warmup, RUN_STEPS - args.steps)
    if start:
        ctx.step([sid], start, dt, 0)
    ctx.sync()
    ctx.scene_totals(sid, clear=True)
    log("untimed: %d steps (warm-up %d)" % (start, args.warmup))

    region = ProfilerRegion()
    region.resume()
    t0 = time.perf_counter()
    ctx.step([sid], args.steps, dt, native.STEP_TIMING)
    ctx.sync()
    sec = time.perf_counter() - t0
    region.pause()
    tot = ctx.scene_totals(sid)
    tm = ctx.get_timings()
    lg = ctx.get_log(sid, eng.params.max_iterations)
    end_state = ctx.get_state(sid, len(eng.bodies))
    log("timed steps %d..%d in %.3f s; totals %s" % (start + 1, start + args.steps, sec, tot))
    assert tot["steps"] == args.steps, tot

    # one more step, outside the timed region, for the solve's latency figures:
    # its sweeps times the colours of its longest sweep lane are dependent phases
    ctx.step([sid], 1, dt, native.STEP_TIMING)
    ctx.sync()
    lat_tm = ctx.get_timings()
    lat_lg = ctx.get_log(sid, eng.params.max_iterations)
    phases = max(1, lat_lg["colors"] * lat_lg["sweeps"])

    citers = tot["contact_iters"]
    value = citers / sec
    tm = {k: 1e3 * v for k, v in tm.items()}          # ms per step, mean over the timed steps
    solve_ms = tm["solver_time"]
    bytes_launch = BYTES_PER_CONTACT_ITER * citers / args.steps
    achieved = bytes_launch / (solve_ms * 1e-3) / 1e9 if solve_ms > 0 else 0.0
    fw_bytes = (48 * tot["candidates"] + 32 * tot["fw_cell_evals"]) / args.steps
    fw_ms = tm["contact_determination"]
    result = {
        "metric": "contact-iters/s",
        "value": value,
        "unit": "contact-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * sec / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 merit/residual)",
        "data": "synthetic (seeded scene of the reference's own shapes)",
        "config": {"workload": WORKLOAD, "bodies": len(eng.bodies), "dt": dt,
                   "gs_sweeps": eng.params.max_iterations, "mu": 0.3, "sdf_resolution": 32,
                   "run_steps": max(RUN_STEPS, start + args.steps),
                   "steps_timed": "%d..%d" % (start + 1, start + args.steps)},
        "roofline": {"kernel": "k_solve_multi", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(),
                     "bytes_per_launch": bytes_launch, "launch_ms": solve_ms,
                     "latency": {"note": "the solve is bound by dependent colour phases, not by HBM",
                                 "colours": lat_lg["colors"], "sweeps": lat_lg["sweeps"],
                                 "phases": phases, "launch_ms": round(1e3 * lat_tm["solver_time"], 4),
                                 "us_per_phase": 1e6 * lat_tm["solver_time"] / phases}},
        "work": {"contacts_per_step": tot["contacts"] / args.steps,
                 "sweeps_per_step": tot["sweeps"] / args.steps,
                 "candidates_per_step": tot["candidates"] / args.steps,
                 "last_step": {"contacts": lg["contact_points"], "colors": lg["colors"],
                               "sweeps": lg["sweeps"], "natural_residual": lg["natural_residual"]}},
        "stages_ms": {k: round(v, 4) for k, v in tm.items()},
        "narrow_phase": {"kernel": "k_frank_wolfe", "bytes_per_launch": fw_bytes,
                         "achieved_GBs": fw_bytes / (fw_ms * 1e-3) / 1e9 if fw_ms > 0 else 0.0,
                         "frac": fw_bytes / (fw_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if fw_ms > 0 else 0.0,
                         "launch_ms": fw_ms,
                         "note": "48 B per candidate triangle + 32 B per trilinear cell evaluation"},
    }
    if not args.no_cpu:
        log("cpu baseline: one oracle step ...")
        cb = cpu_baseline(eng, end_state, start + args.steps)
        result["cpu_baseline"] = cb
        log("cpu baseline: %.3g contact-iters/s (%.1f s); faithful %s" % (
            cb["value"], cb["seconds"], cb["faithful"] and "%.3g" % cb["faithful"]["value"]))
    else:
        result["cpu_baseline"] = None
    if not args.no_batched:
        log("batched leg ...")
        result["batched"] = batched_leg(args.batched_scenes, 20, 150, 64, 5, cpu=not args.no_cpu)
        log("batched: %.0f scene-steps/s (plain %.0f)" % (result["batched"]["scene_steps_per_s"],
                                                          result["batched"]["plain"]["scene_steps_per_s"]))
    if not args.no_serving:
        log("serving leg ...")
        result["serving"] = serving_leg(args.serving_requests, rates, 64, 3, cpu=not args.no_cpu)
    if not args.no_stress:
        log("stress leg ...")
        result["stress"] = stress_leg(args.stress_steps)
        log("stress: %.3g contact-iters/s, %.2f ms/step" % (result["stress"]["contact_iters_per_s"],
                                                            result["stress"]["ms_per_step"]))
    if not args.no_hinges:
        log("hinges leg ...")
        result["hinges"] = hinges_leg(50, graph=not args.no_graph)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
