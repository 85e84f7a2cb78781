"""Continuous-batching server for scene-simulation requests on one GPU.

A request is one scene (an Engine built with the reference's API) and a number
of steps.  Requests arrive over time (Poisson arrivals at a fixed rate in the
benchmark).  The server keeps every admitted scene's state resident in HBM and
advances all of them together: each tick is one `rbh_step` over the current
batch.  Requests that arrive while others run are admitted between ticks, in
arrival order, as soon as the scenes' device footprint fits the stated HBM
budget.  Finished scenes are read back and their device memory is released.

Scenes are independent and every per-scene computation (colouring seeds,
step index, solve decisions) depends only on that scene, so a request's result
is bitwise the same alone and in any mix.  `run_alone` replays a request by
itself to check this.

The reference has no serving layer; this module is the build's own caller of
the step (api.py:676-693 advances one engine at a time).
"""
import time
from dataclasses import dataclass, field

import numpy as np

from . import native, xml_io
from .stepper import engine_desc, gather_state, shared_context, upload_hinges


@dataclass
class Request:
    """A scene to advance `n_steps` steps; `arrival` in seconds from the start."""
    rid: int
    engine: object
    n_steps: int
    arrival: float = 0.0
    # filled by the server
    admitted: float = -1.0
    finished: float = -1.0
    state: tuple = None
    _desc: object = field(default=None, repr=False)
    _keep: object = field(default=None, repr=False)
    _state0: tuple = field(default=None, repr=False)
    footprint: int = 0


def poisson_arrivals(rate, n, seed):
    """Arrival times of n requests of a Poisson process with `rate` per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def _percentile(x, q):
    return float(np.percentile(np.asarray(x, dtype=np.float64), q)) if len(x) else 0.0


class Server:
    """Admits requests against an HBM budget and steps the admitted ones as one batch.

    ctx: a native.Context (FP32 by default); budget: bytes of device memory
    the admitted scenes may occupy (rbh_scene_footprint per scene);
    graph: replay the step as a hipGraph while the batch composition is
    unchanged; max_batch: cap on concurrently admitted scenes; tuning: launch
    shapes (rbh_tuning fields) set on the context for the served batches.
    """

    def __init__(self, ctx=None, budget=64 << 30, graph=True, max_batch=None, device=0, precision="fp32",
                 tuning=None):
        self.ctx = ctx or shared_context(device, precision)
        self.budget = int(budget)
        self.graph = bool(graph)
        self.max_batch = max_batch
        self._active = {}          # during run(): scene id -> [request, steps taken]
        self.used_bytes = 0        # during run(): footprint of the admitted requests
        if tuning:
            self.ctx.set_tuning(**tuning)

    def _create(self, req):
        sid = self.ctx.scene_create(req._desc)
        self.ctx.set_state(sid, *req._state0)
        self.ctx.set_step_index(sid, int(getattr(req.engine, "step_index", 0)))
        if getattr(req.engine, "hinges", None):
            upload_hinges(self.ctx, sid, req.engine)
        return sid

    def checkpoint(self, req, xml_filename):
        """Write a request's scene as a lossless scene file: a finished request
        at its final state, or, between ticks of ``run`` (from ``on_tick``), an
        admitted one at its current state.  ``xml_io.read_xml`` of the file is
        a scene that continues bitwise from that state (its step count is in
        the file); returns the number of steps the request had taken."""
        done = None
        for sid, (r_, n_) in self._active.items():
            if r_ is req:
                state, done = self.ctx.get_state(sid, len(req.engine.bodies)), n_
        if done is None:
            if req.state is None:
                raise RuntimeError("checkpoint(): request %d is neither running nor finished" % req.rid)
            state, done = req.state, req.n_steps
        eng = req.engine
        r, q, v, w = state
        saved = [(b.r, b.q, b.v, b.w) for b in eng.bodies.values()]
        step0 = getattr(eng, "step_index", 0)
        try:
            for k, b in enumerate(eng.bodies.values()):
                b.r, b.q, b.v, b.w = r[k].copy(), q[k].copy(), v[k].copy(), w[k].copy()
            eng.step_index = step0 + done
            xml_io.write_xml(eng, xml_filename, lossless=True)
        finally:
            for b, (r0, q0, v0, w0) in zip(eng.bodies.values(), saved):
                b.r, b.q, b.v, b.w = r0, q0, v0, w0
            eng.step_index = step0
        return done

    def prepare(self, req):
        """Build the device description of a request's scene (host work done ahead of arrival)."""
        req._desc, req._keep = engine_desc(req.engine, self.ctx)
        bodies = list(req.engine.bodies.values())
        req._state0 = gather_state(bodies)
        req.footprint = self.ctx.scene_footprint(req._desc)
        if req.footprint > self.budget:
            raise ValueError("request %d needs %d bytes, more than the budget %d" % (req.rid, req.footprint,
                                                                                   self.budget))
        return req

    def run(self, requests, clock=time.perf_counter, on_tick=None):
        """Serve `requests` (prepared, sorted or not) until all finish; returns a report dict."""
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting, active = [], {}       # active: sid -> [request, steps done]
        self._active = active
        used = peak = max_active = 0
        lat = []                        # per-step latency samples (s), one per scene per tick
        ticks = 0
        flags = native.STEP_GRAPH if self.graph else 0
        t0 = clock()
        while pending or waiting or active:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.pop(0))
            # admit in arrival order while the budget allows
            while waiting and used + waiting[0].footprint <= self.budget and \
                    (self.max_batch is None or len(active) < self.max_batch):
                req = waiting.pop(0)
                sid = self._create(req)
                used += req.footprint
                req.admitted = clock() - t0
                active[sid] = [req, 0]
            self.used_bytes = used
            peak = max(peak, used)
            max_active = max(max_active, len(active))
            if not active:
                if pending:
                    wait = pending[0].arrival - (clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                continue
            ids = sorted(active)
            ts = clock()
            self.ctx.step(ids, 1, 0.0, flags)
            self.ctx.sync()
            dt_tick = clock() - ts
            ticks += 1
            lat.extend([dt_tick] * len(ids))
            done = []
            for sid in ids:
                active[sid][1] += 1
                if active[sid][1] >= active[sid][0].n_steps:
                    done.append(sid)
            if on_tick:   # every admitted request's step count includes this tick
                on_tick(ticks, len(ids), dt_tick)
            for sid in done:
                req = active.pop(sid)[0]
                n = len(req.engine.bodies)
                req.state = self.ctx.get_state(sid, n)
                req.finished = clock() - t0
                self.ctx.scene_destroy(sid)
                used -= req.footprint
            self.used_bytes = used
        wall = clock() - t0
        self._active = {}
        steps = sum(r.n_steps for r in requests)
        waits = [r.admitted - r.arrival for r in requests]
        return dict(requests=len(requests), scene_steps=steps, wall_s=wall, scene_steps_per_s=steps / wall,
                    ticks=ticks, p50_step_ms=1e3 * _percentile(lat, 50), p99_step_ms=1e3 * _percentile(lat, 99),
                    p50_wait_ms=1e3 * _percentile(waits, 50), p99_wait_ms=1e3 * _percentile(waits, 99),
                    budget_bytes=self.budget, peak_used_bytes=peak, max_active=max_active)

    def run_alone(self, req):
        """The request by itself from its initial state; returns its final state (r, q, v, w)."""
        sid = self._create(req)
        try:
            self.ctx.step([sid], req.n_steps, 0.0, 0)
            return self.ctx.get_state(sid, len(req.engine.bodies))
        finally:
            self.ctx.scene_destroy(sid)
