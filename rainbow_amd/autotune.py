"""Autotuner for the launch shapes of the step.

The time step of ``steppers.py:55-100`` has no launch shapes in the reference
(it is NumPy on one core).  Here a batch of scenes is stepped by kernels
whose shapes are free parameters (``rbh_tuning`` in ``include/rbh.h``): the
contact positions per colouring unit, the sweep lanes per batched scene, the
wavefronts per solve workgroup, and the lanes per workgroup of a lone scene's
cooperative grid.  None of them changes a result -- every tuning gives
bitwise the same contacts, impulses and states -- so choosing among them is
a question of time only, and the answer depends on the batch (how many
scenes, how many contacts each) and on where the scenes are in their
motion, since the contact count grows as bodies come to rest.

``autotune`` measures each candidate on the caller's own batch.  Every trial
starts from the same body states and step indices (saved before the first
trial and restored before each one), takes one untimed step (which builds
the workspace and, in graph mode, the captured step), is restored again, and
then times ``steps`` steps.  The context is left on the chosen tuning and
every scene on the state it had before the call.  The per-scene totals
(``rbh_scene_totals``) count the trial steps; clear them afterwards when
they are used for rates.
"""
import itertools
import time

from . import native

DEFAULT = dict(unit_size=0, lanes_per_scene=0, resident_contacts=0, resident_bodies=0,
               lanes_per_group=0, waves_per_group=0, color_lds_contacts=0, color_body_bits=0,
               solve_groups_per_cu=0)


def batch_candidates():
    """Shapes for a batch of several scenes (one workgroup per scene; default first)."""
    out = []
    for groups, waves, lanes, unit in itertools.product((0, 1), (4, 8), (0, 1, 2), (64, 128)):
        out.append(dict(waves_per_group=waves, lanes_per_scene=lanes, unit_size=unit, solve_groups_per_cu=groups))
    return out


def single_candidates():
    """Shapes for one scene stepped alone by the cooperative grid (default first)."""
    out = []
    for lpg, unit in itertools.product((4, 2, 1), (64, 128, 256)):
        out.append(dict(lanes_per_group=lpg, unit_size=unit))
    return out


class _Snapshot:
    """Body states and step indices of a set of scenes."""

    def __init__(self, ctx, sids, n_bodies):
        self.ctx = ctx
        self.items = []
        for sid, n in zip(sids, n_bodies):
            self.items.append((sid, ctx.get_state(sid, n), ctx.get_log(sid, 1)["step_index"]))

    def restore(self):
        for sid, st, k in self.items:
            self.ctx.set_state(sid, *st)
            self.ctx.set_step_index(sid, k)


def autotune(ctx, sids, n_bodies, dt=0.0, steps=3, candidates=None, graph=None, min_gain=0.03, log=None):
    """Time each candidate tuning on the scenes ``sids`` and keep the fastest.

    n_bodies: bodies of each scene (to save and restore its state).
    dt: time step passed to ``rbh_step`` (0 = each scene's own).
    graph: step in graph mode (default: for batches of more than one scene,
    which is how the batched form runs; a lone scene runs eagerly on its
    cooperative grid).
    min_gain: the first candidate is the baseline; another one is chosen
    only if it is faster than the baseline by more than this fraction (a few
    timed steps differ by about that much from run to run, and a scene's
    work changes as it evolves, so a smaller margin chases noise).
    Returns ``{"best": tuning, "ms_per_step": t, "trials": [(tuning, ms), ...]}``
    with the chosen tuning applied to ``ctx``.
    """
    sids = [int(s) for s in sids]
    if len(sids) != len(n_bodies):
        raise ValueError("one body count per scene")
    if not sids:
        raise ValueError("no scenes to tune on")
    if steps < 1:
        raise ValueError("steps must be positive")
    if graph is None:
        graph = len(sids) > 1
    if candidates is None:
        candidates = batch_candidates() if len(sids) > 1 else single_candidates()
    flags = native.STEP_GRAPH if graph else 0
    start = ctx.get_tuning()
    snap = _Snapshot(ctx, sids, n_bodies)
    trials = []
    try:
        for cand in candidates:
            tune = dict(DEFAULT)
            tune.update(cand)
            ctx.set_tuning(**tune)
            snap.restore()
            ctx.step(sids, 1, dt, flags)            # workspace sizing and graph capture, untimed
            ctx.sync()
            snap.restore()
            ctx.sync()
            t0 = time.perf_counter()
            ctx.step(sids, steps, dt, flags)
            ctx.sync()
            ms = 1e3 * (time.perf_counter() - t0) / steps
            trials.append((tune, ms))
            if log:
                log("autotune %s: %.3f ms/step" % (cand, ms))
    except Exception:
        ctx.set_tuning(**start)
        snap.restore()
        ctx.sync()
        raise
    best, best_ms = min(trials, key=lambda t: t[1])
    if best_ms >= trials[0][1] * (1.0 - min_gain):
        best, best_ms = trials[0]
    ctx.set_tuning(**best)
    snap.restore()
    ctx.sync()
    return {"best": best, "ms_per_step": best_ms, "trials": trials}
