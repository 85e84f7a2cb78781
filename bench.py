#!/usr/bin/env python3
"""Throughput of the prox rigid-body time step on one MI355X.

Workload: 4096 spheres (create_sphere(0.5, 16, 16), 32^3 SDF) dropped from
[-9.5, 9.5] x [1, 40] x [-9.5, 9.5] onto a fixed ground box, mu = 0.3,
dt = 0.005, 50 coloured parallel Gauss-Seidel sweeps per step (tolerances 0),
FP32 state, FP64 merit/residual.  The reference's run is 200 steps; the
default --warmup 20 --steps 180 times steps 21..200 of that run.

metric: contact-iterations per second over the whole step (collision
detection + assembly + colouring + solve + integration), inputs resident in
HBM.  A contact-iteration is one projection of one contact's 4-row block in
one sweep, rejected sweeps included (gauss_seidel_solver.py:38-71).

roofline: the solve kernel (the largest single kernel of the step) against
HBM peak, with 288 algorithmic bytes per contact-iteration; its duration is
the mean over the timed steps of HIP events recorded on the library's stream.
traffic: HBM bytes per solve launch from rocprofv3 FETCH_SIZE/WRITE_SIZE runs
of this same command, read from the committed summary under profiles/.

cpu_baseline: the FP64 NumPy restatement of the step, single-threaded, timed
on one step of the same scene from the GPU's end state.

--copy-calibration BYTES runs only a plain device-to-device copy of BYTES
bytes (the reference kernel for scaling the profiler's byte counters).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BYTES_PER_CONTACT_ITER = 288
HBM_PEAK_GBS = 8000.0
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_solve_multi_spheres4096_pmc.txt")
WORKLOAD = "sphere_rain_4096"


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


def cpu_baseline(eng, stepper, steps_done):
    """One oracle step of the same scene from the GPU's current state."""
    from threadpoolctl import threadpool_limits
    from oracle import scene as OS, stepper as OST
    stepper._finish(eng)
    sc = OS.from_engine(eng)
    sc.step_index = steps_done
    dt = eng.params.time_step
    with threadpool_limits(1):
        t0 = time.perf_counter()
        out = OST.step(sc, dt)
        t = time.perf_counter() - t0
    lg = out["log"]
    K = int(lg["contact_points"])
    sweeps = int(lg.get("iterations", 0)) + (1 if lg.get("exitcode", 0) else 0) if K else 0
    return dict(value=K * sweeps / t, unit="contact-iters/s", cores=1, kind="port",
                sample="one full time step (collision detection, assembly, colouring, %d GS sweeps) of the "
                       "same 4096-sphere scene from the GPU state after step %d, FP64 NumPy restatement, "
                       "1 thread: K=%d contacts, %.2f s" % (sweeps, steps_done, K, t),
                seconds=t, contacts=K, sweeps=sweeps,
                natural_residual=float(lg.get("natural_residual", 0.0)))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=180)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--bodies", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES")
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

    from rainbow_amd import native, scenes
    t0 = time.perf_counter()
    bank = scenes.ShapeBank(resolution=32)
    eng = scenes.sphere_rain(args.bodies, bank)
    st = eng.stepper
    st._prepare(eng)
    ctx, sid = st.ctx, st.scene
    dt = eng.params.time_step
    log("scene built: %d bodies in %.1f s" % (len(eng.bodies), time.perf_counter() - t0))

    if args.warmup:
        ctx.step([sid], args.warmup, dt, 0)
    ctx.sync()
    ctx.scene_totals(sid, clear=True)
    log("warmup done: %d steps" % args.warmup)

    t0 = time.perf_counter()
    ctx.step([sid], args.steps, dt, native.STEP_TIMING)
    ctx.sync()
    sec = time.perf_counter() - t0
    tot = ctx.scene_totals(sid)
    tm = ctx.get_timings()
    lg = ctx.get_log(sid, eng.params.max_iterations)
    log("timed %d steps in %.3f s; totals %s" % (args.steps, sec, tot))
    assert tot["steps"] == args.steps, tot

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
                   "steps_timed": "%d..%d" % (args.warmup + 1, args.warmup + args.steps)},
        "roofline": {"kernel": "k_solve_multi", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(),
                     "bytes_per_launch": bytes_launch, "launch_ms": solve_ms},
        "work": {"contacts_per_step": tot["contacts"] / args.steps,
                 "sweeps_per_step": tot["sweeps"] / args.steps,
                 "candidates_per_step": tot["candidates"] / args.steps,
                 "last_step": {"contacts": lg["contact_points"], "colors": lg["colors"],
                               "sweeps": lg["sweeps"], "natural_residual": lg["natural_residual"]}},
        "stages_ms": {k: round(v, 4) for k, v in tm.items()},
        "narrow_phase": {"kernel": "k_frank_wolfe + reduction", "bytes_per_launch": fw_bytes,
                         "achieved_GBs": fw_bytes / (fw_ms * 1e-3) / 1e9 if fw_ms > 0 else 0.0},
    }
    if not args.no_cpu:
        log("cpu baseline: one oracle step ...")
        cb = cpu_baseline(eng, st, args.warmup + args.steps)
        result["cpu_baseline"] = cb
        log("cpu baseline: %.3g contact-iters/s (%.1f s)" % (cb["value"], cb["seconds"]))
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
