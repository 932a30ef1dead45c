#!/usr/bin/env python3
"""Throughput of the fvmic2d marker step on one MI355X.

The default line is the batched ensemble: 64 independent convection instances
(grids 65^2..257^2 nodes, 16-36 markers per cell, viscosity contrast 1e2-1e6,
~47M markers) stepped together -- findMarkerCells, the density loop,
projection of T, eta, rho, k onto all four staggered node families plus Cp
(17 nodal fields), subgridTemperatureChanges and advectMarkersRK -- one launch
per phase over all instances, replayed from a captured HIP graph, inputs
resident in HBM.  ``sequential`` is the same instances stepped one
fvm_step_thermal call after another.  Beside it the line carries the
single-instance steps (``convection_tdep``: 201x201 nodes, 1.44M markers;
``vep_subduction``: 801x401 tanh-stretched, 8M markers, the full
visco-elasto-plastic step) and a small learned Stokes initial guess run
(``learned_guess``: host solver iterations to 1e-8 from the plain and the
learned guess).

Prints ONE JSON line.  ``value`` = active markers x steps / seconds, timed with
HIP events on the library's stream between two synchronisations.  ``roofline``
describes the kernel with the largest share of the step time, measured in a
separate pass with per-launch HIP events; ``cpu_baseline`` times the CPU
restatement of the same step (the oracle) on this host.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "markers/s per MIC step (proj+advect+subgrid); KSP iters to 1e-8 vs plain guess"
HBM_PEAK_GBS = 8000.0

PHASES = ["locate", "sort", "partials", "nodes", "subgrid_t1", "subgrid_t2", "advect", "subgrid_s1",
          "marker_update", "other", "persistent"]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(w):
    """Compulsory state traffic per launch of each phase of one convection_tdep step.

    Per active marker (FP64 = 8 B, int32 = 4 B) and per node.  Sort keys,
    permutations and partial sums are implementation traffic and not counted."""
    n, nodes = w.n_markers, w.NX * w.NY
    return {
        # findMarkerCells + density loop: read X, Y, T, Mat; write cellX, cellY, cpu, rho, rhodot
        "locate": n * (16 + 8 + 4 + 12 + 16),
        # subgrid T pass 1: read X, Y, T, cellX, cellY; write dTsub; gather lastT, rho, Cp, k
        "subgrid_t1": n * (24 + 8 + 8) + nodes * 32,
        # subgrid T pass 2 + RK4: read X, Y, cellX, cellY, T, dTsub, s(3), e(3), Eii;
        # write T, Tdot, X, Y, VX, VY, wxy, s(3), Eii; gather deltaT, vx, vy, wxy
        "advect": n * (96 + 88) + nodes * 32,
    }


STEP_BYTES_PER_MARKER = 268  # locate + projections 60, subgrid T pass 1 32, pass 2 + RK4 176
STEP_BYTES_PER_NODE = 224


# ---- host CPU and the CPU baselines (the oracle, forked before the GPU is touched) ----
def host_cpu():
    """The host CPU model (/proc/cpuinfo) and the size of this process's affinity set."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = os.cpu_count() or 1
    return model, aff


def cpu_workers():
    """Worker processes for the all-core baselines: the affinity set, or the
    thread share the environment grants this job (OMP_NUM_THREADS) if smaller."""
    _, aff = host_cpu()
    n = aff
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n), aff


def _tdep_steps(budget_s):
    """Oracle steps of convection_tdep for budget_s seconds: (markers stepped, s)."""
    import oracle
    from fvmic2d_amd import workloads
    w = workloads.convection_tdep()
    st = oracle.State(w, extra=[p[4] for p in w.projections])
    steps, t0 = 0, time.perf_counter()
    while True:
        st.step_thermal()
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            return w.n_markers * steps, el


def _ensemble_steps(args):
    """Oracle steps of ensemble instances first, first + stride, ... (one step
    each, generation not timed) until budget_s of stepping: (markers stepped, s)."""
    first, stride, budget_s = args
    import oracle
    from fvmic2d_amd import workloads
    markers, el, i = 0, 0.0, first
    while el < budget_s:
        w = workloads.ensemble(1, seed=4, nmin=65, nmax=257, first=i % 64)[0]
        st = oracle.State(w, extra=[p[4] for p in w.projections])
        t0 = time.perf_counter()
        st.step_thermal()
        el += time.perf_counter() - t0
        markers += w.n_markers
        i += stride
    return markers, el


def cpu_baseline_ensemble(budget_s):
    m, el = _ensemble_steps((0, 1, budget_s))
    return {"value": m / el, "unit": "markers/s", "cores": 1, "kind": "port",
            "sample": "one oracle thermal step of ensemble instances 0, 1, 2, ... in turn, %d markers "
                      "stepped in %.1f s (C restatement, gcc -O3 -ffp-contract=off, one thread)" % (m, el),
            "cpu": host_cpu()[0]}


def cpu_baseline_tdep(budget_s):
    m, el = _tdep_steps(budget_s)
    return {"value": m / el, "unit": "markers/s", "cores": 1, "kind": "port",
            "sample": "%d full convection_tdep steps (1.44M markers each) of the C restatement, "
                      "gcc -O3 -ffp-contract=off, one thread, %.1f s" % (m // 1440000, el),
            "cpu": host_cpu()[0]}


def cpu_all_cores(kind, budget_s):
    """The oracle on every core of this job at once, one independent problem per
    process (the reference's ranks each step their own share; leaving out their
    communication bounds the reference's multi-rank rate from above)."""
    import multiprocessing
    n, aff = cpu_workers()
    with multiprocessing.get_context("fork").Pool(n) as pool:
        if kind == "ensemble":
            res = pool.map(_ensemble_steps, [(k, n, budget_s) for k in range(n)])
            sample = ("%d processes, worker k stepping ensemble instances k, k+%d, ... with the C restatement "
                      "for %.0f s" % (n, n, budget_s))
        else:
            res = pool.map(_tdep_steps, [budget_s] * n)
            sample = ("%d processes, each stepping its own convection_tdep copy (1.44M markers) with the C "
                      "restatement for %.0f s" % (n, budget_s))
    return {"value": sum(m / t for m, t in res), "unit": "markers/s", "cores": n, "affinity_cores": aff,
            "kind": "port", "sample": sample, "cpu": host_cpu()[0]}


def traffic_per_launch(kernel):
    """HBM bytes per launch of `kernel` from its committed counter summary, if any:
    profiles/traffic_<kernel>.md holds 'bytes_per_launch: N' (FETCH_SIZE +
    WRITE_SIZE of a rocprofv3 --pmc run, scaled by a copy of known size) and a
    'measured:' line.  This run measures no counters; the figure carries its source."""
    tf = os.path.join(ROOT, "profiles", "traffic_%s.md" % kernel)
    if not os.path.exists(tf):
        return None
    b, when = None, "?"
    for line in open(tf):
        t = line.strip()
        if t.startswith("bytes_per_launch:"):
            b = float(t.split(":", 1)[1])
        elif t.startswith("measured:"):
            when = t.split(":", 1)[1].strip()
    if b is None:
        return None
    return {"bytes": b, "source": "profiles/traffic_%s.md (counter run, %s; not measured in this run)" % (
        kernel, when)}


def roofline(kernel_label, traffic_key, ab, ms):
    achieved = ab / (ms / 1e3) / 1e9
    tr = traffic_per_launch(traffic_key)
    return {"bound": "hbm", "kernel": kernel_label, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": tr and tr["bytes"], "traffic_source": tr and tr["source"],
            "algorithmic_bytes_per_launch": ab, "avg_launch_ms": ms}


def phase_profile(ctx, step, nsteps):
    """Per-phase ms per step and launches per step, HIP events around every launch."""
    import ctypes as C
    msp = (C.c_double * len(PHASES))()
    lau = (C.c_int32 * len(PHASES))()
    ctx.call("fvm_profile_enable", C.c_int32(1))
    for _ in range(nsteps):
        step()
    ctx.call("fvm_profile_read", msp, lau)
    ctx.call("fvm_profile_enable", C.c_int32(0))
    ms = {PHASES[i]: msp[i] / nsteps for i in range(len(PHASES)) if lau[i]}
    la = {PHASES[i]: lau[i] // nsteps for i in range(len(PHASES)) if lau[i]}
    return ms, la


STORE_ORDER = "cell order: fvm_markers_reorder once before the timed steps (a driver reorders every few tens of steps)"


def reorder_stores(paths):
    """Put every path's marker store in cell order (fvm_markers_reorder after
    locating the cells).  Returns the GPU time of the reorders in ms."""
    ctx = paths[0].ctx
    total = 0.0
    for p in paths:
        p.find_marker_cells()
        ctx.timer_start()
        p.call("fvm_markers_reorder", p.grid.ref, p.markers.ref, None)
        total += ctx.timer_stop()
    return total


def time_steps(ctx, step, warmup, steps):
    for _ in range(warmup):
        step()
    ctx.sync()
    ctx.timer_start()
    for _ in range(steps):
        step()
    return ctx.timer_stop() / steps


def initial_order(a, ctx, step, n0, steps, paths):
    """The step on the workloads' own marker order (skipped with --no-extras),
    then the stores put in cell order for the timed steps."""
    init = None if a.no_extras else time_steps(ctx, step, a.warmup, steps)
    out = {"reorder_ms": reorder_stores(paths)}
    if init is not None:
        out.update(ms_per_step=init, markers_per_s=n0 / (init / 1e3))
    return out


def run_ensemble(a):
    """The batched ensemble: 64 convection_tdep-style instances, grids 65^2..257^2
    (--ens-nmin/--ens-nmax), 16-36 markers per cell, viscosity contrast 1e2-1e6,
    against the same instances stepped one fvm_step_thermal call after another."""
    from fvmic2d_amd import workloads
    from fvmic2d_amd.host import Ensemble, MarkerPath
    from fvmic2d_amd.lib import Context

    t0 = time.perf_counter()
    ctx = Context()
    paths, shapes = [], []
    for i in range(0, a.instances, 8):
        for w in workloads.ensemble(min(8, a.instances - i), seed=4, nmin=a.ens_nmin, nmax=a.ens_nmax, first=i):
            paths.append(MarkerPath(w, ctx=ctx))
            shapes.append((w.NX, w.NY))
            w.markers = None  # host copy no longer needed
        log("ensemble: %d instances generated and uploaded (%.0f s)" % (len(paths), time.perf_counter() - t0))
    ctx.sync()
    n0 = sum(p.markers.n for p in paths)
    nodes = sum(nx * ny for nx, ny in shapes)

    # the batched step on the workloads' own marker order, then the stores in
    # cell order for everything below
    ens = Ensemble(paths)
    if not a.no_extras:  # the initial order at its own autotuned shape
        ens.autotune(reps=3)
    b0, l0, p0 = ens.get_tuning()
    init = initial_order(a, ctx, ens.launch, n0, max(3, a.steps // 2), paths)
    init["reorder_ms_per_step_every_50_steps"] = init["reorder_ms"] / 50
    init["tuning"] = {"block": b0, "lanes": l0, "persistent": p0}
    log("ensemble: initial order %s" % init)

    def seq_step():
        for p in paths:
            p.step_thermal()
    for _ in range(a.warmup):
        seq_step()
    ctx.sync()
    ctx.timer_start()
    for _ in range(a.steps):
        seq_step()
    seq_ms = ctx.timer_stop() / a.steps
    log("ensemble: sequential %.3f ms/step" % seq_ms)

    (block, lanes, persistent), tuned_ms = ens.autotune(reps=3)
    log("ensemble: autotuned block=%d lanes=%d persistent=%d (%.3f ms)" % (block, lanes, persistent, tuned_ms))

    def timed(nsteps):
        for _ in range(a.warmup):
            ens.launch()
        ctx.sync()
        ctx.timer_start()
        for _ in range(nsteps):
            ens.launch()
        return ctx.timer_stop() / nsteps

    ens_ms = timed(a.steps)
    log("ensemble: batched %.3f ms/step" % ens_ms)
    nout, nbad = ens.counters()
    active = n0 - nout
    # both launch modes at the tuned lane count, for the record; the per-phase
    # mode last, for the phase profile
    modes = {}
    ens.set_tuning(256, max(lanes, 8), 1)  # the persistent step has no 4-lane form
    ens.capture()
    modes["persistent_ms_per_step"] = timed(max(3, a.steps // 4))
    ens.set_tuning(256 if persistent else block, lanes, 0)
    ens.capture()
    modes["per_phase_ms_per_step"] = timed(max(3, a.steps // 4))
    ens.counters()  # raises if a persistent step stalled
    log("ensemble: persistent %.3f ms, per-phase %.3f ms" % (modes["persistent_ms_per_step"],
                                                              modes["per_phase_ms_per_step"]))
    phase_ms, phase_la = phase_profile(ctx, ens.step, 3)
    step_bytes = n0 * STEP_BYTES_PER_MARKER + nodes * STEP_BYTES_PER_NODE
    ab = n0 * (96 + 88) + nodes * 32
    adv_ms = phase_ms["advect"] / phase_la["advect"]
    value = active / (ens_ms / 1e3)
    out = {
        "metric": METRIC, "value": value, "unit": "markers/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ens_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "ensemble%d" % len(paths),
                   "grids": "%dx%d..%dx%d" % (min(s[0] for s in shapes), min(s[0] for s in shapes),
                                              max(s[0] for s in shapes), max(s[0] for s in shapes)),
                   "markers": n0, "markers_per_cell": "16-36", "viscosity_contrast": "1e2-1e6",
                   "tuning": {"block": block, "lanes": lanes, "persistent": persistent}, "graph": True,
                   "store": STORE_ORDER},
        "initial_order": init,
        "sequential": {"ms_per_step": seq_ms, "markers_per_s": active / (seq_ms / 1e3)},
        "speedup_vs_sequential": seq_ms / ens_ms,
        "launch_modes": modes,
        "roofline": roofline("advect (kb_advect: subgrid T pass 2, RK4, stress rotation)", "kb_advect", ab,
                             adv_ms),
        "step": {"algorithmic_bytes": step_bytes, "achieved_GBs": step_bytes / (ens_ms / 1e3) / 1e9,
                 "phase_ms_per_step": phase_ms, "phase_launches_per_step": phase_la, "bad_cells": nbad},
    }
    ens.close()
    del ens, paths
    return out


def run_thermal(a):
    """convection_tdep (201x201, 1.44M markers): the single-instance step, its
    per-phase times, the dominant kernel's roofline and the PCIe-inclusive rate."""
    import ctypes as C
    from fvmic2d_amd import workloads
    from fvmic2d_amd.host import MarkerPath

    w = workloads.convection_tdep()
    mp = MarkerPath(w)
    if a.lanes != 8:
        mp.ctx.call("fvm_ctx_set_lanes", C.c_int32(a.lanes))
    n0 = w.n_markers
    for _ in range(a.warmup):
        mp.step_thermal()
    init = initial_order(a, mp.ctx, mp.step_thermal, n0, max(5, a.steps // 2), [mp])
    for _ in range(a.warmup):
        mp.step_thermal()
    mp.ctx.sync()
    t0 = time.perf_counter()
    mp.ctx.timer_start()
    for _ in range(a.steps):
        mp.step_thermal()
    ms = mp.ctx.timer_stop()
    wall = time.perf_counter() - t0
    nout = mp.find_marker_cells()
    active = n0 - nout
    value = active * a.steps / (ms / 1e3)
    log("convection_tdep: %.3f ms/step" % (ms / a.steps))

    # end-to-end: every step also takes the host solves' outputs in (vx, vy, wxy,
    # thisT host-to-device) and hands the 17 projected nodal fields back
    # (device-to-host), as a driver with host-side PETSc solves would
    in_names = ("vx", "vy", "wxy", "thisT")
    outs = [p[4] for p in w.projections]
    nodes = w.NX * w.NY
    h_in = mp.ctx.pinned((len(in_names), nodes))
    h_out = mp.ctx.pinned((len(outs), nodes))
    for k, n in enumerate(in_names):
        h_in[k] = mp.get_nodal(n).ravel()
    d_in = (C.c_void_p * len(in_names))(*[mp.nodal.ptr(n) for n in in_names])
    d_out = (C.c_void_p * len(outs))(*[mp.nodal.ptr(n) for n in outs])
    e2e_steps = min(a.steps, 20)

    def e2e_loop(packed):
        mp.ctx.sync()
        t0 = time.perf_counter()
        for _ in range(e2e_steps):
            if packed:
                mp.ctx.call("fvm_nodal_upload", C.c_int64(nodes), len(in_names), C.cast(d_in, C.c_void_p),
                            h_in.ctypes.data_as(C.c_void_p))
            else:
                for k, n in enumerate(in_names):
                    mp.set_nodal(n, h_in[k].reshape(w.NY, w.NX))
            mp.step_thermal()
            if packed:
                mp.ctx.call("fvm_nodal_download", C.c_int64(nodes), len(outs), C.cast(d_out, C.c_void_p),
                            h_out.ctypes.data_as(C.c_void_p))
            else:
                for n in outs:
                    mp.get_nodal(n)
        return (time.perf_counter() - t0) / e2e_steps * 1e3
    e2e_ms = e2e_loop(True)
    e2e_each_ms = e2e_loop(False)

    phase_ms, phase_launches = phase_profile(mp.ctx, mp.step_thermal, 5)
    ab = algorithmic_bytes(w)
    cands = {k: v for k, v in phase_ms.items() if ab.get(k)}
    dom = max(cands, key=cands.get)
    per_launch_ms = phase_ms[dom] / phase_launches[dom]
    step_bytes = n0 * STEP_BYTES_PER_MARKER + nodes * STEP_BYTES_PER_NODE
    return {
        "value": value, "unit": "markers/s", "steps": a.steps, "ms_per_step": ms / a.steps,
        "config": {"workload": "convection_tdep", "grid": "201x201 uniform", "markers": n0,
                   "markers_per_cell": 36, "active_markers_after": active, "projections": 17,
                   "viscosity_contrast": 1e5, "store": STORE_ORDER},
        "initial_order": init,
        "roofline": roofline(dom, dom, ab[dom], per_launch_ms),
        "step": {"algorithmic_bytes": step_bytes, "achieved_GBs": step_bytes / (ms / a.steps / 1e3) / 1e9,
                 "wall_ms_per_step": wall / a.steps * 1e3,
                 "phase_ms_per_step": phase_ms, "phase_launches_per_step": phase_launches},
        "end_to_end": {"ms_per_step": e2e_ms, "markers_per_s": n0 * 1e3 / e2e_ms, "steps": e2e_steps,
                       "bytes_h2d_per_step": len(in_names) * nodes * 8, "bytes_d2h_per_step": len(outs) * nodes * 8,
                       "transfer": "pinned, one packed copy each way (fvm_nodal_upload / fvm_nodal_download)",
                       "ms_per_step_one_copy_per_array": e2e_each_ms},
    }


def run_store_order(a, steps=10):
    """convection_tdep from a randomly permuted marker store (the spatial order a
    long run drifts towards): the step before and after one fvm_markers_reorder,
    and the reorder's own time (cells current; its bucket sort included)."""
    from fvmic2d_amd import workloads
    from fvmic2d_amd.host import MarkerPath
    w = workloads.convection_tdep()
    w.markers = w.markers[np.random.default_rng(7).permutation(w.n_markers)].copy()
    mp = MarkerPath(w)

    def timed():
        for _ in range(2):
            mp.step_thermal()
        mp.ctx.sync()
        mp.ctx.timer_start()
        for _ in range(steps):
            mp.step_thermal()
        return mp.ctx.timer_stop() / steps
    scrambled = timed()
    reorder = []
    for _ in range(3):
        mp.find_marker_cells()
        mp.ctx.timer_start()
        mp.call("fvm_markers_reorder", mp.grid.ref, mp.markers.ref, None)
        reorder.append(mp.ctx.timer_stop())
    ordered = timed()
    log("store order: scrambled %.3f ms/step, reordered %.3f ms/step, reorder %.3f ms" % (
        scrambled, ordered, reorder[0]))
    return {"workload": "convection_tdep, markers randomly permuted", "steps": steps,
            "scrambled_ms_per_step": scrambled, "reordered_ms_per_step": ordered,
            "reorder_ms": reorder[0], "reorder_ms_ordered_store": min(reorder[1:])}


def copy_calibration(nbytes=512 << 20, reps=5):
    """A device-to-device copy of known size, for scaling the FETCH_SIZE and
    WRITE_SIZE counters (profile this run and divide nbytes by the counter)."""
    import ctypes as C
    from fvmic2d_amd.lib import Context
    ctx = Context()
    a, b = ctx.alloc(nbytes), ctx.alloc(nbytes)
    ctx.call("fvm_memset", C.c_void_p(a), 1, nbytes)
    ctx.sync()
    ctx.timer_start()
    for _ in range(reps):
        ctx.call("fvm_memcpy_dtod", C.c_void_p(b), C.c_void_p(a), nbytes)
    ms = ctx.timer_stop() / reps
    print(json.dumps({"copy_bytes": nbytes, "reps": reps, "ms_per_copy": ms,
                      "GBs_read_plus_write": 2 * nbytes / (ms / 1e3) / 1e9}))
    ctx.free(a)
    ctx.free(b)
    return 0


def run_learned_guess(a, N, ntrain, ntest, epochs):
    """The learned Stokes initial guess: plain-method data, U-Net training on the
    GPU kernels, and the host Stokes solve's FGMRES iterations to 1e-8 from the plain,
    previous-solution and learned guesses on held-out states."""
    from fvmic2d_amd import learned_guess as lg
    from fvmic2d_amd.lib import Context
    from fvmic2d_amd.unet import Buf
    workers = a.lg_workers or cpu_workers()[0]
    t0 = time.perf_counter()
    # spawned workers: in the default run this process has used the GPU already
    prog = lambda d, t: log("learned guess: %d of %d host solves done (%.0f s)" % (d, t, time.perf_counter() - t0))
    raw_tr = lg.generate(N, list(range(ntrain)), workers=workers, start="spawn", progress=prog)
    raw_te = lg.generate(N, list(range(10 ** 6, 10 ** 6 + ntest)), workers=workers, start="spawn", progress=prog)
    t_gen = time.perf_counter() - t0
    log("learned guess: %d + %d states at %dx%d solved on the host in %.0f s" % (ntrain, ntest, N, N, t_gen))
    ctx = Context()
    t0 = time.perf_counter()
    tr, te = lg.Dataset(ctx, N, raw_tr), lg.Dataset(ctx, N, raw_te)
    t_pg = time.perf_counter() - t0
    X, T = tr.arrays()
    model = lg.LearnedGuess(ctx, levels=4, base=32, seed=0)
    hist = model.train(X, T, epochs=epochs, batch=8, lr=2e-3, log=lambda m: log("learned guess: " + m))
    steps = epochs * max(1, len(X) // 8)
    dX = Buf(ctx, (8, N - 1, N - 1, model.net.cin), np.uint16)
    model.net.forward(dX)
    ctx.sync()
    ctx.timer_start()
    for _ in range(10):
        model.net.forward(dX)
    fwd_ms = ctx.timer_stop() / 10
    t0 = time.perf_counter()
    rows = lg.evaluate(te, model, progress=lambda k, t: log("learned guess: %d of %d held-out states solved" % (k, t)))
    t_eval = time.perf_counter() - t0
    summ = lg.summarize(rows)
    flops_fwd = model_flops(model.net, N - 1) * 8
    return {
        "metric": "KSP iterations to 1e-8 relative residual: learned vs plain initial guess",
        "value": summ["mean_iterations"]["learned"], "unit": "iterations", "higher_is_better": False,
        "dtype": "bf16 (fp32 accumulate)", "steps": steps,
        "config": {"workload": "learned_guess_box%d" % N, "grid": "%dx%d" % (N, N),
                   "viscosity_contrast": lg.CONTRAST, "train_states": ntrain, "test_states": ntest,
                   "unet": "4 levels, 32 base channels, 3x3 convs", "epochs": epochs,
                   "stokes_solve": "FGMRES(100) + u/p field split (LU velocity block, diagonal Schur), rtol 1e-8"},
        "learned_guess": summ,
        "train_loss_first_last": [hist[0], hist[-1]],
        "seconds": {"host_solves": t_gen, "gpu_pressure_guesses": t_pg, "training": model.train_seconds,
                    "evaluation": t_eval},
        "unet_forward_batch8_ms": fwd_ms,
        "roofline": {"bound": "mfma", "kernel": "unet forward (batch 8)",
                     "achieved": flops_fwd / (fwd_ms / 1e3) / 1e12, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": flops_fwd / (fwd_ms / 1e3) / 1e12 / 2500.0, "traffic": None},
    }


def bench_pressure_guess(a):
    """The plain initial pressure guess (initialPressureGuess): GPU BiCGStab to
    rtol 1e-10 on seeded blob densities, against the host direct solve of the
    same assembled system (the oracle's assembly, SuperLU) on the same states."""
    import scipy.sparse.linalg as sla
    from fvmic2d_amd import learned_guess as lg
    from fvmic2d_amd.host import initial_pressure_guess
    from fvmic2d_amd.lib import Context, DeviceArray, DeviceGrid
    ctx = Context()
    rows = []
    for N in a.pg_grids:
        g = lg.BoxGrid(N)
        rho = lg.fields(lg.blob_field(N, 11))[3]
        dg = DeviceGrid(ctx, g)
        d_rho = DeviceArray(ctx, rho.ravel())
        d_p = DeviceArray(ctx, np.zeros(N * N))
        initial_pressure_guess(ctx, dg, d_rho, lg.G, rtol=1e-10, maxit=200000, p_out=d_p)  # warm-up
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            its, rel = initial_pressure_guess(ctx, dg, d_rho, lg.G, rtol=1e-10, maxit=200000, p_out=d_p)
        gpu_ms = (time.perf_counter() - t0) / reps * 1e3
        row = {"grid": "%dx%d" % (N, N), "iterations": its, "relres": rel, "gpu_ms": gpu_ms}
        if not a.no_cpu:
            import oracle.stokes as OS  # the host restatement: assembly of the same system
            A, b = OS.pressure_guess_system(g, rho, lg.G)
            Ac = A.tocsc()
            t0 = time.perf_counter()
            ref = sla.spsolve(Ac, b)
            row["cpu_direct_ms"] = (time.perf_counter() - t0) * 1e3
            got = d_p.get()
            row["max_rel_diff_vs_direct"] = float(np.abs(got - ref).max() / np.abs(ref).max())
        rows.append(row)
        print("pressure guess %s: %d iterations, %.2f ms on the GPU%s" % (
            row["grid"], its, gpu_ms, ", %.1f ms host direct" % row["cpu_direct_ms"] if "cpu_direct_ms" in row
            else ""), file=sys.stderr, flush=True)
    main_row = next((r for r in rows if r["grid"] == "257x257"), rows[-1])
    out = {"metric": "plain initial pressure guess (initialPressureGuess) solve time", "value": main_row["gpu_ms"],
           "unit": "ms", "n_gpus": 1, "higher_is_better": False, "vs_baseline": None, "dtype": "f64",
           "data": "synthetic", "scaling": "weak", "steps": 5, "warmup": 1,
           "config": {"workload": "pressure_guess_box", "grid": main_row["grid"],
                      "solver": "Jacobi BiCGStab, rtol 1e-10, five kernels per iteration"},
           "grids": rows,
           # one iteration streams 37 FP64 values per node through its five kernels
           # (p: r, p, v, Dinv -> p, y; v: y, 5 coefficients, rhat -> v; s: r, v, y, x,
           # Dinv -> x, s, z; t: z, 5 coefficients, s -> t; x: x, z, s, t, rhat -> x, r)
           "roofline": {"bound": "hbm", "kernel": "BiCGStab iteration (5 kernels)",
                        "achieved": 37 * 8 * int(main_row["grid"].split("x")[0]) ** 2 * main_row["iterations"]
                        / (main_row["gpu_ms"] * 1e-3) / 1e9,
                        "peak": 8000.0, "unit": "GB/s",
                        "frac": 37 * 8 * int(main_row["grid"].split("x")[0]) ** 2 * main_row["iterations"]
                        / (main_row["gpu_ms"] * 1e-3) / 1e9 / 8000.0, "traffic": None},
           "cpu_baseline": None if a.no_cpu else {
               "value": main_row["cpu_direct_ms"], "unit": "ms", "cores": 1, "kind": "port",
               "sample": "the same assembled system solved by scipy sparse LU (SuperLU), as the reference "
                         "solves it with a direct LU"}}
    print(json.dumps(out))
    return 0


VEP_STEP_BYTES_PER_MARKER = 640  # P1 76 + A11 132 + A10/A12 80 + A8 80 + A9 64 + P2 32 + P3 176
VEP_STEP_BYTES_PER_NODE = 376


def vep_algorithmic_bytes(w):
    """Compulsory state traffic per launch of the single-kernel phases of one
    vep_subduction step (FP64 = 8 B, int32 = 4 B); phases made of several kernels
    are left out."""
    n, nodes = w.n_markers, w.NX * w.NY
    return {
        # findMarkerCells + density loop: read X, Y, T, Mat; write cellX, cellY, cpu, rho, rhodot
        "locate": n * (16 + 8 + 4 + 12 + 16),
        # subgrid T pass 2: read X, Y, cellX, cellY, T, dTsub; write T, Tdot; gather deltaT
        "subgrid_t2": n * (16 + 8 + 16 + 16) + nodes * 8,
        # RK4: read X, Y, cellX, cellY, s(3), e(3), Eii; write X, Y, VX, VY, wxy, s(3), Eii;
        # gather vx, vy, wxy
        "advect": n * (80 + 72) + nodes * 24,
    }


def cpu_baseline_vep(w, budget_s):
    """The oracle's visco-elasto-plastic step on the same workload, one thread,
    whole steps until budget_s (at least one)."""
    import oracle
    st = oracle.State(w, extra=oracle.VEP_PROJ)
    steps, t0 = 0, time.perf_counter()
    while True:
        st.step_vep()
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    return {"value": w.n_markers * steps / el, "unit": "markers/s", "cores": 1, "kind": "port",
            "sample": "%d full vep_subduction steps (%d markers each) of the C restatement, "
                      "gcc -O3 -ffp-contract=off, one thread, %.1f s" % (steps, w.n_markers, el),
            "cpu": host_cpu()[0]}


def run_vep(a, extras=True):
    """The visco-elasto-plastic marker step at full size: 801 x 401 tanh-stretched
    grid, 25 markers per cell (8M markers), five materials with Drucker-Prager
    yielding -- locate + density, yielding, viscosity limits, projections,
    strain/pressure update and residual, subgrid stress, stress update, subgrid
    temperature and RK4 advection.  The two scalar results (yield flag,
    residual) are read back by the host every step, as the reference does."""
    from fvmic2d_amd import workloads
    from fvmic2d_amd.host import MarkerPath
    t0 = time.perf_counter()
    w = workloads.vep_subduction()
    log("vep: workload generated in %.0f s" % (time.perf_counter() - t0))
    mp = MarkerPath(w)
    n0 = w.n_markers
    for _ in range(a.warmup):
        mp.step_vep()
    init = initial_order(a, mp.ctx, mp.step_vep, n0, max(3, a.steps // 2), [mp])
    for _ in range(a.warmup):
        mp.step_vep()
    mp.ctx.sync()
    t0 = time.perf_counter()
    mp.ctx.timer_start()
    for _ in range(a.steps):
        y, r = mp.step_vep()
    ms = mp.ctx.timer_stop()
    wall = time.perf_counter() - t0
    nout = mp.find_marker_cells()
    active = n0 - nout
    value = active * a.steps / (ms / 1e3)
    log("vep: %.3f ms per step" % (ms / a.steps))
    phase_ms, phase_launches = phase_profile(mp.ctx, mp.step_vep, 3)
    ab = vep_algorithmic_bytes(w)
    cands = {k: v / phase_launches[k] for k, v in phase_ms.items() if ab.get(k)}
    dom = max(cands, key=cands.get)
    step_bytes = n0 * VEP_STEP_BYTES_PER_MARKER + w.NX * w.NY * VEP_STEP_BYTES_PER_NODE
    out = {
        "value": value, "unit": "markers/s", "steps": a.steps, "ms_per_step": ms / a.steps,
        "config": {"workload": "vep_subduction", "grid": "801x401 tanh-stretched", "markers": n0,
                   "markers_per_cell": 25, "active_markers_after": active, "materials": 5,
                   "plasticity": "Drucker-Prager", "yielding_last_step": y, "residual_last_step": r,
                   "store": STORE_ORDER},
        "initial_order": init,
        "roofline": roofline(dom, dom + "_vep", ab[dom], cands[dom]),
        "step": {"algorithmic_bytes": step_bytes, "achieved_GBs": step_bytes / (ms / a.steps / 1e3) / 1e9,
                 "wall_ms_per_step": wall / a.steps * 1e3,
                 "phase_ms_per_step": phase_ms, "phase_launches_per_step": phase_launches},
    }
    if extras:
        out["other_entry_points"] = bench_other_entry_points(mp, w)
        out["nodal_step"] = bench_nodal_step(a)
        out["thermal_solve"] = bench_thermal_solve(a)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline_vep(w, a.vep_cpu_budget)
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    return out


def bench_other_entry_points(mp, w, reps=10):
    """The entry points the step does not call, timed on the same 8M-marker state
    (HIP events on the library's stream, ms per call): projectMarkersNodesFromScalar
    (the marker eta column onto the basic nodes), projectNodalFieldToMarkersS (the
    nodal lastT onto marker T) and the step glue's max T / max |sxx| (each reads its
    result back)."""
    res = {}
    scalar = mp.markers.col("eta")

    def timed(name, fn):
        fn()
        mp.ctx.sync()
        mp.ctx.timer_start()
        for _ in range(reps):
            fn()
        res[name] = mp.ctx.timer_stop() / reps
    timed("projectMarkersNodesFromScalar_ms", lambda: mp.project_from_scalar(scalar, "etaS"))
    timed("projectNodalFieldToMarkersS_ms", lambda: mp.project_nodal_to_markers("lastT", "T"))
    timed("maxMarkerT_ms", mp.max_marker_T)
    timed("maxAbsSxx_ms", mp.max_abs_sxx)
    n = w.n_markers
    # algorithmic bytes: scalar projection reads X, Y, the scalar column and the
    # cells (28 B per marker); nodal-to-markers reads X, Y, cells, writes T (32 B);
    # the max passes read cellX and one column (12 B)
    res["achieved_GBs"] = {"projectMarkersNodesFromScalar": n * 28 / (res["projectMarkersNodesFromScalar_ms"] * 1e6),
                           "projectNodalFieldToMarkersS": n * 32 / (res["projectNodalFieldToMarkersS_ms"] * 1e6),
                           "maxMarkerT": n * 12 / (res["maxMarkerT_ms"] * 1e6)}
    return res


def bench_thermal_solve(a, reps=5):
    """formThermalSystem + the energy solve (rtol 1e-12) on the 801 x 401
    tanh-stretched grid with the subduction boundary set-up, against the host's
    sparse LU (SuperLU) of the same assembled system, as the reference solves it
    with a direct LU."""
    from fvmic2d_amd import workloads
    from fvmic2d_amd.host import MarkerPath
    from fvmic2d_amd.lib import DeviceArray
    w = workloads.thermal_box(801, 401)
    mp = MarkerPath(w)
    nn = w.NX * w.NY
    coef = DeviceArray(mp.ctx, np.zeros(5 * nn))
    rhs = DeviceArray(mp.ctx, w.nodal["heating"])
    T = DeviceArray(mp.ctx, np.zeros(nn))

    def once():
        rhs.set(w.nodal["heating"])
        mp.form_thermal_system(w.dt, False, coef, rhs)
        return mp.solve_thermal_system(coef, rhs, T, rtol=1e-12, maxit=20000)
    once()
    mp.ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        its, rel = once()
    ms = (time.perf_counter() - t0) / reps * 1e3
    res = {"grid": "801x401 tanh-stretched", "nodes": nn, "gpu_ms": ms, "iterations": its, "relres": rel}
    if not a.no_cpu:
        import scipy.sparse.linalg as sla
        from oracle import thermal as TH
        A, b = TH.thermal_system(w.x, w.y, w.nodal, w.materials.materialRho[0], w.options, w.dt, False,
                                 w.nodal["heating"])
        Ac = A.tocsc()
        t0 = time.perf_counter()
        ref = sla.spsolve(Ac, b)
        res["cpu_direct_ms"] = (time.perf_counter() - t0) * 1e3
        res["max_rel_diff_vs_direct"] = float(np.abs(T.get().ravel() - ref).max() / np.abs(ref).max())
    print("thermal solve: %.3f ms on the GPU (%d iterations)" % (ms, its), file=sys.stderr, flush=True)
    return res


# algorithmic bytes per node of the nodal step: retrieveSolutions (read S 24,
# write S.p 8, vx, vy, p 24, shift p 16), nodalStressStrain (read vx, vy, etaN,
# muN, etaS, muS, s(3) 72; write e(3), wxy, s(3), ds(3) 80; then read s(3),
# etaN, etaS, e(3) 64, write heating, eii, sii 24), limitDisplacementTimestep
# (read vx, vy, rho, rhodot 32)
NODAL_STEP_BYTES_PER_NODE = 72 + 152 + 88 + 32


def bench_nodal_step(a, reps=20):
    """retrieveSolutions + nodalStressStrain + limitDisplacementTimestep on the
    801 x 401 tanh-stretched grid, in the reference's call order; the time step
    is read back every repetition as the reference does.  Kcont = 1 so that the
    repeated in-place scaling of the solution vector leaves it unchanged."""
    from fvmic2d_amd import workloads
    from fvmic2d_amd.host import MarkerPath
    from fvmic2d_amd.lib import DeviceArray
    w = workloads.nodal_box(801, 401)
    mp = MarkerPath(w)
    rng = np.random.default_rng(0)
    S = DeviceArray(mp.ctx, np.stack([w.nodal["vx"], w.nodal["vy"], rng.normal(size=(w.NY, w.NX))], -1))
    heat = DeviceArray(mp.ctx, np.zeros((w.NY, w.NX)))

    def step():
        mp.retrieve_solutions(S, 1.0)
        mp.nodal_stress_strain(w.dt, heat)
        return mp.limit_displacement_timestep(w.dt)
    step()
    mp.ctx.sync()
    mp.ctx.timer_start()
    for _ in range(reps):
        step()
    ms = mp.ctx.timer_stop() / reps
    nodes = w.NX * w.NY
    res = {"grid": "801x401 tanh-stretched", "nodes": nodes, "ms_per_call_sequence": ms,
           "algorithmic_bytes": nodes * NODAL_STEP_BYTES_PER_NODE,
           "achieved_GBs": nodes * NODAL_STEP_BYTES_PER_NODE / (ms / 1e3) / 1e9}
    if not a.no_cpu:
        import oracle
        st = oracle.State(w)
        Sh = np.stack([w.nodal["vx"], w.nodal["vy"], rng.normal(size=(w.NY, w.NX))], -1).ravel().copy()
        hh = np.zeros((w.NY, w.NX))
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 2.0:
            st.retrieve_solutions(Sh, 1.0)
            st.nodal_stress_strain(w.dt, hh)
            st.limit_displacement_timestep(w.dt)
            n += 1
        res["cpu_oracle_ms"] = (time.perf_counter() - t0) / n * 1e3
        res["speedup_vs_cpu"] = res["cpu_oracle_ms"] / ms
    print("nodal step: %.3f ms per call sequence on the GPU" % ms, file=sys.stderr, flush=True)
    return res


def model_flops(net, H):
    """Multiply-add flops (x2) of one forward pass of `net` on one H x H sample."""
    f = 0
    for name, K, ci, co, _ in net.layers:
        lvl = 0 if name == "out" else int(name[1])
        h = H >> lvl
        f += 2 * h * h * K * K * ci * co
    return f


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--vep-cpu-budget", type=float, default=5.0)
    ap.add_argument("--lanes", type=int, default=8, help="lanes per lattice cell in the projections")
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--ens-nmin", type=int, default=65)
    ap.add_argument("--ens-nmax", type=int, default=257)
    ap.add_argument("--no-extras", action="store_true",
                    help="default line: the ensemble only (no single-instance or learned-guess keys); every "
                         "line: no initial-order timing and no scrambled-store run (for profiling)")
    ap.add_argument("--thermal", action="store_true",
                    help="the single-instance convection_tdep step as the line (201x201, 1.44M markers)")
    ap.add_argument("--ensemble", action="store_true", help="the ensemble line only (no extra keys)")
    ap.add_argument("--vep", action="store_true",
                    help="the visco-elasto-plastic step at 801x401, 8M markers, as the line")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 512 MiB device copy (counter calibration)")
    ap.add_argument("--learned-guess", action="store_true",
                    help="train and evaluate the learned Stokes initial guess (own JSON line)")
    ap.add_argument("--lg-grid", type=int, default=129)
    ap.add_argument("--lg-train", type=int, default=512)
    ap.add_argument("--lg-test", type=int, default=24)
    ap.add_argument("--lg-epochs", type=int, default=40)
    ap.add_argument("--lg-workers", type=int, default=0, help="host solver processes (0: this job's cores)")
    ap.add_argument("--pressure-guess", action="store_true",
                    help="time the plain initial pressure guess (own JSON line)")
    ap.add_argument("--pg-grids", type=int, nargs="+", default=[129, 257, 513])
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: every GPU box here has one MI355X, and this library has no "
              "multi-GPU path (the design for one is described, not built).", file=sys.stderr)
        return 2
    if a.copy_calibration:
        return copy_calibration()
    if a.learned_guess:
        out = run_learned_guess(a, a.lg_grid, a.lg_train, a.lg_test, a.lg_epochs)
        out.update({"n_gpus": 1, "warmup": 0, "vs_baseline": None, "data": "synthetic", "scaling": "weak",
                    "cpu_baseline": None})
        print(json.dumps(out))
        return 0
    if a.pressure_guess:
        return bench_pressure_guess(a)

    common = {"metric": METRIC, "n_gpus": 1, "warmup": a.warmup, "higher_is_better": True, "scaling": "weak",
              "vs_baseline": None, "dtype": "f64", "data": "synthetic"}
    if a.thermal or a.vep:
        cpu_all = None if a.no_cpu or a.vep else cpu_all_cores("tdep", a.cpu_budget)
        out = run_thermal(a) if a.thermal else run_vep(a)
        line = dict(common, **out)
        if a.thermal and not a.no_extras:
            line["store_order"] = run_store_order(a)
        if a.thermal and not a.no_cpu:
            line["cpu_baseline"] = cpu_baseline_tdep(a.cpu_budget)
            line["cpu_baseline_all_cores"] = cpu_all
        line.setdefault("cpu_baseline", None)
        print(json.dumps(line))
        return 0

    # default: the batched ensemble; the CPU baselines fork before the GPU is touched
    cpu_all = None if a.no_cpu else cpu_all_cores("ensemble", a.cpu_budget)
    cpu_one = None if a.no_cpu else cpu_baseline_ensemble(a.cpu_budget)
    line = dict(common, **run_ensemble(a))
    line["cpu_baseline"] = cpu_one
    if cpu_one:
        line["speedup_vs_cpu"] = line["value"] / cpu_one["value"]
        line["cpu_baseline_all_cores"] = cpu_all
        line["speedup_vs_cpu_all_cores"] = line["value"] / cpu_all["value"]
    if not (a.no_extras or a.ensemble):
        line["convection_tdep"] = run_thermal(a)
        log("single instance convection_tdep done")
        line["convection_tdep"]["store_order"] = run_store_order(a)
        line["vep_subduction"] = run_vep(a, extras=False)
        log("vep_subduction done")
        try:
            line["learned_guess"] = run_learned_guess(a, 65, 48, 6, 20)
        except Exception as e:  # the marker-step figures stand on their own
            line["learned_guess"] = {"error": "%s: %s" % (type(e).__name__, e)}
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
