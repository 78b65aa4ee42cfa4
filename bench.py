"""Benchmark of the explicit solid MPM step on one MI355X.

A step is one pass of the hot path over the whole particle set:
cell binning (counting sort by cell key), P2G with the grid update,
G2P velocity + MUSL remap, and the dF/F/J/x/V/stress update.
Workload (default): the 3D Neo-Hookean cube on a 128^3 grid at 2x4x2
points per cell = 4,194,304 particles, synthetic (seeded) data, FP32
arithmetic with compensated node sums, FP64 det F and volume, positions
as cell index + FP32 offset.

Prints one JSON line: particle-steps/s over the timed steps, the roofline of
the dominant kernel (algorithmic bytes per launch / HIP-event time on the
engine's stream), and the FP64 CPU oracle timed on the host over a bounded
sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hpc-mpm-code_amd"))

import numpy as np  # noqa: E402

from mpm_mi355x import capi, scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
# The headline metric: particle-steps/s of the plain method on the 4.2M-particle
# cube (value), with the surrogate-vs-plain ratio at the stated stress-error bound
# reported beside it (surrogate).
METRIC = "particle-steps/sec on 1 MI355X, surrogate vs plain at \u22641e-3 rel stress error"

# Algorithmic bytes per particle and per active node of each FP32 kernel
# (compulsory loads/stores of that kernel's inputs and outputs in the
# engine's storage format; the cell-partial buffer between a cells kernel
# and its node pass is not counted).
BYTES = {
    # perm 4 + cell offset 12 + v 12 + m 4 + V 4 + stress 24
    "p2g_cells": (60, 0, 0),
    # node m 4 + v 12 + a 12 + (v + dt a) 12 written
    "p2g_nodes": (0, 40, 0),
    # perm 4 + cell offset 12 + m 4 read, v 12 written; node v + dt a (PIC) or a (FLIP) read 12
    "musl_cells": (32, 12, 0),
    # node m 4 read, MUSL v 12 written
    "musl_nodes": (0, 16, 0),
    # key 4 + offset 12 + H 36 + V0 4 read; offset 12 + H 36 + stress 24 + V 4 written (the key only for
    # particles that change cell); node v 12
    "update": (132, 12, 0),
    # counting sort: keys read twice, perm written (12 per particle); per cell: count read twice,
    # start and end written (16)
    "bin": (12, 0, 16),
}


# HBM traffic per launch measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE on the 4,194,304-particle cube (bytes; copy-kernel calibrated,
# profiles/hbm_traffic_cfg2_4M_fp32.md). Reported as roofline.traffic for
# that workload only.
MEASURED_TRAFFIC = {
    "cfg2-3d-neohookean-16ppc": {"update": 573017160, "p2g_cells": 291583874, "musl_cells": 168259878,
                                 "p2g_nodes": 54485120, "musl_nodes": 30084090},
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="cfg2-4M",
                    choices=["cfg1", "cfg2", "cfg2-4M", "cfg2-q", "cfg2-q212", "cfg2-q4M", "cfg3", "cfg5", "cfg5-16M"])
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--relayout", type=int, default=64, help="relayout interval in steps (0 = never)")
    ap.add_argument("--fused-count", type=int, default=1, help="count cells in the update kernel (1) or not (0)")
    ap.add_argument("--graph-block", type=int, default=0,
                    help="steps per recorded HIP graph block (0 = eager launches)")
    ap.add_argument("--cells-grid", type=int, default=0, help="workgroups of the cells kernels (0 = automatic)")
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lanes-per-cell", type=int, default=0, help="override (1, 2, 4, 8); 0 = automatic")
    ap.add_argument("--no-probe", action="store_true", help="skip the copy-bandwidth probe")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the J2 surrogate vs plain comparison")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 256-scene ensemble measurement")
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="FP32 vs FP64-exact comparison after this many steps (0 = skip)")
    ap.add_argument("--ens-scenes", type=int, default=256)
    ap.add_argument("--ens-steps", type=int, default=500)
    ap.add_argument("--sur-collect", type=int, default=200, help="logged plain steps before the fit")
    ap.add_argument("--sur-samples", type=int, default=5_000_000)
    ap.add_argument("--sur-steps", type=int, default=100, help="timed steps of each method")
    ap.add_argument("--sur-epochs", type=int, default=40)
    ap.add_argument("--sur-batch", type=int, default=4096)
    ap.add_argument("--sur-tol", type=float, default=1e-4, help="the stated fallback tolerance")
    ap.add_argument("--sur-sweep", default="1.5e-4,2e-4,3e-4,1e-3,3e-3,1e-2,1e-1,1e30",
                    help="further tolerances reported")
    ap.add_argument("--no-closure", action="store_true", help="skip the cfg5 pressure-closure measurement")
    ap.add_argument("--closure-collect", type=int, default=100)
    ap.add_argument("--closure-steps", type=int, default=100)
    ap.add_argument("--no-quadratic", action="store_true",
                    help="skip the reference-shape-function (idimc 2 1 2) workload")
    ap.add_argument("--no-long-run", action="store_true",
                    help="skip the mean over the configured 1000 steps and the figure after 1500 steps")
    return ap.parse_args()


def make_scene(name):
    if name == "cfg1":
        return scenes.cfg1()
    if name == "cfg3":
        return scenes.cfg3()
    if name == "cfg5":
        return scenes.cfg5()
    if name == "cfg5-16M":
        return scenes.cfg5(ppc=(4, 2, 4))
    if name == "cfg2":
        return scenes.cfg2(ppc=(2, 2, 2))
    if name == "cfg2-q":  # the cfg2 geometry on the quadratic B-spline (27 control points per particle)
        return scenes.cfg2(ppc=(2, 2, 2), idimc=(2, 2, 2))
    if name == "cfg2-q212":  # idimc 2 1 2, the shape functions of the reference's shipped cases (18 points)
        return scenes.cfg2(ppc=(2, 2, 2), idimc=(2, 1, 2))
    if name == "cfg2-q4M":  # the quadratic B-spline at 2x4x2 points per cell (4,194,304 particles)
        return scenes.cfg2(ppc=(2, 4, 2), idimc=(2, 2, 2))
    return scenes.cfg2(ppc=(2, 4, 2))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, nsteps):
    """The step on the host cores, on the same workload, step loop only: the
    OpenMP port of the FP64 oracle (oracle/mpm_oracle_omp.c, -O3
    -march=native, cell colouring) on up to 16 threads, the box's CPU share,
    and the scalar oracle on one thread for one step beside it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    threads = min(16, affinity)
    o = pyoracle.Oracle(scene, fast="omp")
    t0 = time.perf_counter()
    rc = o.run_omp(1, nsteps, threads)
    dt = time.perf_counter() - t0
    if rc != 0:
        return None
    del o
    o1 = pyoracle.Oracle(scene, fast=True)
    t1 = time.perf_counter()
    rc1 = o1.run(1, 1)
    dt1 = time.perf_counter() - t1
    return {"value": scene.n * nsteps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": "%s, %d particles, %d steps, FP64 OpenMP port of the step (cell colouring), %d threads, "
                      "%.1f s" % (scene.name, scene.n, nsteps, threads, dt),
            "cpu_model": cpu_model(), "host_cores": os.cpu_count(), "cores_available_to_this_process": affinity,
            "single_thread": None if rc1 else {"value": scene.n / dt1, "steps": 1, "seconds": round(dt1, 2)}}


def timed_steps(eng, istep0, nsteps):
    capi.check(eng.lib.mpm_sync(eng.h))
    t0 = time.perf_counter()
    capi.check(eng.step(istep0, nsteps))
    capi.check(eng.lib.mpm_sync(eng.h))
    return time.perf_counter() - t0


def sample_rate(scene, a, collect_steps, target):
    """Logging rate that makes `target` samples over the first collect_steps
    steps: a counting pass (log capacity 1, every candidate offered) gives
    the number of candidate states (J2 plastic trial states, or Tait
    states) over those steps."""
    e = capi.Engine(scene, precision=capi.PREC_FP32, relayout_every=a.relayout)
    e.upload(scene.parts)
    e.set_option(capi.OPT_FULL_NODE_OUTPUT, 0)
    e.surrogate_collect(True, capacity=1, rate=1.0)
    capi.check(e.step(1, collect_steps))
    offered = e.surrogate_stats()["offered"]
    e.close()
    return min(1.0, target / max(1.0, float(offered))), offered


def surrogate_study(a, scene, collect_steps, timed_steps_n, tols, epochs, seed):
    """Plain steps log candidate states (rate set so the log holds
    a.sur_samples of them), the network is fitted on the GPU, then the plain
    method and the surrogate at each tolerance advance the same state; the
    stress error is that of the surrogate run against the plain run
    (Frobenius over all particles)."""
    n = scene.n
    rate, offered = sample_rate(scene, a, collect_steps, a.sur_samples)
    eng = capi.Engine(scene, precision=capi.PREC_FP32, relayout_every=a.relayout)
    eng.upload(scene.parts)
    eng.set_option(capi.OPT_FULL_NODE_OUTPUT, 0)
    eng.surrogate_collect(True, capacity=a.sur_samples, rate=rate)
    t_collect = timed_steps(eng, 1, collect_steps)
    t0 = time.perf_counter()
    loss = eng.surrogate_fit(epochs=epochs, batch=a.sur_batch, lr=2e-3, lr_final=5e-5, seed=seed)
    t_fit = time.perf_counter() - t0
    st = eng.surrogate_stats()
    params, scaling = eng.surrogate_params()
    state = eng.download(n)
    eng.close()
    i0 = 1 + collect_steps

    def advance(tol, fused=True):
        e = capi.Engine(scene, precision=capi.PREC_FP32, relayout_every=a.relayout)
        e.upload(state)
        e.set_option(capi.OPT_FULL_NODE_OUTPUT, 0)
        e.set_option(capi.OPT_SURROGATE_FUSED, int(fused))
        if tol is not None:
            e.surrogate_set_params(params, scaling)
            e.surrogate_enable(True, tol)
        e.set_timing(True)
        wall = timed_steps(e, i0, timed_steps_n)
        kt = e.kernel_times()
        ss = e.surrogate_stats() if tol is not None else None
        out = e.download(n)
        e.close()
        upd = round(1e3 * kt["g2p_upd_ms"] / timed_steps_n, 2)
        return n * timed_steps_n / wall, ss, out.stress, upd

    plain_rate, _, sp, plain_upd = advance(None)
    runs = []
    fused_su = {}
    for tol in tols:
        rate_s, ss, su, upd = advance(tol)
        fused_su[tol] = su
        runs.append({"tol": tol, "particle_steps_per_s": rate_s, "ratio": rate_s / plain_rate,
                     "rel_stress_err": float(np.linalg.norm(su - sp) / np.linalg.norm(sp)),
                     "max_abs_stress_err": float(np.max(np.abs(su - sp))), "evaluated": ss["evaluated"],
                     "fallback_rate": ss["fallbacks"] / max(1, ss["evaluated"]), "update_us": upd})
    # the calibrated tolerance: the largest whose stress error stays within
    # 1e-3 of the plain method (lowest fallback rate at that error bound)
    ok = [r for r in runs if r["rel_stress_err"] <= 1e-3]
    cal = max(ok, key=lambda r: r["tol"]) if ok else None
    # the same tolerance with the network as a separate pass over queued
    # candidates (the fused kernel's results are the same bit for bit)
    ref_tol = (cal or runs[0])["tol"]
    q_rate, q_ss, q_su, q_upd = advance(ref_tol, fused=False)
    fused_run = next(r for r in runs if r["tol"] == ref_tol)
    queued = {"tol": ref_tol, "particle_steps_per_s": q_rate, "ratio": q_rate / plain_rate, "update_us": q_upd,
              "fused_update_us": fused_run["update_us"], "fused_ratio": fused_run["ratio"],
              "same_stress_as_fused": bool(np.array_equal(q_su, fused_su[ref_tol]))}
    return {
        "workload": scene.name, "particles": n, "grid": "x".join(map(str, scene.nelem)),
        "plain_particle_steps_per_s": plain_rate, "plain_update_us": plain_upd,
        "candidates_in_collect_window": offered, "samples": st["samples"], "sample_rate": rate,
        "collect_steps": collect_steps, "collect_s": round(t_collect, 3), "fit_s": round(t_fit, 3),
        "fit_epochs": epochs, "fit_batch": a.sur_batch, "fit_loss": loss, "timed_steps": timed_steps_n,
        "runs": runs, "calibrated": cal, "network_as_separate_pass": queued,
    }


def surrogate_vs_plain(a):
    """J2 column (192x96x96, 2.43 M particles): the surrogate for the plastic
    corrector against the plain return mapping, at the stated tolerance and
    a sweep."""
    tols = [a.sur_tol] + [float(t) for t in a.sur_sweep.split(",") if t]
    r = surrogate_study(a, scenes.cfg3(), a.sur_collect, a.sur_steps, tols, a.sur_epochs, 1)
    head = r["runs"][0]
    r.update({"surrogate_particle_steps_per_s": head["particle_steps_per_s"], "ratio": head["ratio"],
              "rel_stress_err": head["rel_stress_err"], "tol": head["tol"], "fallback_rate": head["fallback_rate"],
              "surrogate_update_us": head["update_us"]})
    return r


def closure_vs_plain(a):
    """cfg5 dam break (4.19 M particles): the learned pressure closure
    p(J, rate) against the exact Tait pressure, which also checks it."""
    tols = [1e-3, 3e-3, 1e-2, 3e-2, 1e30]
    return surrogate_study(a, scenes.cfg5(), a.closure_collect, a.closure_steps, tols, a.sur_epochs, 2)


def quadratic_run(a, name="cfg2-q212", steps=100, warmup=20):
    """The reference's own shape functions (idimc 2 1 2: quadratic B-spline
    in x and z, tent in y; data/generate/solid/input.txt) on the cfg2 cube,
    2,097,152 particles, FP32, same kernels and options as the headline;
    per-kernel times with HIP events."""
    sc = make_scene(name)
    e = capi.Engine(sc, precision=capi.PREC_FP32, relayout_every=a.relayout)
    e.upload(sc.parts)
    e.set_option(capi.OPT_FULL_NODE_OUTPUT, 0)
    capi.check(e.step(1, warmup))
    t = timed_steps(e, 1 + warmup, steps)
    e.set_timing(True)
    capi.check(e.step(1 + warmup + steps, 20))
    e.set_timing(False)
    kt = e.kernel_times()
    e.close()
    keys = {"bin": "bin_ms", "p2g_cells": "p2g_cells_ms", "p2g_nodes": "p2g_nodes_ms", "musl_cells": "musl_cells_ms",
            "musl_nodes": "musl_nodes_ms", "update": "g2p_upd_ms"}
    return {"workload": sc.name, "idimc": list(sc.idimc), "particles": sc.n, "steps": steps,
            "ms_per_step": 1e3 * t / steps, "particle_steps_per_s": sc.n * steps / t,
            "kernels_avg_us": {k: round(1e3 * kt[v] / 20, 2) for k, v in keys.items()}}


def long_run(a, scene_name):
    """Mean ms per step over the configured 1000 steps from the initial
    state (relayouts included), and the rate after 1500 steps, when the
    particles have left their initial lattice."""
    sc = make_scene(scene_name)
    e = capi.Engine(sc, precision=capi.PREC_FP32, relayout_every=a.relayout)
    e.upload(sc.parts)
    e.set_option(capi.OPT_FULL_NODE_OUTPUT, 0)
    t1000 = timed_steps(e, 1, 1000)
    capi.check(e.step(1001, 500))
    t_late = timed_steps(e, 1501, 300)
    e.close()
    return {"workload": sc.name, "steps_1_1000_ms_per_step": 1e3 * t1000 / 1000,
            "steps_1_1000_particle_steps_per_s": sc.n * 1000 / t1000,
            "steps_1501_1800_ms_per_step": 1e3 * t_late / 300,
            "steps_1501_1800_particle_steps_per_s": sc.n * 300 / t_late}


def accuracy_run(a, scene_name):
    """Per-field error of the FP32 engine against the FP64-exact engine (the
    oracle's arithmetic in the reference's summation order) on the same
    workload after the same number of steps."""
    out = {}
    t = {}
    for name, prec, rel in (("fp64", capi.PREC_FP64_EXACT, 0), ("fp32", capi.PREC_FP32, a.relayout)):
        sc = make_scene(scene_name)
        e = capi.Engine(sc, precision=prec, relayout_every=rel)
        e.upload(sc.parts)
        t0 = time.perf_counter()
        capi.check(e.step(1, a.accuracy_steps))
        t[name] = time.perf_counter() - t0
        out[name] = (e.download(sc.n), e.nodes())
        e.close()
    (p64, n64), (p32, n32) = out["fp64"], out["fp32"]
    L = 1.0
    return {
        "workload": sc.name, "steps": a.accuracy_steps, "reference": "FP64-exact engine",
        "position_max_abs_over_L": float(np.max(np.abs(p32.coord - p64.coord)) / L),
        "stress_rel_fro": float(np.linalg.norm(p32.stress - p64.stress) / np.linalg.norm(p64.stress)),
        "stress_max_abs_Pa": float(np.max(np.abs(p32.stress - p64.stress))),
        "velocity_rel_l2": float(np.linalg.norm(p32.vel - p64.vel) / np.linalg.norm(p64.vel)),
        "J_max_abs": float(np.max(np.abs(p32.det_def_grad - p64.det_def_grad))),
        "node_momentum_rel_l2": float(np.linalg.norm(n32.nmome - n64.nmome) / np.linalg.norm(n64.nmome)),
        "fp64_engine_s": round(t["fp64"], 2), "fp32_engine_s": round(t["fp32"], 2),
    }


def ensemble_run(a):
    """cfg4: 256 independent 2D drop scenes in three grid-size buckets. Each
    bucket's scenes are packed into one atlas grid (one engine per bucket,
    a tile per scene); the autotuner picks the cells-kernel shape per bucket,
    and the three engines advance in one captured HIP graph per 32-step
    block. For comparison the same scenes as 256 separate engines (one graph
    branch per scene) are timed over two blocks."""
    t_setup = time.perf_counter()
    atl = scenes.cfg4_atlases(a.ens_scenes)
    ntot = sum(x[0].n for x in atl)
    ens = capi.Ensemble([x[0] for x in atl], bucket=list(range(len(atl))), precision=capi.PREC_FP32,
                        relayout_every=16)
    t_setup = time.perf_counter() - t_setup
    # 500 steps: 2, the autotuner's 9 x 4, then eager steps up to the first
    # whole number of 32-step graph blocks that ends at step 500
    block = 32
    step = 1
    ens.step(step, 2)
    step += 2
    tune = ens.autotune(step, trial_steps=4)
    step += 9 * 4
    n_eager = (a.ens_steps - (step - 1)) % block or block
    t_eager = ens.step(step, n_eager)
    step += n_eager
    ens.capture(step, block)
    nlaunch = (a.ens_steps - (step - 1)) // block
    t_graph = ens.launch(nlaunch)
    step += nlaunch * block
    ens.close()
    per_bucket = []
    for r, (sc, lay, idx) in zip(tune, atl):
        G = int(sc.name.rsplit("-", 1)[1])
        per_bucket.append({"grid": "%dx%d" % (G, G),
                           "scenes": len(idx), "particles": sc.n, "atlas_cells": sc.nelem_total,
                           "dt_per_scene_range": [min(sc.mat_dt), max(sc.mat_dt)] if sc.mat_dt else [sc.dt, sc.dt],
                           "lanes_per_cell": r["lanes_per_cell"], "relayout_every": r["relayout_every"],
                           "particle_steps_per_s_eager": sc.n / (r["ms_per_step"] * 1e-3),
                           "autotune_gain": r["ms_per_step_worst"] / r["ms_per_step"]})
    # the same scenes as separate engines
    scs, bucket = scenes.cfg4(a.ens_scenes)
    sep = capi.Ensemble(scs, bucket=bucket, precision=capi.PREC_FP32, relayout_every=16)
    sep.step(1, 2)
    sep.capture(3, 32)
    t_sep = sep.launch(2)
    sep.close()
    return {"workload": "cfg4-ensemble-%d-scenes" % a.ens_scenes, "particles": ntot, "steps": step - 1,
            "layout": "one atlas engine per grid-size bucket", "setup_s": round(t_setup, 2),
            "graph_block_steps": block, "graph_launches": nlaunch,
            "particle_steps_per_s_graph": ntot * nlaunch * block / (t_graph * 1e-3),
            "particle_steps_per_s_eager": ntot * n_eager / (t_eager * 1e-3),
            "ms_per_step_graph": t_graph / (nlaunch * block), "ms_per_step_eager": t_eager / n_eager,
            "time_step": "each scene its own CFL dt (per-material time steps)",
            "separate_engines_particle_steps_per_s_graph": ntot * 64 / (t_sep * 1e-3),
            "per_bucket": per_bucket}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: this engine has no multi-GPU path; rerun with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    scene = make_scene(a.config)
    prec = capi.PREC_FP32 if a.precision == "fp32" else capi.PREC_FP64_EXACT
    eng = capi.Engine(scene, precision=prec, relayout_every=a.relayout if prec == capi.PREC_FP32 else 0)
    eng.upload(scene.parts)
    eng.set_option(capi.OPT_FULL_NODE_OUTPUT, 0)
    if a.lanes_per_cell:
        eng.set_option(capi.OPT_LANES_PER_CELL, a.lanes_per_cell)
    if a.graph_block:
        eng.set_option(capi.OPT_GRAPH_BLOCK, a.graph_block)
    if a.cells_grid:
        eng.set_option(capi.OPT_CELLS_GRID, a.cells_grid)
    eng.set_option(capi.OPT_FUSED_COUNT, a.fused_count)
    capi.check(eng.step(1, a.warmup))
    capi.check(eng.lib.mpm_sync(eng.h))
    t0 = time.perf_counter()
    capi.check(eng.step(1 + a.warmup, a.steps))
    capi.check(eng.lib.mpm_sync(eng.h))
    wall = time.perf_counter() - t0
    # per-kernel breakdown: a separate eager run with HIP events between the kernels
    kt_steps = max(10, min(a.steps, 100))
    eng.set_timing(True)
    capi.check(eng.step(1 + a.warmup + a.steps, kt_steps))
    eng.set_timing(False)
    kt = eng.kernel_times()
    n = scene.n
    value = n * a.steps / wall
    ms_step = 1e3 * wall / a.steps

    nact = kt["active_nodes"]
    per = {
        "bin": kt["bin_ms"], "p2g_cells": kt["p2g_cells_ms"], "p2g_nodes": kt["p2g_nodes_ms"],
        "musl_cells": kt["musl_cells_ms"], "musl_nodes": kt["musl_nodes_ms"], "update": kt["g2p_upd_ms"],
    }
    kernels = {}
    for k, ms in per.items():
        bp, bn, bc = BYTES[k]
        byts = bp * n + bn * nact + bc * scene.nelem_total
        avg_ms = ms / kt_steps
        gbs = byts / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
        kernels[k] = {"avg_us": round(avg_ms * 1e3, 2), "alg_bytes": byts, "GBps": round(gbs, 1),
                      "frac": round(gbs / HBM_PEAK_GBS, 4)}
        measured = MEASURED_TRAFFIC.get(scene.name, {}).get(k)
        if measured:
            kernels[k]["traffic"] = measured
            kernels[k]["traffic_over_alg"] = round(measured / byts, 3)
    dom = max((k for k in kernels if k != "bin"), key=lambda k: kernels[k]["avg_us"])
    d = kernels[dom]
    traffic = MEASURED_TRAFFIC.get(scene.name, {}).get(dom)
    roofline = {"bound": "hbm", "achieved": d["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": d["frac"], "traffic": traffic, "kernel": dom,
                "alg_bytes": d["alg_bytes"],
                "traffic_source": "profiles/hbm_traffic_cfg2_4M_fp32.md" if traffic else None}
    # P2G and G2P as whole passes
    def phase(keys):
        b = sum(kernels[k]["alg_bytes"] for k in keys)
        us = sum(kernels[k]["avg_us"] for k in keys)
        out = {"avg_us": round(us, 2), "GBps": round(b / (us * 1e-6) / 1e9, 1),
               "frac": round(b / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)}
        if all("traffic" in kernels[k] for k in keys):
            out["traffic_over_alg"] = round(sum(kernels[k]["traffic"] for k in keys) / b, 3)
        return out

    phases = {"p2g": phase(["p2g_cells", "p2g_nodes"]), "g2p": phase(["musl_cells", "musl_nodes", "update"])}

    copy_gbps = None if a.no_probe else round(capi.bandwidth_probe(1 << 30, 20), 1)
    roofline["copy_ceiling_GBps"] = copy_gbps
    # the update kernel's own stream mix (15 fields read, 19 written) as a
    # plain streaming kernel: the ceiling for a kernel that moves bytes so
    roofline["stream_mix_ceiling_GBps"] = None if a.no_probe else round(capi.stream_mix_probe(1 << 22, 20), 1)

    sur = None
    if not a.no_surrogate:
        sur = surrogate_vs_plain(a)
    ens = None
    if not a.no_ensemble:
        ens = ensemble_run(a)
    clo = None
    if not a.no_closure:
        clo = closure_vs_plain(a)
    acc = None
    if a.accuracy_steps > 0:
        acc = accuracy_run(a, a.config)
    quad = None
    if not a.no_quadratic and prec == capi.PREC_FP32 and a.config == "cfg2-4M":
        quad = quadratic_run(a)
    lr = None
    if not a.no_long_run and prec == capi.PREC_FP32:
        lr = long_run(a, a.config)

    cpu = None
    if not a.no_cpu:
        cpu_scene = make_scene(a.config)
        cpu = cpu_baseline(cpu_scene, a.cpu_steps)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if prec == capi.PREC_FP32 else "fp64",
        "data": "synthetic",
        "config": {"workload": scene.name, "particles": n, "grid": "x".join(map(str, scene.nelem)),
                   "active_cells": kt["active_cells"], "active_nodes": nact, "method": "plain (no surrogate)",
                   "scheme": "PIC/MUSL", "shape_function": "quadratic B-spline %s" % (scene.idimc,)
                   if tuple(scene.idimc) != (1, 1, 1) else "linear tent", "relayout_every": a.relayout,
                   "graph_block": a.graph_block,
                   "kernel_timing_steps": kt_steps},
        "roofline": roofline,
        "phases": phases,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "surrogate": sur,
        "closure": clo,
        "ensemble": ens,
        "accuracy": acc,
        "long_run": lr,
        "reference_shape_functions": quad,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
