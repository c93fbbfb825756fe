#!/usr/bin/env python3
"""Headline benchmark of the MI355X Barnes-Hut engine.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode learned|plain]

Workload: 4,194,304 bodies in 64 Gaussian clusters (seeded), softening 0.01,
dt 1e-3, FP32 traversal (Kahan-compensated accumulation at theta = 0.4;
the coarse traversal keeps bodies, monopoles and quadrupoles in three plain
FP32 sums), FP64 summarize and FP64 leapfrog positions, vote width 64. A step is one full
pass of the pipeline (bounding box, build, summarize, sort, force,
kick-drift-kick) over all bodies; inputs are resident in HBM before the
timed region starts. The metric compares two force methods, each of which
must reach a median relative force error <= 1e-3 against the FP64 direct sum
on 4096 sampled bodies:
  plain   Barnes-Hut at theta = 0.4 (the reference's opening test)
  learned coarse traversal (monopole + softened quadrupole; a cell is taken
          when |d| >= b / theta_c for its extent b, theta_c = 0.99, and its
          left-out term m b^3 / s^5 is within alpha of the body's previous
          |a|) plus the learned correction, trained on the GPU on 50 steps of
          theta = 0.4 residuals before the timed region; alpha is the largest
          of 0.03, 0.025, 0.0225, 0.02, 0.0175, 0.015, 0.0125 whose calibration
          error (a separate sample) is <= 0.9e-3. The sampled FP64 check (1024
          bodies every 10 steps, tolerance 1e-3, plain theta as the fallback
          for a failing step, which also tightens alpha)
          runs inside the timed steps and counts in `value` (its direct sum runs
          on a second stream beside the step's other work); `config` also gives
          its device time and the steps/s of steps between checks.
`value` is the steps/s of the requested mode (learned by default) over the
K timed steps; the other mode's steps/s, both errors, each stage's fraction
of the HBM roof and the median of three consecutive runs of K steps are in
`config`. Also in `config`: the coarse traversal alone (no network) at its
own calibrated alpha, timed the same way (what the network gains at equal
error), and 1000 sustained learned steps after the timed ones.

Prints one JSON line. One GPU only: --gpus N > 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-nbody_amd"))

METRIC = "steps/s @4M bodies: learned coarse-θ vs plain θ=0.4 at equal 1e-3 force error"
N_BODIES = 4_194_304
THETA_PLAIN = 0.4
THETA_COARSE = 0.99
ALPHA_CANDIDATES = (0.03, 0.025, 0.0225, 0.02, 0.0175, 0.015, 0.0125)
CALIBRATION_TARGET = 0.9e-3
ERROR_BAR = 1e-3
TRAIN = dict(steps=50, groups_per_step=512, iters=4000, batch=4096, seed=1)
CHECK = dict(check_tol=1e-3, check_every=10, check_samples=1024, check_adapt=1)
SOFTENING = 0.01
DT = 1e-3
VOTE = 64
FLOP_PER_INTERACTION = 20  # GPU Gems 3 ch. 31 convention (Nyland, Harris, Prins 2007)
# a cell of the coarse traversal adds its softened quadrupole to those 20: Q d (3 mul + 6 fma = 15),
# d.Q.d (1 mul + 2 fma = 5), s^-5 (1), the radial factor (1 sub + 2 mul = 3) and the two extra
# terms of the sum (6 fma = 12): 36 more
FLOP_PER_CELL_INTERACTION = 56
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 (vector packed = matrix dense) peak
PEAK_HBM_GBS = 8000.0
SUSTAINED_STEPS = 1000
CPU_VOTE = 16  # the CPU baseline runs the reference's own work-group vote


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_sim(state, mode, alpha=0.02, fused=True):
    from nbody_amd import HipBarnesHutNBodySimulation, Mode, _lib
    from nbody_amd.universe import ArrayUniverse

    kw = dict(theta=THETA_PLAIN, softening=SOFTENING, dt=DT, vote_width=VOTE,
              flags=_lib.NB_FP64_POSITIONS | (_lib.NB_FUSED_UPDATE if fused else 0))
    if mode == "learned":
        kw.update(method=_lib.NB_MODE_LEARNED, coarse_theta=THETA_COARSE, coarse_alpha=alpha, **CHECK)
    elif mode == "coarse":
        kw.update(method=_lib.NB_MODE_COARSE, coarse_theta=THETA_COARSE, coarse_alpha=alpha)
    sim = HipBarnesHutNBodySimulation(Mode.DEFAULT, len(state[0]), ArrayUniverse(state), **kw)
    sim.init(None)
    return sim


def train_learned(state, candidates, train):
    """Train one network per candidate alpha (on the GPU, on the simulation's
    own pairs) and keep the coarsest whose error on a calibration sample meets
    the target."""
    chosen = None
    for al in candidates:
        sim = make_sim(state, "learned", al)
        t0 = time.perf_counter()
        rep = sim.train(**train)
        rep["wall_s"] = time.perf_counter() - t0
        cal = sim.force_error(4096, seed=4242)
        log(f"[learned] theta_c {THETA_COARSE} alpha {al}: calibration median {cal['median']:.3e}, train {rep}")
        if chosen is not None:
            chosen[0].close()
        chosen = (sim, al, rep, cal)
        if cal["median"] <= CALIBRATION_TARGET:
            break
    return chosen


def calibrate_coarse(state, candidates):
    """The coarse traversal alone (no network, no check): after the same 50
    steps of simulated time as the learned run's training, the coarsest alpha
    whose calibration error meets the same target."""
    chosen = None
    for al in candidates:
        sim = make_sim(state, "coarse", al)
        sim.steps(TRAIN["steps"])
        cal = sim.force_error(4096, seed=4242)
        log(f"[coarse] theta_c {THETA_COARSE} alpha {al}: calibration median {cal['median']:.3e}")
        if chosen is not None:
            chosen[0].close()
        chosen = (sim, al, cal)
        if cal["median"] <= CALIBRATION_TARGET:
            break
    return chosen


def timed_run(sim, steps, warmup):
    """W untimed steps, then exactly K steps between two synchronizations
    (nb_step returns after its stream has drained). Also returns the device
    time and the number of the sampled FP64 checks inside the K steps."""
    sim.steps(warmup)
    c0 = sim.counters()
    t0 = time.perf_counter()
    sim.steps(steps)
    t1 = time.perf_counter()
    c1 = sim.counters()
    return (steps / (t1 - t0), (t1 - t0) * 1e3 / steps, c1.check_ms - c0.check_ms,
            int(c1.checks - c0.checks))


def timed_between_checks(sim):
    """Steps/s over the check_every - 1 steps of a check period that hold no
    sampled check: single steps until one runs a check, then the rest of the
    period timed (the counter of checks must not move in it)."""
    c0 = sim.counters().checks
    for _ in range(CHECK["check_every"]):
        sim.steps(1)
        if sim.counters().checks != c0:
            break
    c1 = sim.counters().checks
    k = CHECK["check_every"] - 1
    t0 = time.perf_counter()
    sim.steps(k)
    t1 = time.perf_counter()
    if sim.counters().checks != c1:
        return None
    return k / (t1 - t0)


def stage_rooflines(stages, n, cells, fused, learned):
    """Per-stage fraction of the HBM roof from the algorithmic bytes of each
    stage in the reference's FP32 layout (4-byte words; C = live cells): bbox
    12N, build 16N + 52C, summarize 16N + 68C, sort 4N + 44C, integrate 96N
    with FP64 positions. The stage times are the HIP-event times of the whole
    stage (every launch of it and the gaps between them). The sort stage
    launches nothing: the build's sorted body order is the octant-path order,
    and its buffer becomes sorted[]."""
    algo = {"bbox": 12 * n, "build": 16 * n + 52 * cells, "summarize": 16 * n + 68 * cells,
            "sort": 4 * n + 44 * cells, "integrate": 96 * n}
    out = {}
    for k, b in algo.items():
        ms = stages[k]
        if k == "integrate" and fused:
            out[k] = {"algorithmic_bytes": b, "note": "fused into the force kernel's epilogue"}
            continue
        if k == "sort":
            out[k] = {"algorithmic_bytes": b, "ms": ms,
                      "note": "no launch: the build's sorted order buffer becomes sorted[]"}
            continue
        gbs = b / (ms * 1e-3) / 1e9 if ms > 0 else None
        out[k] = {"algorithmic_bytes": b, "ms": ms, "GB/s": round(gbs, 1) if gbs else None,
                  "frac": round(gbs / PEAK_HBM_GBS, 4) if gbs else None}
    if learned:
        out["summarize"]["note"] = ("counts the reference's monopole bytes; the learned mode also "
                                    "sums second moments and extents (FP64)")
    return out


def stage_profile(sim, steps, checked=False):
    """Per-stage device time (HIP events on the engine's stream) over `steps`.
    With the learned mode's sampled check on (checked), the profiled steps are
    ones without a check: single steps until one runs the check, then at most
    check_every - 1 steps, and the check counter must not move in them (a
    check step's direct sum on the second stream would share the GPU with
    its tree stages)."""
    if checked:
        c0 = sim.counters().checks
        for _ in range(CHECK["check_every"]):
            sim.steps(1)
            if sim.counters().checks != c0:
                break
        steps = min(steps, CHECK["check_every"] - 1)
    c_before = sim.counters().checks
    sim.set_timing(True)
    sim.steps(steps)
    sim.set_timing(False)
    c = sim.counters()
    if checked and c.checks != c_before:
        raise RuntimeError("a sampled check ran inside the stage profile")
    names = ("bbox", "build", "summarize", "sort", "force", "integrate")
    return ({k: round(c.stage_ms[i], 4) for i, k in enumerate(names)}, c.interactions / steps,
            c.cell_interactions / steps, int(c.live_cells))


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    pyoracle.build()
    return pyoracle


def _cpu_step(state, theta, softening, vote, flags):
    """One step of the reference algorithm on the host: the oracle's bounding
    box, build, summarize and sort (sequential C), the FP32 force pass of
    oracle/cpu_fp32.c over every vote group on the OpenMP threads, and the
    oracle's integrate. Returns (tree s, force s, integrate s, interactions)."""
    pyoracle = _oracle()
    n = len(state[0])
    o = pyoracle.Oracle(n, theta=theta, softening=softening, dt=DT, vote=vote, flags=flags)
    o.set_state(*state)
    t0 = time.perf_counter()
    assert o.run_stages(1, 4) == 0
    t_tree = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, ni = o.force_fp32()
    t_force = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.run_stages(6, 6)
    t_int = time.perf_counter() - t0
    o.close()
    return t_tree, t_force, t_int, ni


def host_cpu():
    """Threads the baseline uses and what the host reports: the GPU box gives
    a process 16 cores per GPU, whatever os.cpu_count() shows."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model, avx512 = "unknown", False
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                if line.startswith("flags") and " avx512f" in line:
                    avx512 = True
    except OSError:
        pass
    return {"threads": min(16, avail), "os_cpu_count": os.cpu_count(), "affinity_cpus": avail,
            "model": model, "avx512f": avx512}


def cpu_baseline(state, host):
    """The reference step restated on the host cores (C, FP32 force, the
    reference's vote width of 16) on the benchmark's workload: one full step,
    every vote group."""
    pyoracle = _oracle()
    threads = host["threads"]
    pyoracle.lib().orc_set_threads(threads)
    n = len(state[0])
    t_tree, t_force, t_int, ni = _cpu_step(state, THETA_PLAIN, SOFTENING, CPU_VOTE, pyoracle.FP64_POS)
    step_s = t_tree + t_force + t_int
    return {"value": round(1.0 / step_s, 6), "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": (f"one full step of the workload (N={n}, plain theta {THETA_PLAIN}, vote {CPU_VOTE}): "
                       f"bbox+build+summarize+sort of the oracle, sequential C ({t_tree:.2f} s); force on "
                       f"every vote group in FP32 (oracle/cpu_fp32.c, calculateforce.cl:143-151 arithmetic, "
                       f"{threads} OpenMP threads, {ni} interactions, {t_force:.2f} s); integrate "
                       f"({t_int:.2f} s)"),
            "cpu": host, "force_interactions_per_s": round(ni / t_force, 1)}


def config_table(host):
    """The other configurations of the benchmark description with plain
    Barnes-Hut, each timed on the GPU (HIP engine) and one step on the host
    cores (the oracle's tree stages and the FP32 force pass of every vote
    group): C1, C2, and C3 at theta 0.3, 0.4 and 0.6 (the host at 0.4)."""
    import nbody_amd
    from nbody_amd import HipBarnesHutNBodySimulation, Mode, _lib
    from nbody_amd import universe as U
    from nbody_amd.universe import ArrayUniverse

    pyoracle = _oracle()
    pyoracle.lib().orc_set_threads(host["threads"])
    nbody_amd.load()
    c3 = dict(softening=0.01, vote_width=64, flags=_lib.NB_FUSED_UPDATE | _lib.NB_FP64_POSITIONS)
    cfgs = [
        ("C1", "uniform ball in rotation, N=16384", lambda: U.uniform_ball(16384, seed=1),
         dict(theta=0.5, softening=0.05, vote_width=16, flags=0), 10, 1, True),
        ("C2", "Plummer sphere, N=1048576", lambda: U.plummer(1 << 20, seed=2),
         dict(theta=0.5, softening=0.01, vote_width=64, flags=_lib.NB_FUSED_UPDATE), 20, 3, True),
        ("C3", "two Plummer spheres, N=4194304, FP64 positions", lambda: U.two_plummer(1 << 22, seed=3),
         dict(theta=0.4, **c3), 20, 3, True),
        ("C3_theta0.3", "two Plummer spheres, N=4194304, FP64 positions", None, dict(theta=0.3, **c3), 20, 3,
         False),
        ("C3_theta0.6", "two Plummer spheres, N=4194304, FP64 positions", None, dict(theta=0.6, **c3), 20, 3,
         False),
    ]
    out = {}
    st = None
    for name, desc, gen, kw, steps, warm, on_cpu in cfgs:
        if gen is not None:
            st = gen()
        n = len(st[0])
        sim = HipBarnesHutNBodySimulation(Mode.DEFAULT, n, ArrayUniverse(st), dt=DT, **kw)
        sim.init(None)
        sim.steps(warm)
        t0 = time.perf_counter()
        sim.steps(steps)
        gpu_s = (time.perf_counter() - t0) / steps
        err = sim.force_error(4096, seed=11) if name.startswith("C3") else None
        sim.close()
        row = {"workload": desc, "theta": kw["theta"], "softening": kw["softening"],
               "vote_width": kw["vote_width"], "gpu_ms_per_step": round(gpu_s * 1e3, 3),
               "gpu_steps_per_s": round(1.0 / gpu_s, 2), "gpu_steps_timed": steps}
        if err is not None:
            row["median_rel_force_err"] = round(err["median"], 8)
        if on_cpu:
            oflags = pyoracle.FP64_POS if kw["flags"] & _lib.NB_FP64_POSITIONS else 0
            t_tree, t_force, t_int, _ = _cpu_step(st, kw["theta"], kw["softening"], CPU_VOTE, oflags)
            cpu_s = t_tree + t_force + t_int
            row.update({"cpu_s_per_step": round(cpu_s, 4),
                        "cpu_sample": f"1 full step, vote {CPU_VOTE}, FP32 force on {host['threads']} threads",
                        "gpu_over_cpu": round(cpu_s / gpu_s, 1)})
            log(f"[{name}] GPU {gpu_s * 1e3:.3f} ms/step, CPU {cpu_s:.3f} s/step")
        else:
            log(f"[{name}] GPU {gpu_s * 1e3:.3f} ms/step")
        out[name] = row
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=N_BODIES)
    ap.add_argument("--mode", choices=["learned", "plain"], default=None)
    ap.add_argument("--samples", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-config-table", action="store_true",
                    help="skip the C1-C3 table (GPU and CPU seconds per step of the other configurations)")
    ap.add_argument("--alpha", type=float, default=None,
                    help="use this relative-test alpha instead of calibrating one")
    ap.add_argument("--train-iters", type=int, default=None, help="Adam iterations (default 4000)")
    ap.add_argument("--unfused", action="store_true",
                    help="plain mode with the integrate stage as its own kernel (for profiling it)")
    ap.add_argument("--no-ablation", action="store_true",
                    help="skip the coarse-only run at its own calibrated alpha")
    ap.add_argument("--no-sustained", action="store_true",
                    help=f"skip the {SUSTAINED_STEPS} sustained learned steps after the timed ones")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run the vertex copy kernel once per mode (a copy of known size for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import nbody_amd
    from nbody_amd import universe as U

    nbody_amd.load()
    mode = args.mode or "learned"

    log(f"generating {args.n} clustered bodies")
    state = U.gaussian_clusters(args.n, seed=4)

    results = {}
    learned_info = None
    for m in (["plain", "learned"] if mode == "learned" else ["plain"]):
        if m == "learned":
            train = dict(TRAIN)
            if args.train_iters is not None:
                train["iters"] = args.train_iters
            cands = (args.alpha,) if args.alpha else ALPHA_CANDIDATES
            sim, al, rep, cal = train_learned(state, cands, train)
            learned_info = dict(alpha=al, train=rep, calibration_median=cal["median"])
        else:
            sim = make_sim(state, "plain", fused=not args.unfused)
            sim.steps(TRAIN["steps"])  # the same simulated time as the learned run's training
        log(f"[{m}] timing {args.steps} steps after {args.warmup} warm-up steps")
        sps, ms, check_ms, checks = timed_run(sim, args.steps, args.warmup)
        fallbacks = sim.counters().fallbacks
        # two more runs of K steps right after: the median of three runs
        runs = [sps] + [timed_run(sim, args.steps, 0)[0] for _ in range(2)]
        sps_nc = timed_between_checks(sim) if m == "learned" else None
        stages, inter, cell_inter, live = stage_profile(sim, max(3, min(args.steps, 10)), checked=m == "learned")
        err = sim.force_error(args.samples, seed=1000)
        results[m] = dict(steps_per_s=sps, ms_per_step=ms, stages=stages, interactions=inter,
                          cell_interactions=cell_inter, err=err, fallbacks=fallbacks, check_ms=check_ms,
                          steps_per_s_between_checks=sps_nc, runs=runs,
                          checks=checks, live_cells=live,
                          rooflines=stage_rooflines(stages, args.n, live, not args.unfused, m == "learned"))
        if args.copy_calibration:
            # k_copy_vertices reads 2 x 16 B and writes 2 x 16 B per body
            pos4 = np.zeros((args.n, 4), np.float32)
            vel4 = np.zeros((args.n, 4), np.float32)
            sim.copy_vertices(pos4, vel4)
        log(f"[{m}] {sps:.3f} steps/s, {ms:.2f} ms/step, median err {err['median']:.2e}, stages {stages}")
        if m == "learned" and not args.no_sustained:
            # the learned run carries on: SUSTAINED_STEPS more steps in one call,
            # checks and check_adapt live, then the error at the end
            c0 = sim.counters()
            t0 = time.perf_counter()
            sim.steps(SUSTAINED_STEPS)
            t1 = time.perf_counter()
            c1 = sim.counters()
            e_end = sim.force_error(args.samples, seed=1001)
            results[m]["sustained"] = {
                "steps": SUSTAINED_STEPS, "steps_per_s": round(SUSTAINED_STEPS / (t1 - t0), 4),
                "checks": int(c1.checks - c0.checks), "fallbacks": int(c1.fallbacks - c0.fallbacks),
                "alpha_start": round(c0.coarse_alpha, 6), "alpha_end": round(c1.coarse_alpha, 6),
                "median_rel_force_err_end": round(e_end["median"], 8)}
            log(f"[{m}] sustained {results[m]['sustained']}")
        sim.close()
    if mode == "learned" and not args.no_ablation:
        # the coarse traversal alone at its own calibrated alpha: what the
        # network buys at the same error target
        csim, cal_alpha, ccal = calibrate_coarse(state, ALPHA_CANDIDATES)
        csps, cms, _, _ = timed_run(csim, args.steps, args.warmup)
        cruns = [csps] + [timed_run(csim, args.steps, 0)[0] for _ in range(2)]
        cerr = csim.force_error(args.samples, seed=1000)
        results["coarse"] = dict(steps_per_s=csps, ms_per_step=cms, runs=cruns, err=cerr, alpha=cal_alpha,
                                 calibration_median=ccal["median"])
        log(f"[coarse] alpha {cal_alpha}: {csps:.3f} steps/s, median err {cerr['median']:.2e}")
        csim.close()

    main_r = results[mode]
    force_ms = main_r["stages"]["force"]  # the traversal (and, learned, the network and the body update)
    cells = main_r["cell_interactions"]
    inter = main_r["interactions"]
    # the convention of the scope table: 20 FLOP per interaction, cell or body
    achieved = FLOP_PER_INTERACTION * inter / (force_ms * 1e-3) / 1e12
    # beside it: cells of the coarse traversal priced with their quadrupole (56)
    achieved_q = (FLOP_PER_INTERACTION * (inter - cells) + FLOP_PER_CELL_INTERACTION * cells) / (force_ms * 1e-3) / 1e12
    # HBM bytes per launch of the dominant kernel: the rocprofv3 counter runs
    # (FETCH_SIZE, WRITE_SIZE, copy-kernel calibrated) summarized in
    # profiles/force_traffic.json, with the command and kernel they came from
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", "force_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic = tj.get("bytes_per_launch", {}).get("k_coarse64" if mode == "learned" else "k_force64")
        traffic_src = {"file": "profiles/force_traffic.json", "measured_on": tj.get("measured_on"),
                       "command": tj.get("command")}

    line = {
        "metric": METRIC,
        "value": round(main_r["steps_per_s"], 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_r["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded mixture of 64 Gaussian clusters",
        "config": {
            "workload": f"clustered universe, N={args.n}, {mode} force method",
            "n_bodies": args.n,
            "mode": mode,
            "theta_plain": THETA_PLAIN,
            "softening": SOFTENING,
            "dt": DT,
            "vote_width": VOTE,
            "precision": ("FP32 traversal (Kahan), FP64 summarize, FP64 positions" if mode == "plain" else
                          "FP32 traversal (three plain sums: bodies, monopoles, quadrupoles), FP64 summarize, FP64 positions"),
            "median_rel_force_err": round(main_r["err"]["median"], 8),
            "p90_rel_force_err": round(main_r["err"]["p90"], 8),
            "stage_ms": main_r["stages"],
            "live_cells": main_r["live_cells"],
            "stage_roofline": main_r["rooflines"],
            "interactions_per_step": int(main_r["interactions"]),
            "cell_interactions_per_step": int(main_r["cell_interactions"]),
            "interactions_per_s": round(main_r["interactions"] / (force_ms * 1e-3), 1),
            # `value` is the first run of K steps; two more runs follow it
            "steps_per_s_runs": [round(r, 4) for r in main_r["runs"]],
            "steps_per_s_median_of_3_runs": round(float(np.median(main_r["runs"])), 4),
        },
        "roofline": {
            "kernel": "force traversal (" + ("k_coarse64, coarse traversal" if mode == "learned" else "k_force64") + ")",
            "bound": "valu",
            "engine": ("FP32 vector ALU; peak is MI355X FP32 157.3 TFLOP/s" +
                       ("; the BF16 matrix-core network in the kernel's epilogue is neither counted nor "
                        "bounded here" if mode == "learned" else "")),
            "achieved": round(achieved, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "flop_per_interaction": FLOP_PER_INTERACTION,
            "interactions_per_launch": int(inter),
            "launch_ms": force_ms,
            "flop_note": "20 FLOP per interaction (cell or body), the GPU Gems 3 ch. 31 convention",
        },
    }
    if mode == "learned":
        line["roofline"].update({
            "achieved_quadrupole_priced": round(achieved_q, 3),
            "frac_quadrupole_priced": round(achieved_q / PEAK_FP32_TFLOPS, 4),
            "quadrupole_note": ("cells of the coarse traversal priced at 56 FLOP (monopole 20 + softened "
                                "quadrupole 36), bodies at 20"),
            "cell_interactions_per_launch": int(cells)})
        p = results["plain"]
        tr = learned_info["train"]
        line["config"].update({
            "theta_coarse": THETA_COARSE,
            "alpha_coarse": learned_info["alpha"],
            "plain_steps_per_s": round(p["steps_per_s"], 4),
            "plain_ms_per_step": round(p["ms_per_step"], 3),
            "speedup_vs_plain": round(main_r["steps_per_s"] / p["steps_per_s"], 3),
            "plain_steps_per_s_median_of_3_runs": round(float(np.median(p["runs"])), 4),
            "speedup_vs_plain_median_of_3_runs": round(float(np.median(main_r["runs"]) / np.median(p["runs"])), 3),
            "plain_median_rel_force_err": round(p["err"]["median"], 8),
            "plain_p90_rel_force_err": round(p["err"]["p90"], 8),
            "plain_stage_ms": p["stages"],
            "both_within_1e-3": bool(main_r["err"]["median"] <= ERROR_BAR and p["err"]["median"] <= ERROR_BAR),
            "learned_calibration_median": round(learned_info["calibration_median"], 8),
            "check": dict(CHECK, fallbacks_in_timed_steps=int(main_r["fallbacks"]),
                          checks_in_timed_steps=main_r["checks"],
                          # device time of the direct sums, which run on a second
                          # stream beside the check steps' other work
                          check_device_ms_in_timed_steps=round(main_r["check_ms"], 3)),
            # steps/s over a stretch of steps that holds no sampled check
            "steps_per_s_between_checks": (round(main_r["steps_per_s_between_checks"], 4)
                                           if main_r["steps_per_s_between_checks"] else None),
            "speedup_vs_plain_between_checks": (round(main_r["steps_per_s_between_checks"] / p["steps_per_s"], 3)
                                                if main_r["steps_per_s_between_checks"] else None),
            "stage_ms_note": "HIP-event stage times over steps without a sampled check",
            "training": {"steps": TRAIN["steps"], "pairs": tr["samples"], "iters": TRAIN["iters"],
                         "batch": TRAIN["batch"], "pair_generation_ms": round(tr["gen_ms"], 1),
                         "adam_ms": round(tr["train_ms"], 1), "loss_first": round(tr["loss_first"], 4),
                         "loss_last": round(tr["loss_last"], 4)},
        })
        if "sustained" in main_r:
            line["config"]["sustained"] = main_r["sustained"]
        if "coarse" in results:
            c = results["coarse"]
            line["config"]["ablation_coarse_only"] = {
                "what": ("the coarse traversal alone (no network, no check) at its own calibrated alpha, "
                         "same calibration target, timed the same way"),
                "alpha": c["alpha"], "calibration_median": round(c["calibration_median"], 8),
                "steps_per_s": round(c["steps_per_s"], 4),
                "steps_per_s_median_of_3_runs": round(float(np.median(c["runs"])), 4),
                "median_rel_force_err": round(c["err"]["median"], 8),
                "learned_over_coarse_only": round(float(np.median(main_r["runs"]) / np.median(c["runs"])), 3),
                # the coarse mode has no sampled check: the learned steps between checks beside it
                "learned_between_checks_over_coarse_only": (
                    round(main_r["steps_per_s_between_checks"] / float(np.median(c["runs"])), 3)
                    if main_r["steps_per_s_between_checks"] else None)}
    if not args.no_cpu_baseline:
        host = host_cpu()
        log(f"cpu baseline on {host['threads']} threads ({host['model']})")
        line["cpu_baseline"] = cpu_baseline(state, host)
        if not args.no_config_table:
            line["cpu_baseline"]["configs"] = config_table(host)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
