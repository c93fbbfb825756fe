#!/usr/bin/env python3
"""Headline benchmark of the MI355X Barnes-Hut engine.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode learned|plain]

Workload: 4,194,304 bodies in 64 Gaussian clusters (seeded), softening 0.01,
dt 1e-3, FP32 traversal with Kahan-compensated accumulation, FP64
summarize and FP64 leapfrog positions, vote width 64. A step is one full
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
`config`.

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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_sim(state, mode, alpha=0.02, fused=True):
    from nbody_amd import HipBarnesHutNBodySimulation, Mode, _lib
    from nbody_amd.universe import ArrayUniverse

    kw = dict(theta=THETA_PLAIN, softening=SOFTENING, dt=DT, vote_width=VOTE,
              flags=_lib.NB_FP64_POSITIONS | (_lib.NB_FUSED_UPDATE if fused else 0))
    if mode == "learned":
        kw.update(method=_lib.NB_MODE_LEARNED, coarse_theta=THETA_COARSE, coarse_alpha=alpha, **CHECK)
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
    stage (every launch of it and the gaps between them)."""
    algo = {"bbox": 12 * n, "build": 16 * n + 52 * cells, "summarize": 16 * n + 68 * cells,
            "sort": 4 * n + 44 * cells, "integrate": 96 * n}
    out = {}
    for k, b in algo.items():
        ms = stages[k]
        if k == "integrate" and fused:
            out[k] = {"algorithmic_bytes": b, "note": "fused into the force kernel's epilogue"}
            continue
        gbs = b / (ms * 1e-3) / 1e9 if ms > 0 else None
        out[k] = {"algorithmic_bytes": b, "ms": ms, "GB/s": round(gbs, 1) if gbs else None,
                  "frac": round(gbs / PEAK_HBM_GBS, 4) if gbs else None}
    if learned:
        out["summarize"]["note"] = ("counts the reference's monopole bytes; the learned mode also "
                                    "sums second moments and extents (FP64)")
    return out


def stage_profile(sim, steps):
    """Per-stage device time (HIP events on the engine's stream) over `steps`."""
    sim.set_timing(True)
    sim.steps(steps)
    sim.set_timing(False)
    c = sim.counters()
    names = ("bbox", "build", "summarize", "sort", "force", "integrate")
    return ({k: round(c.stage_ms[i], 4) for i, k in enumerate(names)}, c.interactions / steps,
            c.cell_interactions / steps, int(c.live_cells))


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    pyoracle.build()
    return pyoracle


def _oracle_sampled_step(state, theta, softening, vote, flags, every=4):
    """One oracle step: full bbox, build, summarize, sort and integrate, the
    force on one vote group in `every`, scaled to all groups."""
    import numpy as np

    pyoracle = _oracle()
    n = len(state[0])
    o = pyoracle.Oracle(n, theta=theta, softening=softening, dt=DT, vote=vote, flags=flags)
    o.set_state(*state)
    t0 = time.perf_counter()
    assert o.run_stages(1, 4) == 0
    t_tree = time.perf_counter() - t0
    groups = (n + vote - 1) // vote
    sample = np.arange(0, groups, every, dtype=np.int32)
    t0 = time.perf_counter()
    assert o.force_groups(sample) == 0
    t_force = (time.perf_counter() - t0) * groups / len(sample)
    t0 = time.perf_counter()
    o.run_stages(6, 6)
    t_int = time.perf_counter() - t0
    o.close()
    return t_tree, t_force, t_int, len(sample), groups


def cpu_baseline(state, threads):
    """The oracle (plain-C restatement of the reference step, OpenMP force) on
    the host cores: a full tree build of the workload plus the force pass on
    one vote group in four, scaled to a whole step."""
    pyoracle = _oracle()
    pyoracle.lib().orc_set_threads(threads)
    n = len(state[0])
    t_tree, t_force, t_int, ns, groups = _oracle_sampled_step(state, THETA_PLAIN, SOFTENING, VOTE,
                                                              pyoracle.FP64_POS)
    step_s = t_tree + t_force + t_int
    return {"value": round(1.0 / step_s, 6), "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": (f"oracle/bh_oracle.c (reference step restated in C, OpenMP force), plain theta "
                       f"{THETA_PLAIN}, vote {VOTE}, N={n}: full bbox+build+summarize+sort "
                       f"({t_tree:.2f} s) and integrate ({t_int:.2f} s), force on 1 of every 4 vote groups "
                       f"({ns} groups) scaled x{groups / ns:.0f} to {t_force:.1f} s")}


def config_table(threads):
    """The other configurations of the benchmark description with plain
    Barnes-Hut, each timed on the GPU (HIP engine) and on the host cores (the
    oracle): C1 in full (10 steps), one step of C2, and one step of C3 at
    theta 0.4 with the force on one vote group in four, scaled."""
    import nbody_amd
    from nbody_amd import HipBarnesHutNBodySimulation, Mode, _lib
    from nbody_amd import universe as U
    from nbody_amd.universe import ArrayUniverse

    pyoracle = _oracle()
    pyoracle.lib().orc_set_threads(threads)
    nbody_amd.load()
    cfgs = [
        ("C1", "uniform ball in rotation, N=16384", lambda: U.uniform_ball(16384, seed=1),
         dict(theta=0.5, softening=0.05, vote_width=16, flags=0), 10, 1),
        ("C2", "Plummer sphere, N=1048576", lambda: U.plummer(1 << 20, seed=2),
         dict(theta=0.5, softening=0.01, vote_width=64, flags=_lib.NB_FUSED_UPDATE), 20, 3),
        ("C3", "two Plummer spheres, N=4194304, FP64 positions", lambda: U.two_plummer(1 << 22, seed=3),
         dict(theta=0.4, softening=0.01, vote_width=64, flags=_lib.NB_FUSED_UPDATE | _lib.NB_FP64_POSITIONS),
         20, 3),
    ]
    out = {}
    for name, desc, gen, kw, steps, warm in cfgs:
        st = gen()
        n = len(st[0])
        sim = HipBarnesHutNBodySimulation(Mode.DEFAULT, n, ArrayUniverse(st), dt=DT, **kw)
        sim.init(None)
        sim.steps(warm)
        t0 = time.perf_counter()
        sim.steps(steps)
        gpu_s = (time.perf_counter() - t0) / steps
        sim.close()
        oflags = pyoracle.FP64_POS if kw["flags"] & _lib.NB_FP64_POSITIONS else 0
        if name == "C1":
            o = pyoracle.Oracle(n, theta=kw["theta"], softening=kw["softening"], dt=DT,
                                vote=kw["vote_width"], flags=oflags)
            o.set_state(*st)
            t0 = time.perf_counter()
            assert o.step(steps) == 0
            cpu_s = (time.perf_counter() - t0) / steps
            o.close()
            how = f"{steps} full steps"
        else:
            every = 1 if name == "C2" else 4
            t_tree, t_force, t_int, ns, groups = _oracle_sampled_step(st, kw["theta"], kw["softening"],
                                                                      kw["vote_width"], oflags, every)
            cpu_s = t_tree + t_force + t_int
            how = "1 full step" if every == 1 else f"1 step, force on {ns} of {groups} vote groups scaled"
        out[name] = {"workload": desc, "theta": kw["theta"], "softening": kw["softening"],
                     "vote_width": kw["vote_width"], "gpu_ms_per_step": round(gpu_s * 1e3, 3),
                     "gpu_steps_timed": steps, "cpu_s_per_step": round(cpu_s, 4), "cpu_sample": how,
                     "gpu_over_cpu": round(cpu_s / gpu_s, 1)}
        log(f"[{name}] GPU {gpu_s * 1e3:.3f} ms/step, CPU {cpu_s:.3f} s/step ({how})")
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
        stages, inter, cell_inter, live = stage_profile(sim, max(3, min(args.steps, 10)))
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
        sim.close()

    main_r = results[mode]
    force_ms = main_r["stages"]["force"]  # the traversal (and, learned, the network and the body update)
    cells = main_r["cell_interactions"]
    flop = FLOP_PER_INTERACTION * (main_r["interactions"] - cells) + FLOP_PER_CELL_INTERACTION * cells
    achieved = flop / (force_ms * 1e-3) / 1e12
    # HBM bytes per launch of the dominant kernel, from the rocprofv3 counter runs
    # summarized in profiles/ (FETCH_SIZE and WRITE_SIZE, copy-kernel calibrated)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "force_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch", {}).get("k_coarse64" if mode == "learned" else "k_force64")

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
            "precision": "FP32 traversal (Kahan), FP64 summarize, FP64 positions",
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
            "bound": "mfma",
            "engine": "FP32 VALU (no matrix instructions); peak is MI355X FP32 157.3 TFLOP/s",
            "achieved": round(achieved, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": traffic,
            "flop_per_interaction": FLOP_PER_INTERACTION,
            "flop_per_cell_interaction": FLOP_PER_CELL_INTERACTION if mode == "learned" else None,
            "flop_note": ("20 per body-body interaction (monopole convention), 56 per cell interaction of "
                          "the coarse traversal (monopole + softened quadrupole); the network and the "
                          "features are not counted") if mode == "learned" else "20 per interaction",
            "launch_ms": force_ms,
        },
    }
    if mode == "learned":
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
            "training": {"steps": TRAIN["steps"], "pairs": tr["samples"], "iters": TRAIN["iters"],
                         "batch": TRAIN["batch"], "pair_generation_ms": round(tr["gen_ms"], 1),
                         "adam_ms": round(tr["train_ms"], 1), "loss_first": round(tr["loss_first"], 4),
                         "loss_last": round(tr["loss_last"], 4)},
        })
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {threads} threads")
        line["cpu_baseline"] = cpu_baseline(state, threads)
        if not args.no_config_table:
            line["cpu_baseline"]["configs"] = config_table(threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
