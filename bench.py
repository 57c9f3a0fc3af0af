"""Throughput of the MI355X P3M / PM force step (particle-updates/s).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

One step = one iteration of the reference run loop (source/pmMethod.cpp:
79-122, source/p3mMethod.cpp:93-154) without I/O: drift, force (deposit,
FFT x G, field, interpolation, short range for P3M), kick.  Inputs are
resident in HBM before timing starts; the timed region is exactly K steps
between two stream synchronisations.  Prints one JSON line on stdout.

The line carries, for the headline workload (the 1M-particle P3M disk on
128^3 with a 3-cell cutoff):
  roofline      the dominant kernel's achieved rate vs the MI355X peak,
                measured with HIP events on the library's stream,
  cpu_baseline  the CPU restatement of the reference (oracle/, OpenMP) timed
                on this host: the whole PM step and a quarter of the
                short-range pass (every 4th particle of each cell), scaled
                to the whole pass by its pair count,
  force_error   RMS and mean relative error of the GPU force against an FP64
                direct sum (host CPU) on a seeded subsample of targets,
and two more measured blocks:
  at_error_target   the same disk and grid at the smallest cutoff whose
                force error is below 1e-3 (textbook TSC, 4-point gradient,
                a = rc = 3.5 cells): throughput, roofline, error, CPU baseline;
                at_error_target_reference_tsc the same with the reference's
                floor-anchored TSC and 2-point gradient (a = rc = 5 cells),
  learned_vs_plain  the learned influence-function correction trained on the
                GPU within this run, then plain PM, learned PM and P3M on the
                4M-particle clustered workload on 256^3: error against the
                periodic direct sum and throughput.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-updates/s, P3M 1M on 128^3 at fixed force err; learned vs plain"
PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 vector peak
PEAK_HBM_GBS = 8000.0
SR_FLOP_PER_PAIR = 20      # per ordered pair within the cutoff
SR_SAMPLE_STRIDE = 4       # CPU baseline: every 4th particle of each cell as the home particle
# the stride-4 estimate of the 1M disk's pass against the timed whole pass
SR_SAMPLE_ERROR = ("stride-4 estimates were 0.94x the timed whole pass on 8 threads (206 s vs 220 s) and "
                   "0.94-0.97x on 16 (61.6-64.1 s vs 65.9 s): the whole pass loses more to its densest cells' "
                   "tail, so the estimate errs toward a faster CPU")
SCHEMES = ["NGP", "CIC", "TSC", "TSC_CENTRED"]
FDS = ["TWO_POINT", "FOUR_POINT"]
GREENS = ["DISCRETE_LAPLACIAN", "S1_OPTIMAL", "S2_OPTIMAL", "POOR_MAN"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load():
    import particlesimulation_amd as P
    if not os.path.exists(P.LIB_PATH):
        from particlesimulation_amd import build as B
        B.build_all()
    P.load_library()
    return P


def host_threads():
    """Threads for the host-side work: this process's CPU share (the affinity
    mask, capped by OMP_NUM_THREADS when that is set), and the counts behind it."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = nproc
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = affinity
    if omp and omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    return threads, {"nproc": nproc, "affinity": affinity, "omp_num_threads": omp}


def workload_config(w, seed):
    return {"workload": w.name, "n": w.n, "grid": list(w.grid), "box": w.box,
            "interpolation": SCHEMES[w.interpolation], "finite_diff": FDS[w.finite_diff],
            "greens_function": GREENS[w.greens_function], "a_cells": w.a_cells, "cutoff_cells": w.rc_cells,
            "dt": w.dt, "seed": seed, "p3m": w.p3m, "periodic": w.periodic}


def make_context(P, w, pos, vel, m):
    box = (0, 0, 0) if w.periodic else w.grid
    pm = P.PM(w.grid, interpolation=w.interpolation, finite_diff=w.finite_diff,
              greens_function=w.greens_function, particle_diameter=w.a_cells, box=box)
    pm.set_particles(pos, vel, m)
    pm.init_greens()
    if w.p3m:
        pm.p3m_configure(cutoff=w.rc_cells, particle_diameter=w.a_cells, softening=w.softening,
                         cloud_shape=w.cloud_shape, use_table=w.sr_table, box=w.grid)
    return pm


class DirectSum:
    """Direct-sum accelerations of a seeded target subsample, computed once
    per particle set and shared by the workloads that use it.  Isolated
    workloads: FP64 on the host CPU over every particle.  Periodic ones: the
    GPU's periodic sum (nearest image + Ewald correction, FP32 with
    compensated sums) on every target, cross-checked by the FP64 host sum on
    the first `fp64_check` targets."""

    def __init__(self, P, pos, m, n_targets, seed=7, box=None, fp64_check=256):
        threads, _ = host_threads()
        rng = np.random.default_rng(seed)
        self.targets = np.sort(rng.choice(len(m), size=min(n_targets, len(m)), replace=False))
        t0 = time.perf_counter()
        if box is None:
            self.ref = P.direct_sum_host(pos, m, pos[self.targets], threads=threads)
            self.kind, self.check = "fp64 direct sum (host)", None
        else:
            self.ref = P.direct_sum(pos, m, pos[self.targets], box=box).astype(np.float64)
            k = self.targets[:fp64_check]
            ref64 = P.direct_sum_host(pos, m, pos[k], box=box, threads=threads)
            self.check = {"targets": int(len(k)),
                          "rms_rel_fp32_vs_fp64": float(np.linalg.norm(self.ref[:len(k)] - ref64)
                                                        / np.linalg.norm(ref64))}
            self.kind = "gpu fp32 periodic (Ewald) direct sum, checked against the fp64 host sum"
        self.seconds = time.perf_counter() - t0
        self.threads = threads

    def error(self, acc):
        """e_rms = sqrt(sum|F - F_ds|^2 / sum|F_ds|^2) and the reference's mean
        relative error (script/pm_accuracy.py:44-48)."""
        diff = np.asarray(acc, np.float64)[self.targets] - self.ref
        e_rms = float(np.sqrt((diff ** 2).sum() / (self.ref ** 2).sum()))
        mean_rel = float(np.mean(np.linalg.norm(diff, axis=1) / (np.linalg.norm(self.ref, axis=1) + 1e-30)))
        out = {"e_rms": e_rms, "mean_rel": mean_rel, "targets": int(len(self.targets)), "reference": self.kind,
               "direct_sum_s": self.seconds, "host_threads": self.threads}
        if self.check:
            out["fp64_check"] = self.check
        return out


def force_of(pm, w):
    pm.pm_step()
    if w.p3m:
        pm.p3m_short_range()
    return pm.get_particles()[2]


def measured_traffic(workload):
    """HBM bytes per launch of the short-range kernel on this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE measurement (profiles/), or None."""
    path = os.path.join(ROOT, "profiles", f"k_sr_{workload}_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    return d["traffic_bytes_per_launch"], f"profiles/k_sr_{workload}_traffic.json ({d['kernel']})"


def measure(pm, w, steps, warmup, pos, vel, m):
    """Warm-up, then exactly `steps` run-loop iterations timed between two
    stream synchronisations.  In a P3M run the short-range pair kernel alone
    is bracketed by HIP events inside the timed region (its live launch time
    for the roofline; the run loop then enqueues every iteration directly);
    a PM run is timed without events, replaying the run loop's graph.  Then
    the same iterations again with every kernel bracketed by events (the
    per-stage times), and for P3M once more without events (the
    graph-replayed rate); each of the three starts from the initial
    particles and the same warm-up.  Exits when the run escaped."""

    def start():
        pm.set_particles(pos, vel, m)
        pm.run(warmup, initial=True, p3m=w.p3m)

    start()
    pairs0, tested0 = pm.p3m_count_pairs(tested=True) if w.p3m else (0, 0)

    def timed(events, fresh=True):
        # events: None, "short_range" or "all"
        if not fresh:
            start()
        if events is None:
            # four untimed iterations capture the graph that the timed run
            # replays (switching events on or off retires the graphs; an even
            # count ends on the buffer the capture did not start from, which
            # is the one the timed run's graph starts from)
            pm.run(4, initial=False, p3m=w.p3m, sync=True)
        else:
            pm.profile(True, short_range_only=events == "short_range")
            pm.profile_reset()
        pm.synchronize()
        t0 = time.perf_counter()
        # the run ends with a 4-byte read of the escape word (the first
        # iteration whose drift left the box) and a stream synchronisation
        done = pm.run(steps, initial=False, p3m=w.p3m, sync=True)
        t1 = time.perf_counter()
        stages = {}
        if events is not None:
            stages = pm.profile_read()
            pm.profile(False)
        return t1 - t0, done, stages

    elapsed, done, sr_stage = timed("short_range" if w.p3m else None)
    escaped = pm.escaped()
    if escaped or done != steps:
        # a particle outside the box stops every later kernel of the run, so
        # the timed steps would not all have been computed
        log(f"[bench] invalid run of {w.name}: {done} of {steps} iterations done, escaped={escaped}")
        sys.exit(3)
    pairs1, tested1 = pm.p3m_count_pairs(tested=True) if w.p3m else (0, 0)
    t_st, done_st, stages = timed("all", fresh=False)
    with_stage_events = None
    if not pm.escaped() and done_st == steps:
        with_stage_events = {"value": w.n * steps / t_st, "ms_per_step": t_st / steps * 1e3}
    without_events = None
    if w.p3m:
        t_ne, done_ne, _ = timed(None, fresh=False)
        if not pm.escaped() and done_ne == steps:
            without_events = {"value": w.n * steps / t_ne, "ms_per_step": t_ne / steps * 1e3}
    log(f"[bench] {w.name} with every stage bracketed by events: {with_stage_events}; "
        f"without events: {without_events}")
    pairs = 0.5 * (pairs0 + pairs1)
    tested = 0.5 * (tested0 + tested1)
    value = w.n * steps / elapsed
    log(f"[bench] {w.name}: {steps} steps in {elapsed:.4f}s -> {value:.4e} upd/s; iterations={done} "
        f"escaped={escaped}")
    per_step = {k: v[0] / max(1, steps) for k, v in stages.items()}
    roofline = None
    if w.p3m:
        ms, cnt = sr_stage["sr_pairs"]
        avg_s = ms / cnt * 1e-3
        achieved = SR_FLOP_PER_PAIR * pairs / avg_s / 1e12
        roofline = {"bound": "valu", "kernel": "k_sr (short-range pairs)", "achieved": achieved,
                    "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                    "traffic": None, "avg_launch_ms": avg_s * 1e3, "ordered_pairs": pairs,
                    "tested_pairs": tested, "flop_per_pair": SR_FLOP_PER_PAIR,
                    "sr_table_mode": w.sr_table, "sr_cell_division": pm.get_tuning()[3],
                    "sr_cells": dict(zip(("div_x", "div_y", "div_z", "cells"), pm.p3m_cell_geometry()))}
        roofline["traffic"], roofline["traffic_source"] = measured_traffic(w.name)
    # algorithmic-byte rates of the bandwidth-bound stages (per launch); the
    # interpolation reads x (16 B) and writes a (16 B) per particle, and in a
    # PM step also reads and writes v and writes x (48 B more)
    nx, ny, nz = w.grid
    M, K, n = nx * ny * nz, (nx // 2 + 1) * ny * nz, w.n
    alg = {"deposit": 16 * n + 4 * M, "fft_x_r2c": 4 * M + 8 * K, "fft_y_fwd": 16 * K,
           "fft_z_green": 20 * K, "fft_y_inv": 16 * K, "fft_x_c2r": 8 * K + 4 * M,
           "interp": 4 * M + (32 if w.p3m else 80) * n}
    hbm = {}
    for k, b in alg.items():
        if k in stages and stages[k][1]:
            s = stages[k][0] / stages[k][1] * 1e-3
            hbm[k] = {"ms": s * 1e3, "GB/s": b / s / 1e9, "frac": b / s / 1e9 / PEAK_HBM_GBS}
    if not w.p3m and hbm:
        # a PM step's longest kernel, from the run with every stage bracketed
        # (the timed PM run carries no events)
        k = max(hbm, key=lambda s: hbm[s]["ms"])
        roofline = {"bound": "hbm", "kernel": k, "achieved": hbm[k]["GB/s"], "peak": PEAK_HBM_GBS,
                    "unit": "GB/s", "frac": hbm[k]["frac"], "traffic": None, "avg_launch_ms": hbm[k]["ms"],
                    "algorithmic_bytes": alg[k], "measured_in": "the run with every stage bracketed by events"}
    return {"value": value, "ms_per_step": elapsed / steps * 1e3, "steps": steps, "iterations": done,
            "escaped": escaped, "roofline": roofline, "stage_ms_per_step": per_step, "hbm_stages": hbm,
            "pairs": pairs, "with_stage_events": with_stage_events, "without_events": without_events}


def cpu_baseline(w, pos, m, sr_pairs=None, sr_scale_from=None):
    """The oracle (CPU restatement of the reference path, OpenMP) on this
    host, on a bounded sample of one step: the whole PM force, and the
    short-range pass (the reference's chaining mesh with y-sorted cells and
    the accurate table) with every SR_SAMPLE_STRIDE-th particle of each
    cell's list as the home particle.  The pass's time is its part that does
    not grow with the pairs (the same pass with one home particle per cell:
    cell lists, y-sort, the slot sums) plus the sample's time per ordered
    pair times the whole step's pair count (sr_pairs, counted on the device
    on the same particles; the sample's count times the stride without it).
    Sampling within every cell keeps the pass's balance over the threads:
    a sample of whole cells is bounded by the densest cell's single thread.
    With sr_scale_from = (seconds, pairs, fixed seconds, this workload's
    pairs) of that estimate on the same particles, the short-range time is
    its time per pair times this workload's pair count instead."""
    from oracle import oracle as O
    threads, counts = host_threads()
    O.set_num_threads(threads)
    grid = w.grid
    t0 = time.perf_counter()
    G = O.green_grid(w.greens_function, w.interpolation, w.finite_diff, w.a_cells, grid)
    t_green = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.pm_accel(pos, m, grid, w.interpolation, w.finite_diff, G, wrap=True)
    t_pm = time.perf_counter() - t0
    t_sr, t_fixed, pairs, extra = 0.0, 0.0, 0, {}
    sample = f"one whole PM step ({t_pm:.2f}s)"
    if w.p3m and sr_scale_from is None:
        H = w.H
        box_orig = np.array([w.box] * 3, np.float32)
        args = (pos, m, box_orig, w.rc_cells * H, H, w.rc_cells, w.a_cells, w.softening, w.cloud_shape)
        t0 = time.perf_counter()
        O.p3m_short_range(*args, use_table=w.sr_table, ysort=True, sample_stride=1 << 30)
        t_fixed = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, p_s = O.p3m_short_range(*args, use_table=w.sr_table, ysort=True, sample_stride=SR_SAMPLE_STRIDE)
        t_s = time.perf_counter() - t0
        pairs = sr_pairs if sr_pairs else p_s * SR_SAMPLE_STRIDE
        t_sr = t_fixed + (t_s - t_fixed) * pairs / max(p_s, 1)
        sample = (f"the whole PM step ({t_pm:.2f}s) + the short-range pass on every {SR_SAMPLE_STRIDE}th particle "
                  f"of each chaining cell ({p_s:.3e} ordered pairs, {t_s:.1f}s) scaled to the step's {pairs:.3e} "
                  f"pairs ({t_sr:.1f}s; {t_fixed:.2f}s of it the part with one home particle per cell)")
        extra = {"sr_sample_stride": SR_SAMPLE_STRIDE, "sr_sample_pairs": p_s, "sr_sample_s": t_s,
                 "sr_fixed_s": t_fixed, "sample_error": SR_SAMPLE_ERROR}
    elif w.p3m:
        t_ref, pairs_ref, t_fixed, pairs = sr_scale_from
        t_sr = t_fixed + (t_ref - t_fixed) * pairs / pairs_ref
        sample = (f"PM step measured ({t_pm:.2f}s) + short-range pass estimated as {pairs:.3e} ordered pairs at "
                  f"the time per pair of the 3-cell pass on the same particles ({t_ref:.1f}s for "
                  f"{pairs_ref:.3e} pairs, itself from its sample)")
    t_step = t_pm + t_sr
    return {"value": len(m) / t_step, "unit": "particle-updates/s", "cores": threads, "kind": "port",
            "sample": sample, "s_per_step": t_step, "pm_s": t_pm, "sr_s": t_sr, "sr_pairs": pairs,
            "sr_fixed_s": t_fixed, "green_init_s": t_green, **extra, **counts}


def learned_live(P, seed, budget_steps=150, snapshots=24, n_targets=65536):
    """Train the correction on the GPU in this run and compare plain PM,
    learned PM and P3M on the 4M-particle clustered workload on 256^3."""
    from particlesimulation_amd import learned as Lm
    t0 = time.perf_counter()
    trainers = []
    for gi, g in enumerate((64, 128)):
        seeds = range(1000 * (gi + 1), 1000 * (gi + 1) + snapshots)
        trainers.append(Lm.build_trainer((g, g, g), seeds, 65536))
    t_data = time.perf_counter() - t0
    params = P.mlp_init(seed)
    t0 = time.perf_counter()
    params, hist = P.train(trainers, params, budget_steps, 2e-3)
    t_train = time.perf_counter() - t0
    for t in trainers:
        t.close()
    large = Lm.evaluate_large(params, n_targets=n_targets)
    pm_, lp, p3 = large["plain_pm"], large["learned_pm"], large["plain_p3m"]
    out = {"training": {"grids": [64, 128], "snapshots_per_grid": snapshots, "particles": 65536,
                        "adam_steps": budget_steps, "loss_first": float(hist[0]), "loss_last": float(hist[-1]),
                        "data_s": t_data, "train_s": t_train},
           **large,
           "learned_error_over_plain": lp["e_rms"] / pm_["e_rms"],
           "learned_throughput_over_p3m": lp["updates_per_s"] / p3["updates_per_s"],
           "goal_met": bool(lp["e_rms"] <= p3["e_rms"] and lp["updates_per_s"] >= 2 * p3["updates_per_s"])}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="p3m_disk_1M_128")
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--targets", type=int, default=65536, help="direct-sum targets of the force error")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--no-at-error", action="store_true", help="skip the block at the 1e-3 error target")
    ap.add_argument("--no-learned", action="store_true", help="skip the live learned-vs-plain block")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (the reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings) and exit")
    ap.add_argument("--sr-div", type=int, default=None,
                    help="short-range cell division (cells of side >= cutoff/div); default: the library's")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path shards by "
              "slab/pencil decomposition, which this build does not run)", file=sys.stderr)
        sys.exit(2)

    P = load()
    from particlesimulation_amd import ics
    w = ics.WORKLOADS[args.workload]
    log(f"[bench] workload {w.name}: n={w.n} grid={w.grid} p3m={w.p3m}")
    t0 = time.perf_counter()
    pos, vel, m = w.code_units(seed=args.seed)
    log(f"[bench] initial conditions {time.perf_counter() - t0:.1f}s")

    pm = make_context(P, w, pos, vel, m)
    if args.sr_div is not None:
        k = list(pm.get_tuning())
        k[3] = args.sr_div
        pm.set_tuning(*k)
    ds = None if args.no_error else DirectSum(P, pos, m, args.targets, box=w.grid if w.periodic else None)
    err = ds.error(force_of(pm, w)) if ds else None
    if err:
        log(f"[bench] force error vs direct sum: {err}")
    # ordered pairs within the cutoff in the initial state (the CPU baseline's sample is scaled to it)
    sr_pairs = pm.p3m_count_pairs() if w.p3m else None
    res = measure(pm, w, args.steps, args.warmup, pos, vel, m)
    pm.close()

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(w, pos, m, sr_pairs=sr_pairs)
            log(f"[bench] cpu baseline: {cpu}")
        except Exception as e:  # the baseline must not hide the GPU measurement
            log(f"[bench] cpu baseline failed: {e!r}")

    def at_error(we):
        pe = make_context(P, we, pos, vel, m)
        e_err = ds.error(force_of(pe, we)) if ds else None
        log(f"[bench] {we.name} force error: {e_err}")
        r = measure(pe, we, args.steps, args.warmup, pos, vel, m)
        pe.close()
        c = None
        if cpu and cpu["sr_pairs"]:
            c = cpu_baseline(we, pos, m, sr_scale_from=(cpu["sr_s"], cpu["sr_pairs"], cpu["sr_fixed_s"],
                                                        r["pairs"]))
        return {"config": workload_config(we, args.seed), "value": r["value"], "unit": "particle-updates/s",
                "ms_per_step": r["ms_per_step"], "force_error": e_err,
                "error_target_met": bool(e_err and e_err["e_rms"] <= 1e-3), "roofline": r["roofline"],
                "stage_ms_per_step": r["stage_ms_per_step"], "without_events": r["without_events"],
                "cpu_baseline": c,
                "gpu_vs_cpu": (r["value"] / c["value"]) if c else None, "iterations": r["iterations"],
                "escaped": r["escaped"]}

    at_err = at_err_ref = None
    if not args.no_at_error and w.name == ics.C2_P3M_DISK.name:
        at_err = at_error(ics.C2_P3M_DISK_ERR)
        at_err_ref = at_error(ics.C2_P3M_DISK_ERR_REF)

    learned = None
    if not args.no_learned and w.name == ics.C2_P3M_DISK.name:
        try:
            learned = learned_live(P, args.seed)
            log(f"[bench] learned vs plain: {learned}")
        except Exception as e:
            log(f"[bench] learned block failed: {e!r}")

    out = {
        "metric": METRIC, "value": res["value"], "unit": "particle-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": workload_config(w, args.seed),
        "roofline": res["roofline"],
        "cpu_baseline": cpu,
        "force_error": err,
        "at_error_target": at_err,
        "at_error_target_reference_tsc": at_err_ref,
        "learned_vs_plain": learned,
        "stage_ms_per_step": res["stage_ms_per_step"],
        "hbm_stages": res["hbm_stages"],
        "with_stage_events": res["with_stage_events"],
        "without_events": res["without_events"],
        "gpu_vs_cpu": (res["value"] / cpu["value"]) if cpu else None,
        "iterations": res["iterations"],
        "escaped": res["escaped"],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
