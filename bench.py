"""Throughput of the MI355X P3M / PM force step (particle-updates/s).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

One step = one iteration of the reference run loop (source/pmMethod.cpp:
79-122, source/p3mMethod.cpp:93-154) without I/O: drift, force (deposit,
FFT x G, field, interpolation, short range for P3M), kick.  Inputs are
resident in HBM before timing starts; the timed region is exactly K steps
between two stream synchronisations.  Prints one JSON line on stdout.

The line also carries
  roofline      the dominant kernel's achieved rate vs the MI355X peak,
                measured with HIP events on the library's stream,
  cpu_baseline  the CPU restatement of the reference (oracle/, OpenMP) timed
                on a bounded sample of the same workload on this host,
  force_error   RMS and mean relative error of the GPU force against an FP64
                direct sum on a seeded subsample,
  learned_vs_plain  the recorded learned-vs-plain PM comparison (profiles/).
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
GC = 1.0 / (4.0 * np.pi)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load():
    import particlesimulation_amd as P
    if not os.path.exists(P.LIB_PATH):
        from particlesimulation_amd import build as B
        B.build_all()
    P.load_library()
    return P


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


def force_error(P, w, pm, pos, m, n_targets=512, seed=7):
    """e_rms = sqrt(sum|F - F_ds|^2 / sum|F_ds|^2) and the reference's mean relative
    error (script/pm_accuracy.py:44-48) against a direct sum over all particles, on a
    seeded subsample of targets.  Isolated workloads: FP64 Newtonian direct sum (numpy).
    Periodic workloads: the GPU's periodic direct sum (nearest image + Ewald
    correction, FP32 with compensated sums; checked against the FP64 oracle)."""
    pm.pm_step()
    if w.p3m:
        pm.p3m_short_range()
    _, _, acc = pm.get_particles()
    rng = np.random.default_rng(seed)
    t = rng.choice(len(m), size=min(n_targets, len(m)), replace=False)
    if w.periodic:
        ref = P.direct_sum(pos, m, pos[t], box=w.grid).astype(np.float64)
        kind = "gpu fp32 periodic (Ewald) direct sum"
    else:
        x = pos.astype(np.float64)
        mm = m.astype(np.float64)
        ref = np.zeros((len(t), 3))
        for k, i in enumerate(t):
            d = x[i] - x
            r2 = (d * d).sum(1)
            r2[i] = np.inf
            ref[k] = -GC * ((mm / (r2 * np.sqrt(r2)))[:, None] * d).sum(0)
        kind = "fp64 direct sum"
    diff = acc[t].astype(np.float64) - ref
    e_rms = float(np.sqrt((diff ** 2).sum() / (ref ** 2).sum()))
    mean_rel = float(np.mean(np.linalg.norm(diff, axis=1) / (np.linalg.norm(ref, axis=1) + 1e-30)))
    return {"e_rms": e_rms, "mean_rel": mean_rel, "targets": int(len(t)), "reference": kind}


def measured_traffic(workload):
    """HBM bytes per launch of the short-range kernel on this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE measurement (profiles/), or None."""
    path = os.path.join(ROOT, "profiles", f"k_sr_{workload}_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    return d["traffic_bytes_per_launch"], f"profiles/k_sr_{workload}_traffic.json ({d['kernel']})"


def measured_learned():
    """The learned-vs-plain comparison (error against the periodic direct sum
    and throughput of plain PM, learned PM and P3M on the 4M-particle clumps
    workload), from the committed measurement in profiles/, or None.  Training
    takes tens of seconds, so the bench reports the recorded run."""
    path = os.path.join(ROOT, "profiles", "learned_pm_clumps_4M_256.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    large = d["large"]
    return {"source": f"profiles/learned_pm_clumps_4M_256.json ({d['command']})", "workload": large["workload"],
            "targets": large["targets"], "reference": large["reference"],
            **{k: large[k] for k in ("plain_pm", "learned_pm", "plain_p3m") if k in large}}


def host_threads():
    """Threads for the CPU baseline: this process's CPU share (the affinity
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


def cpu_baseline(w, pos, vel, m, pairs_total, budget_s=15.0, nsamples=3):
    """The oracle (CPU restatement of the reference path) on this host: the
    full PM step, plus the short-range pass on `nsamples` disjoint random sets
    of chaining cells, each scaled to the whole pass by its pair count; the
    spread of the per-sample estimates is the sampling error."""
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
    t_sr, sample, sr_err = 0.0, "full PM step", None
    if w.p3m:
        H = w.H
        box_orig = np.array([w.box] * 3, np.float32)
        dims = O.p3m_mesh_dims(box_orig, w.rc_cells * H)
        ncell = dims[0] * dims[1] * dims[2]
        perm = np.random.default_rng(11).permutation(ncell)

        def run(cells):
            mask = np.zeros(ncell, np.uint8)
            mask[cells] = 1
            t0 = time.perf_counter()
            _, p = O.p3m_short_range(pos, m, box_orig, w.rc_cells * H, H, w.rc_cells, w.a_cells, w.softening,
                                     w.cloud_shape, use_table=w.sr_table, ysort=True, cell_mask=mask)
            return time.perf_counter() - t0, p

        # pilot: grow a sample until it takes a measurable time, to size the others
        k = max(1, ncell // 1000)
        while True:
            dt, p = run(perm[:k])
            if (dt > 0.5 and p > 0) or k >= ncell // (nsamples + 1):
                break
            k = min(ncell // (nsamples + 1), k * 4)
        if p == 0:
            raise RuntimeError("short-range sample found no pairs")
        per_sample = budget_s / nsamples
        k = int(min(ncell // (nsamples + 1), max(1, k * per_sample / max(dt, 1e-3))))
        est, used_pairs, used_cells, used_s = [], 0, 0, 0.0
        for j in range(nsamples):
            cells = perm[ncell - (j + 1) * k: ncell - j * k]  # disjoint from each other and from the pilot
            dt, p = run(cells)
            if p == 0:
                continue
            est.append(dt * pairs_total / p)
            used_pairs, used_cells, used_s = used_pairs + p, used_cells + len(cells), used_s + dt
        if not est:
            raise RuntimeError("short-range samples found no pairs")
        est = np.array(est)
        t_sr = float(est.mean())
        sr_err = float(est.std(ddof=1) / np.sqrt(len(est)) / t_sr) if len(est) > 1 else None
        sample = (f"full PM step ({t_pm:.2f}s) + short range on {len(est)} disjoint random samples of {k} of "
                  f"{ncell} chaining cells ({used_pairs:.3e} of {pairs_total:.3e} ordered pairs in all, "
                  f"{used_s:.1f}s), each scaled to the whole pass by its pair count")
    t_step = t_pm + t_sr
    return {"value": len(m) / t_step, "unit": "particle-updates/s", "cores": threads, "kind": "port",
            "sample": sample, "s_per_step": t_step, "pm_s": t_pm, "sr_s": t_sr,
            "sr_sampling_rel_err": sr_err, "green_init_s": t_green, **counts}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="p3m_disk_1M_128")
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-error", action="store_true")
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
    err = None if args.no_error else force_error(P, w, pm, pos, m)
    if err:
        log(f"[bench] force error vs direct sum: {err}")
    pm.set_particles(pos, vel, m)
    pm.run(args.warmup, initial=True, p3m=w.p3m)
    pairs0, tested0 = pm.p3m_count_pairs(tested=True) if w.p3m else (0, 0)
    pm.profile(True)
    pm.profile_reset()
    pm.synchronize()
    t0 = time.perf_counter()
    # the run ends with a 4-byte read of the escape word (the first iteration
    # whose drift left the box) and a stream synchronisation
    done = pm.run(args.steps, initial=False, p3m=w.p3m, sync=True)
    t1 = time.perf_counter()
    stages = pm.profile_read()
    pm.profile(False)
    escaped = pm.escaped()
    if escaped or done != args.steps:
        # a particle outside the box stops every later kernel of the run, so
        # the timed steps would not all have been computed
        log(f"[bench] invalid run: {done} of {args.steps} iterations done, escaped={escaped}")
        sys.exit(3)
    pairs1, tested1 = pm.p3m_count_pairs(tested=True) if w.p3m else (0, 0)
    pairs = 0.5 * (pairs0 + pairs1)
    tested = 0.5 * (tested0 + tested1)
    elapsed = t1 - t0
    value = w.n * args.steps / elapsed
    log(f"[bench] {args.steps} steps in {elapsed:.4f}s -> {value:.4e} upd/s; iterations={done} escaped={escaped}")

    per_step = {k: v[0] / max(1, args.steps) for k, v in stages.items()}
    if w.p3m:
        ms, cnt = stages["sr_pairs"]
        avg_s = ms / cnt * 1e-3
        achieved = SR_FLOP_PER_PAIR * pairs / avg_s / 1e12
        roofline = {"bound": "valu", "kernel": "k_sr (short-range pairs)", "achieved": achieved,
                    "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                    "traffic": None, "avg_launch_ms": avg_s * 1e3, "ordered_pairs": pairs,
                    "tested_pairs": tested, "flop_per_pair": SR_FLOP_PER_PAIR,
                    "sr_table_mode": w.sr_table, "sr_cell_division": pm.get_tuning()[3],
                    "sr_cells": dict(zip(("div_x", "div_y", "div_z", "cells"), pm.p3m_cell_geometry()))}
        roofline["traffic"], roofline["traffic_source"] = measured_traffic(w.name)
    else:
        roofline = None
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

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(w, pos, vel, m, pairs if w.p3m else 0)
            log(f"[bench] cpu baseline: {cpu}")
        except Exception as e:  # the baseline must not hide the GPU measurement
            log(f"[bench] cpu baseline failed: {e!r}")
    out = {
        "metric": METRIC, "value": value, "unit": "particle-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": w.name, "n": w.n, "grid": list(w.grid), "box": w.box,
                   "interpolation": ["NGP", "CIC", "TSC"][w.interpolation],
                   "finite_diff": ["TWO_POINT", "FOUR_POINT"][w.finite_diff],
                   "greens_function": ["DISCRETE_LAPLACIAN", "S1_OPTIMAL", "S2_OPTIMAL", "POOR_MAN"][
                       w.greens_function],
                   "a_cells": w.a_cells, "cutoff_cells": w.rc_cells, "dt": w.dt, "seed": args.seed,
                   "p3m": w.p3m},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "force_error": err,
        "learned_vs_plain": measured_learned(),
        "stage_ms_per_step": per_step,
        "hbm_stages": hbm,
        "gpu_vs_cpu": (value / cpu["value"]) if cpu else None,
        "iterations": done,
        "escaped": escaped,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
