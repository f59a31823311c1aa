"""Benchmark of the gfx950 RDF / kNN spatial-statistics path.

One step = one pass of the hot path over the resident 20M-ion clustered
frustum tip: cell-list build, center/tile preparation, the query kernel for
all 9 ordered type pairs (RDF + MKNN k=1..100, Rmax 5 nm, bin 0.01 nm), and
the int64 histograms back on the host.  Inputs are in HBM before the timed
region starts.

Prints ONE JSON line.  `value` is reference-equivalent pair-distance
evaluations per second: the sum over considered centers of the neighbours
within Rmax (all types; exactly the size of the reference's per-center
neighbour list), divided by the step time.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pair-distance evals/s at 20M ions; served jobs/s + p99 latency vs plain loop"
VALU_PEAK_OPS = 78.6e12     # MI355X FP32 VALU element-ops/s (157.3 TFLOP/s counting FMA as 2)
OPS_PER_EVAL = 9            # 3 sub + 3 mul + 2 add + 1 compare, no FMA


def host_cpu():
    model, cores = "unknown", os.cpu_count() or 1
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                model = ln.split(":", 1)[1].strip()
    except Exception:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        pass
    return model, cores


def headline_job():
    from paraprobe_amd import FAM_MKNN, FAM_RDF, Job
    from paraprobe_amd.synth import ALL_PAIRS

    return Job(ALL_PAIRS, 0.0, 0.01, 5.0, FAM_RDF | FAM_MKNN, kvalues=list(range(1, 101)))


def cpu_baseline(tip, job, zwin, cores):
    """The reference's algorithm on the host cores, on the centers of a z
    window of the same tip (the other centers excluded through ion2surf)."""
    from oracle import cpu_baseline as cb

    s2 = tip.ion2surf_sqr.copy()
    z = tip.xyz[:, 2]
    s2[(z < zwin[0]) | (z >= zwin[1])] = 0.0
    t0 = time.perf_counter()
    hist, st = cb.spatstat(tip.xyz, tip.types, s2, job.tasks, job.rmin, job.rincr, job.rmax,
                           job.families, kvalues=job.kvalues, nthreads=cores)
    dt = time.perf_counter() - t0
    return hist, st, dt, s2


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=20_000_000)
    ap.add_argument("--cpu-window", type=float, default=1.0,
                    help="thickness (nm) of the z window of centers timed on the CPU")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib", action="store_true",
                    help="also run a 1 GiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (every box has one MI355X); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from paraprobe_amd import SpatStat, synth

    t = time.perf_counter()
    tip = synth.frustum_20m(n=args.n)
    print(f"# generated {tip.n} ions in {time.perf_counter() - t:.1f} s", file=sys.stderr,
          flush=True)
    job = headline_job()

    # The library calls are synchronous; the device-wide synchronisation
    # around the timed region goes through the same HIP runtime the library
    # uses (torch is not imported: its wheel carries a second HIP runtime).
    hip = ctypes.CDLL("libamdhip64.so.7")

    def sync():
        if hip.hipDeviceSynchronize() != 0:
            raise SystemExit("bench: hipDeviceSynchronize failed")

    with SpatStat(0) as ctx:
        cloud = ctx.upload(tip.xyz, tip.types, tip.ion2surf_sqr)
        for i in range(args.warmup):
            r = ctx.run(cloud, job)
            print(f"# warmup {i}: {r.ms_total:.1f} ms (query {r.ms_query:.1f} ms)",
                  file=sys.stderr, flush=True)
        sync()
        t0 = time.perf_counter()
        res = []
        for i in range(args.steps):
            res.append(ctx.run(cloud, job))
            print(f"# step {i}: {res[-1].ms_total:.1f} ms (query {res[-1].ms_query:.1f} ms)",
                  file=sys.stderr, flush=True)
        sync()
        dt = time.perf_counter() - t0
        last = res[-1]
        for r in res[1:]:
            for k, v in r.arrays().items():
                if not np.array_equal(v, getattr(res[0], k)):
                    raise SystemExit("bench: results differ between steps")
        # GPU result of the CPU sample window, for the parity check
        zwin = (125.0 - args.cpu_window / 2, 125.0 + args.cpu_window / 2)
        s2w = tip.ion2surf_sqr.copy()
        zz = tip.xyz[:, 2]
        s2w[(zz < zwin[0]) | (zz >= zwin[1])] = 0.0
        cw = ctx.upload(tip.xyz, tip.types, s2w)
        gwin = ctx.run(cw, job)
        cw.free()
        cloud.free()
        if args.calib:
            # 1 GiB device-to-device copy: its FETCH_SIZE / WRITE_SIZE per
            # dispatch scale the counters of the query kernel to bytes
            a, b = ctypes.c_void_p(), ctypes.c_void_p()
            nb = 1 << 30
            if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nb)) == 0 and \
                    hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nb)) == 0:
                hip.hipMemset(a, 1, ctypes.c_size_t(nb))
                hip.hipMemcpy(b, a, ctypes.c_size_t(nb), 3)  # hipMemcpyDeviceToDevice
                sync()
                hip.hipFree(a)
                hip.hipFree(b)

    ms_step = dt * 1e3 / args.steps
    inball = last.pairs_inball
    value = inball / (ms_step / 1e3)
    q_ms = float(np.mean([r.ms_query for r in res]))
    cand = last.pairs_candidate
    achieved = cand * OPS_PER_EVAL / (q_ms / 1e3) / 1e12
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-distance evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "frustum-20M-clustered: 20,000,000 ions, r 40->20 nm, h 250 nm, "
                        "types A/B/C 0.95/0.04/0.01, 50,000 clusters r~U(1,2) nm B=40%; "
                        "9 type pairs RDF + MKNN k=1..100, Rmax 5 nm, bin 0.01 nm, int64",
            "n_ions": tip.n,
            "tasks": 9,
            "rmax_nm": 5.0,
            "bin_nm": 0.01,
            "kmax": 100,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": VALU_PEAK_OPS / 1e12,
            "unit": "Tops/s (FP32 VALU element-ops, 9 per candidate pair)",
            "frac": achieved / (VALU_PEAK_OPS / 1e12),
            "traffic": None,
            "kernel": last.kernel,
            "kernel_ms": q_ms,
            "candidate_evals": cand,
        },
        "centers_considered": last.centers_considered,
        "pairs_inball": inball,
        "candidate_evals_per_s": cand / (q_ms / 1e3),
        "ms_index": float(np.mean([r.ms_index for r in res])),
        "ms_prep": float(np.mean([r.ms_prep for r in res])),
        "ms_query": q_ms,
    }
    if not args.no_cpu:
        model, cores = host_cpu()
        hist, st, cdt, _ = cpu_baseline(tip, job, zwin, cores)
        parity = all(np.array_equal(getattr(gwin, k), v) for k, v in hist.items())
        line["cpu_baseline"] = {
            "value": st["inball"] / st["compute_s"],
            "unit": "pair-distance evals/s",
            "cores": cores,
            "kind": "port",
            "sample": f"same tip, centers with z in [{zwin[0]}, {zwin[1]}) nm "
                      f"({st['considered']} centers, {st['inball']} in-ball pairs); compute "
                      f"{st['compute_s']:.1f} s timed (the DescrStatsThreadparallelCompute "
                      f"span), z-slab + KD-tree build {st['index_s']:.1f} s not timed; {model}",
            "parity_with_gpu": parity,
        }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
