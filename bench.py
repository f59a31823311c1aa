"""Benchmark of the gfx950 RDF / kNN spatial-statistics path.

One step = one pass of the hot path over the resident 20M-ion clustered
frustum tip: center/tile preparation, the query kernel for all 9 ordered type
pairs (RDF + MKNN k=1..100, Rmax 5 nm, bin 0.01 nm), and the int64 histograms
back on the host.  Inputs are in HBM before the timed region starts, and so
is the cloud's cell list: the first (warmup) job builds it and every step
reuses it, as the reference's KD-trees are built once by
init_spatialindexing before compute_generic_spatstat3.  The index build alone
(the cell list and the candidate lists of the symmetric pair counting) is
reported as ms_index_build (a fresh copy of the cloud, timed).

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
FP32_PEAK_FLOPS = 157.3e12  # MI355X published FP32 vector peak (an FMA counts as 2 FLOP)
VALU_PEAK_OPS = 78.6e12     # the same peak as FP32 element-ops (one FMA = one op)
PLAIN_PEAK_OPS = 39.3e12    # without packed FP32: one wave-64 VALU instruction per CU per clock
FLOP_PER_EVAL = 8           # 3 sub + 3 mul + 2 add (no FMA: fp-contract off)
OPS_PER_EVAL = 9            # the same plus the in-ball compare
# HBM bytes of the query kernel per step of the 20M-ion headline job, from two
# rocprofv3 --pmc runs of this script (FETCH_SIZE x 2 and WRITE_SIZE x 1, the
# factors a 1 GiB device copy reads): 93.0 GB read + 10.5 GB written
TRAFFIC_HEADLINE = 1.035e11
TRAFFIC_SOURCE = "profiles/k_query_frustum20M_rdf_mknn100.md"
# the same for the 100M-ion job's query launches: 161.3 GB read + 11.9 GB
# written
TRAFFIC_LARGE = 1.732e11
TRAFFIC_LARGE_SOURCE = "profiles/k_query_frustum100M_rdf_mknn20.md"



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


def roofline(cand, kernel_ms, kernel, traffic=None, traffic_source=None):
    """The query kernel against MI355X's published FP32 peak: 8 FLOP per
    candidate pair evaluation over the kernel's own time (HIP events around
    its launches)."""
    s = kernel_ms / 1e3
    achieved = cand * FLOP_PER_EVAL / s / 1e12
    ops = cand * OPS_PER_EVAL / s / 1e12
    return {
        "bound": "valu",
        "achieved": achieved,
        "peak": FP32_PEAK_FLOPS / 1e12,
        "unit": "TFLOP/s (FP32 VALU, 8 FLOP per candidate pair evaluation)",
        "frac": achieved / (FP32_PEAK_FLOPS / 1e12),
        "frac_published": achieved / (FP32_PEAK_FLOPS / 1e12),
        "elementops_achieved": ops,
        "frac_elementops": ops / (VALU_PEAK_OPS / 1e12),
        "frac_practical": ops / (PLAIN_PEAK_OPS / 1e12),
        "practical_note": "9 element-ops per evaluation (the 8 FLOP and the in-ball compare) "
                          "against 78.6 T element-ops/s; packed FP32 (v_pk_*_f32) is disabled "
                          "in this build, so plain FP32 issues one wave-64 instruction per CU "
                          "per clock: 39.3 T element-ops/s is the practical ceiling",
        "traffic": traffic,
        "traffic_unit": "bytes per step (HBM, all query launches)",
        "traffic_source": traffic_source,
        "kernel": kernel,
        "kernel_ms": kernel_ms,
        "candidate_evals": cand,
        "candidate_note": "pair evaluations the kernel performs; with the symmetric pair "
                          "counting a pair of two centers of the dominant type is evaluated "
                          "once for both",
    }


def index_build_ms(ctx, tip, job, reps=3):
    """The index build on its own (the cell list and the job's candidate
    lists): the first job on a freshly uploaded copy of the cloud builds them
    (later jobs reuse them); median of reps."""
    out = []
    for _ in range(reps):
        c = ctx.upload(tip.xyz, tip.types, tip.ion2surf_sqr)
        try:
            out.append(ctx.run(c, job).ms_index)
        finally:
            c.free()
    return float(np.median(out))


def binedge(tip, job, stride, cores):
    """The FP64 bin-edge report (oracle) on every stride-th ion as a center
    (the ions are in random order), all of the job's pairs at once: how many
    in-ball pairs lie within one float ulp of an ideal bin edge, how many
    FP32-pipeline bins differ from the exact distance's, how many pairs the
    FP32 and exact in-ball tests disagree on."""
    from oracle import oracle as O

    trg = sorted({t for ts, _ in job.tasks for t in ts})
    env = sorted({e for _, es in job.tasks for e in es})
    t0 = time.perf_counter()
    r = O.binedge_report(tip.xyz, tip.types, tip.ion2surf_sqr, trg, env, job.rmin, job.rincr,
                         job.rmax, center_stride=stride, nthreads=cores)
    r["sample"] = (f"every {stride}th ion as a center ({100.0 / stride:g}% of the centers), "
                   f"all {len(job.tasks)} type pairs; {time.perf_counter() - t0:.1f} s on "
                   f"{cores} host cores")
    print(f"# binedge: {r['pairs']} pairs, {r['within_one_ulp_of_edge']} within one ulp of an "
          f"edge, {r['bin_differs_from_fp64']} bins differ from FP64, "
          f"{r['ball_membership_flips']} ball flips", file=sys.stderr, flush=True)
    return r


def slab(tip, zwin, R):
    """The ions of z in [zwin - R, zwin + R) with the surface distance kept
    only for z in zwin (the others cannot be centers): every neighbour of a
    window center is in the slab, so its histograms equal the whole tip's
    for those centers."""
    from paraprobe_amd import synth

    z = tip.xyz[:, 2]
    keep = (z >= zwin[0] - R - 0.01) & (z < zwin[1] + R + 0.01)
    s2 = tip.ion2surf_sqr[keep].copy()
    zk = z[keep]
    s2[(zk < zwin[0]) | (zk >= zwin[1])] = 0.0
    return synth.Tip(np.ascontiguousarray(tip.xyz[keep]), np.ascontiguousarray(tip.types[keep]),
                     s2, f"slab z in [{zwin[0]}, {zwin[1]}) +- {R} nm")


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


def _same(a, b) -> bool:
    return all(np.array_equal(v, getattr(b, k)) for k, v in a.arrays().items())


class Phase:
    """A named ROCTx range around one phase of the benchmark, so that
    `rocprofv3 --kernel-trace --marker-trace` traces show where each phase
    starts and ends (`bench.py --busy DIR` turns them into GPU-busy
    fractions).  A no-op when the ROCTx library is absent."""

    _lib = None
    _tried = False

    def __init__(self, name: str):
        self.name = name.encode()

    @classmethod
    def lib(cls):
        if not cls._tried:
            cls._tried = True
            try:
                cls._lib = ctypes.CDLL("librocprofiler-sdk-roctx.so.1")
                cls._lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
            except OSError:
                cls._lib = None
        return cls._lib

    def __enter__(self):
        if self.lib() is not None:
            self._lib.roctxRangePushA(self.name)
        return self

    def __exit__(self, *exc):
        if self._lib is not None:
            self._lib.roctxRangePop()
        return False


def busy_report(trace_dir: str) -> str:
    """GPU-busy fraction of every named phase of a rocprofv3 trace
    (`--kernel-trace --marker-trace --output-format csv`): the union of the
    kernel intervals inside the phase over its wall time, with the kernels
    that take the most time.  Returns a markdown table."""
    import csv
    import glob

    def rows(pattern):
        out = []
        for f in sorted(glob.glob(os.path.join(trace_dir, "**", pattern), recursive=True)):
            with open(f, newline="") as fh:
                out.extend(csv.DictReader(fh))
        return out

    kern = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"])
            for r in rows("*kernel_trace.csv")]
    kern.sort()
    phases = []
    for r in rows("*marker_api_trace.csv"):
        name = r.get("Function") or r.get("Message") or ""
        s, e = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        if e > s:
            phases.append((s, e, name))
    phases.sort()
    lines = ["| phase | wall ms | kernel-busy ms | busy fraction | kernels | top kernels (ms) |",
             "|---|---|---|---|---|---|"]
    for s, e, name in phases:
        busy, cur_s, cur_e, n = 0, None, None, 0
        per = {}
        for ks, ke, kn in kern:
            if ke <= s or ks >= e:
                continue
            ks, ke = max(ks, s), min(ke, e)
            n += 1
            short = kn.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "")
            per[short] = per.get(short, 0) + (ke - ks)
            if cur_e is None or ks > cur_e:
                if cur_e is not None:
                    busy += cur_e - cur_s
                cur_s, cur_e = ks, ke
            else:
                cur_e = max(cur_e, ke)
        if cur_e is not None:
            busy += cur_e - cur_s
        top = sorted(per.items(), key=lambda kv: -kv[1])[:4]
        tops = ", ".join(f"{k} {v / 1e6:.1f}" for k, v in top)
        wall = (e - s) / 1e6
        lines.append(f"| {name} | {wall:.1f} | {busy / 1e6:.1f} | "
                     f"{busy / 1e6 / wall if wall > 0 else 0:.3f} | {n} | {tops} |")
    return "\n".join(lines)


def batch_leg(ctx, args):
    """256 clouds of uneven size, each with its own job: one persistent launch
    (pp_batch_run) against the plain per-job launch loop (pp_spatstat_run)."""
    from paraprobe_amd import workloads

    t = time.perf_counter()
    spec = workloads.batch_spec(args.batch_jobs)
    tips = workloads.batch_clouds(spec)
    gen_s = time.perf_counter() - t
    clouds = [ctx.upload(tp.xyz, tp.types, tp.ion2surf_sqr) for tp in tips]
    jobs = [it.job for it in spec]
    nions = sum(tp.n for tp in tips)
    del tips
    if args.autotune:
        ctx.set_tuning(-1, -1)
    t = time.perf_counter()
    ctx.run_batch(clouds, jobs)              # warmup: cell lists, typed lists, tuning
    warm_s = time.perf_counter() - t
    print(f"# batch: {len(jobs)} clouds, {nions} ions, generated in {gen_s:.1f} s, "
          f"warmup {warm_s:.1f} s", file=sys.stderr, flush=True)
    with Phase("batch_plain_loop"):
        t = time.perf_counter()
        plain = [ctx.run(c, j) for c, j in zip(clouds, jobs)]
        plain_s = time.perf_counter() - t
    with Phase("batch_batched"):
        t = time.perf_counter()
        batched = ctx.run_batch(clouds, jobs)
        batch_s = time.perf_counter() - t
    parity = all(_same(a, b) for a, b in zip(plain, batched))
    ctx.set_tuning(0, 0)
    for c in clouds:
        c.free()
    print(f"# batch: plain loop {plain_s:.2f} s, batched {batch_s:.2f} s, parity {parity}",
          file=sys.stderr, flush=True)
    return {
        "workload": f"{len(jobs)} clouds, ions log-uniform in [50k, 4M] ({nions} total), "
                    "uniform or clustered frustum tips, one type pair each, Rmax U(1,5) nm, "
                    "RDF + MKNN k=1..kmax (kmax in 1,10,50), bin 0.01 nm",
        "autotune": bool(args.autotune),
        "graph_launch": bool(args.graph),
        "plain_jobs_per_s": len(jobs) / plain_s,
        "batched_jobs_per_s": len(jobs) / batch_s,
        "speedup": plain_s / batch_s,
        "parity_batched_vs_plain": parity,
    }


def serve_leg(ctx, args):
    """8 resident tips and a Poisson stream of 10,000 requests: the
    continuous-batching server against the plain per-request loop.

    The plain loop runs back to back over a sample (the first
    `--serve-plain-sample` requests): its service times give the plain
    capacity and, through a first-come first-served replay of the same
    arrivals, its p50/p99; the server runs the same sample submitted at once
    (capacity), then the whole stream as an open loop of Poisson arrivals at
    0.7 x the plain capacity (p50/p99 over all requests, and over the sample
    for a like-for-like comparison)."""
    import threading

    from paraprobe_amd import workloads
    from paraprobe_amd.capi import Server

    t = time.perf_counter()
    tips = workloads.serve_tips()
    clouds = []
    for tp in tips:
        c = ctx.upload(tp.xyz, tp.types, tp.ion2surf_sqr, cell_width=1.0)
        c.add_cell_list(1.3)
        c.add_cell_list(0.6)
        clouds.append(c)
    nions = sum(tp.n for tp in tips)
    del tips
    M = args.serve_requests
    S = min(args.serve_plain_sample, M)
    stream = workloads.serve_stream(10_000)[:M]
    print(f"# serve: 8 resident tips ({nions} ions) ready in {time.perf_counter() - t:.1f} s; "
          f"{M} requests, plain loop on the first {S}", file=sys.stderr, flush=True)
    # build every (cloud, cell width, neighbour type) candidate list once, with
    # jobs whose central type holds no ion (so no query work is done), and the
    # lists of the symmetric pair counting of same-type pair counts (one RDF
    # job per cloud, width and type), so that neither the plain loop nor the
    # server builds lists while it is timed
    from paraprobe_amd import FAM_RDF, Job

    for c in clouds:
        for R in (1.5, 3.0, 5.0):
            for e in (1, 2, 3):
                ctx.run(c, Job([((7,), (e,))], 0.0, 0.01, R, FAM_RDF))
                ctx.run(c, Job([((e,), (e,))], 0.0, 0.01, R, FAM_RDF))
    # plain loop, back to back: per-request service times
    plain, svc = [], []
    with Phase("serve_plain_loop"):
        t0 = time.perf_counter()
        for r in stream[:S]:
            a = time.perf_counter()
            plain.append(ctx.run(clouds[r.cloud], r.job))
            svc.append(time.perf_counter() - a)
        plain_s = time.perf_counter() - t0
    plain_rate = S / plain_s
    print(f"# serve: plain loop {plain_rate:.1f} jobs/s on {S} requests", file=sys.stderr,
          flush=True)

    def served(reqs, rate, keep):
        """Runs reqs through the server; rate None submits everything at once
        (capacity), otherwise Poisson arrivals at `rate` per second.  Returns
        (completions per second, latencies in ms, results of the first `keep`
        requests, statuses)."""
        m = len(reqs)
        gaps = np.random.default_rng(99).exponential(1.0 / rate, size=m) if rate else np.zeros(m)
        arrive = np.cumsum(gaps)
        done, results, status = {}, {}, {}
        with Server(ctx, slots=args.serve_slots, max_batch_jobs=args.serve_batch,
                    batch_work=args.serve_batch_work) as srv:
            start = time.perf_counter()
            tick = {}
            stop = threading.Event()

            def poller():
                while not stop.is_set() or len(done) < m:
                    for tk, st, res in srv.poll():
                        done[tk] = time.perf_counter()
                        status[tk] = st
                        if tick[tk][0] < keep:
                            results[tk] = res
                    if len(done) >= m:
                        break
                    time.sleep(0.0002)

            th = threading.Thread(target=poller)
            th.start()
            for i, r in enumerate(reqs):
                now = time.perf_counter() - start
                if arrive[i] > now:
                    time.sleep(arrive[i] - now)
                tk = srv.submit(clouds[r.cloud], r.job)
                tick[tk] = (i, start + arrive[i])
                if rate and (i + 1) % 1000 == 0:
                    print(f"# serve: {i + 1} of {m} requests submitted, {len(done)} done",
                          file=sys.stderr, flush=True)
            stop.set()
            th.join()
            end = max(done.values())
        lat = np.zeros(m)
        out = [None] * keep
        ok = True
        for tk, (i, a) in tick.items():
            lat[i] = done[tk] - a
            ok &= status[tk] == 0
            if i < keep:
                out[i] = results[tk] if status[tk] == 0 else None
        return m / (end - start), lat * 1e3, out, ok

    with Phase("serve_capacity"):
        cap_rate, _, res_cap, ok_cap = served(stream[:S], None, S)
    lam = 0.7 * plain_rate
    print(f"# serve: capacity {cap_rate:.1f} jobs/s; open loop of {M} requests at "
          f"{lam:.1f}/s", file=sys.stderr, flush=True)
    with Phase("serve_open_loop"):
        rate, lat_ms, res_open, ok_open = served(stream, lam, S)
    parity = ok_cap and ok_open and \
        all(o is not None and _same(p, o) for p, o in zip(plain, res_cap)) and \
        all(o is not None and _same(p, o) for p, o in zip(plain, res_open))
    # the plain loop under the same arrivals (the first S of them):
    # first come first served, one request at a time
    arrive = np.cumsum(np.random.default_rng(99).exponential(1.0 / lam, size=M))[:S]
    fin, plat = 0.0, []
    for a, s_ in zip(arrive, svc):
        fin = max(fin, a) + s_
        plat.append(fin - a)
    plat = np.array(plat) * 1e3
    for c in clouds:
        c.free()
    p50, p99 = float(np.percentile(lat_ms, 50)), float(np.percentile(lat_ms, 99))
    s50, s99 = float(np.percentile(lat_ms[:S], 50)), float(np.percentile(lat_ms[:S], 99))
    # query time of each request's call (plain) or batch (served), HIP events
    # around its query launches: separates GPU service time from queueing
    srun = np.array([o.ms_query for o in res_open if o is not None])
    prun = np.array([o.ms_query for o in plain])
    print(f"# serve: query ms p50/p99, first {S}: served {np.percentile(srun, 50):.1f} / "
          f"{np.percentile(srun, 99):.1f}, plain {np.percentile(prun, 50):.1f} / "
          f"{np.percentile(prun, 99):.1f}", file=sys.stderr, flush=True)
    if args.serve_dump:
        # per request of the sample: arrival, served and plain latencies, query times
        with open(args.serve_dump, "w") as fh:
            json.dump({"arrive_s": arrive.tolist(), "served_ms": lat_ms[:S].tolist(),
                       "plain_ms": plat.tolist(), "plain_svc_ms": (np.array(svc) * 1e3).tolist(),
                       "served_query_ms": srun.tolist(), "plain_query_ms": prun.tolist(),
                       "cloud": [r.cloud for r in stream[:S]],
                       "rmax": [r.job.rmax for r in stream[:S]]}, fh)
    print(f"# serve: plain {plain_rate:.1f} jobs/s, served capacity {cap_rate:.1f} jobs/s, "
          f"open loop at {lam:.1f}/s: p50 {p50:.1f} ms p99 {p99:.1f} ms (first {S}: "
          f"{s50:.1f} / {s99:.1f}; plain {np.percentile(plat, 50):.1f} / "
          f"{np.percentile(plat, 99):.1f}), parity {parity}", file=sys.stderr, flush=True)
    return {
        "workload": f"8 resident clustered frustum tips (1M-50M ions, {nions} total), a "
                    f"{M}-request Poisson stream: one type pair, Rmax U(1,5) nm, RDF or "
                    "MKNN k=1..kmax (kmax U(1,100)), bin 0.01 nm",
        "requests": M,
        "plain_sample": S,
        "slots": args.serve_slots,
        "max_batch_jobs": args.serve_batch,
        "batch_work": args.serve_batch_work or "library default (5e9 candidate evaluations)",
        "plain_jobs_per_s": plain_rate,
        "served_capacity_jobs_per_s": cap_rate,
        "speedup": cap_rate / plain_rate,
        "poisson_rate_per_s": lam,
        "served_jobs_per_s": rate,
        "served_p50_ms": p50,
        "served_p99_ms": p99,
        "served_p50_ms_sample": s50,
        "served_p99_ms_sample": s99,
        "plain_p50_ms_same_arrivals_sample": float(np.percentile(plat, 50)),
        "plain_p99_ms_same_arrivals_sample": float(np.percentile(plat, 99)),
        "served_query_p99_ms_sample": float(np.percentile(srun, 99)),
        "plain_query_p99_ms_sample": float(np.percentile(prun, 99)),
        "parity_served_vs_plain": parity,
        "parity_scope": f"the first {S} requests, capacity and open-loop runs; "
                        f"all {M} completed with PP_OK" if parity else "failed",
    }


def large_leg(ctx, args, sync):
    """The large configuration: 100M ions in a frustum r 60 -> 30 nm (22
    ions/nm^3, 300,000 clusters), all 9 pair RDFs at Rmax 3 nm, bin 0.005 nm
    (600 bins) plus MKNN k = 1..20, timed like the headline; the reference's
    algorithm on the host cores for the centers of a 1 nm z-window (on the
    slab that holds their neighbours, the GPU compared on the same slab);
    the FP64 bin-edge report on 1% of the centers."""
    from paraprobe_amd import FAM_MKNN, FAM_RDF, Job, synth

    t = time.perf_counter()
    tip = synth.frustum_100m(n=args.large_n)
    print(f"# large: generated {tip.n} ions in {time.perf_counter() - t:.1f} s",
          file=sys.stderr, flush=True)
    job = Job(synth.ALL_PAIRS, 0.0, 0.005, 3.0, FAM_RDF | FAM_MKNN, kvalues=list(range(1, 21)))
    cloud = ctx.upload(tip.xyz, tip.types, tip.ion2surf_sqr)
    try:
        first = ctx.run(cloud, job)
        sync()
        with Phase("large_steps"):
            t0 = time.perf_counter()
            res = [ctx.run(cloud, job) for _ in range(args.large_steps)]
            sync()
            dt = time.perf_counter() - t0
    finally:
        cloud.free()
    for r in res[1:]:
        if not _same(r, res[0]):
            raise SystemExit("bench: large results differ between steps")
    last = res[-1]
    ms_step = dt * 1e3 / args.large_steps
    kms = float(np.mean([r.ms_kernel for r in res]))
    print(f"# large: {ms_step:.1f} ms per job (query kernel {kms:.1f} ms), "
          f"{last.pairs_inball} in-ball pairs", file=sys.stderr, flush=True)
    out = {
        "workload": "frustum-100M-clustered: 100,000,000 ions, r 60->30 nm, h 689 nm "
                    "(22 ions/nm^3), types A/B/C 0.95/0.04/0.01, 300,000 clusters r~U(1,2) nm "
                    "B=40%; 9 type pairs RDF + MKNN k=1..20, Rmax 3 nm, bin 0.005 nm, int64",
        "n_ions": tip.n,
        "steps": args.large_steps,
        "ms_per_step": ms_step,
        "value": last.pairs_inball / (ms_step / 1e3),
        "unit": "pair-distance evals/s",
        "centers_considered": last.centers_considered,
        "pairs_inball": last.pairs_inball,
        "ms_index_build": first.ms_index,
        "ms_query": float(np.mean([r.ms_query for r in res])),
        "roofline": roofline(last.pairs_candidate, kms, last.kernel,
                             TRAFFIC_LARGE if args.large_n == 100_000_000 else None,
                             TRAFFIC_LARGE_SOURCE if args.large_n == 100_000_000 else None),
    }
    mid = float(np.median(tip.xyz[:, 2]))
    zwin = (round(mid, 1) - 0.5, round(mid, 1) + 0.5)
    sl = slab(tip, zwin, 3.0)
    cw = ctx.upload(sl.xyz, sl.types, sl.ion2surf_sqr)
    try:
        gwin = ctx.run(cw, job)
    finally:
        cw.free()
    model, cores = host_cpu()
    if not args.no_cpu:
        hist, st, _, _ = cpu_baseline(sl, job, zwin, cores)
        out["cpu_baseline"] = {
            "value": st["inball"] / st["compute_s"],
            "unit": "pair-distance evals/s",
            "cores": cores,
            "kind": "port",
            "sample": f"centers with z in [{zwin[0]:.1f}, {zwin[1]:.1f}) nm ({st['considered']} "
                      f"centers, {st['inball']} in-ball pairs) on the slab of their neighbours "
                      f"({sl.n} ions); compute {st['compute_s']:.1f} s timed, z-slab + KD-tree "
                      f"build {st['index_s']:.1f} s not; {model}",
            "parity_with_gpu": all(np.array_equal(getattr(gwin, k), v) for k, v in hist.items()),
        }
        out["binedge"] = binedge(tip, job, 100, cores)
    return out


def surf_leg(ctx, tip, hip, args):
    """ion2surf, the producer of the per-ion surface distances: the 20M ions
    against a 512 x 256 triangulated frustum hull (263,168 triangles), R =
    max(SpatStatRadiusMax 5, SurfaceCellsCarvingRadius 2) nm, inputs and output
    resident in HBM; the reference's structure (triangle-box tree query per
    ion) timed on the host cores on every 4th ion."""
    from paraprobe_amd import synth

    tri = synth.hull("frustum", 40.0, 20.0, 250.0, 512, 256)
    R = 5.0
    n, nt = tip.n, tri.shape[0]
    ptr = [ctypes.c_void_p() for _ in range(3)]
    sizes = [12 * n, 36 * nt, 4 * n]
    for p_, b in zip(ptr, sizes):
        if hip.hipMalloc(ctypes.byref(p_), ctypes.c_size_t(b)) != 0:
            raise SystemExit("bench: hipMalloc failed")
    try:
        hip.hipMemcpy(ptr[0], tip.xyz.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(sizes[0]), 1)
        hip.hipMemcpy(ptr[1], tri.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(sizes[1]), 1)
        ctx.ion2surf_device(ptr[0].value, n, ptr[1].value, nt, R, ptr[2].value)
        hip.hipDeviceSynchronize()
        reps = max(1, args.steps)
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.ion2surf_device(ptr[0].value, n, ptr[1].value, nt, R, ptr[2].value)
        hip.hipDeviceSynchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        got = np.empty(n, np.float32)
        hip.hipMemcpy(got.ctypes.data_as(ctypes.c_void_p), ptr[2], ctypes.c_size_t(sizes[2]), 2)
    finally:
        for p_ in ptr:
            hip.hipFree(p_)
    t0 = time.perf_counter()
    host = ctx.ion2surf(tip.xyz, tri, R)
    host_ms = (time.perf_counter() - t0) * 1e3
    out = {
        "workload": "frustum-20M ions vs 512x256 frustum hull (263,168 triangles), R 5 nm",
        "ions_per_s": n / (ms / 1e3),
        "ms": ms,
        "ms_host_buffers": host_ms,
        "ions_within_R": int((got < np.float32(R * R)).sum()),
        "parity_device_vs_host_entry": bool(np.array_equal(got.view(np.uint32),
                                                           host.view(np.uint32))),
    }
    print(f"# ion2surf: {ms:.1f} ms resident, {host_ms:.1f} ms with host buffers",
          file=sys.stderr, flush=True)
    if not args.no_cpu:
        from oracle import cpu_baseline as cb

        model, cores = host_cpu()
        stride = 4
        v, build_s, query_s = cb.ion2surf(tip.xyz, tri, R, nthreads=cores, stride=stride)
        out["cpu_baseline"] = {
            "value": v.size / query_s,
            "unit": "ions/s",
            "cores": cores,
            "kind": "port",
            "sample": f"every {stride}th ion ({v.size} ions), triangle-box tree built in "
                      f"{build_s:.2f} s (not timed), query {query_s:.1f} s; {model}",
            "parity_with_gpu": bool(np.array_equal(v.view(np.uint32),
                                                   got[::stride].view(np.uint32))),
        }
    return out


POS_RANGES = (  # Unknown's dummy range, then Al, Fe, Si, Sc as in a four-range RRNG file
    np.array([0.0, 26.7, 55.9, 27.8, 44.8], np.float32),
    np.array([0.001, 26.9, 56.0, 28.0, 50.0], np.float32),
    np.array([0, 1, 2, 3, 4], np.uint8),
)


def pos_leg(ctx, tip, hip, args):
    """POS records -> positions, mass-to-charge and ranged types: the 20M ions
    as big-endian 16-byte records resident in HBM, decoded into device
    buffers; the reference's sequential per-record loop (byte reversal +
    mq2ionname) timed on one host core."""
    mq = np.select([tip.types == 1, tip.types == 2], [26.8, 55.95], 47.0).astype(np.float32)
    rec = np.concatenate([tip.xyz, mq[:, None]], axis=1).astype(">f4").tobytes()
    n = tip.n
    ptr = [ctypes.c_void_p() for _ in range(4)]
    sizes = [16 * n, 12 * n, 4 * n, n]
    for p_, b in zip(ptr, sizes):
        if hip.hipMalloc(ctypes.byref(p_), ctypes.c_size_t(b)) != 0:
            raise SystemExit("bench: hipMalloc failed")
    try:
        src = np.frombuffer(rec, np.uint8)
        hip.hipMemcpy(ptr[0], src.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(sizes[0]), 1)
        args_ = (ptr[0].value, 16 * n, POS_RANGES, ptr[1].value, ptr[2].value, ptr[3].value)
        ctx.pos_decode_device(*args_)
        hip.hipDeviceSynchronize()
        reps = max(5, args.steps)
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.pos_decode_device(*args_)
        hip.hipDeviceSynchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        ty = np.empty(n, np.uint8)
        hip.hipMemcpy(ty.ctypes.data_as(ctypes.c_void_p), ptr[3], ctypes.c_size_t(n), 2)
    finally:
        for p_ in ptr:
            hip.hipFree(p_)
    gbs = 33.0 * n / (ms / 1e3) / 1e9
    out = {
        "workload": "frustum-20M ions as POS records (320 MB) -> xyz, mq, type; 5 ranges",
        "ions_per_s": n / (ms / 1e3),
        "ms": ms,
        "gb_per_s": gbs,
        "roofline_frac_hbm": gbs / 8000.0,
        "bytes_per_ion": 33,
    }
    print(f"# pos decode: {ms:.2f} ms, {gbs:.0f} GB/s", file=sys.stderr, flush=True)
    if not args.no_cpu:
        from oracle import oracle as O

        m = n
        t0 = time.perf_counter()
        _, _, oty = O.pos_decode(rec[:16 * m], POS_RANGES)
        cs = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": m / cs,
            "unit": "ions/s",
            "cores": 1,
            "kind": "port",
            "sample": f"all {m} records, the reference's sequential loop (byte reversal, "
                      f"first matching range) in C, {cs:.2f} s",
            "parity_with_gpu": bool(np.array_equal(oty, ty[:m])),
        }
    return out


def npcorr_leg(ctx, tip, args):
    """NPCORR with the reference's settings fixture values (tasks A-A; B-B;
    A-B; B-A, SpatStatRadiusMax 2.5, SpatStatRadiusIncr 0.05, SpatStatKNNOrder
    100, i.e. the 100th neighbour's difference vector in a 103^3 cube) on the
    resident 20M-ion tip; the reference's structure (z-slabs, KD-trees,
    nth_element) timed on the host cores for the centers of a 2 nm z-window."""
    tasks = [((1,), (1,)), ((2,), (2,)), ((1,), (2,)), ((2,), (1,))]
    R, w, K = 2.5, 0.05, 99
    cloud = ctx.upload(tip.xyz, tip.types, tip.ion2surf_sqr)
    try:
        ctx.npcorr(cloud, tasks, R, w, K)
        reps = 2
        t0 = time.perf_counter()
        for _ in range(reps):
            got = ctx.npcorr(cloud, tasks, R, w, K)
        ms = (time.perf_counter() - t0) * 1e3 / reps
        zwin = (124.0, 126.0)
        s2w = tip.ion2surf_sqr.copy()
        zz = tip.xyz[:, 2]
        s2w[(zz < zwin[0]) | (zz >= zwin[1])] = 0.0
        cw = ctx.upload(tip.xyz, tip.types, s2w)
        gwin = ctx.npcorr(cw, tasks, R, w, K)
        cw.free()
    finally:
        cloud.free()
    considered = int((tip.ion2surf_sqr >= np.float32(R) * np.float32(R)).sum())
    out = {
        "workload": "frustum-20M, NPCORR A-A;B-B;A-B;B-A, Rmax 2.5 nm, bin 0.05 nm (103^3 cube), "
                    "100th neighbour",
        "ms": ms,
        "centers_per_s": considered / (ms / 1e3),
        "vectors_binned": int(got.sum()),
    }
    print(f"# npcorr: {ms:.1f} ms", file=sys.stderr, flush=True)
    if not args.no_cpu:
        from oracle import cpu_baseline as cb

        model, cores = host_cpu()
        cnt, idx_s, comp_s = cb.npcorr(tip.xyz, tip.types, s2w, tasks, R, w, K, nthreads=cores)
        nwin = int(((zz >= zwin[0]) & (zz < zwin[1]) &
                    (tip.ion2surf_sqr >= np.float32(R) * np.float32(R))).sum())
        out["cpu_baseline"] = {
            "value": nwin / comp_s,
            "unit": "centers/s",
            "cores": cores,
            "kind": "port",
            "sample": f"centers with z in [{zwin[0]}, {zwin[1]}) nm ({nwin} considered); compute "
                      f"{comp_s:.1f} s timed, z-slab + KD-tree build {idx_s:.1f} s not; {model}",
            "parity_with_gpu": bool(np.array_equal(cnt, gwin)),
        }
    return out


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
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--large-n", type=int, default=100_000_000)
    ap.add_argument("--large-steps", type=int, default=3)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-surf", action="store_true")
    ap.add_argument("--no-pos", action="store_true")
    ap.add_argument("--no-npcorr", action="store_true")
    ap.add_argument("--batch-jobs", type=int, default=256)
    ap.add_argument("--autotune", type=int, default=1)
    ap.add_argument("--serve-requests", type=int, default=10_000)
    ap.add_argument("--serve-plain-sample", type=int, default=1000,
                    help="requests of the stream the plain loop is timed on")
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--serve-batch", type=int, default=32)
    ap.add_argument("--serve-batch-work", type=float, default=0.0,
                    help="estimated candidate evaluations per serving batch (0: the library's)")
    ap.add_argument("--serve-dump", default=None,
                    help="write the sample's per-request latencies to this JSON file")
    ap.add_argument("--graph", type=int, default=1,
                    help="1: batch launch groups captured into HIP graphs; 0: eager launches")
    ap.add_argument("--busy", metavar="TRACE_DIR", default=None,
                    help="print the GPU-busy fraction of each phase of a rocprofv3 "
                         "--kernel-trace --marker-trace csv trace and exit (no GPU)")
    args = ap.parse_args()
    if args.busy:
        print(busy_report(args.busy))
        return
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
        ctx.set_graph_launch(bool(args.graph))
        cloud = ctx.upload(tip.xyz, tip.types, tip.ion2surf_sqr)
        for i in range(args.warmup):
            r = ctx.run(cloud, job)
            print(f"# warmup {i}: {r.ms_total:.1f} ms (query {r.ms_query:.1f} ms)",
                  file=sys.stderr, flush=True)
        sync()
        with Phase("headline_steps"):
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
        ms_index_build = index_build_ms(ctx, tip, job)
        print(f"# index build {ms_index_build:.2f} ms", file=sys.stderr, flush=True)
        # GPU result of the CPU sample window, for the parity check
        zwin = (125.0 - args.cpu_window / 2, 125.0 + args.cpu_window / 2)
        s2w = tip.ion2surf_sqr.copy()
        zz = tip.xyz[:, 2]
        s2w[(zz < zwin[0]) | (zz >= zwin[1])] = 0.0
        cw = ctx.upload(tip.xyz, tip.types, s2w)
        gwin = ctx.run(cw, job)
        cw.free()
        cloud.free()
        extra = {}
        if not args.no_large:
            extra["large"] = large_leg(ctx, args, sync)
        if not args.no_surf:
            extra["ion2surf"] = surf_leg(ctx, tip, hip, args)
        if not args.no_pos:
            extra["pos_decode"] = pos_leg(ctx, tip, hip, args)
        if not args.no_npcorr:
            extra["npcorr"] = npcorr_leg(ctx, tip, args)
        if not args.no_batch:
            extra["batch"] = batch_leg(ctx, args)
        if not args.no_serve:
            extra["serve"] = serve_leg(ctx, args)
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
    k_ms = float(np.mean([r.ms_kernel for r in res]))
    cand = last.pairs_candidate
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
        "roofline": roofline(cand, k_ms, last.kernel,
                             TRAFFIC_HEADLINE if tip.n == 20_000_000 else None,
                             TRAFFIC_SOURCE if tip.n == 20_000_000 else None),
        "centers_considered": last.centers_considered,
        "pairs_inball": inball,
        "candidate_evals_per_s": cand / (k_ms / 1e3),
        "ms_index": float(np.mean([r.ms_index for r in res])),
        "ms_index_note": "each step reuses the cloud's cell list and candidate lists (built by the first job)",
        "ms_index_build": ms_index_build,
        "ms_prep": float(np.mean([r.ms_prep for r in res])),
        "ms_query": q_ms,
        "ms_query_note": "query kernel launches, their k-th neighbour reductions and the "
                         "full-range rerun; roofline.kernel_ms is the query kernel alone",
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
        line["binedge"] = binedge(tip, job, 100, cores)
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
