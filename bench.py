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
# HBM bytes of the query kernel per step of the 20M-ion headline job, from two
# rocprofv3 --pmc runs of this script (FETCH_SIZE x 2 and WRITE_SIZE x 1, the
# factors a 1 GiB device copy reads): 141.0 GB read + 15.7 GB written
TRAFFIC_HEADLINE = 1.567e11
TRAFFIC_SOURCE = "profiles/k_query_frustum20M_rdf_mknn100.md"



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


def _same(a, b) -> bool:
    return all(np.array_equal(v, getattr(b, k)) for k, v in a.arrays().items())


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
    t = time.perf_counter()
    plain = [ctx.run(c, j) for c, j in zip(clouds, jobs)]
    plain_s = time.perf_counter() - t
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
    """8 resident tips and a Poisson request stream: the continuous-batching
    server against the plain per-request loop."""
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
    stream = workloads.serve_stream(10_000)[:M]
    print(f"# serve: 8 resident tips ({nions} ions) ready in {time.perf_counter() - t:.1f} s; "
          f"{M} requests of the 10,000-request stream", file=sys.stderr, flush=True)
    # build every (cloud, cell width, neighbour type) candidate list once, with
    # jobs whose central type holds no ion (so no query work is done)
    from paraprobe_amd import FAM_RDF, Job

    for c in clouds:
        for R in (1.5, 3.0, 5.0):
            for e in (1, 2, 3):
                ctx.run(c, Job([((7,), (e,))], 0.0, 0.01, R, FAM_RDF))
    # plain loop, back to back: per-request service times
    plain, svc = [], []
    t0 = time.perf_counter()
    for r in stream:
        a = time.perf_counter()
        plain.append(ctx.run(clouds[r.cloud], r.job))
        svc.append(time.perf_counter() - a)
    plain_s = time.perf_counter() - t0
    plain_rate = M / plain_s

    def served(rate):
        """Runs the stream through the server; rate None submits everything at
        once (capacity), otherwise Poisson arrivals at `rate` per second."""
        gaps = np.random.default_rng(99).exponential(1.0 / rate, size=M) if rate else np.zeros(M)
        arrive = np.cumsum(gaps)
        done, lat, results = {}, {}, {}
        with Server(ctx, slots=args.serve_slots, max_batch_jobs=args.serve_batch) as srv:
            start = time.perf_counter()
            tick = {}
            stop = threading.Event()

            def poller():
                while not stop.is_set() or len(done) < M:
                    for tk, st, res in srv.poll():
                        done[tk] = time.perf_counter()
                        results[tk] = (st, res)
                    if len(done) >= M:
                        break
                    time.sleep(0.0002)

            th = threading.Thread(target=poller)
            th.start()
            for i, r in enumerate(stream):
                now = time.perf_counter() - start
                if arrive[i] > now:
                    time.sleep(arrive[i] - now)
                tk = srv.submit(clouds[r.cloud], r.job)
                tick[tk] = (i, start + arrive[i])
            stop.set()
            th.join()
            end = max(done.values())
        out = [None] * M
        for tk, (i, a) in tick.items():
            lat[i] = done[tk] - a
            st, res = results[tk]
            out[i] = res if st == 0 else None
        return M / (end - start), np.array([lat[i] for i in range(M)]) * 1e3, out

    cap_rate, _, res_cap = served(None)
    lam = 0.7 * cap_rate
    rate, lat_ms, res_open = served(lam)
    parity = all(o is not None and _same(p, o) for p, o in zip(plain, res_cap)) and \
        all(o is not None and _same(p, o) for p, o in zip(plain, res_open))
    # the plain loop under the same arrivals: FIFO, one request at a time
    arrive = np.cumsum(np.random.default_rng(99).exponential(1.0 / lam, size=M))
    fin, plat = 0.0, []
    for a, s_ in zip(arrive, svc):
        fin = max(fin, a) + s_
        plat.append(fin - a)
    plat = np.array(plat) * 1e3
    for c in clouds:
        c.free()
    print(f"# serve: plain {plain_rate:.1f} jobs/s, served capacity {cap_rate:.1f} jobs/s, "
          f"open loop at {lam:.1f}/s: p50 {np.percentile(lat_ms, 50):.1f} ms "
          f"p99 {np.percentile(lat_ms, 99):.1f} ms, parity {parity}", file=sys.stderr, flush=True)
    return {
        "workload": f"8 resident clustered frustum tips (1M-50M ions, {nions} total), first {M} "
                    "requests of a 10,000-request stream: one type pair, Rmax U(1,5) nm, RDF or "
                    "MKNN k=1..kmax (kmax U(1,100)), bin 0.01 nm",
        "slots": args.serve_slots,
        "max_batch_jobs": args.serve_batch,
        "served_capacity_jobs_per_s": cap_rate,
        "plain_jobs_per_s": plain_rate,
        "speedup": cap_rate / plain_rate,
        "poisson_rate_per_s": lam,
        "served_jobs_per_s": rate,
        "served_p50_ms": float(np.percentile(lat_ms, 50)),
        "served_p99_ms": float(np.percentile(lat_ms, 99)),
        "plain_p50_ms_same_arrivals": float(np.percentile(plat, 50)),
        "plain_p99_ms_same_arrivals": float(np.percentile(plat, 99)),
        "parity_served_vs_plain": parity,
    }


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
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-surf", action="store_true")
    ap.add_argument("--no-pos", action="store_true")
    ap.add_argument("--no-npcorr", action="store_true")
    ap.add_argument("--batch-jobs", type=int, default=256)
    ap.add_argument("--autotune", type=int, default=1)
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-slots", type=int, default=1)
    ap.add_argument("--serve-batch", type=int, default=32)
    ap.add_argument("--graph", type=int, default=1,
                    help="1: batch launch groups captured into HIP graphs; 0: eager launches")
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
        ctx.set_graph_launch(bool(args.graph))
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
        extra = {}
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
            "traffic": TRAFFIC_HEADLINE if tip.n == 20_000_000 else None,
            "traffic_unit": "bytes per step (HBM, all query launches)",
            "traffic_source": TRAFFIC_SOURCE if tip.n == 20_000_000 else None,
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
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
