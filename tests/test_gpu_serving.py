"""Batching and serving on the GPU: every result must be identical alone and
in the mix (integer histograms -> exact), admission control must reject
oversized requests and keep serving the rest, and the autotuner and
multi-width resident clouds must not change any result."""
import time

import numpy as np
import pytest

from oracle import oracle as O
from paraprobe_amd import FAM_1NN, FAM_MKNN, FAM_RDF, FAM_RIPK, Job, PPError
from paraprobe_amd.capi import Server

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _cloud(n, seed, L):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0, L, size=(n, 3)).astype(np.float32)
    ty = rng.choice([1, 2, 3], size=n, p=[0.7, 0.2, 0.1]).astype(np.uint8)
    s2 = rng.uniform(0, 6, size=n).astype(np.float32)
    return xyz, ty, s2


def _random_job(rng):
    a, b = int(rng.integers(1, 4)), int(rng.integers(1, 4))
    R = float(np.float32(rng.uniform(0.5, 1.5)))
    fam = [FAM_RDF, FAM_RDF | FAM_MKNN, FAM_MKNN | FAM_1NN, FAM_RDF | FAM_RIPK | FAM_1NN][
        int(rng.integers(0, 4))]
    ks = sorted(set(int(k) for k in rng.integers(1, 21, size=3)))
    return Job([((a,), (b,))], 0.0, 0.01, R, fam, kvalues=ks if fam & FAM_MKNN else [])


def _same(a, b):
    for k, v in a.arrays().items():
        np.testing.assert_array_equal(getattr(b, k), v)
    assert a.pairs_inball == b.pairs_inball and a.centers_considered == b.centers_considered


def test_serving_alone_equals_mix(ctx):
    _say("test_serving_alone_equals_mix")
    raw = [_cloud(20000, 1, 9.0), _cloud(60000, 2, 13.0), _cloud(5000, 3, 6.0)]
    clouds = []
    for xyz, ty, s2 in raw:
        c = ctx.upload(xyz, ty, s2, cell_width=0.4)
        c.add_cell_list(0.25)
        clouds.append(c)
    rng = np.random.default_rng(5)
    reqs = [(int(rng.integers(0, 3)), _random_job(rng)) for _ in range(48)]
    try:
        with Server(ctx, slots=3, max_batch_jobs=8) as srv:
            tickets = {}
            for i, (ci, job) in enumerate(reqs):
                tickets[srv.submit(clouds[ci], job)] = i
                if i % 12 == 11:
                    time.sleep(0.01)  # later requests arrive while earlier ones run
            srv.drain()
            got = {}
            for t, st, res in srv.poll():
                assert st == 0
                got[tickets[t]] = res
        assert len(got) == len(reqs)
        for i, (ci, job) in enumerate(reqs):
            alone = ctx.run(clouds[ci], job)
            _same(alone, got[i])
        # and a few against the oracle
        for i in (0, 7, 21):
            ci, job = reqs[i]
            xyz, ty, s2 = raw[ci]
            ref, st = O.spatstat(xyz, ty, s2, job.tasks, job.rmin, job.rincr, job.rmax,
                                 job.families, kvalues=job.kvalues)
            for k, v in ref.items():
                np.testing.assert_array_equal(getattr(got[i], k), v)
    finally:
        for c in clouds:
            c.free()


def test_admission_rejects_and_keeps_serving(ctx):
    _say("test_admission_rejects_and_keeps_serving")
    xyz, ty, s2 = _cloud(30000, 9, 10.0)
    cl = ctx.upload(xyz, ty, s2, cell_width=0.4)
    job = Job([((1,), (2,))], 0.0, 0.01, 1.0, FAM_RDF)
    try:
        with Server(ctx, slots=2, max_batch_jobs=4, hbm_budget_bytes=4096) as srv:
            with pytest.raises(PPError) as e:
                srv.submit(cl, job)
            assert e.value.status == 4  # PP_ERR_ADMISSION_REJECTED
        # a budget for roughly one request at a time still serves everything
        with Server(ctx, slots=2, max_batch_jobs=4, hbm_budget_bytes=8 << 20) as srv:
            ts = [srv.submit(cl, job) for _ in range(6)]
            srv.drain()
            res = srv.poll()
            assert sorted(t for t, _, _ in res) == sorted(ts)
            alone = ctx.run(cl, job)
            for _, st, r in res:
                assert st == 0
                _same(alone, r)
        # batches of one request each (batch_work below any request's estimate)
        # give the same results; a negative batch_work is refused
        with Server(ctx, slots=2, max_batch_jobs=4, batch_work=1.0) as srv:
            ts = [srv.submit(cl, job) for _ in range(3)]
            srv.drain()
            res = srv.poll()
            assert sorted(t for t, _, _ in res) == sorted(ts)
            for _, st, r in res:
                assert st == 0
                _same(alone, r)
        with pytest.raises(PPError):
            Server(ctx, slots=1, max_batch_jobs=4, batch_work=-1.0)
        # malformed requests are refused at submit; the server keeps going
        with Server(ctx, slots=1, max_batch_jobs=4) as srv:
            with pytest.raises(PPError):
                srv.submit(cl, Job([((1,), (2,))], 0.0, 1e-9, 1.0, FAM_RDF))
            t = srv.submit(cl, job)
            srv.drain()
            (tt, st, r), = srv.poll()
            assert tt == t and st == 0
            _same(alone, r)
    finally:
        cl.free()


def test_autotune_and_widths_keep_results(ctx):
    _say("test_autotune_and_widths_keep_results")
    xyz, ty, s2 = _cloud(40000, 11, 12.0)
    job = Job([((1,), (1,)), ((2,), (1, 3))], 0.0, 0.01, 1.2, FAM_RDF | FAM_MKNN, kvalues=[1, 5])
    cl = ctx.upload(xyz, ty, s2)
    try:
        base = ctx.run(cl, job)
        ctx.set_tuning(-1, -1)
        try:
            a1 = ctx.run(cl, job)   # tunes this bucket
            a2 = ctx.run(cl, job)   # uses the tuned shape
        finally:
            ctx.set_tuning(0, 0)
        _same(base, a1)
        _same(base, a2)
    finally:
        cl.free()
    cl = ctx.upload(xyz, ty, s2, cell_width=0.3)
    try:
        cl.add_cell_list(0.9)
        for R in (0.6, 1.2, 2.5):
            j = Job(job.tasks, 0.0, 0.01, R, job.families, kvalues=job.kvalues)
            g = ctx.run(cl, j)
            ref, _ = O.spatstat(xyz, ty, s2, j.tasks, j.rmin, j.rincr, j.rmax, j.families,
                                kvalues=j.kvalues)
            for k, v in ref.items():
                np.testing.assert_array_equal(getattr(g, k), v)
    finally:
        cl.free()
