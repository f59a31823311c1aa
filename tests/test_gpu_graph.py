"""Graph-captured query launches on the GPU: a run whose launch groups are
captured into HIP graphs must give the same integer histograms as the same
launches issued one by one, on every repetition, and both must equal the
oracle.  The batch mixes the three kernel modes (pair counts only, kNN-range
histograms with the full-range rerun of short centers, full-range per-center
histograms) so that one run captures several launch groups."""
import numpy as np
import pytest

from oracle import oracle as O
from paraprobe_amd import FAM_1NN, FAM_CNTNB, FAM_MKNN, FAM_RDF, FAM_RIPK, Job

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _cloud(n, seed, L, p=(0.6, 0.3, 0.1)):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0, L, size=(n, 3)).astype(np.float32)
    ty = rng.choice([1, 2, 3], size=n, p=list(p)).astype(np.uint8)
    s2 = rng.uniform(0, 3, size=n).astype(np.float32)
    return xyz, ty, s2


SPECS = [
    # pair counts only (mode 0)
    (4000, 31, 7.0, Job([((1,), (1,)), ((2,), (1, 3))], 0.0, 0.02, 1.3, FAM_RDF | FAM_RIPK)),
    # kNN-range histograms; the rare type 3 leaves centers short of k
    (6000, 32, 8.0, Job([((1,), (3,)), ((2,), (2,))], 0.0, 0.01, 1.6,
                        FAM_RDF | FAM_MKNN | FAM_1NN, kvalues=[1, 4, 20])),
    # full-range per-center histograms (CNTNB)
    (2500, 33, 6.0, Job([((1, 2), (2,)), ((3,), (1, 2, 3))], 0.1, 0.05, 1.2,
                        FAM_RDF | FAM_1NN | FAM_CNTNB)),
]


def _same(a, b):
    for k, v in a.arrays().items():
        np.testing.assert_array_equal(getattr(b, k), v)
    assert a.pairs_inball == b.pairs_inball and a.centers_considered == b.centers_considered


def test_graph_launch_equals_eager_launch(ctx):
    _say("test_graph_launch_equals_eager_launch: batch of 3 jobs, modes 0/1/2")
    raw = [_cloud(n, seed, L) for n, seed, L, _ in SPECS]
    jobs = [j for *_, j in SPECS]
    clouds = [ctx.upload(*r) for r in raw]
    try:
        ctx.set_graph_launch(False)
        eager = ctx.run_batch(clouds, jobs)
        eager1 = [ctx.run(c, j) for c, j in zip(clouds, jobs)]
        ctx.set_graph_launch(True)
        for rep in range(3):
            _say(f"  graph batch, repetition {rep}")
            got = ctx.run_batch(clouds, jobs)
            for e, g in zip(eager, got):
                _same(e, g)
            # the eager run of one job between the graph runs
            _same(eager1[rep], ctx.run(clouds[rep], jobs[rep]))
        for e, e1 in zip(eager, eager1):
            _same(e, e1)
        for (xyz, ty, s2), j, g in zip(raw, jobs, eager):
            o, st = O.spatstat(xyz, ty, s2, j.tasks, j.rmin, j.rincr, j.rmax, j.families,
                               kvalues=j.kvalues, cntnb_cap=65536)
            for k, v in o.items():
                np.testing.assert_array_equal(getattr(g, k), v)
    finally:
        ctx.set_graph_launch(True)
        for c in clouds:
            c.free()


def test_graph_batch_with_empty_clouds(ctx):
    _say("test_graph_batch_with_empty_clouds: batches of empty clouds, alone and mixed")
    xyz, ty, s2 = _cloud(3000, 41, 6.0)
    empty = ctx.upload(np.zeros((0, 3), np.float32), np.zeros(0, np.uint8),
                       np.zeros(0, np.float32))
    full = ctx.upload(xyz, ty, s2)
    job = Job([((1,), (2,))], 0.0, 0.02, 1.0, FAM_RDF | FAM_MKNN | FAM_1NN, kvalues=[1, 3])
    try:
        ctx.set_graph_launch(True)
        outs = ctx.run_batch([empty, empty], [job, job])
        for o in outs:
            assert o.rdf.sum() == 0 and o.centers_considered == 0
        mixed = ctx.run_batch([empty, full, empty], [job, job, job])
        _same(ctx.run(full, job), mixed[1])
        assert mixed[0].rdf.sum() == 0 and mixed[2].rdf.sum() == 0
    finally:
        empty.free()
        full.free()
