"""pp_spatstat_run_noalloc through the C ABI: every device temporary in the
library's fixed scratch, results equal to pp_spatstat_run's.

Two calls in a row, the second with a smaller input (the scratch keeps the
first call's contents, so whatever must start at zero has to be cleared by the
call itself); then the error paths the host turns into status codes (a cloud
whose job needs more than the scratch, a cloud without the cell list of the
job's width)."""
import numpy as np
import pytest

from oracle import oracle as O
from paraprobe_amd import FAM_1NN, FAM_CNTNB, FAM_MKNN, FAM_RDF, FAM_RIPK, Job, PPError, synth

pytestmark = pytest.mark.gpu

ALL = FAM_RDF | FAM_1NN | FAM_RIPK | FAM_MKNN | FAM_CNTNB
TASKS = [((1,), (1,)), ((2,), (1, 3)), ((1, 2), (2,)), ((3,), (1, 2, 3))]


def _same(a, b, candidates=True):
    for k, v in a.arrays().items():
        np.testing.assert_array_equal(getattr(b, k), v, err_msg=k)
    assert a.centers_considered == b.centers_considered
    assert a.pairs_inball == b.pairs_inball
    if candidates:
        assert a.pairs_candidate == b.pairs_candidate


def test_noalloc_two_calls_equal_run(ctx):
    # a resident 400k-ion clustered tip (hybrid kNN mode with its rerun), then a
    # smaller cube; kNN + RDF + CNTNB fill every part of the scratch
    print("noalloc: 400k-ion clustered tip, resident", flush=True)
    tip = synth.scaled_frustum(400_000, 7, True)
    job = Job(synth.ALL_PAIRS, 0.0, 0.01, 3.0, FAM_RDF | FAM_MKNN, kvalues=list(range(1, 51)))
    cl = ctx.upload(tip.xyz, tip.types, tip.ion2surf_sqr, cell_width=0.75)
    try:
        # the no-allocation entry point builds no candidate lists of its own,
        # so it counts without the symmetric pair lists: equal to a plain run
        # without them, candidate evaluations included, and to one with them
        # in every result
        a = ctx.run(cl, job)
        ctx.set_pair_symmetry(False)
        try:
            a0 = ctx.run(cl, job)
        finally:
            ctx.set_pair_symmetry(True)
        b = ctx.run_noalloc(cl, job)
        _same(a0, b)
        _same(a, b, candidates=False)
        assert b.centers_considered > 10_000
    finally:
        cl.free()
    print("noalloc: 6000-ion cube, all families", flush=True)
    rng = np.random.default_rng(5)
    xyz = rng.uniform(0, 7, size=(6000, 3)).astype(np.float32)
    ty = rng.choice([1, 2, 3], size=6000, p=[0.6, 0.3, 0.1]).astype(np.uint8)
    s2 = rng.uniform(0, 3, size=6000).astype(np.float32)
    job2 = Job(TASKS, 0.0, 0.05, 1.2, ALL, kvalues=[1, 2, 5, 17])
    cl = ctx.upload(xyz, ty, s2, cell_width=0.4)
    try:
        b2 = ctx.run_noalloc(cl, job2)
        a2 = ctx.run(cl, job2)
        _same(a2, b2)
        b3 = ctx.run_noalloc(cl, job2)  # again, on the scratch the call before left
        _same(a2, b3)
    finally:
        cl.free()
    o, st = O.spatstat(xyz, ty, s2, job2.tasks, job2.rmin, job2.rincr, job2.rmax,
                       job2.families, kvalues=job2.kvalues)
    for k, v in o.items():
        np.testing.assert_array_equal(getattr(b2, k), v, err_msg=k)


def test_noalloc_refusals(ctx):
    # a job whose temporaries exceed the scratch (about 32 bytes per ion): refused
    print("noalloc: 10M-ion cube, refused", flush=True)
    rng = np.random.default_rng(9)
    n = 10_000_000
    xyz = rng.uniform(0, 70, size=(n, 3)).astype(np.float32)
    ty = np.ones(n, np.uint8)
    s2 = np.full(n, 100.0, np.float32)
    job = Job([((1,), (1,))], 0.0, 0.05, 1.0, FAM_RDF)
    cl = ctx.upload(xyz, ty, s2, cell_width=0.5)
    try:
        with pytest.raises(PPError) as e:
            ctx.run_noalloc(cl, job)
        assert "no-allocation scratch" in str(e.value)
    finally:
        cl.free()
    # a non-resident cloud: refused until a plain run has built the list of
    # this job's width, then served from it
    print("noalloc: non-resident cloud", flush=True)
    xyz = rng.uniform(0, 5, size=(3000, 3)).astype(np.float32)
    ty = rng.choice([1, 2], size=3000).astype(np.uint8)
    s2 = np.full(3000, 100.0, np.float32)
    job = Job([((1,), (1, 2))], 0.0, 0.02, 1.0, FAM_RDF | FAM_1NN)
    cl = ctx.upload(xyz, ty, s2)
    try:
        with pytest.raises(PPError):
            ctx.run_noalloc(cl, job)
        a = ctx.run(cl, job)
        # (without the symmetric pair lists, as the no-allocation call)
        _same(a, ctx.run_noalloc(cl, job), candidates=False)
    finally:
        cl.free()
