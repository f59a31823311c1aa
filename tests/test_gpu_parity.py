"""GPU parity: the gfx950 path through the C ABI against the CPU oracle.

Integer histograms must be bit-identical (the contract is the reference's FP32
distance expression with fp-contract off and histogram::add binning).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from paraprobe_amd import FAM_1NN, FAM_CNTNB, FAM_MKNN, FAM_RDF, FAM_RIPK, Job
from paraprobe_amd import capi, synth

pytestmark = pytest.mark.gpu

ALL = FAM_RDF | FAM_1NN | FAM_RIPK | FAM_MKNN | FAM_CNTNB


def _say(msg):
    print(msg, flush=True)


def _cube(n, seed, L=6.0, p=(0.6, 0.3, 0.1)):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0, L, size=(n, 3)).astype(np.float32)
    ty = rng.choice([1, 2, 3], size=n, p=list(p)).astype(np.uint8)
    s2 = rng.uniform(0, 3, size=n).astype(np.float32)
    return xyz, ty, s2


def _compare(ctx, xyz, ty, s2, job: Job, cloud=None, cntnb_cap=65536):
    if cloud is None:
        cloud = ctx.upload(xyz, ty, s2)
        own = True
    else:
        own = False
    try:
        g = ctx.run(cloud, job, cntnb_cap=cntnb_cap)
    finally:
        if own:
            cloud.free()
    o, st = O.spatstat(xyz, ty, s2, job.tasks, job.rmin, job.rincr, job.rmax, job.families,
                       kvalues=job.kvalues, cntnb_cap=cntnb_cap)
    for name, ref in o.items():
        got = getattr(g, name)
        assert got is not None, name
        if not np.array_equal(got, ref):
            bad = np.argwhere(got != ref)[:8]
            raise AssertionError(f"{name} differs at {bad.tolist()}: gpu "
                                 f"{got[tuple(bad.T)].tolist()} oracle {ref[tuple(bad.T)].tolist()}")
    assert g.centers_considered == st["considered"]
    assert g.ions_discarded == st["discarded"]
    # -1: only the job's neighbour types were scanned (typed candidate list)
    assert g.pairs_inball in (st["inball"], -1)
    if all(t in {e for _, env in job.tasks for e in env} for t in set(np.unique(ty).tolist())):
        assert g.pairs_inball == st["inball"]
    return g, o, st


TASKS_MIX = [((1,), (1,)), ((2,), (1, 3)), ((1, 2), (2,)), ((3,), (1, 2, 3))]


@pytest.mark.parametrize("seed", [1, 2])
def test_small_all_families(ctx, seed):
    _say(f"test_small_all_families seed={seed}")
    xyz, ty, s2 = _cube(3000, seed)
    job = Job(TASKS_MIX, 0.0, 0.05, 1.2, ALL, kvalues=[1, 2, 5, 17])
    _compare(ctx, xyz, ty, s2, job)


@pytest.mark.parametrize("factor", [0.0, 0.3, 3.0, 50.0])
def test_knn_range_does_not_change_results(ctx, factor):
    """kNN jobs keep per-center histograms over a range where factor * k
    neighbours are expected, re-running the centers short of k there; 0 means
    full-range histograms.  Every factor must give the oracle's histograms;
    0.3 forces many re-runs on this clustered tip."""
    _say(f"test_knn_range factor={factor}")
    base = [0.85, 0.12, 0.03]
    tip = synth.tip(60_000, 5, "frustum", 12.0, 7.0, 35.0, base, nclusters=250, crmin=1.0,
                    crmax=2.0, cfrac=synth.cluster_fracs(base))
    jobs = [Job(synth.ALL_PAIRS + [((1, 2), (2, 3))], 0.0, 0.02, 2.5,
                FAM_RDF | FAM_MKNN | FAM_1NN, kvalues=list(range(1, 41))),
            Job(synth.ALL_PAIRS, 0.1, 0.01, 2.0, FAM_1NN)]
    ctx.set_knn_range(factor)
    try:
        for job in jobs:
            g, _, _ = _compare(ctx, tip.xyz, tip.types, tip.ion2surf_sqr, job)
            if factor == 0.0:
                assert g.percenter == 1
            elif factor == 0.3 and job.families & FAM_MKNN:
                assert g.percenter == 2 and g.centers_rerun > 0
    finally:
        ctx.set_knn_range(4.0)


@pytest.mark.parametrize("case", ["clustered_frustum", "uniform_random_ion2surf"])
def test_pair_symmetry_does_not_change_results(ctx, case):
    """Symmetric pair counting (the dominant center type's tile counts a pair
    of two of its centers for both when the other center's cell comes later
    in cell order, and reads the earlier cells only where they hold non-centers
    or lie in its kNN range) against the oracle and against the same jobs with
    it switched off.  On the frustum most rows before the own cell hold
    centers only; with ion2surf uniform in [0, 3) nm^2 every row mixes
    centers and non-centers."""
    _say(f"test_pair_symmetry {case}")
    if case == "clustered_frustum":
        base = [0.85, 0.12, 0.03]
        tip = synth.tip(60_000, 5, "frustum", 12.0, 7.0, 35.0, base, nclusters=250, crmin=1.0,
                        crmax=2.0, cfrac=synth.cluster_fracs(base))
        xyz, ty, s2 = tip.xyz, tip.types, tip.ion2surf_sqr
    else:
        xyz, ty, _ = _cube(40_000, 5, L=9.0)
        s2 = np.random.default_rng(6).uniform(0, 12, size=len(ty)).astype(np.float32)
    # type 1 dominates: its tiles count symmetrically, reading type 1 from
    # its own list and (on the frustum, and in the last job) types 2 and 3
    # from another; type 2's tiles (second job) count as usual
    jobs = [Job([((1,), (1,))], 0.0, 0.02, 2.5, FAM_RDF | FAM_MKNN,
                kvalues=list(range(1, 11))),
            Job([((1,), (1,)), ((2,), (1,))], 0.0, 0.01, 1.5, FAM_RDF | FAM_RIPK | FAM_MKNN,
                kvalues=[1, 3, 10]),
            Job([((1,), (1, 2))], 0.0, 0.02, 2.0, FAM_RDF | FAM_1NN),
            # pair counts only (the kernel's direct mode)
            Job([((1,), (1,))], 0.0, 0.02, 2.5, FAM_RDF),
            Job([((1,), (1,)), ((3,), (1, 3))], 0.05, 0.01, 1.5, FAM_RDF | FAM_RIPK)]
    # the switch takes 0 or 1 only
    assert ctx.lib.pp_set_pair_symmetry(ctx.ctx, 2) == 1  # PP_ERR_INVALID_ARG
    # (a kNN range of about k neighbours keeps these small clouds' kNN jobs
    # in the hybrid mode; the symmetric counting applies there and in the
    # pair-count mode)
    ctx.set_knn_range(1.0)
    try:
        for job in jobs:
            g, _, _ = _compare(ctx, xyz, ty, s2, job)
            ctx.set_pair_symmetry(False)
            try:
                h, _, _ = _compare(ctx, xyz, ty, s2, job)
            finally:
                ctx.set_pair_symmetry(True)
            _check_symmetry(case, g, h)
    finally:
        ctx.set_knn_range(4.0)


def _check_symmetry(case, g, h):
    assert g.percenter == h.percenter and g.percenter in (0, 2)
    # fewer candidate evaluations with it on where rows of centers only
    # exist (elsewhere about as many: the own-type and other-type lists trim
    # their rows' end cells at their own sub-cell granularity)
    if case == "clustered_frustum":
        assert g.pairs_candidate < h.pairs_candidate, (g.pairs_candidate, h.pairs_candidate)
    _say(f"  candidates {g.pairs_candidate} symmetric vs {h.pairs_candidate}")


@pytest.mark.parametrize("T,r", [(8, 1), (16, 2), (32, 3), (16, 4), (8, 2), (32, 1), (8, 6),
                                 (8, 8)])
def test_tuning_does_not_change_results(ctx, T, r):
    _say(f"test_tuning T={T} cells_per_rmax={r}")
    xyz, ty, s2 = _cube(4000, 11)
    ctx.set_tuning(T, r)
    try:
        job = Job(TASKS_MIX, 0.0, 0.02, 1.0, ALL, kvalues=[1, 3, 9, 30])
        _compare(ctx, xyz, ty, s2, job)
        job2 = Job(TASKS_MIX, 0.0, 0.02, 1.0, FAM_RDF)
        _compare(ctx, xyz, ty, s2, job2)
    finally:
        ctx.set_tuning(0, 0)


@pytest.mark.parametrize("r,ay", [(1, 2), (2, 2), (2, 4), (3, 3)])
def test_cell_aspect_does_not_change_results(ctx, r, ay):
    # cells finer along y and z than along x (ay y/z cells per x cell width)
    _say(f"test_cell_aspect cells_per_rmax={r} yz_per_x={ay}")
    xyz, ty, s2 = _cube(4000, 12)
    ctx.set_tuning(8, r)
    ctx.set_cell_aspect(ay)
    try:
        _compare(ctx, xyz, ty, s2, Job(TASKS_MIX, 0.0, 0.02, 1.0, ALL, kvalues=[1, 3, 9, 30]))
        _compare(ctx, xyz, ty, s2, Job(TASKS_MIX, 0.0, 0.02, 1.0, FAM_RDF))
    finally:
        ctx.set_cell_aspect(0)
        ctx.set_tuning(0, 0)


def test_density_rule_finest_cells(ctx):
    # 125 ions/nm^3 and Rmax 2.6 nm: the density rule picks its finest width,
    # Rmax/4, which it takes as Rmax/2 along x and Rmax/4 along y and z (the
    # other tests' jobs at Rmax/2 take Rmax by Rmax/2)
    _say("test_density_rule_finest_cells")
    xyz, ty, s2 = _cube(8000, 13, L=4.0)
    s2 = s2 * 4.0  # squared ion2surf up to 12 nm^2: 44 % of the ions are centers
    job = Job([((1,), (1, 2)), ((2, 3), (1,))], 0.0, 0.05, 2.6, ALL, kvalues=[1, 10, 100])
    _compare(ctx, xyz, ty, s2, job)
    _compare(ctx, xyz, ty, s2, Job([((1,), (1, 2))], 0.0, 0.05, 2.6, FAM_RDF | FAM_MKNN,
                                   kvalues=[1, 50]))


def test_rmin_positive_and_truncated_bins(ctx):
    _say("test_rmin_positive_and_truncated_bins")
    xyz, ty, s2 = _cube(3000, 3)
    # 2.0/0.1 -> 19 bins of 2/19; start 0.3 gives a populated lowest category
    _compare(ctx, xyz, ty, s2, Job(TASKS_MIX, 0.0, 0.1, 2.0, ALL, kvalues=[1, 4]))
    _compare(ctx, xyz, ty, s2, Job(TASKS_MIX, 0.3, 0.01, 1.5, ALL, kvalues=[2, 6]))


def test_duplicates_and_sentinels(ctx):
    _say("test_duplicates_and_sentinels")
    rng = np.random.default_rng(4)
    base = rng.uniform(0, 3, size=(200, 3)).astype(np.float32)
    xyz = np.concatenate([base, base, base[:50]])  # coincident points: d = 0 pairs
    ty = rng.choice([1, 2], size=xyz.shape[0]).astype(np.uint8)
    s2 = np.full(xyz.shape[0], 100.0, np.float32)
    job = Job([((1,), (1,)), ((2,), (1, 2))], 0.0, 0.01, 0.6, ALL, kvalues=[1, 2, 3, 50, 200])
    _compare(ctx, xyz, ty, s2, job)


def test_constructed_fp32_cases(ctx):
    _say("test_constructed_fp32_cases")
    t = 2.0 ** -12
    # summation order and no-FMA cases (see the oracle pins)
    for nb, R in (((1.0, t, t), 1.0), ((1.0 + t, t, 0.0), 1.0 + t)):
        xyz = np.array([(0, 0, 0), nb], np.float32)
        ty = np.array([1, 1], np.uint8)
        s2 = np.array([1e6, 0.0], np.float32)
        g, o, st = _compare(ctx, xyz, ty, s2, Job([((1,), (1,))], 0.0, 0.5, R, FAM_RDF))
        assert st["inball"] == 1 and g.pairs_inball == 1 and g.rdf.sum() == 1


def test_fcc_lattice_closed_form(ctx):
    _say("test_fcc_lattice_closed_form")
    a = 0.404
    xyz = synth.lattice(True, a, 9, 9, 9)
    L = a * 9
    d = np.minimum(xyz, (L - a / 2) - xyz).min(axis=1).astype(np.float64)
    s2 = (np.maximum(d, 0) ** 2).astype(np.float32)
    ty = np.ones(xyz.shape[0], np.uint8)
    g, o, st = _compare(ctx, xyz, ty, s2, Job([((1,), (1,))], 0.0, 0.01, 0.65, ALL,
                                               kvalues=list(range(1, 90))))
    assert g.pairs_inball == 78 * g.centers_considered


def test_edge_cases(ctx):
    _say("test_edge_cases")
    # empty cloud
    e = np.zeros((0, 3), np.float32)
    g, _, _ = _compare(ctx, e, np.zeros(0, np.uint8), np.zeros(0, np.float32),
                       Job([((1,), (1,))], 0.0, 0.1, 1.0, ALL, kvalues=[1]))
    assert g.centers_considered == 0
    # one ion
    one = np.array([[1, 2, 3]], np.float32)
    _compare(ctx, one, np.array([1], np.uint8), np.array([10], np.float32),
             Job([((1,), (1,))], 0.0, 0.1, 1.0, ALL, kvalues=[1, 2]))
    # every center excluded
    xyz, ty, s2 = _cube(2000, 5)
    g, _, _ = _compare(ctx, xyz, ty, np.zeros_like(s2), Job(TASKS_MIX, 0.0, 0.05, 1.0, ALL,
                                                             kvalues=[1]))
    assert g.centers_considered == 0
    # a type with no ions at all, and a flat (2-D) cloud
    xyz2 = xyz.copy()
    xyz2[:, 2] = 1.0
    _compare(ctx, xyz2, ty, s2, Job([((7,), (1,)), ((1,), (7,)), ((2,), (1, 2, 3))], 0.0, 0.05,
                                    0.8, ALL, kvalues=[1, 2]))


def test_invalid_jobs_raise(ctx):
    from paraprobe_amd import PPError

    xyz, ty, s2 = _cube(100, 6)
    cl = ctx.upload(xyz, ty, s2)
    try:
        with pytest.raises(PPError):
            ctx.run(cl, Job([((1,), (1,))], 0.0, 1e-7, 1.0, FAM_RDF))
        with pytest.raises(PPError):
            ctx.run(cl, Job([((1,), (1,))], 1.0, 0.1, 0.5, FAM_RDF))
        with pytest.raises(PPError):
            ctx.run(cl, Job([((1,), (1,))], 0.0, 0.1, 1.0, 1 << 4))  # NPCORR
        with pytest.raises(PPError):
            ctx.run(cl, Job([((1,), (1,))], 0.0, 0.1, 1.0, 0))
    finally:
        cl.free()


def test_cylinder_1m_rdf_and_knn(ctx):
    """The CPU-runnable workload: 1M ions in a cylinder, B-B RDF (R=2, 0.01)
    and B-B kNN k=1..10 (bin 0.005)."""
    _say("test_cylinder_1m generate")
    tip = synth.cylinder_1m()
    cl = ctx.upload(tip.xyz, tip.types, tip.ion2surf_sqr)
    try:
        _say("test_cylinder_1m rdf")
        g, o, st = _compare(ctx, tip.xyz, tip.types, tip.ion2surf_sqr,
                            Job([((2,), (2,))], 0.0, 0.01, 2.0, FAM_RDF), cloud=cl)
        assert 50_000 < g.centers_considered < 62_000
        _say("test_cylinder_1m knn")
        _compare(ctx, tip.xyz, tip.types, tip.ion2surf_sqr,
                 Job([((2,), (2,))], 0.0, 0.005, 2.0, FAM_MKNN, kvalues=list(range(1, 11))),
                 cloud=cl)
    finally:
        cl.free()


def test_batch_equals_single_runs(ctx):
    _say("test_batch_equals_single_runs")
    clouds, jobs, raw = [], [], []
    specs = [(2500, 21, Job(TASKS_MIX, 0.0, 0.05, 1.2, ALL, kvalues=[1, 2, 5])),
             (1800, 22, Job([((2,), (3,))], 0.0, 0.01, 0.7, FAM_RDF)),
             (3000, 23, Job([((1,), (1, 2))], 0.1, 0.02, 1.5, FAM_MKNN | FAM_1NN,
                            kvalues=[1, 10, 50]))]
    for n, seed, job in specs:
        xyz, ty, s2 = _cube(n, seed)
        raw.append((xyz, ty, s2))
        clouds.append(ctx.upload(xyz, ty, s2))
        jobs.append(job)
    try:
        outs = ctx.run_batch(clouds, jobs)
        for (xyz, ty, s2), c, j, gb in zip(raw, clouds, jobs, outs):
            gs = ctx.run(c, j)
            for k, v in gs.arrays().items():
                np.testing.assert_array_equal(getattr(gb, k), v)
            o, st = O.spatstat(xyz, ty, s2, j.tasks, j.rmin, j.rincr, j.rmax, j.families,
                               kvalues=j.kvalues)
            for k, v in o.items():
                np.testing.assert_array_equal(getattr(gb, k), v)
    finally:
        for c in clouds:
            c.free()


def test_resident_cell_list_any_rmax(ctx):
    """A cloud uploaded with a fixed cell width serves jobs of any Rmax."""
    _say("test_resident_cell_list_any_rmax")
    xyz, ty, s2 = _cube(5000, 31, L=8.0)
    cl = ctx.upload(xyz, ty, s2, cell_width=0.45)
    try:
        for R in (0.3, 0.9, 1.7):
            _compare(ctx, xyz, ty, s2, Job(TASKS_MIX, 0.0, 0.01, R, ALL, kvalues=[1, 4]), cloud=cl)
    finally:
        cl.free()


EDGE_GEOMS = [(0.0, 0.01, 5.0), (0.0, 0.1, 2.0), (0.3, 0.01, 1.5), (0.0, 0.005, 3.0),
              (4.8, 0.002, 5.0)]
# query-kernel modes: RDF alone runs the pair-count mode (0); RDF + MKNN over
# the full range the per-center mode (1); a kNN range of 0.3 expected
# neighbours for k = 1 (factor 0.07 x (1 + 6 + 10) / 4) the hybrid mode (2),
# whose short centers re-run in mode 1
EDGE_MODES = {"paircount": (FAM_RDF, 3.0, 0), "percenter": (FAM_RDF | FAM_MKNN, 0.0, 1),
              "hybrid": (FAM_RDF | FAM_MKNN, 0.07, 2)}


@pytest.mark.parametrize("mode", list(EDGE_MODES))
@pytest.mark.parametrize("rmin,rincr,rmax", EDGE_GEOMS)
def test_distances_packed_around_every_bin_edge(ctx, rmin, rincr, rmax, mode):
    """Isolated pairs whose distances lie within 40 float steps of every bin
    edge (the ideal edge positions start + c / tmp): the category of each
    pair, near-edge path included, must equal the literal expression's, in
    every kernel mode (the floor form with its near-edge fallback in modes 0
    and 2, the nearest-edge form in mode 1).  4.8 / 0.002 / 5.0 has the
    widest margin (start far from 0, 500 bins per nm)."""
    fams, factor, want = EDGE_MODES[mode]
    if want == 2 and rmin > 1.0:
        pytest.skip("the kNN range of these sparse pairs ends below rmin: no hybrid plan")
    _say(f"bin edges {rmin}/{rincr}/{rmax} mode {mode}")
    nb = capi.hist_geometry(rmin, rincr, rmax)
    r = np.array([rmin + c * (rmax - rmin) / nb for c in range(1, nb + 1)], np.float32)
    # k float steps from each edge: positive floats order like their bit patterns
    bits = r.view(np.int32)[:, None] + np.arange(-40, 41, dtype=np.int32)[None, :]
    xs = bits.reshape(-1).view(np.float32)
    xs = xs[(xs > 0) & (xs <= np.float32(rmax))]
    n = xs.shape[0]
    side = int(np.ceil(np.sqrt(n)))
    sp = np.float32(2.0 * (rmax + 1.0))
    j, k = np.divmod(np.arange(n), side)
    a = np.stack([np.zeros(n, np.float32), j * sp, k * sp], axis=1).astype(np.float32)
    b = a.copy()
    b[:, 0] = xs
    xyz = np.concatenate([a, b])
    ty = np.ones(2 * n, np.uint8)
    s2 = np.full(2 * n, 1e6, np.float32)
    ctx.set_knn_range(factor)
    try:
        g, _, st = _compare(ctx, xyz, ty, s2, Job([((1,), (1,))], rmin, rincr, rmax, fams,
                                                   kvalues=[1] if fams & FAM_MKNN else []))
    finally:
        ctx.set_knn_range(4.0)
    assert st["inball"] == 2 * n
    assert g.percenter == want, (g.percenter, want)
    if want == 2:
        assert g.centers_rerun > 0


def test_headline_parameters_clustered_frustum(ctx):
    """The headline job's parameters on a 400k-ion clustered frustum of the
    headline shape: all 9 ordered type pairs, Rmax 5 nm, bin 0.01 nm, RDF +
    MKNN k = 1..100, bit-exact against the oracle."""
    _say("headline parameters: 400k-ion clustered frustum, 9 pairs, R 5, k 1..100")
    tip = synth.scaled_frustum(400_000, 7, True)
    job = Job(synth.ALL_PAIRS, 0.0, 0.01, 5.0, FAM_RDF | FAM_MKNN, kvalues=list(range(1, 101)))
    g, _, st = _compare(ctx, tip.xyz, tip.types, tip.ion2surf_sqr, job)
    assert st["considered"] > 50_000 and g.pairs_inball == st["inball"]
    _say(f"  kernel mode {g.percenter}, {g.centers_rerun} centers rerun")


def test_large_configuration_shape(ctx):
    """The large configuration's job (9 pair RDFs at Rmax 3 nm, bin 0.005 nm,
    600 bins, plus MKNN k = 1..20) on its tip shape (frustum r 60 -> 30 nm,
    22 ions/nm^3, 300,000 clusters per 100M ions) scaled to 300k ions."""
    _say("large configuration shape: 300k ions, 9 pairs, R 3, bin 0.005, k 1..20")
    tip = synth.frustum_100m(n=300_000)
    job = Job(synth.ALL_PAIRS, 0.0, 0.005, 3.0, FAM_RDF | FAM_MKNN, kvalues=list(range(1, 21)))
    g, _, st = _compare(ctx, tip.xyz, tip.types, tip.ion2surf_sqr, job)
    assert g.nbins == 600 and st["considered"] > 80_000


def test_sparse_cloud_in_a_large_box(ctx):
    """Clusters far apart in a large box: the grid is capped near 1M cells,
    where the cell list keeps each cell's ions unordered (no octant order),
    next to a dense cloud whose cells are octant-ordered; both match the
    oracle."""
    rng = np.random.default_rng(17)
    parts = [rng.normal(loc=c, scale=0.6, size=(400, 3))
             for c in rng.uniform(0, 400, size=(6, 3))]
    xyz = np.concatenate(parts).astype(np.float32)
    ty = rng.choice([1, 2], size=xyz.shape[0], p=[0.7, 0.3]).astype(np.uint8)
    s2 = np.full(xyz.shape[0], 50.0, np.float32)
    job = Job([((1,), (1, 2)), ((2,), (1,))], 0.0, 0.05, 1.0, FAM_RDF | FAM_MKNN | FAM_1NN,
              kvalues=[1, 2, 5])
    _say("sparse cloud in a 400 nm box")
    _compare(ctx, xyz, ty, s2, job)
    xyz, ty, s2 = _cube(8000, 5, L=6.0)
    _say("dense cube")
    _compare(ctx, xyz, ty, s2, job)
