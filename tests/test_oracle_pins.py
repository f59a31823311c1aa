"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no tests and no golden vectors, so the oracle is pinned
here by closed forms (histogram geometry quirks, fcc coordination shells,
constructed FP32 rounding cases), by brute force against the grid search, and
by invariants.  Every case is chosen so that a plausible slip in the oracle
(a dropped term, a different summation order, FMA contraction, an off-by-one
bin, a wrong sentinel, counting a center twice per task) fails.
"""
import math

import numpy as np
import pytest

from oracle import oracle

F = oracle


def _tip(n, seed, L=6.0):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0, L, size=(n, 3)).astype(np.float32)
    ty = rng.choice([1, 2, 3], size=n, p=[0.6, 0.3, 0.1]).astype(np.uint8)
    return xyz, ty


# --------------------------------------------------------------------------
# histogram geometry and binning (src/PARAPROBE_Histogram.h:54-116)

def test_nbins_truncation_quirks():
    # float(0.1) = 0.100000001 -> 2/0.100000001 = 19.99... -> 19 bins
    assert F.hist_geometry(0.0, 0.1, 2.0)[:1] == (19,)
    assert F.hist_geometry(0.0, 0.05, 2.5)[:1] == (49,)
    # the measured configs give exact counts
    assert F.hist_geometry(0.0, 0.01, 2.0)[0] == 200
    assert F.hist_geometry(0.0, 0.005, 2.0)[0] == 400
    assert F.hist_geometry(0.0, 0.01, 5.0)[0] == 500
    assert F.hist_geometry(0.0, 0.005, 3.0)[0] == 600
    # tmp = nbins / (end - start): 19 bins over 2.0 means bins of 2/19
    nb, tmp, _ = F.hist_geometry(0.0, 0.1, 2.0)
    assert tmp == pytest.approx(19 / 2.0, rel=0, abs=1e-12)


def test_category_closed_forms():
    # categories: 0 lowest, 1 + b bin b, nbins + 1 highest
    assert F.category(0.0, 0.01, 2.0, 0.0) == 1
    assert F.category(0.0, 0.01, 2.0, 2.0) == 201          # d == Rmax -> cnts_highest
    assert F.category(0.0, 0.01, 2.0, 2.5) == 201          # above end
    for b in (0, 1, 17, 99, 150, 199):
        assert F.category(0.0, 0.01, 2.0, (b + 0.5) * 0.01) == 1 + b
    # start > 0: below start -> lowest
    assert F.category(0.5, 0.01, 2.0, 0.25) == 0
    assert F.category(0.5, 0.01, 2.0, 0.505) == 1
    # 19-bin quirk: bins are 2/19 wide, so d = 1.95 is in bin 18 (1.8947..2.0)
    assert F.category(0.0, 0.1, 2.0, 1.95) == 1 + 18
    assert F.category(0.0, 0.1, 2.0, 0.11) == 1 + 1       # 0.11 * 9.5 = 1.045
    # R + EPSILON sentinel is above end
    R = np.float32(2.0)
    sent = np.float32(float(R) + 1e-6)
    assert F.category(0.0, 0.01, 2.0, float(sent)) == 201


@pytest.mark.parametrize("rincr,nbins", [(0.01, 200), (0.1, 19)])
def test_category_sweep_every_float(rincr, nbins):
    # every float d in [0, 2.02] through histogram::add: the category starts at
    # 1 (bin 0), rises by exactly one at each change, and changes only at the
    # ideal edges start + b / tmp (tmp = nbins / (end - start), so 2/19-wide
    # bins for the truncated 2.0 / 0.1 geometry), to within one float ulp; the
    # last change is to cnts_highest at d == Rmax exactly.
    first, cat = F.category_sweep(0.0, rincr, 2.0, 0.0, 2.02)
    assert F.category(0.0, rincr, 2.0, 0.0) == 1
    assert len(first) == nbins
    np.testing.assert_array_equal(cat, np.arange(2, nbins + 2))
    tmp = nbins / 2.0
    for b, t in enumerate(first[:-1].astype(np.float64), start=1):
        ulp = float(np.spacing(np.float32(t)))
        assert abs(t - b / tmp) <= ulp, (b, t)
    assert first[-1] == np.float32(2.0)
    # the float just below each change is still in the category before it
    below = np.nextafter(first, np.float32(0), dtype=np.float32)
    for c, d in zip(cat, below):
        assert F.category(0.0, rincr, 2.0, float(d)) == c - 1


# --------------------------------------------------------------------------
# neighbour set: constructed FP32 rounding cases pin the expression

def _pair_case(center, nbor, R):
    xyz = np.array([center, nbor], dtype=np.float32)
    ty = np.array([1, 1], dtype=np.uint8)
    s2 = np.array([1e6, 0.0], dtype=np.float32)  # only ion 0 is a center
    h, st = F.spatstat(xyz, ty, s2, [((1,), (1,))], 0.0, 0.5, R, F.FAM_RDF, brute=True)
    return st["inball"]


def test_summation_order_is_left_to_right():
    # (dx^2 + dy^2) + dz^2 with dx^2 = 1, dy^2 = dz^2 = 2^-24:
    # left to right gives 1 (two ties to even), dx^2 + (dy^2 + dz^2) gives 1 + 2^-23
    t = 2.0 ** -12
    assert _pair_case((0, 0, 0), (1.0, t, t), 1.0) == 1


def test_no_fma_contraction():
    # dx = 1 + 2^-12: dx^2 = 1 + 2^-11 + 2^-24 rounds (tie, even) to 1 + 2^-11;
    # plus dy^2 = 2^-24 stays 1 + 2^-11 <= fl(R*R).  An FMA would keep the
    # 2^-24 of the product and land on 1 + 2^-11 + 2^-23 > fl(R*R).
    t = 2.0 ** -12
    R = 1.0 + t
    assert np.float32(R) * np.float32(R) == np.float32(1 + 2.0 ** -11)
    assert _pair_case((0, 0, 0), (1.0 + t, t, 0.0), R) == 1


def test_self_excluded_by_index_duplicates_count():
    # coincident points: d = 0 neighbours count (self is excluded by index only)
    xyz = np.zeros((4, 3), np.float32)
    ty = np.ones(4, np.uint8)
    s2 = np.full(4, 1e6, np.float32)
    h, st = F.spatstat(xyz, ty, s2, [((1,), (1,))], 0.0, 0.1, 1.0, F.FAM_RDF, brute=True)
    assert st["inball"] == 12
    assert h["rdf"][0, 1] == 12  # bin 0


# --------------------------------------------------------------------------
# fcc lattice: closed-form coordination shells

def _fcc(nc=7, a=0.404):
    from paraprobe_amd import synth

    xyz = synth.lattice(True, a, nc, nc, nc)
    L = a * nc
    # distance to the block faces (the block hull), as ion2surf
    d = np.minimum(xyz, (L - a / 2) - xyz).min(axis=1).astype(np.float64)
    s2 = np.maximum(d, 0.0) ** 2
    return xyz, s2.astype(np.float32)


def test_fcc_shells_rdf_and_knn():
    a = 0.404
    xyz, s2 = _fcc(7, a)
    ty = np.ones(xyz.shape[0], np.uint8)
    R = 0.65
    ks = list(range(1, 80))
    h, st = F.spatstat(xyz, ty, s2, [((1,), (1,))], 0.0, 0.01, R,
                       F.FAM_RDF | F.FAM_MKNN | F.FAM_1NN | F.FAM_CNTNB, kvalues=ks)
    n = st["considered"]
    assert n > 0
    shells = [(a / math.sqrt(2), 12), (a, 6), (a * math.sqrt(1.5), 24), (a * math.sqrt(2), 12),
              (a * math.sqrt(2.5), 24)]
    rdf = h["rdf"][0]
    # float(0.65) / float(0.01) = 64.99999... -> 64 bins of 0.65 / 64 (the
    # reference's truncation quirk); bin of distance r is floor(r * 64 / 0.65)
    nb, tmp, _ = F.hist_geometry(0.0, 0.01, R)
    assert nb == 64
    for r, z in shells:
        b = int(r * tmp)
        assert rdf[1 + b] == z * n, (r, z)
    assert rdf.sum() == 78 * n
    assert st["inball"] == 78 * n
    # 1NN: every center's nearest neighbour is on the first shell
    assert h["nn1"][0, 1 + int(shells[0][0] * tmp)] == n
    # MKNN: k-th neighbour shell from the cumulative coordination 12, 18, 42, 54, 78
    kd = sorted((k - 1 for k in ks), reverse=True)
    cum = [12, 18, 42, 54, 78]
    for kth, km1 in enumerate(kd):
        k = km1 + 1
        if k > 78:  # fewer than k neighbours: R + EPSILON -> cnts_highest
            assert h["mknn"][0, kth, nb + 1] == n, k
            continue
        s = next(i for i, c in enumerate(cum) if k <= c)
        b = int(shells[s][0] * tmp)
        assert h["mknn"][0, kth, 1 + b] == n, k
    # CNTNB: every center has 78 neighbours
    assert h["cntnb"][0, 78] == n and h["cntnb"][0].sum() == n


def test_knn_sentinel_when_too_few():
    # an isolated pair: k = 1 exists, k = 2 does not -> R + EPSILON -> highest
    xyz = np.array([[0, 0, 0], [0.3, 0, 0], [10, 10, 10]], np.float32)
    ty = np.ones(3, np.uint8)
    s2 = np.full(3, 1e6, np.float32)
    h, st = F.spatstat(xyz, ty, s2, [((1,), (1,))], 0.0, 0.01, 1.0, F.FAM_MKNN | F.FAM_1NN,
                       kvalues=[1, 2])
    nb = 100
    # descending k list: [k=2, k=1]
    assert h["mknn"][0, 0, nb + 1] == 3            # k=2: all three centers hit the sentinel
    assert h["mknn"][0, 1, 1 + 30] == 2            # k=1 for the pair
    assert h["mknn"][0, 1, nb + 1] == 1            # the isolated ion
    assert h["nn1"][0, nb + 1] == 1 and h["nn1"][0, 1 + 30] == 2


# --------------------------------------------------------------------------
# brute force vs grid, invariants and task semantics

FAMS = F.FAM_RDF | F.FAM_1NN | F.FAM_RIPK | F.FAM_MKNN | F.FAM_CNTNB


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_grid_equals_brute_force(seed):
    xyz, ty = _tip(1500, seed)
    rng = np.random.default_rng(seed + 10)
    s2 = rng.uniform(0, 3, size=xyz.shape[0]).astype(np.float32)
    tasks = [((1,), (1,)), ((2,), (1, 3)), ((1, 2), (2,)), ((3,), (1, 2, 3))]
    args = (xyz, ty, s2, tasks, 0.0, 0.05, 1.2, FAMS)
    a, sa = F.spatstat(*args, kvalues=[1, 2, 5, 17], brute=True)
    b, sb = F.spatstat(*args, kvalues=[1, 2, 5, 17], brute=False)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])
    assert sa == sb


def test_pair_symmetry_without_exclusion():
    xyz, ty = _tip(3000, 5)
    s2 = np.full(xyz.shape[0], 1e6, np.float32)
    h, _ = F.spatstat(xyz, ty, s2, [((1,), (2,)), ((2,), (1,))], 0.0, 0.02, 1.0, F.FAM_RDF)
    # every A-B pair is counted once from each side with the same distance
    np.testing.assert_array_equal(h["rdf"][0], h["rdf"][1])


def test_task_sets_count_each_center_once():
    xyz, ty = _tip(2000, 7)
    s2 = np.full(xyz.shape[0], 1e6, np.float32)
    single = [((a,), (b,)) for a in (1, 2) for b in (1, 2)]
    h1, _ = F.spatstat(xyz, ty, s2, single, 0.0, 0.02, 1.0, F.FAM_RDF)
    h2, _ = F.spatstat(xyz, ty, s2, [((1, 2), (1, 2))], 0.0, 0.02, 1.0, F.FAM_RDF)
    np.testing.assert_array_equal(h1["rdf"].sum(axis=0), h2["rdf"][0])
    # RIPK feeds the same distances as RDF (S.cpp:4041-4056)
    h3, _ = F.spatstat(xyz, ty, s2, single, 0.0, 0.02, 1.0, F.FAM_RDF | F.FAM_RIPK)
    np.testing.assert_array_equal(h3["rdf"], h3["ripk"])


def test_cntnb_totals_match_rdf():
    xyz, ty = _tip(2000, 8)
    s2 = np.full(xyz.shape[0], 1e6, np.float32)
    tasks = [((1,), (2,)), ((2,), (1, 3))]
    h, st = F.spatstat(xyz, ty, s2, tasks, 0.0, 0.02, 1.0, F.FAM_RDF | F.FAM_CNTNB)
    for t in range(2):
        cnt = h["cntnb"][t]
        assert (cnt * np.arange(cnt.shape[0])).sum() == h["rdf"][t].sum()


def test_exclusion_threshold_is_inclusive():
    # considered iff ion2surf >= fl32(R*R)  (S.cpp:3859)
    xyz, ty = _tip(500, 9)
    R = np.float32(1.1)
    R2 = R * R
    s2 = np.full(xyz.shape[0], R2, np.float32)
    s2[::2] = np.nextafter(R2, np.float32(0), dtype=np.float32)
    h, st = F.spatstat(xyz, ty, s2, [((1, 2, 3), (1,))], 0.0, 0.1, float(R), F.FAM_RDF)
    assert st["considered"] == xyz.shape[0] // 2


# --------------------------------------------------------------------------
# FP64 bin-edge report (pp_oracle_binedge_report): the exact real distance of
# the two stored points, binned in double, against the FP32 pipeline

def _one_pair(nbor, rmin=0.0, rincr=0.125, rmax=2.0):
    xyz = np.array([(0.0, 0.0, 0.0), nbor], dtype=np.float32)
    ty = np.array([1, 1], np.uint8)
    s2 = np.array([1e6, 0.0], np.float32)  # only ion 0 is a center
    return F.binedge_report(xyz, ty, s2, (1,), (1,), rmin, rincr, rmax)


def test_binedge_fp32_rounds_onto_edge():
    # d = (1 - 2^-24, 2^-12, 2^-12 (1 - 2^-24)): the float squares and sums
    # round to exactly 1.0 (1 - 2^-23, then + 2^-24, then + 2^-24 - 2^-47),
    # so sqrtf gives 1.0 = the edge of bin 8 (bins of 1/8: tmp = 8), while the
    # exact d^2 = 1 - 2^-48 + 2^-72 < 1 puts the real distance in bin 7
    t = 2.0 ** -12
    p = (1.0 - 2.0 ** -24, t, t * (1.0 - 2.0 ** -24))
    assert np.float32(p[2]) == p[2]
    r = _one_pair(p)
    assert r == {"pairs": 1, "bin_differs_from_fp64": 1, "within_one_ulp_of_edge": 1,
                 "ball_membership_flips": 0}
    # the same pair 1/4 of a bin further out: no edge nearby, same bin either way
    r = _one_pair((1.0 + 2.0 ** -5, 0.0, 0.0))
    assert r == {"pairs": 1, "bin_differs_from_fp64": 0, "within_one_ulp_of_edge": 0,
                 "ball_membership_flips": 0}


def test_binedge_ball_flip_beyond_R():
    # d = (2 - 2^-23, 2^-11, 2^-11): the float sum is 4 - 2^-21 (the square's
    # 2^-46 rounded off), + 2^-22, + 2^-22 = 4.0 = fl(R^2), so the FP32 test
    # keeps the pair, while the exact d^2 = 4 + 2^-46 > R^2 puts it outside
    p = (2.0 - 2.0 ** -23, 2.0 ** -11, 2.0 ** -11)
    r = _one_pair(p)
    assert r["ball_membership_flips"] == 1
    assert r["pairs"] == 1                      # examined: inside by FP32
    assert r["bin_differs_from_fp64"] == 0      # d == Rmax and d > Rmax are both highest
    assert r["within_one_ulp_of_edge"] == 1     # at the last edge, Rmax itself
    # (2, 2^-11, 2^-11) flips too: 4 + 2^-22 is a tie that rounds to even, 4.0
    assert _one_pair((2.0, 2.0 ** -11, 2.0 ** -11))["ball_membership_flips"] == 1
    # one float ulp beyond Rmax on an axis, d^2 = 4 + 2^-20 is exact: outside either way
    r = _one_pair((2.0 + 2.0 ** -22, 0.0, 0.0))
    assert r == {"pairs": 0, "bin_differs_from_fp64": 0, "within_one_ulp_of_edge": 0,
                 "ball_membership_flips": 0}


def test_binedge_edges_follow_tmp_not_width():
    # 2.0 / 0.1 -> 19 bins of 2/19: the ideal edges are multiples of 2/19
    r = _one_pair((np.float32(10.0 / 19.0), 0.0, 0.0), 0.0, 0.1, 2.0)
    assert r["within_one_ulp_of_edge"] == 1
    r = _one_pair((0.5, 0.0, 0.0), 0.0, 0.1, 2.0)   # 0.026 from the nearest edge 10/19
    assert r["within_one_ulp_of_edge"] == 0


def test_binedge_fcc_shells_on_edges():
    # fcc lattice, a = 0.4 nm, one center at the origin: the shells within
    # R = 1 are d^2 = 0.08 m for m = 1..12 with 12, 6, 24, 12, 24, 8, 48, 6,
    # 36, 24, 24, 24 sites (248 in all).  With bins of 0.01 only the shells
    # m = 2 (d = 0.4, the six (+-0.4, 0, 0)) and m = 8 (d = 0.8, the six
    # (+-0.8, 0, 0)) sit on ideal edges, and their float coordinates put
    # them within one ulp of the edge: 12 near-edge pairs.  Their FP32 bins
    # equal the exact ones (sqrtf(fl(x^2)) = x), so nothing differs.
    g = np.arange(-6, 7)
    ijk = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    ijk = ijk[(ijk.sum(axis=1) % 2) == 0]
    xyz = (ijk * 0.2).astype(np.float32)
    order = np.argsort((ijk ** 2).sum(axis=1), kind="stable")
    xyz = xyz[order]                              # the origin first
    ty = np.ones(xyz.shape[0], np.uint8)
    s2 = np.zeros(xyz.shape[0], np.float32)
    s2[0] = 1e6
    r = F.binedge_report(xyz, ty, s2, (1,), (1,), 0.0, 0.01, 1.0)
    assert r == {"pairs": 248, "bin_differs_from_fp64": 0, "within_one_ulp_of_edge": 12,
                 "ball_membership_flips": 0}


def test_binedge_pairs_equal_oracle_inball_count():
    # examined pairs = in-ball env-type pairs of the considered centers, which
    # is the RDF total of the same one-task job; center_stride samples centers
    xyz, ty = _tip(4000, 21)
    rng = np.random.default_rng(22)
    s2 = rng.uniform(0, 3, size=xyz.shape[0]).astype(np.float32)
    h, _ = F.spatstat(xyz, ty, s2, [((1, 2), (2, 3))], 0.0, 0.01, 1.0, F.FAM_RDF)
    r = F.binedge_report(xyz, ty, s2, (1, 2), (2, 3), 0.0, 0.01, 1.0)
    assert r["pairs"] == h["rdf"][0].sum()
    keep = np.zeros(xyz.shape[0], bool)
    keep[::7] = True
    h7, _ = F.spatstat(xyz, ty, np.where(keep, s2, 0).astype(np.float32), [((1, 2), (2, 3))],
                       0.0, 0.01, 1.0, F.FAM_RDF)
    r7 = F.binedge_report(xyz, ty, s2, (1, 2), (2, 3), 0.0, 0.01, 1.0, center_stride=7)
    assert r7["pairs"] == h7["rdf"][0].sum()


def test_mknn_partial_selection_equals_full_sort():
    # the oracle settles the kmax smallest by partition when they are a small
    # part of the neighbour list, and sorts the whole list otherwise; the k-th
    # values must not depend on which (duplicates included)
    xyz, ty = _tip(3000, 31, L=5.0)
    xyz[::9] = xyz[1::9]                     # coincident pairs: ties at d = 0 and beyond
    s2 = np.full(xyz.shape[0], 1e6, np.float32)
    tasks = [((1,), (1,)), ((2, 3), (1, 2))]
    a, _ = F.spatstat(xyz, ty, s2, tasks, 0.0, 0.01, 1.5, F.FAM_MKNN, kvalues=[1, 3, 10])
    b, _ = F.spatstat(xyz, ty, s2, tasks, 0.0, 0.01, 1.5, F.FAM_MKNN,
                      kvalues=[1, 3, 10, 400])
    # descending k: a holds k = 10, 3, 1; b holds k = 400, 10, 3, 1
    np.testing.assert_array_equal(a["mknn"], b["mknn"][:, 1:])
    assert b["mknn"][:, 0, -1].sum() > 0     # k = 400 exceeds some lists: the sentinel
