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


def test_category_monotone_and_edges_dense():
    # exhaustive over every float d in [0, 2.02): category is monotone and
    # changes exactly at the bin edges start + b / tmp (to within one ulp)
    lo = np.float32(0.0).view(np.uint32)
    hi = np.float32(2.02).view(np.uint32)
    u = np.arange(lo, hi, 97, dtype=np.uint32)  # stride keeps it fast; edges checked below
    d = u.view(np.float32)
    cats = F.categories_d2(0.0, 0.01, 2.0, (d * d).astype(np.float32))
    assert np.all(np.diff(cats) >= 0)
    for b in range(1, 200):
        e = np.float32(b / 100.0)
        below = np.nextafter(e, np.float32(0), dtype=np.float32)
        above = np.nextafter(e, np.float32(3), dtype=np.float32)
        cb = F.category(0.0, 0.01, 2.0, float(below))
        ca = F.category(0.0, 0.01, 2.0, float(above))
        assert cb in (b, b + 1) and ca in (b + 1, b + 2)
        assert ca - cb >= 1 - 1 and cb <= b + 1


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
