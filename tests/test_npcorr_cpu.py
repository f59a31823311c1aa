"""Pins of the NPCORR oracle (n-point correlations, S.cpp:3899-3960,
npc3d src/PARAPROBE_NPointCorrelations3D.cpp:50-207) without a GPU: the
histogram geometry, hand-computed bins of constructed neighbourhoods (rounding
half away from zero), the K-th pick, the neighbour-count conditions, and the
total count against an independent neighbour search (scipy's KD-tree)."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from oracle import oracle as O
from paraprobe_amd import capi


def test_geometry_matches_npc3d():
    g = capi.load_library().pp_npcorr_geometry
    assert g(2.5, 0.05) == 103     # the reference's settings fixture: ceil(50) + 1 = 51
    assert g(2.0, 0.25) == 19      # ceil(8) + 1 = 9
    assert g(1.0, 0.3) == 11       # ceil(3.33) + 1 = 5
    assert g(5.0, 0.02) == 503     # 251 * 2 + 1, the largest allowed edge
    assert g(5.0, 0.01) == -1      # 1003^3 > 503^3: the reference skips
    assert g(0.0, 0.1) == -1


def _one_center(neigh, K, width=0.25, rmax=2.0, s2=100.0, types=None):
    xyz = np.array([[0, 0, 0]] + neigh, np.float32)
    ty = np.array(types if types is not None else [1] + [2] * len(neigh), np.uint8)
    s = np.full(len(xyz), s2, np.float32)
    s[1:] = 0.0  # only the first ion is a considered center
    return O.npcorr(xyz, ty, s, [((1,), (2,))], rmax, width, K)


def _bin(counts):
    nz = np.argwhere(counts[0])
    assert nz.shape[0] == 1 and counts[0][tuple(nz[0])] == 1
    half = (counts.shape[1] - 1) // 2
    z, y, x = nz[0] - half
    return int(x), int(y), int(z)


def test_hand_computed_bins():
    # width 0.25 -> fl(1 / 0.25) = 4 exactly; components chosen on exact floats
    neigh = [[0.5, -0.375, 0.125], [1.0, 1.0, 1.0], [-1.25, 0.0, 0.0]]
    # K = 0: the nearest, |(0.5, -0.375, 0.125)| = 0.637: t = (2, -1.5, 0.5)
    assert _bin(_one_center(neigh, 0)) == (2, -2, 1)      # halves go away from zero
    # K = 1: (-1.25, 0, 0) at 1.25 -> t = -5
    assert _bin(_one_center(neigh, 1)) == (-5, 0, 0)
    # K = 2: (1, 1, 1) at 1.732 -> t = (4, 4, 4)
    assert _bin(_one_center(neigh, 2)) == (4, 4, 4)


def test_k_conditions_and_exclusions():
    neigh = [[0.5, 0.0, 0.0], [0.0, 0.75, 0.0]]
    # only one env neighbour: K = 1 needs two
    c = _one_center(neigh, 1, types=[1, 2, 3])
    assert c.sum() == 0
    # more than K ions of any type, but K + 1 env neighbours are not there
    c = _one_center(neigh + [[0, 0, 0.25]], 1, types=[1, 2, 3, 3])
    assert c.sum() == 0
    # beyond R: not a neighbour
    c = _one_center([[2.5, 0, 0]], 0)
    assert c.sum() == 0
    # a center with ion2surf below R^2 is not considered
    c = _one_center(neigh, 0, s2=3.99)
    assert c.sum() == 0
    c = _one_center(neigh, 0, s2=4.0)
    assert _bin(c) == (2, 0, 0)


def test_ties_take_the_smallest_vector():
    # six neighbours at the same distance and squared distance: (-1, 0, 0) is
    # the smallest (u, v, w)
    neigh = [[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1]]
    assert _bin(_one_center(neigh, 0)) == (-4, 0, 0)
    assert _bin(_one_center(neigh, 1)) == (0, -4, 0)


def test_equal_distances_take_the_smaller_square_first():
    # two neighbours whose float lengths sqrtf(d2) are equal while their
    # squared lengths differ by one ulp: the smaller d2 comes first although
    # the other has the smaller u
    f = np.float32
    a = [f("0.380425"), f("0.5418678"), f(0)]
    b = [f("-0.380425"), f("0.54186785"), f(0)]
    d2a = f(f(a[0] * a[0]) + f(a[1] * a[1]))
    d2b = f(f(b[0] * b[0]) + f(b[1] * b[1]))
    assert d2a < d2b and np.sqrt(d2a) == np.sqrt(d2b)
    assert _bin(_one_center([a, b], 0)) == (2, 2, 0)
    assert _bin(_one_center([a, b], 1)) == (-2, 2, 0)


def test_total_counts_against_kdtree():
    rng = np.random.default_rng(3)
    n = 3000
    xyz = rng.uniform(0, 6, size=(n, 3)).astype(np.float32)
    ty = rng.choice([1, 2, 3], size=n, p=[0.5, 0.3, 0.2]).astype(np.uint8)
    s2 = rng.uniform(0, 3, size=n).astype(np.float32)
    R, K = 1.0, 2
    tasks = [((1,), (1, 2)), ((2, 3), (3,))]
    got = O.npcorr(xyz, ty, s2, tasks, R, 0.1, K)
    tree = cKDTree(xyz.astype(np.float64))
    want = [0, 0]
    for i in range(n):
        if s2[i] < np.float32(R) * np.float32(R):
            continue
        nb = [j for j in tree.query_ball_point(xyz[i], R * (1 - 1e-5)) if j != i]
        if len(nb) <= K:
            continue
        for t, (trg, env) in enumerate(tasks):
            if ty[i] in trg and sum(1 for j in nb if ty[j] in env) > K:
                want[t] += 1
    # points within 1e-5 of R can flip either way in float; none at this seed
    assert [int(got[t].sum()) for t in range(2)] == want


def test_cpu_baseline_equals_oracle():
    """The reference-structured CPU baseline (z-slabs, KD-trees, nth_element)
    gives the oracle's counts at 1 and 4 threads."""
    from oracle import cpu_baseline as cb

    rng = np.random.default_rng(4)
    n = 4000
    xyz = rng.uniform(0, 6, size=(n, 3)).astype(np.float32)
    ty = rng.choice([1, 2], size=n, p=[0.7, 0.3]).astype(np.uint8)
    s2 = rng.uniform(0, 4, size=n).astype(np.float32)
    tasks = [((1,), (1, 2)), ((2,), (2,))]
    want = O.npcorr(xyz, ty, s2, tasks, 1.3, 0.1, 5)
    for th in (1, 4):
        got, _, _ = cb.npcorr(xyz, ty, s2, tasks, 1.3, 0.1, 5, nthreads=th)
        assert np.array_equal(got, want)
