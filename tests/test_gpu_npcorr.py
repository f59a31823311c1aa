"""GPU parity of NPCORR (n-point correlations) against the brute-force oracle:
identical uint32 counts for random and clustered clouds, several K, tasks and
bin widths, and a lattice where every pick is a tie."""
import numpy as np
import pytest

from oracle import oracle as O
from paraprobe_amd import synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("K,width,R", [(0, 0.05, 1.0), (3, 0.1, 1.2), (40, 0.2, 1.5),
                                       (300, 0.1, 2.5)])
def test_npcorr_matches_oracle(ctx, K, width, R):
    rng = np.random.default_rng(K + 1)
    n = 6000
    xyz = rng.uniform(0, 7, size=(n, 3)).astype(np.float32)
    ty = rng.choice([1, 2, 3], size=n, p=[0.6, 0.3, 0.1]).astype(np.uint8)
    s2 = rng.uniform(0, 8, size=n).astype(np.float32)
    tasks = [((1,), (1,)), ((1, 2), (3,)), ((3,), (1, 2, 3))]
    _say(f"npcorr K={K} width={width} R={R}")
    cloud = ctx.upload(xyz, ty, s2)
    try:
        got = ctx.npcorr(cloud, tasks, R, width, K)
    finally:
        cloud.free()
    want = O.npcorr(xyz, ty, s2, tasks, R, width, K)
    assert got.shape == want.shape
    assert np.array_equal(got, want)
    assert got.sum() > 0


def test_npcorr_lattice_ties_and_clusters(ctx):
    a = 0.404
    xyz = synth.lattice(True, a, 9, 9, 9)
    n = xyz.shape[0]
    ty = np.ones(n, np.uint8)
    ty[::7] = 2
    s2 = np.full(n, 100.0, np.float32)
    tasks = [((1,), (1,)), ((2,), (1, 2))]
    for K in (0, 11, 12):
        _say(f"npcorr fcc K={K}")
        cloud = ctx.upload(xyz, ty, s2)
        try:
            got = ctx.npcorr(cloud, tasks, 1.0, 0.05, K)
        finally:
            cloud.free()
        assert np.array_equal(got, O.npcorr(xyz, ty, s2, tasks, 1.0, 0.05, K))
    t = synth.tip(8000, 5, "cylinder", 6.0, 6.0, 10.0, [0.9, 0.1], 20, 0.5, 1.0, [0.4, 0.6])
    _say("npcorr clustered")
    cloud = ctx.upload(t.xyz, t.types, t.ion2surf_sqr)
    try:
        got = ctx.npcorr(cloud, [((2,), (2,)), ((1,), (2,))], 1.5, 0.1, 2)
    finally:
        cloud.free()
    assert np.array_equal(got, O.npcorr(t.xyz, t.types, t.ion2surf_sqr,
                                        [((2,), (2,)), ((1,), (2,))], 1.5, 0.1, 2))


def test_npcorr_through_the_host_driver(ctx, tmp_path):
    """AnalysisSpatDistrType 15: RDF files as before, NPCORR counts in memory;
    the RDF must not change when NPCORR rides along."""
    from paraprobe_amd import horderdist as H

    rng = np.random.default_rng(9)
    n = 5000
    xyz = rng.uniform(0, 7, size=(n, 3)).astype(np.float32)
    ty = rng.choice([1, 2], size=n, p=[0.7, 0.3]).astype(np.uint8)
    s2 = rng.uniform(0, 5, size=n).astype(np.float32)
    xml = """<Parameters>
  <AnalysisSpatDistrType>{}</AnalysisSpatDistrType>
  <DescrStatTasksCode>Al-Al;Sc-Al</DescrStatTasksCode>
  <SpatStatRadiusMin>0.0</SpatStatRadiusMin>
  <SpatStatRadiusIncr>0.1</SpatStatRadiusIncr>
  <SpatStatRadiusMax>1.2</SpatStatRadiusMax>
  <SpatStatKNNOrder>2</SpatStatKNNOrder>
</Parameters>"""
    cl = ctx.upload(xyz, ty, s2)
    try:
        for code, sub in (("1", "a"), ("15", "b")):
            d = tmp_path / sub
            d.mkdir()
            _say(f"host driver AnalysisSpatDistrType={code}")
            assert H.run_spatstat(ctx, cl, xml.format(code), 1, {"Al": 1, "Sc": 2}, n, 343.0,
                                  str(d)) == 0
        for f in sorted(p.name for p in (tmp_path / "a").iterdir()):
            assert (tmp_path / "a" / f).read_bytes() == (tmp_path / "b" / f).read_bytes()
        got = ctx.npcorr(cl, [((1,), (1,)), ((2,), (1,))], 1.2, 0.1, 1)
    finally:
        cl.free()
    assert np.array_equal(got, O.npcorr(xyz, ty, s2, [((1,), (1,)), ((2,), (1,))], 1.2, 0.1, 1))


def test_npcorr_crowded_bucket(ctx):
    """More neighbours in one distance bucket than the wave's list holds:
    700 ions stacked on three points at the same distance from a few centers."""
    rng = np.random.default_rng(11)
    pts = rng.uniform(0, 5, size=(2000, 3)).astype(np.float32)
    stack = np.repeat(np.array([[2.5, 2.5, 3.0], [2.5, 3.0, 2.5], [3.0, 2.5, 2.5]],
                               np.float32), [300, 200, 200], axis=0)
    centers = np.array([[2.5, 2.5, 2.5], [2.6, 2.5, 2.5]], np.float32)
    xyz = np.concatenate([centers, stack, pts])
    ty = np.full(xyz.shape[0], 2, np.uint8)
    ty[:2] = 1
    s2 = np.full(xyz.shape[0], 50.0, np.float32)
    tasks = [((1,), (2,))]
    for K in (0, 250, 400, 447):
        _say(f"npcorr crowded K={K}")
        cloud = ctx.upload(xyz, ty, s2)
        try:
            got = ctx.npcorr(cloud, tasks, 1.0, 0.05, K)
        finally:
            cloud.free()
        want = O.npcorr(xyz, ty, s2, tasks, 1.0, 0.05, K)
        assert np.array_equal(got, want)
        assert got.sum() == 2


def test_npcorr_bucket_edges_and_many_tasks(ctx):
    """Neighbours at the distances that separate the kernel's distance buckets
    (k R / 256, along the axes and along random directions), so that the
    shell each pick is searched in is cut exactly at its candidates; and six
    tasks per center, more than one counting sweep serves."""
    rng = np.random.default_rng(21)
    R = 1.0
    centers = rng.uniform(2.0, 4.0, size=(6, 3)).astype(np.float32)
    parts = [centers]
    for c in centers:
        d = (np.arange(1, 256, dtype=np.float64) * (R / 256.0))
        axis = np.zeros((d.size, 3))
        axis[np.arange(d.size), np.arange(d.size) % 3] = np.where(np.arange(d.size) % 2, d, -d)
        u = rng.normal(size=(d.size, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        parts.append((c + axis).astype(np.float32))
        parts.append((c + u * d[:, None]).astype(np.float32))
    parts.append(rng.uniform(0, 6, size=(3000, 3)).astype(np.float32))
    xyz = np.concatenate(parts)
    ty = rng.choice([1, 2, 3], size=xyz.shape[0], p=[0.5, 0.3, 0.2]).astype(np.uint8)
    ty[:6] = 1
    s2 = np.full(xyz.shape[0], 50.0, np.float32)
    s2[6:] = rng.uniform(0, 2, size=xyz.shape[0] - 6).astype(np.float32)
    tasks = [((1,), (1,)), ((1,), (2,)), ((1,), (3,)), ((1,), (1, 2)), ((1, 2), (2, 3)),
             ((1,), (1, 2, 3))]
    for K in (0, 5, 60):
        _say(f"npcorr bucket edges K={K}")
        cloud = ctx.upload(xyz, ty, s2)
        try:
            got = ctx.npcorr(cloud, tasks, R, 0.02, K)
        finally:
            cloud.free()
        want = O.npcorr(xyz, ty, s2, tasks, R, 0.02, K)
        assert np.array_equal(got, want)
        assert got.sum() > 0


def test_npcorr_sparse_region_of_a_dense_cloud(ctx):
    """Most ions in a dense blob, the centers in a sparse background: the
    first counting radius (set from the cloud's mean density) holds too few
    neighbours there, so the count is redone over R."""
    rng = np.random.default_rng(5)
    blob = rng.normal(loc=(2.0, 2.0, 2.0), scale=0.3, size=(20000, 3)).astype(np.float32)
    back = rng.uniform(0, 12, size=(3000, 3)).astype(np.float32)
    xyz = np.concatenate([blob, back])
    ty = rng.choice([1, 2], size=xyz.shape[0], p=[0.8, 0.2]).astype(np.uint8)
    s2 = np.full(xyz.shape[0], 100.0, np.float32)
    s2[:blob.shape[0]] = 0.0  # the blob's ions are neighbours only
    tasks = [((1,), (1,)), ((2,), (1, 2)), ((1, 2), (2,))]
    for K in (2, 20):
        _say(f"npcorr sparse region K={K}")
        cloud = ctx.upload(xyz, ty, s2)
        try:
            got = ctx.npcorr(cloud, tasks, 2.5, 0.1, K)
        finally:
            cloud.free()
        want = O.npcorr(xyz, ty, s2, tasks, 2.5, 0.1, K)
        assert np.array_equal(got, want)
        assert got.sum() > 0


def test_npcorr_equal_distance_different_square(ctx):
    """Two neighbours with equal float length sqrtf(d2) and squared lengths
    one ulp apart: the smaller d2 is ranked first, as in the oracle."""
    f = np.float32
    xyz = np.array([[0, 0, 0], [f("0.380425"), f("0.5418678"), 0],
                    [f("-0.380425"), f("0.54186785"), 0]], np.float32)
    ty = np.array([1, 2, 2], np.uint8)
    s2 = np.array([100, 0, 0], np.float32)
    for K in (0, 1):
        cloud = ctx.upload(xyz, ty, s2)
        try:
            got = ctx.npcorr(cloud, [((1,), (2,))], 2.0, 0.25, K)
        finally:
            cloud.free()
        assert np.array_equal(got, O.npcorr(xyz, ty, s2, [((1,), (2,))], 2.0, 0.25, K))
        assert got.sum() == 1
