"""Pins of the ion-to-surface oracle (the producer of ion2surf) that need no GPU.

The oracle restates mathHdl::closestPointOnTriangle (src/PARAPROBE_Math.cpp:88-171)
and solver::surface_distancing2 (src/PARAPROBE_SolverHdl.cpp:312-484).  It is
pinned here by closed forms (one point per Voronoi region of a right
triangle), by an independent FP64 computation (plane projection + the three
edge segments) on random triangles, and by the capped minimum over a whole hull.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from paraprobe_amd import synth

TRI = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32)  # right triangle in z = 0


@pytest.mark.parametrize("p,want", [
    ((0.25, 0.25, 0.5), 0.25),   # region 0: above the interior -> height^2
    ((-1.0, -1.0, 0.0), 2.0),    # region 4: nearest the right-angle vertex
    ((0.5, -2.0, 0.0), 4.0),     # region 5: nearest edge 0 (y = 0)
    ((-3.0, 0.5, 0.0), 9.0),     # region 3: nearest edge 1 (x = 0)
    ((1.0, 1.0, 0.0), 0.5),      # region 1: nearest the hypotenuse, foot (0.5, 0.5)
    ((-0.5, 3.0, 0.0), 4.25),    # region 2: nearest vertex (0, 1)
    # region 6 (across y = 0, beyond the hypotenuse's line): the reference's
    # far branch sets s = CLAMP(-e/c) where the geometric answer is
    # CLAMP(-d/a) (src/PARAPROBE_Math.cpp:151-153), so it returns the distance
    # to vertex (0, 0), 9.25, not to (1, 0), 4.25; for this triangle every
    # region-6 point takes that branch
    ((3.0, -0.5, 0.0), 9.25),
    ((0.0, 0.0, 0.0), 0.0),      # on a vertex
])
def test_closest_tri_closed_forms(p, want):
    assert O.closest_tri(TRI, p) == pytest.approx(want, rel=1e-6, abs=1e-7)


def test_closest_tri_region6_near_branch():
    # obtuse at V2: P = (2, -0.1) lies in region 6 and its nearest point is
    # the foot (1.16, 0.32) on edge V2V3, d^2 = 0.84^2 + 0.42^2
    tri = np.array([0, 0, 0, 1, 0, 0, 1.5, 1, 0], np.float32)
    assert O.closest_tri(tri, (2.0, -0.1, 0.0)) == pytest.approx(0.882, rel=1e-5)


def _seg_d2(p, a, b):
    ab = b - a
    t = np.clip(np.dot(p - a, ab) / np.dot(ab, ab), 0.0, 1.0)
    q = a + t * ab
    return float(np.dot(p - q, p - q))


def _tri_d2_fp64(tri, p):
    """Exact squared distance in FP64 by a different route: the foot of the
    perpendicular when it falls inside (barycentric test), else the nearest
    of the three edge segments."""
    a, b, c = (tri[0:3].astype(np.float64), tri[3:6].astype(np.float64),
               tri[6:9].astype(np.float64))
    p = np.asarray(p, np.float64)
    n = np.cross(b - a, c - a)
    nn = np.dot(n, n)
    best = min(_seg_d2(p, a, b), _seg_d2(p, b, c), _seg_d2(p, c, a))
    if nn > 0:
        h = np.dot(p - a, n) / nn
        f = p - h * n
        # barycentric coordinates of the foot
        u = np.dot(np.cross(b - f, c - f), n) / nn
        v = np.dot(np.cross(c - f, a - f), n) / nn
        w = 1.0 - u - v
        if u >= 0 and v >= 0 and w >= 0:
            best = min(best, h * h * nn)
    return best


def test_closest_tri_matches_fp64_geometry():
    """Every value is the exact distance (FP64, another route) except in the
    reference's region-6 far branch, where it is the distance to the point of
    edge V1V2 at s = clamp(-(e1.v0) / (e1.e1)); that branch is reached only
    when P lies across line V1V2 from V3 and its projection on V2V3 falls
    before V2."""
    rng = np.random.default_rng(5)
    worst = 0.0
    quirks = 0
    for _ in range(4000):
        tri = rng.normal(0, 1, 9).astype(np.float32)
        p = rng.normal(0, 2, 3).astype(np.float32)
        got = O.closest_tri(tri, p)
        want = _tri_d2_fp64(tri, p)
        scale = max(1.0, float(np.abs(tri).max()), float(np.abs(p).max())) ** 2
        err = abs(got - want) / scale
        if err > 2e-5:
            v1, v2, v3 = (tri[0:3].astype(np.float64), tri[3:6].astype(np.float64),
                          tri[6:9].astype(np.float64))
            q = np.asarray(p, np.float64)
            e0, e1 = v2 - v1, v3 - v1
            n = np.cross(e0, e1)
            # P across line V1V2 from V3, and its projection on V2V3 before V2
            assert np.dot(np.cross(e0, q - v1), n) < 0
            assert np.dot(v3 - v2, q - v2) <= 0
            s = np.clip(-np.dot(e1, v1 - q) / np.dot(e1, e1), 0.0, 1.0)
            c = v1 + s * e0
            err = abs(got - float(np.dot(q - c, q - c))) / scale
            quirks += 1
        worst = max(worst, err)
        assert got >= want - 2e-5 * scale  # always a distance to a point of the triangle
    # a wrong region, sign or swapped barycentric coordinate gives O(1) errors
    assert worst < 2e-5, worst
    assert 0 < quirks < 400, quirks


def test_hull_is_closed_and_encloses_the_solid():
    nseg, nz = 24, 6
    tri = synth.hull("frustum", 4.0, 2.0, 10.0, nseg, nz).astype(np.float64)
    assert tri.shape == (2 * nseg * nz + 2 * nseg, 9)
    # every undirected edge is shared by exactly two triangles
    edges = {}
    for t in tri:
        v = [tuple(np.round(t[3 * k:3 * k + 3], 5)) for k in range(3)]
        for a, b in ((0, 1), (1, 2), (2, 0)):
            e = tuple(sorted((v[a], v[b])))
            edges[e] = edges.get(e, 0) + 1
    assert set(edges.values()) == {2}
    # enclosed volume by the divergence theorem equals the polygonal frustum's
    a, b, c = tri[:, 0:3], tri[:, 3:6], tri[:, 6:9]
    vol = abs(np.einsum("ij,ij->i", a, np.cross(b, c)).sum()) / 6.0
    k = 0.5 * nseg * math.sin(2 * math.pi / nseg)  # polygon area / r^2
    want = 10.0 / 3.0 * k * (16.0 + 8.0 + 4.0)
    assert vol == pytest.approx(want, rel=1e-5)


def _seg_d2_many(p, a, b):
    ab = b - a
    t = np.clip(np.einsum("ij,ij->i", p - a, ab) / np.einsum("ij,ij->i", ab, ab), 0.0, 1.0)
    q = a + t[:, None] * ab
    return np.einsum("ij,ij->i", p - q, p - q)


def _hull_d2_fp64(tri, p):
    """_tri_d2_fp64 over all triangles of a hull at once; returns the minimum."""
    t = tri.astype(np.float64)
    a, b, c = t[:, 0:3], t[:, 3:6], t[:, 6:9]
    q = np.broadcast_to(np.asarray(p, np.float64), a.shape)
    best = np.minimum(np.minimum(_seg_d2_many(q, a, b), _seg_d2_many(q, b, c)),
                      _seg_d2_many(q, c, a))
    n = np.cross(b - a, c - a)
    nn = np.einsum("ij,ij->i", n, n)
    h = np.einsum("ij,ij->i", q - a, n) / nn
    f = q - h[:, None] * n
    u = np.einsum("ij,ij->i", np.cross(b - f, c - f), n) / nn
    v = np.einsum("ij,ij->i", np.cross(c - f, a - f), n) / nn
    inside = (u >= 0) & (v >= 0) & (1.0 - u - v >= 0)
    best = np.where(inside, np.minimum(best, h * h * nn), best)
    return float(best.min())


def test_ion2surf_oracle_is_capped_minimum_over_the_hull():
    tri = synth.hull("cylinder", 5.0, 5.0, 12.0, 32, 8)
    rng = np.random.default_rng(9)
    xyz = rng.uniform([-7, -7, -2], [7, 7, 14], size=(600, 3)).astype(np.float32)
    for R in (1.5, 3.0):
        got = O.ion2surf(xyz, tri, R)
        r2 = np.float32(R) * np.float32(R)
        for i in range(xyz.shape[0]):
            want = _hull_d2_fp64(tri, xyz[i])
            if want < r2 * (1 - 1e-4):
                assert got[i] == pytest.approx(want, rel=2e-5, abs=2e-5), (R, i)
            elif want > r2 * (1 + 1e-4):
                # no triangle within R: the start value fl32(R*R), bit for bit
                assert got[i] == r2, (R, i)


def test_ion2surf_oracle_edge_cases():
    xyz = np.array([[0, 0, 0], [10, 0, 0]], np.float32)
    # no triangles: every ion keeps fl32(R*R)
    out = O.ion2surf(xyz, np.zeros((0, 9), np.float32), 2.0)
    assert np.array_equal(out, np.full(2, 4.0, np.float32))
    # a triangle whose box touches the query box exactly still counts
    t = np.array([[2, -1, -1, 2, 1, -1, 2, 0, 1]], np.float32)
    out = O.ion2surf(xyz, t, 2.0)
    assert out[0] == 4.0 and out[1] == 4.0
    # degenerate triangles (det = 0) fall into region 1 with finite values:
    # three equal vertices give 0/0, which CLAMP turns into s = 1, i.e. the
    # vertex (1, 0, 0); the collinear triple gives s = clamp(0.75 / 0.25) = 1,
    # the middle vertex (1, 0, 0) rather than the nearest point (0.5, 0, 0)
    deg = np.array([[1, 0, 0, 1, 0, 0, 1, 0, 0],
                    [0.5, 0, 0, 1.0, 0, 0, 1.5, 0, 0]], np.float32)
    assert O.closest_tri(deg[0], xyz[0]) == 1.0
    assert O.closest_tri(deg[1], xyz[0]) == 1.0
    out = O.ion2surf(xyz[:1], deg, 2.0)
    assert out[0] == 1.0
