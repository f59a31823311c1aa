"""GPU parity of the ion-to-surface squared distance (pp_ion2surf) against the
oracle's brute force over the hull, bit for bit (same FP32 expression, no
contraction, IEEE division; the grid only prunes candidates)."""
import numpy as np
import pytest

from oracle import oracle as O
from paraprobe_amd import synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32),
                          np.asarray(b, np.float32).view(np.uint32))


def _with_outside(xyz, n_out, seed, pad=3.0):
    rng = np.random.default_rng(seed)
    lo, hi = xyz.min(0) - pad, xyz.max(0) + pad
    out = rng.uniform(lo, hi, size=(n_out, 3)).astype(np.float32)
    return np.concatenate([xyz, out])


@pytest.mark.parametrize("shape,R", [("cylinder", 2.0), ("cylinder", 3.5), ("frustum", 2.0)])
def test_ion2surf_matches_oracle(ctx, shape, R):
    if shape == "cylinder":
        t = synth.tip(100_000, 11, "cylinder", 15.0, 15.0, 64.0, [0.9, 0.1])
        tri = synth.hull("cylinder", 15.0, 15.0, 64.0, 96, 48)
    else:
        t = synth.tip(100_000, 12, "frustum", 20.0, 10.0, 60.0, [0.9, 0.1])
        tri = synth.hull("frustum", 20.0, 10.0, 60.0, 80, 40)
    xyz = _with_outside(t.xyz, 5000, 3)
    _say(f"ion2surf {shape} R={R} n={xyz.shape[0]} ntri={tri.shape[0]}")
    got = ctx.ion2surf(xyz, tri, R)
    want = O.ion2surf(xyz, tri, R)
    bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, (bad[:10], got[bad[:10]], want[bad[:10]])
    r2 = np.float32(R) * np.float32(R)
    assert (got <= r2).all()
    assert 0 < (got < r2).sum() < got.size  # some ions near the hull, some deep inside


def test_ion2surf_degenerate_and_duplicate_triangles(ctx):
    rng = np.random.default_rng(4)
    xyz = rng.uniform(-3, 3, size=(20_000, 3)).astype(np.float32)
    tri = rng.uniform(-3, 3, size=(300, 9)).astype(np.float32)
    deg = tri[:60].copy()
    deg[:20, 3:6] = deg[:20, 0:3]           # two equal vertices
    deg[20:40, 3:9] = np.tile(deg[20:40, 0:3], 2)  # three equal vertices
    deg[40:60, 6:9] = 2 * deg[40:60, 3:6] - deg[40:60, 0:3]  # collinear
    allt = np.concatenate([tri, deg, tri[:50]])  # and exact duplicates
    for R in (0.5, 2.0):
        _say(f"ion2surf degenerate R={R}")
        got = ctx.ion2surf(xyz, allt, R)
        assert _bits_equal(got, O.ion2surf(xyz, allt, R))


def test_ion2surf_boundary_and_empty(ctx):
    # a triangle whose box touches an ion's query box exactly is a candidate
    xyz = np.array([[0, 0, 0], [10, 0, 0], [2.0, 0, 0.5]], np.float32)
    t = np.array([[2, -1, -1, 2, 1, -1, 2, 0, 1]], np.float32)
    got = ctx.ion2surf(xyz, t, 2.0)
    assert _bits_equal(got, O.ion2surf(xyz, t, 2.0))
    assert got[2] == 0.0
    # no triangles: every ion keeps fl32(R * R); no ions: nothing to do
    got = ctx.ion2surf(xyz, np.zeros((0, 9), np.float32), 1.7)
    assert _bits_equal(got, np.full(3, np.float32(1.7) * np.float32(1.7)))
    assert ctx.ion2surf(np.zeros((0, 3), np.float32), t, 2.0).size == 0


def test_ion2surf_headline_tip_properties(ctx):
    """The 20M-ion frustum (r 40 -> 20 nm, h 250 nm) against a 512 x 256 hull:
    a sample of ions bit for bit against the oracle, and every ion against the
    smooth solid's exact distance (the hull is inscribed; its chord sagitta is
    r (1 - cos(pi / 512)) < 8e-4 nm)."""
    t = synth.frustum_20m()
    tri = synth.hull("frustum", 40.0, 20.0, 250.0, 512, 256)
    R = 2.0
    _say(f"ion2surf frustum_20m ntri={tri.shape[0]}")
    got = ctx.ion2surf(t.xyz, tri, R)
    rng = np.random.default_rng(8)
    idx = rng.choice(t.n, 10_000, replace=False)
    # bias the sample to the shell where the hull matters
    shell = np.flatnonzero(got < np.float32(4.0))
    idx = np.concatenate([idx, rng.choice(shell, 10_000, replace=False)])
    want = O.ion2surf(t.xyz[idx], tri, R)
    assert _bits_equal(got[idx], want)
    exact = np.sqrt(t.ion2surf_sqr.astype(np.float64))
    d = np.sqrt(got.astype(np.float64))
    near = exact < R - 2e-3
    far = exact > R + 2e-3
    z = t.xyz[:, 2]
    lateral = near & (z > 3.0) & (z < 247.0)
    assert np.abs(d[lateral] - exact[lateral]).max() < 1e-3
    # next to the caps the nearest triangles are the long thin fan triangles
    # (40 nm x 0.49 nm), where the FP32 expression loses a few hundredths of
    # a nm to the cancellation in det = a c - b^2 (the reference's behaviour)
    assert np.abs(d[near] - exact[near]).max() < 0.1
    assert (got[far] == np.float32(4.0)).all()


def test_ion2surf_device_entry_matches_host_entry(ctx):
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so.7")
    t = synth.tip(50_000, 21, "cylinder", 10.0, 10.0, 30.0, [1.0])
    tri = synth.hull("cylinder", 10.0, 10.0, 30.0, 64, 24)
    n, nt = t.n, tri.shape[0]
    ptr = [ctypes.c_void_p() for _ in range(3)]
    sizes = [12 * n, 36 * nt, 4 * n]
    try:
        for p, b in zip(ptr, sizes):
            assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(b)) == 0
        assert hip.hipMemcpy(ptr[0], t.xyz.ctypes.data_as(ctypes.c_void_p),
                             ctypes.c_size_t(sizes[0]), 1) == 0
        assert hip.hipMemcpy(ptr[1], tri.ctypes.data_as(ctypes.c_void_p),
                             ctypes.c_size_t(sizes[1]), 1) == 0
        got = np.empty(n, np.float32)
        for R, ntri in ((2.0, nt), (3.0, 0)):
            _say(f"ion2surf device R={R} ntri={ntri}")
            ctx.ion2surf_device(ptr[0].value, n, ptr[1].value, ntri, R, ptr[2].value)
            assert hip.hipMemcpy(got.ctypes.data_as(ctypes.c_void_p), ptr[2],
                                 ctypes.c_size_t(sizes[2]), 2) == 0
            assert _bits_equal(got, ctx.ion2surf(t.xyz, tri[:ntri], R))
    finally:
        for p in ptr:
            if p.value:
                hip.hipFree(p)
