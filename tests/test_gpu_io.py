"""GPU parity of the POS decoding and ranging kernel and of the file-driven
pipeline (POS + RRNG + hull -> cloud -> CSV) against the oracle."""
import math
import os
import struct

import numpy as np
import pytest

from conftest import ROOT
from oracle import oracle as O
from paraprobe_amd import FAM_MKNN, FAM_RDF, Job, synth
from paraprobe_amd import horderdist as H

pytestmark = pytest.mark.gpu

RRNG = os.path.join(ROOT, "tests", "golden", "PARAPROBE.Paper14.RangingTest.rrng")


def _say(msg):
    print(msg, flush=True)


def _records(n, seed):
    """n POS records: random coordinates, mass-to-charge values spread over and
    around the fixture's ranges (edges included), and special bit patterns."""
    rng = np.random.default_rng(seed)
    v = rng.normal(0, 40, size=(n, 4)).astype(np.float32)
    edges = np.array([26.7, 26.9, 55.9, 56.0, 27.8, 28.0, 44.8, 50.0, 0.0, 0.001],
                     np.float32)
    v[:, 3] = rng.uniform(0, 60, n).astype(np.float32)
    k = n // 10
    v[:k, 3] = rng.choice(edges, k)
    toward = (np.float32(1e9) * rng.choice([-1, 1], k)).astype(np.float32)
    v[k:2 * k, 3] = np.nextafter(rng.choice(edges, k), toward)
    special = np.array([0.0, -0.0, math.inf, -math.inf, math.nan, 1e-45, -3.4e38],
                       np.float32)
    m = min(64, n - 2 * k)
    v[2 * k:2 * k + m] = rng.choice(special, (m, 4))
    return v.astype(">f4").tobytes(), v


def test_pos_decode_matches_oracle(ctx):
    _, ranges = H.read_rangefile(RRNG)
    for n, seed in ((1, 1), (1000, 2), (1_000_003, 3)):
        _say(f"pos_decode n={n}")
        pos, v = _records(n, seed)
        xyz, mq, ty = ctx.pos_decode(pos, ranges)
        oxyz, omq, oty = O.pos_decode(pos, ranges)
        assert np.array_equal(xyz.view(np.uint32), oxyz.view(np.uint32))
        assert np.array_equal(mq.view(np.uint32), omq.view(np.uint32))
        assert np.array_equal(ty, oty)
        assert np.array_equal(xyz.view(np.uint32), v[:, :3].copy().view(np.uint32))
    # no ranges: every ion Unknown
    xyz, _, ty = ctx.pos_decode(pos, None, with_mq=False)
    assert (ty == 0).all()
    with pytest.raises(Exception):
        ctx.pos_decode(pos[:-3], ranges)


def test_pos_decode_device_matches_host(ctx):
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so.7")
    _, ranges = H.read_rangefile(RRNG)
    pos, _ = _records(200_000, 4)
    n = len(pos) // 16
    ptr = [ctypes.c_void_p() for _ in range(4)]
    sizes = [len(pos), 12 * n, 4 * n, n]
    try:
        for p, b in zip(ptr, sizes):
            assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(b)) == 0
        src = np.frombuffer(pos, np.uint8)
        assert hip.hipMemcpy(ptr[0], src.ctypes.data_as(ctypes.c_void_p),
                             ctypes.c_size_t(sizes[0]), 1) == 0
        _say("pos_decode_device")
        ctx.pos_decode_device(ptr[0].value, len(pos), ranges, ptr[1].value, ptr[2].value,
                              ptr[3].value)
        xyz = np.empty((n, 3), np.float32)
        ty = np.empty(n, np.uint8)
        assert hip.hipMemcpy(xyz.ctypes.data_as(ctypes.c_void_p), ptr[1],
                             ctypes.c_size_t(sizes[1]), 2) == 0
        assert hip.hipMemcpy(ty.ctypes.data_as(ctypes.c_void_p), ptr[3],
                             ctypes.c_size_t(sizes[3]), 2) == 0
        hxyz, _, hty = ctx.pos_decode(pos, ranges)
        assert np.array_equal(xyz.view(np.uint32), hxyz.view(np.uint32))
        assert np.array_equal(ty, hty)
    finally:
        for p in ptr:
            if p.value:
                hip.hipFree(p)


def _tip_pos(seed=7):
    """A 200k-ion cylinder tip as POS bytes, its types mapped to mass-to-charge
    values inside the fixture's ranges (1 Al, 2 Fe, 4 Sc)."""
    t = synth.tip(200_000, seed, "cylinder", 10.0, 10.0, 40.0, [0.8, 0.15, 0.05])
    mq = np.select([t.types == 1, t.types == 2], [26.8, 55.95], 47.0).astype(np.float32)
    rec = np.concatenate([t.xyz, mq[:, None]], axis=1).astype(">f4").tobytes()
    want_types = np.select([t.types == 1, t.types == 2], [1, 2], 4).astype(np.uint8)
    return t, rec, want_types


def test_cloud_from_pos_equals_upload(ctx):
    _, ranges = H.read_rangefile(RRNG)
    t, rec, want_types = _tip_pos()
    tri = synth.hull("cylinder", 10.0, 10.0, 40.0, 64, 32)
    R = 2.0
    job = Job([((1,), (4,)), ((4,), (4,))], 0.0, 0.02, R, FAM_RDF | FAM_MKNN, kvalues=[1, 5])
    _say("cloud_from_pos")
    a = ctx.cloud_from_pos(rec, ranges, tri, R)
    s2 = ctx.ion2surf(t.xyz, tri, R)
    b = ctx.upload(t.xyz, want_types, s2)
    try:
        assert np.array_equal(a.types(), want_types)
        ra, rb = ctx.run(a, job), ctx.run(b, job)
        for k, v in ra.arrays().items():
            np.testing.assert_array_equal(v, getattr(rb, k))
        ref, _ = O.spatstat(t.xyz, want_types, s2, job.tasks, 0.0, 0.02, R, FAM_RDF | FAM_MKNN,
                            kvalues=[1, 5])
        np.testing.assert_array_equal(ra.rdf, ref["rdf"])
        np.testing.assert_array_equal(ra.mknn, ref["mknn"])
    finally:
        a.free()
        b.free()


def test_run_pos_end_to_end(ctx, tmp_path):
    """XML + POS file + RRNG file (+ hull) -> CSV files equal to the oracle."""
    t, rec, want_types = _tip_pos(8)
    pos_path = tmp_path / "tip.pos"
    pos_path.write_bytes(rec)
    xml = """<Parameters>
  <AnalysisSpatDistrType>1</AnalysisSpatDistrType>
  <DescrStatTasksCode>Al-Al;Sc-Al,Fe</DescrStatTasksCode>
  <SpatStatRadiusMin>0.0</SpatStatRadiusMin>
  <SpatStatRadiusIncr>0.05</SpatStatRadiusIncr>
  <SpatStatRadiusMax>1.5</SpatStatRadiusMax>
  <AnalysisVolumeTessellation>1</AnalysisVolumeTessellation>
  <SurfaceCellsCarvingRadius>2.5</SurfaceCellsCarvingRadius>
</Parameters>"""
    tri = synth.hull("cylinder", 10.0, 10.0, 40.0, 64, 32)
    out = tmp_path / "out"
    out.mkdir()
    _say("run_pos")
    st = H.run_pos(ctx, xml, str(pos_path), RRNG, tri, 3, 150_000, 12_000.0, str(out))
    assert st == 0
    R = max(np.float32(1.5), np.float32(2.5))       # the surface distance uses the carving radius
    s2 = O.ion2surf(t.xyz, tri, float(R))
    ref, _ = O.spatstat(t.xyz, want_types, s2, [((1,), (1,)), ((4,), (1, 2))], 0.0, 0.05, 1.5,
                        FAM_RDF)
    nb = ref["rdf"].shape[1] - 2
    for i, (a, b) in enumerate([("Al", "Al"), ("Sc", "AlFe")]):
        rows = [ln.split(";") for ln in
                open(out / f"PARAPROBE.SimID.3.RDF.{a}.{b}.Original.csv").read().splitlines()]
        got = [int(float(rows[7][1]))] + [int(float(r[1])) for r in rows[8:8 + nb]] + \
              [int(float(rows[8 + nb][1]))]
        np.testing.assert_array_equal(got, ref["rdf"][i])
    assert len(os.listdir(out)) == 2


def _epos_of(v):
    """EPOS records for rows v = (x, y, z, mq): the other five floats and the
    two ints filled with recognisable values."""
    n = v.shape[0]
    f = np.zeros((n, 9), np.float32)
    f[:, :4] = v
    f[:, 4:] = np.arange(5, dtype=np.float32) + 100.0
    rec = np.zeros(n, dtype=[("f", ">f4", 9), ("i", ">i4", 2)])
    rec["f"] = f
    rec["i"] = np.stack([np.arange(n), -np.arange(n)], axis=1)
    return rec.tobytes()


def test_epos_decode_and_cloud(ctx, tmp_path):
    _, ranges = H.read_rangefile(RRNG)
    pos, v = _records(300_001, 6)
    ep = _epos_of(v)
    assert len(ep) == 44 * v.shape[0]
    _say("epos decode")
    xyz, mq, ty = ctx.pos_decode(ep, ranges, epos=True)
    oxyz, omq, oty = O.pos_decode(ep, ranges, epos=True)
    assert np.array_equal(xyz.view(np.uint32), oxyz.view(np.uint32))
    assert np.array_equal(mq.view(np.uint32), omq.view(np.uint32))
    assert np.array_equal(ty, oty)
    pxyz, _, pty = ctx.pos_decode(pos, ranges)
    assert np.array_equal(xyz.view(np.uint32), pxyz.view(np.uint32)) and np.array_equal(ty, pty)
    with pytest.raises(Exception):
        ctx.pos_decode(ep[:-4], ranges, epos=True)
    # the cloud and the file-driven run from EPOS equal those from POS
    t, rec, want_types = _tip_pos(9)
    v2 = np.frombuffer(rec, ">f4").reshape(-1, 4).astype(np.float32)
    ep2 = _epos_of(v2)
    tri = synth.hull("cylinder", 10.0, 10.0, 40.0, 64, 32)
    a = ctx.cloud_from_pos(ep2, ranges, tri, 2.0, epos=True)
    b = ctx.cloud_from_pos(rec, ranges, tri, 2.0)
    try:
        job = Job([((1,), (4,)), ((4,), (1, 2))], 0.0, 0.02, 2.0, FAM_RDF | FAM_MKNN,
                  kvalues=[1, 7])
        ra, rb = ctx.run(a, job), ctx.run(b, job)
        for k, val in ra.arrays().items():
            np.testing.assert_array_equal(val, getattr(rb, k))
    finally:
        a.free()
        b.free()
    xml = """<Parameters>
  <AnalysisSpatDistrType>1</AnalysisSpatDistrType>
  <DescrStatTasksCode>Al-Al;Sc-Al,Fe</DescrStatTasksCode>
  <SpatStatRadiusMin>0.0</SpatStatRadiusMin>
  <SpatStatRadiusIncr>0.05</SpatStatRadiusIncr>
  <SpatStatRadiusMax>1.5</SpatStatRadiusMax>
</Parameters>"""
    for name, data in (("tip.pos", rec), ("tip.epos", ep2)):
        (tmp_path / name).write_bytes(data)
        d = tmp_path / ("out_" + name.replace(".", "_"))
        d.mkdir()
        _say(f"run_pos {name}")
        assert H.run_pos(ctx, xml, str(tmp_path / name), RRNG, tri, 4, 150_000, 12_000.0,
                         str(d)) == 0
    for f in sorted(p.name for p in (tmp_path / "out_tip_pos").iterdir()):
        assert (tmp_path / "out_tip_pos" / f).read_bytes() == \
            (tmp_path / "out_tip_epos" / f).read_bytes()
