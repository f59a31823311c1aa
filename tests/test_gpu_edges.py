"""Edge cases of the entry points around the hot path: empty inputs, the
largest accepted and the first rejected arguments, and error returns that
leave the context usable."""
import numpy as np
import pytest

from oracle import oracle as O
from paraprobe_amd import FAM_RDF, Job, PPError, synth
from paraprobe_amd import horderdist as H

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_empty_inputs(ctx):
    _say("edges: empty inputs")
    z3 = np.zeros((0, 3), np.float32)
    xyz, mq, ty = ctx.pos_decode(b"", None)
    assert xyz.shape == (0, 3) and ty.shape == (0,)
    assert ctx.ion2surf(z3, np.zeros((0, 9), np.float32), 1.0).shape == (0,)
    cl = ctx.cloud_from_pos(b"", None, None, 1.0)
    try:
        assert cl.n == 0
        r = ctx.run(cl, Job([((1,), (1,))], 0.0, 0.1, 1.0, FAM_RDF))
        assert r.rdf.sum() == 0
        assert ctx.npcorr(cl, [((1,), (1,))], 1.0, 0.1, 0).sum() == 0
        rel = cl.relabel(np.zeros(0, np.uint8))
        rel.free()
    finally:
        cl.free()


def test_rejections_keep_the_context_usable(ctx):
    _say("edges: rejections")
    rng = np.random.default_rng(2)
    xyz = rng.uniform(0, 4, size=(500, 3)).astype(np.float32)
    ty = rng.choice([1, 2], size=500).astype(np.uint8)
    s2 = np.full(500, 9.0, np.float32)
    cl = ctx.upload(xyz, ty, s2)
    try:
        with pytest.raises(PPError):
            ctx.npcorr(cl, [((1,), (2,))], 1.0, 0.1, 448)        # K beyond the buffer
        with pytest.raises(PPError):
            ctx.npcorr(cl, [((1,), (2,))], 5.0, 0.01, 0)         # cube above 503^3
        with pytest.raises(PPError):
            ctx.pos_decode(b"\0" * 17, None)                      # not whole records
        with pytest.raises(PPError):
            ctx.ion2surf(xyz, np.zeros((1, 9), np.float32), 0.0)  # R must be positive
        with pytest.raises(ValueError):
            cl.relabel(ty[:-1])                                   # one type per ion
        # the largest K accepted, and the context still answers correctly
        got = ctx.npcorr(cl, [((1,), (1, 2))], 1.5, 0.1, 447)
        assert np.array_equal(got, O.npcorr(xyz, ty, s2, [((1,), (1, 2))], 1.5, 0.1, 447))
        r = ctx.run(cl, Job([((1,), (2,))], 0.0, 0.05, 1.0, FAM_RDF))
        ref, _ = O.spatstat(xyz, ty, s2, [((1,), (2,))], 0.0, 0.05, 1.0, FAM_RDF)
        assert np.array_equal(r.rdf, ref["rdf"])
    finally:
        cl.free()


def test_single_ion_and_all_unknown(ctx, tmp_path):
    _say("edges: one ion, no ranges")
    rec = np.array([[1.0, 2.0, 3.0, 27.0]], np.float32).astype(">f4").tobytes()
    cl = ctx.cloud_from_pos(rec, None, synth.hull("cylinder", 5.0, 5.0, 6.0, 8, 2), 2.0)
    try:
        assert cl.types().tolist() == [0]
        r = ctx.run(cl, Job([((0,), (0,))], 0.0, 0.1, 1.0, FAM_RDF))
        assert r.rdf.sum() == 0
    finally:
        cl.free()
    f = tmp_path / "one.rrng"
    f.write_text("[Ranges]\nRange1=26.5 27.5 Vol:0.1 Al:1 Color:F\n")
    _, ranges = H.read_rangefile(str(f))
    _, _, ty = ctx.pos_decode(rec, ranges)
    assert ty.tolist() == [1]
