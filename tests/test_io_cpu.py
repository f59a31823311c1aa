"""POS records, range files and settings at either end of the path, without a
GPU: the range-file parser (read_rangefile2) against the reference's own RRNG
fixture and hand-made files, the settings reader against its XML fixture, and
the oracle's POS decoding against an independent big-endian encoding."""
import math
import os
import struct

import numpy as np
import pytest

from conftest import ROOT
from oracle import oracle as O
from paraprobe_amd import horderdist as H

GOLDEN = os.path.join(ROOT, "tests", "golden")
RRNG = os.path.join(GOLDEN, "PARAPROBE.Paper14.RangingTest.rrng")
XML = os.path.join(GOLDEN, "PARAPROBE.Paper14.SyntheticTest.xml")


def f32(x):
    return float(np.float32(x))


def test_reference_range_file():
    names, (lo, hi, ty) = H.read_rangefile(RRNG)
    # ion types in order of first appearance after Unknown (PeriodicTable.cpp:119-176)
    assert names == ["Unknown", "Al", "Fe", "Si", "Sc"]
    # Unknown's dummy range first, then one range per type, type id ascending
    assert ty.tolist() == [0, 1, 2, 3, 4]
    assert lo.tolist() == [0.0, f32(26.7), f32(55.9), f32(27.8), f32(44.8)]
    assert hi.tolist() == [f32(0.001), f32(26.9), f32(56.0), f32(28.0), f32(50.0)]


def _write(tmp_path, text, name="r.rrng"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_molecular_ions_and_token_rules(tmp_path):
    text = "\n".join([
        "[Ions]", "Number=1", "Ion1=X", "[Ranges]", "Number=6",
        # element names between "Vol:" and the last token, in the table's key order
        "Range1=10.0 11.0 Vol:0.1 O:1 Al:1 Color:FF0000",
        # a second range of the same molecular ion
        "Range2=20.0 21.0 Vol:0.1 Al:1 O:2 Color:FF0000",
        # "Name:Al" also holds "Na": the reference adds Na (a substring match)
        "Range3=12.0 13.0 Vol:0.1 Name:Al Al:1 Color:00FF00",
        # one-letter names need "X:" -- "Sc:1" is Sc, not S
        "Range4=44.0 45.0 Vol:0.1 Sc:1 Color:0000FF",
        # too few tokens, or no "Vol" in the third: not a range line
        "Range5=1.0 2.0 Vol:0.1 H:1",
        "Range6=3.0 4.0 Xol:0.1 H:1 Color:FFFFFF",
        # tab separated, and a run of blanks makes empty tokens
        "Range7=5.0\t6.0\tVol:0.1\tH:1\tColor:FFFFFF",
    ])
    names, (lo, hi, ty) = H.read_rangefile(_write(tmp_path, text))
    assert names == ["Unknown", "AlO", "AlNa", "Sc", "H"]
    assert ty.tolist() == [0, 1, 1, 2, 3, 4]
    assert lo.tolist() == [0.0, 10.0, 20.0, 12.0, 44.0, 5.0]


def test_range_file_missing_and_malformed(tmp_path):
    with pytest.raises(ValueError):
        H.read_rangefile(str(tmp_path / "missing.rrng"))
    with pytest.raises(ValueError):  # std::stof throws on a non-number
        H.read_rangefile(_write(tmp_path, "[Ranges]\nRange1=abc 2.0 Vol:0.1 Al:1 Color:F\n"))
    # no [Ranges] section: loaded, no types but Unknown
    names, (lo, _, _) = H.read_rangefile(_write(tmp_path, "[Ions]\nIon1=Al\n", "e.rrng"))
    assert names == ["Unknown"] and lo.tolist() == [0.0]


def _pos_bytes(rows):
    return b"".join(struct.pack(">4f", *r) for r in rows)


def test_pos_decode_oracle_against_struct_encoding():
    rng = np.random.default_rng(1)
    rows = rng.normal(0, 50, size=(500, 4)).astype(np.float32).tolist()
    rows += [[0.0, -0.0, math.inf, -math.inf], [1e-45, -1e-45, 3.4e38, 1.0]]
    xyz, mq, ty = O.pos_decode(_pos_bytes(rows))
    want = np.array(rows, np.float32)
    assert np.array_equal(xyz.view(np.uint32), want[:, :3].view(np.uint32))
    assert np.array_equal(mq.view(np.uint32), want[:, 3].copy().view(np.uint32))
    assert (ty == 0).all()  # no ranges: Unknown


def test_ranging_rules_on_reference_ranges():
    _, ranges = H.read_rangefile(RRNG)
    lo26, hi26 = np.float32(26.7), np.float32(26.9)
    mqs = [lo26, hi26, np.nextafter(hi26, np.float32(1e9)), np.nextafter(lo26, np.float32(0)),
           np.float32(50.0), np.float32(45.0), np.float32(0.0005), np.float32(0.001),
           np.float32(math.nan), np.float32(-1.0), np.float32(27.9), np.float32(55.95)]
    want = [1, 1, 0, 0, 4, 4, 0, 0, 0, 0, 3, 2]
    pos = _pos_bytes([[0.0, 0.0, 0.0, float(m)] for m in mqs])
    _, mq, ty = O.pos_decode(pos, ranges)
    assert ty.tolist() == want
    # overlapping ranges of two types: the lower type id wins
    lo = np.array([0.0, 10.0, 10.5], np.float32)
    hi = np.array([0.001, 11.0, 12.0], np.float32)
    rt = np.array([0, 1, 2], np.uint8)
    _, _, ty = O.pos_decode(_pos_bytes([[0, 0, 0, 10.7], [0, 0, 0, 11.5]]), (lo, hi, rt))
    assert ty.tolist() == [1, 2]


def test_reference_settings_fixture():
    xml = open(XML).read()
    st, fam, radii, tasks, mknn = H.read_settings(xml)
    assert st == 0
    assert fam == 1 << 4                      # AnalysisSpatDistrType 5: NPCORR only
    assert tasks == "Al-Al;Sc-Sc;Al-Sc;Sc-Al"
    assert radii == [0.0, f32(0.05), f32(2.5)]
    assert mknn == ""                         # SpatStatMKNNCode is read only with MKNN
    assert H.label_randomization(xml) == 0


def test_epos_decode_oracle_against_struct_encoding():
    rng = np.random.default_rng(2)
    rows = rng.normal(0, 30, size=(300, 9)).astype(np.float32)
    ints = rng.integers(-2**31, 2**31 - 1, size=(300, 2))
    data = b"".join(struct.pack(">9f2i", *[float(v) for v in r], *[int(v) for v in q])
                    for r, q in zip(rows, ints))
    _, ranges = H.read_rangefile(RRNG)
    xyz, mq, ty = O.pos_decode(data, ranges, epos=True)
    assert np.array_equal(xyz.view(np.uint32), rows[:, :3].copy().view(np.uint32))
    assert np.array_equal(mq.view(np.uint32), rows[:, 3].copy().view(np.uint32))
    _, _, want = O.pos_decode(_pos_bytes(rows[:, :4].tolist()), ranges)
    assert np.array_equal(ty, want)
