"""The C++ host mirror of the reference's task interface: task codes
(src/PARAPROBE_PeriodicTable.cpp:498-690), MKNN codes (S.cpp:3739-3767),
settings keys (src/PARAPROBE_Settings.cpp:347-393, 473-503) and the CSV
writers (S.cpp:4299-4453)."""
import math
import os

import numpy as np
import pytest

from paraprobe_amd import horderdist as H

REG = {"Al": 1, "Sc": 2, "Mn": 3, "AlH": 4, "HAl": 5, "S": 6}


def test_single_and_multi_type_tasks():
    t = H.parse_tasks("Al-Al;Sc-Sc;Al-Sc;Sc-Al", REG)
    assert [(a, b) for a, b, _, _ in t] == [("Al", "Al"), ("Sc", "Sc"), ("Al", "Sc"), ("Sc", "Al")]
    t = H.parse_tasks("Al,Mn-Sc,Al", REG)
    # names concatenate in map (lexicographic) order
    assert t[0][0] == "AlMn" and t[0][1] == "AlSc"
    trg, env = t[0][2], t[0][3]
    assert trg[0] == (1 << 1) | (1 << 3) and env[0] == (1 << 1) | (1 << 2)


def test_invalid_tasks_are_dropped():
    # no minus, two minuses, unknown names, empty side
    t = H.parse_tasks("AlAl;Al-Sc-Mn;Xx-Al;Al-;Sc-Al", REG)
    assert [(a, b) for a, b, _, _ in t] == [("Sc", "Al")]
    assert H.parse_tasks("", REG) == []


def test_element_matching_rules():
    # "Sc" contains "S" and "Sc": an exact match selects the single element
    assert H.parse_tasks("Sc-S", REG)[0][:2] == ("Sc", "S")
    # a molecular name is the concatenation of every contained symbol in map
    # order ("Al" < "H"): "HAl" and "AlH" both resolve to "AlH"
    t = H.parse_tasks("HAl-Al;AlH-Al", REG)
    assert [a for a, *_ in t] == ["AlH", "AlH"]


def test_mknn_code():
    assert H.parse_mknn_code("1;10;100") == [99, 9, 0]
    assert H.parse_mknn_code("5;2;2") == [4, 1, 1]
    assert H.parse_mknn_code("1;x;3") == [2, 0]          # non-numbers are skipped
    assert H.parse_mknn_code("1-5") == []                 # a minus disables MKNN
    assert H.parse_mknn_code("") == []
    assert H.parse_mknn_code("0;4") is None               # k < 1 aborts


def _xml(mode, extra=""):
    return (f'<?xml version="1.0"?><Parameters><!-- c --><AnalysisSpatDistrType>{mode}'
            f"</AnalysisSpatDistrType><DescrStatTasksCode>Al-Al</DescrStatTasksCode>"
            f"<SpatStatRadiusMin>0.0</SpatStatRadiusMin><SpatStatRadiusIncr>0.01"
            f"</SpatStatRadiusIncr><SpatStatRadiusMax>2.0</SpatStatRadiusMax>"
            f"<SpatStatMKNNCode>1;10</SpatStatMKNNCode>{extra}</Parameters>")


@pytest.mark.parametrize("mode,fam", [("1", 1), ("4", 8), ("6", 32), ("14", 9), ("1236", 39),
                                      ("0", 0), ("12345", 0), ("7", 0), ("104", 0)])
def test_settings_digits(mode, fam):
    st, f, radii, tc, mc = H.read_settings(_xml(mode))
    assert f == fam
    if fam:
        assert radii == [0.0, np.float32(0.01), 2.0] and tc == "Al-Al"
        assert mc == ("1;10" if fam & 8 else "")


def test_settings_checks():
    bad = _xml("1").replace("<SpatStatRadiusMin>0.0", "<SpatStatRadiusMin>3.0")
    assert H.read_settings(bad)[0] == 2
    bad = _xml("1").replace("<SpatStatRadiusIncr>0.01", "<SpatStatRadiusIncr>0.0000001")
    assert H.read_settings(bad)[0] == 2


def test_csv_1nn_layout(tmp_path):
    # width 0.25 over [0, 1]: 4 bins, all values exact in binary
    h = np.array([0, 1, 3, 0, 4, 2], np.int64)
    p = str(tmp_path / "a.csv")
    assert H.write_descrstat2(p, 2, "1NN", 0.0, 0.25, 1.0, h) == 0
    want = ("1NN\n"
            "BinEnd(r);Counts;ECDFon[BinMinBinMax];CDFon[BinMinBinMax]\n"
            "nm;1;1;1\n"
            "BinEnd(r);Counts;ECDFon[BinMinBinMax];CDFon[BinMinBinMax]\n"
            "0;0;;\n"
            "0.25;1;1;0.125\n"
            "0.5;3;4;0.5\n"
            "0.75;0;4;0.5\n"
            "1;4;8;1\n"
            ";2;;\n")
    assert open(p).read() == want


def test_csv_mknn_and_ripk_layout(tmp_path):
    h = np.array([1, 2, 0, 5, 7], np.int64)   # 3 bins of 0.5 over [0, 1.5]
    p = str(tmp_path / "m.csv")
    assert H.write_descrstat2(p, 4, "MKNN10", 0.0, 0.5, 1.5, h) == 0
    lines = open(p).read().splitlines()
    assert lines[0] == "MKNN10" and lines[4] == "0;1;;" and lines[-1] == ";7;;"
    assert lines[5:8] == ["0.5;2;2;0.285714285714285698", "1;0;2;0.285714285714285698", "1.5;5;7;1"]
    p = str(tmp_path / "r.csv")
    assert H.write_descrstat2(p, 3, "RIPK", 0.0, 0.5, 1.5, h) == 0
    assert open(p).read() == ("BinEnd(r);AccumulatedCounts\nnm;1\nBinEnd(r);AccumulatedCounts\n"
                              "0;1\n0.5;2\n1;2\n1.5;7\n;7\n")


def test_csv_rdf_layout(tmp_path):
    h = np.array([0, 4, 10, 3], np.int64)     # 2 bins of 0.5 over [0, 1]
    p = str(tmp_path / "rdf.csv")
    assert H.write_descrstat2(p, 1, "RDF", 0.0, 0.5, 1.0, h, 1000, 500.0) == 0
    lines = open(p).read().splitlines()
    assert lines[:3] == ["NumberOfIonsInside;1000", "VolumeInside;500 (nm^3)",
                         "InsideAtomicDensityEst;2"]
    norm = (1.0 / 2.0) * (1.0 / (float(np.float32(4.0) / np.float32(3.0)) * math.pi))
    assert lines[3] == "NormalizationFactor;%.18g" % norm
    assert lines[4:7] == ["r;AccumulatedCount;nRDF(r);ShellVolume;RDF(r)", "nm;1;1;nm^3;1",
                          "r;AccumulatedCount;nRDF(r);ShellVolume;RDF(r)"]
    assert lines[7] == "0;0;;;"
    # bin 0: r = 0.25, shell 0.5^3 - 0^3 = 0.125, diff 0
    assert lines[8] == "0.25;4;0;0.125;0"
    # bin 1: r = 0.75, diff 6, shell 1 - 0.125 = 0.875
    assert lines[9] == "0.75;10;6;0.875;%.18g" % (norm * 6.0 / 0.875)
    assert lines[10] == ";3;;;"
    # no volume: header lines only (S.cpp:4328)
    assert H.write_descrstat2(p, 1, "RDF", 0.0, 0.5, 1.0, h, 10, 0.0) == 0
    assert open(p).read() == "NumberOfIonsInside;10\nVolumeInside;0 (nm^3)\n"


def test_csv_cntnb_layout(tmp_path):
    p = str(tmp_path / "c.csv")
    assert H.write_descrstat3(p, np.array([0, 2, 0, 6], np.int64)) == 0
    assert open(p).read() == ("NeighborCountInRMax\nValue;Counts;ECDFon[BinMinBinMax];"
                              "CDFon[BinMinBinMax]\n1;1;1;1\nValue;Counts;ECDFon[BinMinBinMax];"
                              "CDFon[BinMinBinMax]\n0;0;0;0\n1;2;2;0.25\n2;0;2;0.25\n3;6;8;1\n")
