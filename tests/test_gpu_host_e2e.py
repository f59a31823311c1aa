"""End to end through the C++ host mirror: XML settings + task code -> gfx950
path -> the reference's CSV files, whose counts must equal the oracle's."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from paraprobe_amd import horderdist as H

pytestmark = pytest.mark.gpu

TYPES = {"Al": 1, "Sc": 2, "Zr": 3}
XML = """<?xml version="1.0" encoding="utf-8"?>
<Parameters>
  <AnalysisSpatDistrType>1246</AnalysisSpatDistrType>
  <DescrStatTasksCode>Al-Al;Sc-Sc;Al,Sc-Zr;Zr-Al,Sc</DescrStatTasksCode>
  <SpatStatRadiusMin>0.0</SpatStatRadiusMin>
  <SpatStatRadiusIncr>0.02</SpatStatRadiusIncr>
  <SpatStatRadiusMax>1.2</SpatStatRadiusMax>
  <SpatStatKNNOrder>1</SpatStatKNNOrder>
  <SpatStatMKNNCode>1;3;10</SpatStatMKNNCode>
</Parameters>
"""


def _read_counts(path, col=1, skip_head=5):
    rows = [ln.split(";") for ln in open(path).read().splitlines()]
    return rows


def test_xml_to_csv(ctx, tmp_path):
    print("test_xml_to_csv", flush=True)
    rng = np.random.default_rng(12)
    n = 6000
    xyz = rng.uniform(0, 8, size=(n, 3)).astype(np.float32)
    ty = rng.choice([1, 2, 3], size=n, p=[0.6, 0.3, 0.1]).astype(np.uint8)
    s2 = rng.uniform(0, 4, size=n).astype(np.float32)
    cl = ctx.upload(xyz, ty, s2)
    try:
        st = H.run_spatstat(ctx, cl, XML, 7, TYPES, 5000, 400.0, str(tmp_path))
    finally:
        cl.free()
    assert st == 0
    files = sorted(os.listdir(tmp_path))
    tasks = [((1,), (1,)), ((2,), (2,)), ((1, 2), (3,)), ((3,), (1, 2))]
    names = [("Al", "Al"), ("Sc", "Sc"), ("AlSc", "Zr"), ("Zr", "AlSc")]
    ref, _ = O.spatstat(xyz, ty, s2, tasks, 0.0, 0.02, 1.2, 1 | 2 | 8 | 32, kvalues=[1, 3, 10])
    nb = ref["rdf"].shape[1] - 2
    for t, (a, b) in enumerate(names):
        f = tmp_path / f"PARAPROBE.SimID.7.RDF.{a}.{b}.Original.csv"
        rows = [ln.split(";") for ln in open(f).read().splitlines()]
        # rows[7] lowest, rows[8 : 8 + nb] bins, rows[8 + nb] highest
        got = [int(float(rows[7][1]))] + [int(float(r[1])) for r in rows[8:8 + nb]] + \
              [int(float(rows[8 + nb][1]))]
        np.testing.assert_array_equal(got, ref["rdf"][t])
        f = tmp_path / f"PARAPROBE.SimID.7.1NN.{a}.{b}.Original.csv"
        rows = [ln.split(";") for ln in open(f).read().splitlines()]
        got = [int(float(rows[4][1]))] + [int(float(r[1])) for r in rows[5:5 + nb]] + \
              [int(float(rows[5 + nb][1]))]
        np.testing.assert_array_equal(got, ref["nn1"][t])
        for kth, k in enumerate([10, 3, 1]):
            f = tmp_path / f"PARAPROBE.SimID.7.MKNN{k}.{a}.{b}.Original.csv"
            rows = [ln.split(";") for ln in open(f).read().splitlines()]
            assert rows[0] == [f"MKNN{k}"]
            got = [int(float(rows[4][1]))] + [int(float(r[1])) for r in rows[5:5 + nb]] + \
                  [int(float(rows[5 + nb][1]))]
            np.testing.assert_array_equal(got, ref["mknn"][t, kth])
        f = tmp_path / f"PARAPROBE.SimID.7.NBCNTinRMAX.{a}.{b}.Original.csv"
        rows = [ln.split(";") for ln in open(f).read().splitlines()][4:]
        got = np.array([int(float(r[1])) for r in rows])
        want = ref["cntnb"][t]
        last = np.nonzero(want)[0].max() + 1
        np.testing.assert_array_equal(got, want[:last])
    assert len(files) == 4 * (1 + 1 + 3 + 1)


def _rdf_from_csv(path, nb):
    rows = [ln.split(";") for ln in open(path).read().splitlines()]
    return [int(float(rows[7][1]))] + [int(float(r[1])) for r in rows[8:8 + nb]] + \
           [int(float(rows[8 + nb][1]))]


def test_label_randomisation_rerun(ctx, tmp_path):
    """SpatStatAdditionalLabelRandomization = 1: the original labels, then the
    labels shuffled by the rndlabeler's std::mt19937 + std::shuffle in upload
    order, written to *.Rndmized.csv; the cloud keeps its own labels."""
    print("test_label_randomisation_rerun", flush=True)
    rng = np.random.default_rng(21)
    n = 8000
    xyz = rng.uniform(0, 9, size=(n, 3)).astype(np.float32)
    # clustered Sc so that the randomised RDF visibly differs
    ty = np.where(np.linalg.norm(xyz - 4.5, axis=1) < 2.0, 2, 1).astype(np.uint8)
    ty[rng.random(n) < 0.05] = 3
    s2 = rng.uniform(0, 4, size=n).astype(np.float32)
    xml = XML.replace("</Parameters>",
                      "  <SpatStatAdditionalLabelRandomization>1"
                      "</SpatStatAdditionalLabelRandomization>\n</Parameters>")
    xml = xml.replace("<AnalysisSpatDistrType>1246</AnalysisSpatDistrType>",
                      "<AnalysisSpatDistrType>14</AnalysisSpatDistrType>")
    cl = ctx.upload(xyz, ty, s2, cell_width=1.2)
    try:
        st = H.run_spatstat(ctx, cl, xml, 9, TYPES, 7000, 700.0, str(tmp_path))
        assert st == 0
        assert np.array_equal(cl.types(), ty)
    finally:
        cl.free()
    tasks = [((1,), (1,)), ((2,), (2,)), ((1, 2), (3,)), ((3,), (1, 2))]
    names = [("Al", "Al"), ("Sc", "Sc"), ("AlSc", "Zr"), ("Zr", "AlSc")]
    shuffled = H.shuffle_types(ty)
    ref, _ = O.spatstat(xyz, ty, s2, tasks, 0.0, 0.02, 1.2, 1 | 8, kvalues=[1, 3, 10])
    rnd, _ = O.spatstat(xyz, shuffled, s2, tasks, 0.0, 0.02, 1.2, 1 | 8, kvalues=[1, 3, 10])
    nb = ref["rdf"].shape[1] - 2
    for t, (a, b) in enumerate(names):
        got = _rdf_from_csv(tmp_path / f"PARAPROBE.SimID.9.RDF.{a}.{b}.Original.csv", nb)
        np.testing.assert_array_equal(got, ref["rdf"][t])
        got = _rdf_from_csv(tmp_path / f"PARAPROBE.SimID.9.RDF.{a}.{b}.Rndmized.csv", nb)
        np.testing.assert_array_equal(got, rnd["rdf"][t])
        for kth, k in enumerate([10, 3, 1]):
            f = tmp_path / f"PARAPROBE.SimID.9.MKNN{k}.{a}.{b}.Rndmized.csv"
            rows = [ln.split(";") for ln in open(f).read().splitlines()]
            got = [int(float(rows[4][1]))] + [int(float(r[1])) for r in rows[5:5 + nb]] + \
                  [int(float(rows[5 + nb][1]))]
            np.testing.assert_array_equal(got, rnd["mknn"][t, kth])
    assert not np.array_equal(ref["rdf"][1], rnd["rdf"][1])  # Sc-Sc loses its clustering
    assert len(os.listdir(tmp_path)) == 2 * 4 * (1 + 3)


def test_relabelled_cloud_equals_fresh_upload(ctx):
    print("test_relabelled_cloud_equals_fresh_upload", flush=True)
    from paraprobe_amd import FAM_1NN, FAM_RDF, Job

    rng = np.random.default_rng(22)
    n = 20000
    xyz = rng.uniform(0, 12, size=(n, 3)).astype(np.float32)
    ty = rng.choice([1, 2, 3], size=n, p=[0.5, 0.3, 0.2]).astype(np.uint8)
    s2 = rng.uniform(0, 6, size=n).astype(np.float32)
    new = H.shuffle_types(ty, seed=5)
    job = Job([((1,), (2,)), ((2, 3), (1, 2, 3))], 0.0, 0.01, 1.5, FAM_RDF | FAM_1NN)
    for width in (0.0, 1.5):
        src = ctx.upload(xyz, ty, s2, cell_width=width)
        if width:
            src.add_cell_list(0.75)
        rel = src.relabel(new)
        fresh = ctx.upload(xyz, new, s2)
        try:
            assert np.array_equal(rel.types(), new)
            a, b = ctx.run(rel, job), ctx.run(fresh, job)
            for k, v in a.arrays().items():
                np.testing.assert_array_equal(v, getattr(b, k))
        finally:
            for c in (src, rel, fresh):
                c.free()


def test_npcorr_to_hdf5(ctx, tmp_path):
    """AnalysisSpatDistrType 5 with label randomisation: the NPCORR cubes of
    the original and the randomised labels go to PARAPROBE.SimID.<id>.Results.h5
    (/DescrSpatStats/TwoPointStats: BinNXNYNZ, BinCenterXYZ and one
    <trg>_<env>_<Orig|Rndm> column per task, report_npc3d_*_hdf5,
    S.cpp:4506-4575), equal to the oracle's counts."""
    from paraprobe_amd import h5

    if not h5.available():
        pytest.skip("libpp_h5.so not built (no HDF5 library)")
    print("test_npcorr_to_hdf5", flush=True)
    rng = np.random.default_rng(31)
    n = 6000
    xyz = rng.uniform(0, 8, size=(n, 3)).astype(np.float32)
    ty = rng.choice([1, 2, 3], size=n, p=[0.6, 0.3, 0.1]).astype(np.uint8)
    s2 = rng.uniform(0, 4, size=n).astype(np.float32)
    xml = XML.replace("<AnalysisSpatDistrType>1246</AnalysisSpatDistrType>",
                      "<AnalysisSpatDistrType>5</AnalysisSpatDistrType>")
    xml = xml.replace("<SpatStatRadiusIncr>0.02</SpatStatRadiusIncr>",
                      "<SpatStatRadiusIncr>0.1</SpatStatRadiusIncr>")
    xml = xml.replace("<SpatStatKNNOrder>1</SpatStatKNNOrder>",
                      "<SpatStatKNNOrder>2</SpatStatKNNOrder>")
    xml = xml.replace("</Parameters>", "  <SpatStatAdditionalLabelRandomization>1"
                      "</SpatStatAdditionalLabelRandomization>\n</Parameters>")
    cl = ctx.upload(xyz, ty, s2)
    try:
        st = H.run_spatstat(ctx, cl, xml, 11, TYPES, 5000, 400.0, str(tmp_path))
    finally:
        cl.free()
    assert st == 0
    fn = str(tmp_path / "PARAPROBE.SimID.11.Results.h5")
    tasks = [((1,), (1,)), ((2,), (2,)), ((1, 2), (3,)), ((3,), (1, 2))]
    names = ["Al_Al", "Sc_Sc", "AlSc_Zr", "Zr_AlSc"]
    # SpatStatKNNOrder 2 in the XML is the 0-based order 1 (Settings.cpp:484-491)
    ref = O.npcorr(xyz, ty, s2, tasks, 1.2, 0.1, 1)
    rnd = O.npcorr(xyz, H.shuffle_types(ty), s2, tasks, 1.2, 0.1, 1)
    edge = ref.shape[1]
    assert h5.read(fn, "/DescrSpatStats/TwoPointStats/BinNXNYNZ").ravel().tolist() == [edge] * 3
    assert h5.read(fn, "/DescrSpatStats/TwoPointStats/BinCenterXYZ").shape == (edge ** 3, 3)
    for t, nm in enumerate(names):
        got = h5.read(fn, f"/DescrSpatStats/TwoPointStats/{nm}_Orig")
        np.testing.assert_array_equal(got.ravel(), ref[t].ravel())
        got = h5.read(fn, f"/DescrSpatStats/TwoPointStats/{nm}_Rndm")
        np.testing.assert_array_equal(got.ravel(), rnd[t].ravel())
    assert ref.sum() > 0 and not np.array_equal(ref, rnd)
