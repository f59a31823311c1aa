"""paraprobe_amd -- MI355X-native RDF / kNN spatial statistics for atom-probe point clouds.

The GPU path replaces the compute step of the reference's
``horderdist::compute_generic_spatstat3`` (src/PARAPROBE_SolverHdl.cpp:3685-4296):
KD-tree range-ball queries become a uniform cell list on gfx950, and the
RDF / 1NN / RIPK / MKNN / CNTNB histograms are binned bit-exactly in LDS.
"""
from .capi import (  # noqa: F401
    FAM_1NN,
    FAM_CNTNB,
    FAM_MKNN,
    FAM_NPCORR,
    FAM_RDF,
    FAM_RIPK,
    Job,
    PPError,
    Result,
    SpatStat,
    hist_geometry,
    load_library,
    type_mask,
)
