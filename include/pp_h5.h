/*
 * pp_h5.h -- optional HDF5 writer of the spatial-statistics results
 * (libpp_h5.so, built only where the HDF5 C library is installed; the GPU
 * library does not depend on it).
 *
 * It writes the part of the reference's results file that the spatial
 * statistics own, with the reference's group and dataset names, types and
 * shapes:
 *   h5Hdl::create_paraprobe_results_file      src/PARAPROBE_HDF5.cpp:369-408
 *     (the /DescrSpatStats groups, PARAPROBE_DESCRSTATS*, src/PARAPROBE_HDF5.h:66-73)
 *   horderdist::report_npc3d_bincenters_hdf5  src/PARAPROBE_SolverHdl.cpp:4506-4548
 *   horderdist::report_npc3d_histvalues_hdf5  src/PARAPROBE_SolverHdl.cpp:4551-4575
 * Every function opens and closes the file itself (as h5Hdl does) and
 * returns 0 on success, a negative value on an HDF5 error.
 */
#ifndef PP_H5_H
#define PP_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Creates (truncates) fn with the group /DescrSpatStats and, when npcorr is
 * non-zero, /DescrSpatStats/TwoPointStats. */
int pph5_create_results(const char *fn, int32_t npcorr);

/* /DescrSpatStats/TwoPointStats/BinNXNYNZ (u64le, 3 x 1: edge, edge, edge)
 * and .../BinCenterXYZ (f32le, edge^3 x 3, x fastest): the center of bin
 * (ix, iy, iz) as npc3d::get_bincenter gives it
 * (src/PARAPROBE_NPointCorrelations3D.cpp:239-249), i.e.
 * float(ix) * width - float((edge - 1) / 2) per axis (the integer half edge
 * is not scaled by the width there). */
int pph5_report_npc3d_bincenters(const char *fn, uint64_t edge, float width);

/* /DescrSpatStats/TwoPointStats/<trg>_<env>_<Orig|Rndm> (u32le, nxyz x 1):
 * one task's npc3d counts; trg and env are the concatenated names of the
 * task's central and neighbour types, randomized selects the _Rndm suffix
 * (the label-randomised rerun). */
int pph5_report_npc3d_histvalues(const char *fn, const char *trg, const char *env,
                                 int32_t randomized, const uint32_t *cnts, uint64_t nxyz);

/* Reads a two-dimensional dataset: *nr, *nc its shape, *kind 1 = u32,
 * 2 = u64, 3 = f32 (0: other); copies it to buf when it fits cap_bytes
 * (else returns -2 with the shape filled in). */
int pph5_read(const char *fn, const char *dset, void *buf, uint64_t cap_bytes, uint64_t *nr,
              uint64_t *nc, int32_t *kind);

#ifdef __cplusplus
}
#endif

#endif /* PP_H5_H */
