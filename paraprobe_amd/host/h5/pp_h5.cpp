// pp_h5.cpp -- HDF5 output of the spatial-statistics results (libpp_h5.so).
//
// The datasets the reference's spatial statistics write into its results
// file: the /DescrSpatStats groups (h5Hdl::create_paraprobe_results_file,
// src/PARAPROBE_HDF5.cpp:369-408, names from src/PARAPROBE_HDF5.h:66-73), and
// the two-point (NPCORR) cube: its bin counts per axis and bin centers
// (horderdist::report_npc3d_bincenters_hdf5, src/PARAPROBE_SolverHdl.cpp:
// 4506-4548) and one u32 column of counts per task
// (report_npc3d_histvalues_hdf5, S.cpp:4551-4575).  Every dataset is a
// contiguous two-dimensional little-endian matrix, as h5Hdl creates them.
#include "../../../include/pp_h5.h"

#include <hdf5.h>

#include <string>
#include <vector>

namespace {

constexpr const char *kDescrStats = "/DescrSpatStats";
constexpr const char *kTwoPoint = "/DescrSpatStats/TwoPointStats";
constexpr const char *kBinning = "/DescrSpatStats/TwoPointStats/BinNXNYNZ";
constexpr const char *kCenters = "/DescrSpatStats/TwoPointStats/BinCenterXYZ";

// HDF5 prints its error stack by default; the status codes say it here
void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

// Creates dataset `name` (nr x nc of file type ftype) in an existing file and
// writes buf (memory type mtype) into it whole.
int write_matrix(const char *fn, const std::string &name, hid_t ftype, hid_t mtype, uint64_t nr,
                 uint64_t nc, const void *buf)
{
    const hid_t f = H5Fopen(fn, H5F_ACC_RDWR, H5P_DEFAULT);
    if (f < 0)
        return -1;
    const hsize_t dims[2] = {(hsize_t)nr, (hsize_t)nc};
    const hid_t sp = H5Screate_simple(2, dims, dims);
    int rc = -1;
    if (sp >= 0) {
        const hid_t d = H5Dcreate2(f, name.c_str(), ftype, sp, H5P_DEFAULT, H5P_DEFAULT,
                                   H5P_DEFAULT);
        if (d >= 0) {
            rc = (nr * nc == 0 || H5Dwrite(d, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) >= 0)
                     ? 0
                     : -1;
            H5Dclose(d);
        }
        H5Sclose(sp);
    }
    H5Fclose(f);
    return rc;
}

} // namespace

extern "C" {

int pph5_create_results(const char *fn, int32_t npcorr)
{
    quiet();
    if (!fn)
        return -1;
    const hid_t f = H5Fcreate(fn, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0)
        return -1;
    int rc = 0;
    hid_t g = H5Gcreate2(f, kDescrStats, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    rc |= g < 0 ? -1 : 0;
    if (g >= 0)
        H5Gclose(g);
    if (npcorr) {
        g = H5Gcreate2(f, kTwoPoint, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        rc |= g < 0 ? -1 : 0;
        if (g >= 0)
            H5Gclose(g);
    }
    H5Fclose(f);
    return rc;
}

int pph5_report_npc3d_bincenters(const char *fn, uint64_t edge, float width)
{
    quiet();
    if (!fn || edge == 0)
        return -1;
    const uint64_t bins[3] = {edge, edge, edge};
    if (write_matrix(fn, kBinning, H5T_STD_U64LE, H5T_NATIVE_UINT64, 3, 1, bins) != 0)
        return -1;
    // npc3d::get_bincenter: float(i) * width - float((n - 1) / 2), the
    // integer half edge unscaled, x fastest
    const float half = (float)((edge - 1) / 2);
    std::vector<float> xyz;
    xyz.reserve(3 * edge * edge * edge);
    for (uint64_t z = 0; z < edge; ++z)
        for (uint64_t y = 0; y < edge; ++y)
            for (uint64_t x = 0; x < edge; ++x) {
                const float px = (float)x * width, py = (float)y * width, pz = (float)z * width;
                xyz.push_back(px - half);
                xyz.push_back(py - half);
                xyz.push_back(pz - half);
            }
    return write_matrix(fn, kCenters, H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, edge * edge * edge, 3,
                        xyz.data());
}

int pph5_report_npc3d_histvalues(const char *fn, const char *trg, const char *env,
                                 int32_t randomized, const uint32_t *cnts, uint64_t nxyz)
{
    quiet();
    if (!fn || !trg || !env || (!cnts && nxyz))
        return -1;
    const std::string name = std::string(kTwoPoint) + "/" + trg + "_" + env +
                             (randomized ? "_Rndm" : "_Orig");
    return write_matrix(fn, name, H5T_STD_U32LE, H5T_NATIVE_UINT32, nxyz, 1, cnts);
}

int pph5_read(const char *fn, const char *dset, void *buf, uint64_t cap_bytes, uint64_t *nr,
              uint64_t *nc, int32_t *kind)
{
    quiet();
    if (!fn || !dset || !nr || !nc || !kind)
        return -1;
    const hid_t f = H5Fopen(fn, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0)
        return -1;
    int rc = -1;
    const hid_t d = H5Dopen2(f, dset, H5P_DEFAULT);
    if (d >= 0) {
        const hid_t sp = H5Dget_space(d);
        hsize_t dims[2] = {0, 0};
        if (sp >= 0 && H5Sget_simple_extent_ndims(sp) == 2 &&
            H5Sget_simple_extent_dims(sp, dims, nullptr) == 2) {
            *nr = dims[0];
            *nc = dims[1];
            const hid_t t = H5Dget_type(d);
            hid_t mem = -1;
            size_t esz = 0;
            *kind = 0;
            if (H5Tequal(t, H5T_STD_U32LE) > 0) {
                *kind = 1;
                mem = H5T_NATIVE_UINT32;
                esz = 4;
            } else if (H5Tequal(t, H5T_STD_U64LE) > 0) {
                *kind = 2;
                mem = H5T_NATIVE_UINT64;
                esz = 8;
            } else if (H5Tequal(t, H5T_IEEE_F32LE) > 0) {
                *kind = 3;
                mem = H5T_NATIVE_FLOAT;
                esz = 4;
            }
            H5Tclose(t);
            const uint64_t bytes = (uint64_t)dims[0] * dims[1] * esz;
            if (mem < 0)
                rc = -3;
            else if (bytes > cap_bytes || !buf)
                rc = -2;
            else
                rc = (bytes == 0 || H5Dread(d, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) >= 0)
                         ? 0
                         : -1;
        }
        if (sp >= 0)
            H5Sclose(sp);
        H5Dclose(d);
    }
    H5Fclose(f);
    return rc;
}

} // extern "C"
