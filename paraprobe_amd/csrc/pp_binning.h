// pp_binning.h -- exact histogram geometry and the squared-distance edge table.
//
// The reference bins a float distance x with histogram::add
// (src/PARAPROBE_Histogram.h:97-116) using a geometry fixed by the constructor
// (src/PARAPROBE_Histogram.h:54-89):
//   nbins = (size_t)((end - start) / width), tmp = 1.0 / (end - start) * nbins,
//   start/width/end being float settings promoted to double; and
//   xx = (double)x; if (start <= xx <= end) { rb = (float)((xx - start) * tmp);
//   b = floor(rb); b < nbins ? cnts[b] : cnts_highest } else lowest / highest.
// The distance itself is sqrtf(d2) of the float squared distance
// (src/PARAPROBE_KDTree.cpp:571).  Category numbering used everywhere here:
// 0 = cnts_lowest, 1 + b = cnts[b], nbins + 1 = cnts_highest.
//
// Because sqrtf, the double product, the float rounding and floor are all
// monotone, the category is a monotone step function of the bit pattern of
// the non-negative float d2.  EdgeTable holds, for every category c >= 1, the
// smallest d2 bit pattern whose category is >= c, found by binary search
// over bit patterns with the literal expression above.  The GPU then needs no
// FP64 and no correctly-rounded sqrt: it guesses the nearest edge from an
// approximate sqrt and settles the category with one integer compare.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace pp {

struct HistGeometry {
    double start = 0, width = 0, end = 0, tmp = 0;
    uint32_t nbins = 0;
    bool valid = false;

    HistGeometry() = default;
    HistGeometry(float rmin, float rincr, float rmax);
    // category of a float distance
    uint32_t category(float x) const;
    // category of a float squared distance: category(sqrtf(d2))
    uint32_t category_d2(float d2) const;
    uint32_t ncat() const { return nbins + 2; }
};

struct EdgeTable {
    std::vector<uint32_t> edges;  // size ncat, edges[0] = 0 (unused)
    float idx_scale = 0, idx_off = 0;
    double max_offset = 0;        // worst distance of an edge from its ideal index
    // Floor form of the same guess: y = fmaf(sqrt(d2), idx_scale, idx_off1)
    // approximates the index coordinate (d - start) * tmp + 1 within
    // y_err(R) (a v_sqrt_f32 of at most 2 ulp, the float scale and offset, one
    // fma rounding).  Wherever y lies at least max_offset + y_err from every
    // integer, floor(y) is the category; closer to an integer, the edge table
    // decides.  amb_margin(R) = max_offset + y_err for distances up to R.
    float idx_off1 = 0;
    // bound on |y - ((d - start) tmp + 1)| for distances up to R
    double y_err(const HistGeometry &g, double R) const;
    // the kernel's binning is exact up to R when max_offset + y_err < 0.5 (the
    // near-edge path rounds y to the nearest index and settles it with the
    // table); jobs beyond that are refused
    bool exact_up_to(const HistGeometry &g, double R) const { return max_offset + y_err(g, R) < 0.5; }
    float amb_margin(const HistGeometry &g, double R) const;
    // Builds the table; returns false (with a reason) when the bins are too
    // narrow for a nearest-edge guess to be provably within one edge.
    bool build(const HistGeometry &g, std::string *why);
};

} // namespace pp
