// pp_binning.cpp -- exact histogram geometry and the squared-distance edge table.
#include "pp_binning.h"

#include <cmath>
#include <cstring>

namespace pp {

static constexpr double kEpsilon = 1.0e-6;  // src/PARAPROBE_Numerics.h:45

static inline double ncat_of(const HistGeometry &g) { return (double)g.nbins + 2.0; }

static inline float f_of_bits(uint32_t u)
{
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

HistGeometry::HistGeometry(float rmin, float rincr, float rmax)
{
    const double s = rmin, w = rincr, e = rmax;
    if (w > kEpsilon) {
        const size_t nb = (size_t)((e - s) / w);
        if (nb <= 0xFFFFFFFFull) {
            start = s;
            width = w;
            end = e;
            nbins = (uint32_t)nb;
            tmp = 1.0 / (e - s) * (double)nb;
            valid = true;
        }
    }
}

uint32_t HistGeometry::category(float x) const
{
    const double xx = x;
    if (xx >= start && xx <= end) {
        const float rb = (float)((xx - start) * tmp);
        const unsigned int b = (unsigned int)std::floor(rb);
        return b < nbins ? 1u + b : nbins + 1u;
    }
    return xx < start ? 0u : nbins + 1u;
}

uint32_t HistGeometry::category_d2(float d2) const { return category(std::sqrt(d2)); }

bool EdgeTable::build(const HistGeometry &g, std::string *why)
{
    if (!g.valid || g.nbins == 0) {
        if (why)
            *why = "invalid histogram geometry (width <= 1e-6 or no bins)";
        return false;
    }
    const uint32_t C = g.ncat();
    edges.assign(C, 0u);
    const uint32_t top = 0x7F800000u;  // +inf: every finite non-negative d2 is below
    for (uint32_t c = 1; c < C; ++c) {
        // smallest u in [0, top] with category(u) >= c (category(top) is
        // highest).  Bracket it by galloping from the ideal edge position
        // (start + (c - 1) / tmp)^2, then bisect: a few evaluations per edge.
        if (g.category_d2(f_of_bits(0)) >= c) {
            edges[c] = 0;
            continue;
        }
        const double dc = g.start + (double)(c - 1) / g.tmp;
        const float guess = (float)(dc * dc);
        uint32_t u0;
        std::memcpy(&u0, &guess, 4);
        if (!(guess >= 0.f) || u0 > top)
            u0 = top / 2;
        uint32_t lo, hi;  // invariant: category(lo) < c <= category(hi)
        if (g.category_d2(f_of_bits(u0)) >= c) {
            hi = u0;
            uint32_t step = 1;
            for (;;) {
                const uint32_t t = hi > step ? hi - step : 0;
                if (t == 0 || g.category_d2(f_of_bits(t)) < c) {
                    lo = t;
                    break;
                }
                hi = t;
                step = step < 0x40000000u ? step * 2 : step;
            }
        } else {
            lo = u0;
            uint32_t step = 1;
            for (;;) {
                const uint32_t t = (top - lo > step) ? lo + step : top;
                if (t == top || g.category_d2(f_of_bits(t)) >= c) {
                    hi = t;
                    break;
                }
                lo = t;
                step = step < 0x40000000u ? step * 2 : step;
            }
        }
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (g.category_d2(f_of_bits(mid)) >= c)
                hi = mid;
            else
                lo = mid;
        }
        edges[c] = hi;
    }
    // Edge c sits at distance start + (c - 1) / tmp in the ideal geometry, so
    // the index space coordinate y(d) = (d - start) * tmp + 1 puts it at c.
    // The device rounds an approximate y to the nearest integer q and answers
    // q - 1 + (bits(d2) >= edges[q]); that is exact whenever every d2 lies
    // within 0.5 of an index between its two bracketing edges, i.e. when
    // every exact edge is within (0.5 - device error) of its ideal index.
    max_offset = 0;
    for (uint32_t c = 1; c < C; ++c) {
        if (c > 1 && edges[c] <= edges[c - 1]) {
            if (why)
                *why = "empty histogram category; bins narrower than float resolution";
            return false;
        }
        if (c == 1 && g.start <= 0.0)
            continue;  // edge 1 at d = 0: every d2 is at or above it, the clamp suffices
        const uint32_t probes[2] = {edges[c], edges[c] > 0 ? edges[c] - 1 : 0};
        for (uint32_t u : probes) {
            const double d = std::sqrt((double)f_of_bits(u));
            const double y = (d - g.start) * g.tmp + 1.0;
            const double off = std::fabs(y - (double)c);
            if (off > max_offset)
                max_offset = off;
        }
    }
    if (max_offset > 0.25) {
        if (why)
            *why = "bin edges too far from their ideal positions for the nearest-edge guess";
        return false;
    }
    // the kernel turns the floor-form index into an integer with a
    // saturating unsigned conversion (negative -> 0, no upper clamp): an
    // in-ball index y stays below ncat as long as its rounding (a few 2^-24
    // of R tmp + |1 - start tmp|) stays far below 1
    if (std::fabs(g.end) * g.tmp + std::fabs(1.0 - g.start * g.tmp) + 2.0 >= 1048576.0) {
        if (why)
            *why = "too many bins for the float index coordinate";
        return false;
    }
    idx_scale = (float)g.tmp;
    idx_off = (float)(1.5 - g.start * g.tmp);
    idx_off1 = (float)(1.0 - g.start * g.tmp);
    return true;
}

double EdgeTable::y_err(const HistGeometry &g, double R) const
{
    // |y - ((d - start) tmp + 1)| for d <= R, with A = R tmp and
    // B = |1 - start tmp|: the square root's error (v_sqrt_f32, at most 2 ulp,
    // 2^-22 relative) scaled, 2.4e-7 A; the float scale's rounding (2^-24
    // relative), 6e-8 A; the fma's rounding of a result below A + B + 1 (the
    // shifted coordinate adds at most 0.5), 6e-8 (A + B + 1); the float
    // offset's rounding, 6e-8 (B + 1)
    const double A = std::fabs(R) * g.tmp, B = std::fabs(1.0 - g.start * g.tmp);
    return 2.4e-7 * A + 6e-8 * A + 6e-8 * (A + B + 1.0) + 6e-8 * (B + 1.0);
}

float EdgeTable::amb_margin(const HistGeometry &g, double R) const
{
    // max_offset plus 1.5 times the rounding bound.  At 0.5 every lane takes
    // the near-edge path, which is exact while exact_up_to(g, R) holds (jobs
    // where it does not are refused before they run), so the cap loses nothing.
    const double m = max_offset + 1.5 * y_err(g, R) + 1e-6;
    return (float)std::min(m, 0.5);
}

} // namespace pp
