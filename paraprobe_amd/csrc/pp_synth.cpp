// pp_synth.cpp -- seeded synthetic atom-probe tips for the spatial-statistics path.
//
// The reference synthesises tips in solverHdl::generate_synthetic_tip
// (src/PARAPROBE_SolverHdl.cpp:6596-6794) from a frustum-with-caps geometry
// (geomodel::is_inside, src/PARAPROBE_Datatypes.cpp:443-553) and spherical
// second-phase clusters (secondphasemodel, src/PARAPROBE_Datatypes.cpp:780-845).
// The workloads measured here are defined differently (uniform random ions in a
// cylinder or a conical frustum, i.i.d. type labels, Poisson-placed clusters
// with an enriched type), so this is a generator of its own:
//   * positions: uniform in the solid by rejection from the bounding cylinder;
//   * types: labels 1..ntypes drawn i.i.d. from `frac`, or from `cfrac` for ions
//     inside any cluster sphere;
//   * ion2surf_sqr: squared distance to the analytic hull (lateral surface and
//     both caps), evaluated in double and rounded to float; this stands in for
//     the CGAL-hull distance the reference computes in surface_distancing2
//     (S.cpp:312-484).
// Every ion is generated from a counter-based stream (splitmix64 over
// seed/chunk/index), so the output does not depend on the thread count.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

struct Stream {
    uint64_t s;
    explicit Stream(uint64_t seed) : s(splitmix64(seed)) {}
    uint64_t next() { return s = splitmix64(s); }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Solid {
    int shape; // 0 = cylinder, 1 = conical frustum
    double r0, r1, h;
    double rmax() const { return shape == 0 ? r0 : std::fmax(r0, r1); }
    double radius_at(double z) const { return shape == 0 ? r0 : r0 + (r1 - r0) * (z / h); }
    bool inside(double x, double y, double z) const
    {
        if (z < 0.0 || z > h)
            return false;
        double r = radius_at(z);
        return x * x + y * y <= r * r;
    }
    // distance of an interior point to the hull: lateral surface and both caps
    double dist(double x, double y, double z) const
    {
        double rho = std::sqrt(x * x + y * y);
        double dcap = std::fmin(z, h - z);
        double dlat;
        if (shape == 0) {
            dlat = r0 - rho;
        } else {
            // distance in the (rho, z) half plane to the segment (r0,0)-(r1,h)
            double ax = r0, az = 0.0, bx = r1, bz = h;
            double vx = bx - ax, vz = bz - az;
            double t = ((rho - ax) * vx + (z - az) * vz) / (vx * vx + vz * vz);
            t = std::fmin(1.0, std::fmax(0.0, t));
            double px = ax + t * vx - rho, pz = az + t * vz - z;
            dlat = std::sqrt(px * px + pz * pz);
        }
        double d = std::fmin(dlat, dcap);
        return d < 0.0 ? 0.0 : d;
    }
};

struct ClusterGrid {
    double cell;
    double lo[3];
    int64_t dim[3];
    std::vector<int64_t> start;
    std::vector<double> cx, cy, cz, cr;
    std::vector<int64_t> order;
    int64_t coord(int a, double v) const
    {
        int64_t c = (int64_t)std::floor((v - lo[a]) / cell);
        return c < 0 ? 0 : (c >= dim[a] ? dim[a] - 1 : c);
    }
    bool covered(double x, double y, double z) const
    {
        if (cr.empty())
            return false;
        int64_t c0 = coord(0, x), c1 = coord(1, y), c2 = coord(2, z);
        for (int64_t k = c2 - 1; k <= c2 + 1; ++k) {
            if (k < 0 || k >= dim[2])
                continue;
            for (int64_t j = c1 - 1; j <= c1 + 1; ++j) {
                if (j < 0 || j >= dim[1])
                    continue;
                for (int64_t i = c0 - 1; i <= c0 + 1; ++i) {
                    if (i < 0 || i >= dim[0])
                        continue;
                    int64_t cc = (k * dim[1] + j) * dim[0] + i;
                    for (int64_t p = start[cc]; p < start[cc + 1]; ++p) {
                        int64_t q = order[p];
                        double dx = x - cx[q], dy = y - cy[q], dz = z - cz[q];
                        if (dx * dx + dy * dy + dz * dz <= cr[q] * cr[q])
                            return true;
                    }
                }
            }
        }
        return false;
    }
};

} // namespace

extern "C" {

// shape: 0 cylinder (radius r0), 1 frustum (radius r0 at z=0, r1 at z=h).
// frac/cfrac: ntypes probabilities (labels 1..ntypes).  Returns 0 on success.
int pp_synth_tip(uint64_t n, uint64_t seed, int32_t shape, double r0, double r1, double h,
                 int32_t ntypes, const double *frac, uint64_t nclusters, double crmin,
                 double crmax, const double *cfrac, float *xyz, uint8_t *type,
                 float *ion2surf_sqr)
{
    if (ntypes < 1 || ntypes > 255 || h <= 0.0 || r0 <= 0.0 || (shape == 1 && r1 <= 0.0))
        return 1;
    Solid s{shape, r0, shape == 0 ? r0 : r1, h};
    const double R = s.rmax();

    std::vector<double> cum(ntypes), ccum(ntypes);
    double acc = 0.0, cacc = 0.0;
    for (int t = 0; t < ntypes; ++t) {
        acc += frac[t];
        cum[t] = acc;
        cacc += (cfrac ? cfrac[t] : frac[t]);
        ccum[t] = cacc;
    }
    for (int t = 0; t < ntypes; ++t) {
        cum[t] /= acc;
        ccum[t] /= cacc;
    }

    // cluster centres: uniform in the solid, radius U(crmin, crmax)
    ClusterGrid cg;
    if (nclusters > 0) {
        Stream cs(seed ^ 0xC1C1C1C1ull);
        cg.cx.resize(nclusters);
        cg.cy.resize(nclusters);
        cg.cz.resize(nclusters);
        cg.cr.resize(nclusters);
        for (uint64_t c = 0; c < nclusters; ++c) {
            double x, y, z;
            do {
                x = (2.0 * cs.uniform() - 1.0) * R;
                y = (2.0 * cs.uniform() - 1.0) * R;
                z = cs.uniform() * h;
            } while (!s.inside(x, y, z));
            cg.cx[c] = x;
            cg.cy[c] = y;
            cg.cz[c] = z;
            cg.cr[c] = crmin + (crmax - crmin) * cs.uniform();
        }
        cg.cell = std::fmax(2.0 * crmax, 1e-3);
        cg.lo[0] = -R - cg.cell;
        cg.lo[1] = -R - cg.cell;
        cg.lo[2] = -cg.cell;
        cg.dim[0] = cg.dim[1] = (int64_t)std::ceil((2.0 * R + 2.0 * cg.cell) / cg.cell) + 1;
        cg.dim[2] = (int64_t)std::ceil((h + 2.0 * cg.cell) / cg.cell) + 1;
        int64_t nc = cg.dim[0] * cg.dim[1] * cg.dim[2];
        cg.start.assign(nc + 1, 0);
        std::vector<int64_t> cell(nclusters);
        for (uint64_t c = 0; c < nclusters; ++c) {
            cell[c] = (cg.coord(2, cg.cz[c]) * cg.dim[1] + cg.coord(1, cg.cy[c])) * cg.dim[0] +
                      cg.coord(0, cg.cx[c]);
            cg.start[cell[c] + 1]++;
        }
        for (int64_t c = 0; c < nc; ++c)
            cg.start[c + 1] += cg.start[c];
        std::vector<int64_t> fill(cg.start.begin(), cg.start.end() - 1);
        cg.order.resize(nclusters);
        for (uint64_t c = 0; c < nclusters; ++c)
            cg.order[fill[cell[c]]++] = (int64_t)c;
    }

    const uint64_t chunk = 1u << 16;
    const int64_t nchunks = (int64_t)((n + chunk - 1) / chunk);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t ch = 0; ch < nchunks; ++ch) {
        Stream st(seed * 0x100000001B3ull + (uint64_t)ch * 0x9E3779B97F4A7C15ull + 17u);
        uint64_t a = (uint64_t)ch * chunk, b = a + chunk < n ? a + chunk : n;
        for (uint64_t i = a; i < b; ++i) {
            double x, y, z;
            do {
                x = (2.0 * st.uniform() - 1.0) * R;
                y = (2.0 * st.uniform() - 1.0) * R;
                z = st.uniform() * h;
            } while (!s.inside(x, y, z));
            float fx = (float)x, fy = (float)y, fz = (float)z;
            xyz[3 * i] = fx;
            xyz[3 * i + 1] = fy;
            xyz[3 * i + 2] = fz;
            bool inc = cg.covered(fx, fy, fz);
            const std::vector<double> &cu = inc ? ccum : cum;
            double u = st.uniform();
            int t = 0;
            while (t < ntypes - 1 && u >= cu[t])
                ++t;
            type[i] = (uint8_t)(t + 1);
            double d = s.dist(fx, fy, fz);
            ion2surf_sqr[i] = (float)(d * d);
        }
    }
    return 0;
}

// Triangulated hull of the cylinder / frustum solid of pp_synth_tip (the
// input surface_distancing2 needs; the reference builds it with CGAL alpha
// shapes, src/PARAPROBE_SolverHdl.cpp:2477-2690): the lateral surface as nseg
// x nz quads split into two triangles, and both caps as fans around their
// centre.  Vertices lie on the exact surface.  Writes 9 floats per triangle;
// returns the triangle count (2 nseg nz + 2 nseg), or -1 for bad parameters.
int64_t pp_synth_hull(int32_t shape, double r0, double r1, double h, int32_t nseg, int32_t nz,
                      float *tri, int64_t cap)
{
    if (nseg < 3 || nz < 1 || h <= 0.0 || r0 <= 0.0 || (shape == 1 && r1 <= 0.0))
        return -1;
    const int64_t ntri = 2 * (int64_t)nseg * nz + 2 * (int64_t)nseg;
    if (!tri)
        return ntri;
    if (cap < ntri)
        return -1;
    Solid sol{shape, r0, shape == 0 ? r0 : r1, h};
    const double two_pi = 6.283185307179586476925;
    auto vert = [&](int i, int k, float *o) {
        const double z = h * (double)k / (double)nz;
        const double r = sol.radius_at(z);
        const double a = two_pi * (double)(i % nseg) / (double)nseg;
        o[0] = (float)(r * std::cos(a));
        o[1] = (float)(r * std::sin(a));
        o[2] = (float)z;
    };
    int64_t w = 0;
    for (int k = 0; k < nz; ++k)
        for (int i = 0; i < nseg; ++i) {
            float a[3], b[3], c[3], d[3];
            vert(i, k, a);
            vert(i + 1, k, b);
            vert(i + 1, k + 1, c);
            vert(i, k + 1, d);
            const float *t0[3] = {a, b, c}, *t1[3] = {a, c, d};
            for (const float *const *t : {t0, t1}) {
                for (int v = 0; v < 3; ++v)
                    for (int q = 0; q < 3; ++q)
                        tri[9 * w + 3 * v + q] = t[v][q];
                ++w;
            }
        }
    for (int cap_k : {0, nz})
        for (int i = 0; i < nseg; ++i) {
            float a[3], b[3];
            vert(i, cap_k, a);
            vert(i + 1, cap_k, b);
            const float ctr[3] = {0.f, 0.f, (float)(h * (double)cap_k / (double)nz)};
            const float *t[3] = {ctr, a, b};
            for (int v = 0; v < 3; ++v)
                for (int q = 0; q < 3; ++q)
                    tri[9 * w + 3 * v + q] = t[v][q];
            ++w;
        }
    return w;
}

// Simple cubic / fcc lattice block for closed-form tests: ions at
// a*(i + basis) for i in [0,nx)x[0,ny)x[0,nz); fcc uses the 4-atom basis.
// Returns the number of points written (at most cap).
int64_t pp_synth_lattice(int32_t fcc, double a, int32_t nx, int32_t ny, int32_t nz, float *xyz,
                         int64_t cap)
{
    static const double basis_fcc[4][3] = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
    int nb = fcc ? 4 : 1;
    int64_t w = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                for (int b = 0; b < nb; ++b) {
                    if (w >= cap)
                        return w;
                    xyz[3 * w] = (float)(a * (i + basis_fcc[b][0]));
                    xyz[3 * w + 1] = (float)(a * (j + basis_fcc[b][1]));
                    xyz[3 * w + 2] = (float)(a * (k + basis_fcc[b][2]));
                    ++w;
                }
    return w;
}

} // extern "C"
