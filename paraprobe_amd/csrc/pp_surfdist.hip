// pp_surfdist.hip -- gfx950 kernels for the ion-to-tip-surface squared
// distance that excludes centers near the surface (ion2surf).
//
// Replaces solver::surface_distancing2 (src/PARAPROBE_SolverHdl.cpp:312-484):
// every ion starts at fl32(R * R), R = max(SpatStatRadiusMax,
// SurfaceCellsCarvingRadius); every triangle of the hull whose vertex
// bounding box overlaps the ion's float box [p - R, p + R]
// (AABB::overlaps, src/PARAPROBE_AABBTree.cpp:112-123) is a candidate, and the
// ion keeps the smallest mathHdl::closestPointOnTriangle value
// (src/PARAPROBE_Math.cpp:88-171) not above that start.  The reference finds
// the candidates with an R-tree; here the triangles are binned into a uniform
// grid (each into every cell its R-inflated box touches) and one thread per
// ion tests the triangles of its own cell with the reference's exact overlap
// test and FP32 expression (no contraction, IEEE division).
#include "pp_launch.h"

#pragma clang fp contract(off)

namespace pp {

namespace {

inline unsigned blocks_for(uint64_t n, unsigned block)
{
    uint64_t g = (n + block - 1) / block;
    if (g > 0x7FFFFFFFull)
        g = 0x7FFFFFFFull;
    return (unsigned)(g ? g : 1);
}

// the reference's MIN / MAX / CLAMP macros (src/PARAPROBE_Abinitio.h:47-49):
// ternaries, so a NaN argument clamps to hi
__device__ inline float ref_min(float x, float y) { return (x < y) ? x : y; }
__device__ inline float ref_max(float x, float y) { return (x < y) ? y : x; }
__device__ inline float ref_clamp01(float x) { return ref_min(ref_max(x, 0.f), 1.f); }

} // namespace

// mathHdl::closestPointOnTriangle (src/PARAPROBE_Math.cpp:88-171): squared
// distance from (sx, sy, sz) to triangle t[0..8], same operations in the same
// order.
__host__ __device__ inline float closest_sqr(const float *t, float sx, float sy, float sz)
{
    const float e0u = t[3] - t[0], e0v = t[4] - t[1], e0w = t[5] - t[2];
    const float e1u = t[6] - t[0], e1v = t[7] - t[1], e1w = t[8] - t[2];
    const float v0u = t[0] - sx, v0v = t[1] - sy, v0w = t[2] - sz;
    const float a = e0u * e0u + e0v * e0v + e0w * e0w;
    const float b = e0u * e1u + e0v * e1v + e0w * e1w;
    const float c = e1u * e1u + e1v * e1v + e1w * e1w;
    const float d = e0u * v0u + e0v * v0v + e0w * v0w;
    const float e = e1u * v0u + e1v * v0v + e1w * v0w;
    const float det = a * c - b * b;
    float s = b * e - c * d;
    float tt = b * d - a * e;
    auto clamp01 = [](float x) {
        const float m = (x < 0.f) ? 0.f : x;
        return (m < 1.f) ? m : 1.f;
    };
    if (s + tt < det) {
        if (s < 0.f) {
            if (tt < 0.f) {  // region 4
                if (d < 0.f) {
                    s = clamp01(-d / a);
                    tt = 0.f;
                } else {
                    s = 0.f;
                    tt = clamp01(-e / c);
                }
            } else {  // region 3
                s = 0.f;
                tt = clamp01(-e / c);
            }
        } else if (tt < 0.f) {  // region 5
            s = clamp01(-d / a);
            tt = 0.f;
        } else {  // region 0
            const float invDet = 1.f / det;
            s *= invDet;
            tt *= invDet;
        }
    } else {
        if (s < 0.f) {  // region 2
            const float tmp0 = b + d;
            const float tmp1 = c + e;
            if (tmp1 > tmp0) {
                const float numer = tmp1 - tmp0;
                const float denom = a - 2.f * b + c;
                s = clamp01(numer / denom);
                tt = 1.f - s;
            } else {
                tt = clamp01(-e / c);
                s = 0.f;
            }
        } else if (tt < 0.f) {  // region 6
            if (a + d > b + e) {
                const float numer = c + e - b - d;
                const float denom = a - 2.f * b + c;
                s = clamp01(numer / denom);
                tt = 1.f - s;
            } else {
                s = clamp01(-e / c);
                tt = 0.f;
            }
        } else {  // region 1
            const float numer = c + e - b - d;
            const float denom = a - 2.f * b + c;
            s = clamp01(numer / denom);
            tt = 1.f - s;
        }
    }
    const float cx = t[0] + s * e0u + tt * e1u;
    const float cy = t[1] + s * e0v + tt * e1v;
    const float cz = t[2] + s * e0w + tt * e1w;
    const float dx = cx - sx, dy = cy - sy, dz = cz - sz;
    return dx * dx + dy * dy + dz * dz;
}

// Triangle boxes (fmin / fmax of the vertices, src/PARAPROBE_SolverHdl.cpp:
// 2716-2721) and the range of grid cells their box inflated by R touches,
// widened by a relative 2^-18 so that every ion passing the float overlap
// test lies inside it.
__device__ inline void tri_cells(const float *box, float R, const GridGeom &g, int lo[3],
                                 int hi[3])
{
    for (int k = 0; k < 3; ++k) {
        const double r = (double)R * (1.0 + 0x1p-18) + 1e-6;
        const double a = ((double)box[k] - r - g.lo[k]) * g.inv_w;
        const double b = ((double)box[3 + k] + r - g.lo[k]) * g.inv_w;
        const int n = k == 0 ? g.nx : (k == 1 ? g.ny : g.nz);
        int l = (int)floor(a), h = (int)floor(b);
        l = l < 0 ? 0 : (l >= n ? n - 1 : l);
        h = h < 0 ? 0 : (h >= n ? n - 1 : h);
        lo[k] = l;
        hi[k] = h;
    }
}

__global__ __launch_bounds__(256) void k_tribox(const float *__restrict__ tri, uint64_t ntri,
                                                float R, GridGeom g, float *__restrict__ box,
                                                uint32_t *__restrict__ count)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntri)
        return;
    const float *t = tri + 9 * i;
    float b[6];
    for (int k = 0; k < 3; ++k) {
        b[k] = fminf(fminf(t[k], t[3 + k]), t[6 + k]);
        b[3 + k] = fmaxf(fmaxf(t[k], t[3 + k]), t[6 + k]);
    }
    for (int k = 0; k < 6; ++k)
        box[6 * i + k] = b[k];
    int lo[3], hi[3];
    tri_cells(b, R, g, lo, hi);
    for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
            for (int x = lo[0]; x <= hi[0]; ++x)
                atomicAdd(&count[((uint64_t)z * g.ny + y) * g.nx + x], 1u);
}

__global__ __launch_bounds__(256) void k_trifill(const float *__restrict__ box, uint64_t ntri,
                                                 float R, GridGeom g, uint32_t *__restrict__ cursor,
                                                 uint32_t *__restrict__ list)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntri)
        return;
    int lo[3], hi[3];
    tri_cells(box + 6 * i, R, g, lo, hi);
    for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
            for (int x = lo[0]; x <= hi[0]; ++x)
                list[atomicAdd(&cursor[((uint64_t)z * g.ny + y) * g.nx + x], 1u)] = (uint32_t)i;
}

// One thread per ion: the candidates of its cell, the reference's overlap
// test against the float box [p - R, p + R], the smallest squared distance
// not above R^2 (NaN values of degenerate triangles are skipped).
__global__ __launch_bounds__(256) void k_ion2surf(const float *__restrict__ xyz, uint64_t n,
                                                  const float *__restrict__ tri,
                                                  const float *__restrict__ box, GridGeom g,
                                                  const uint32_t *__restrict__ cellstart,
                                                  const uint32_t *__restrict__ list, float R,
                                                  float *__restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float px = xyz[3 * i], py = xyz[3 * i + 1], pz = xyz[3 * i + 2];
    const float RSQR = R * R;
    const float qlx = px - R, qly = py - R, qlz = pz - R;
    const float qux = px + R, quy = py + R, quz = pz + R;
    int c[3];
    const float p[3] = {px, py, pz};
    for (int k = 0; k < 3; ++k) {
        const int nk = k == 0 ? g.nx : (k == 1 ? g.ny : g.nz);
        int v = (int)floor(((double)p[k] - g.lo[k]) * g.inv_w);
        c[k] = v < 0 ? 0 : (v >= nk ? nk - 1 : v);
    }
    const uint64_t cell = ((uint64_t)c[2] * g.ny + c[1]) * g.nx + c[0];
    float best = RSQR;
    for (uint32_t q = cellstart[cell]; q < cellstart[cell + 1]; ++q) {
        const uint32_t t = list[q];
        const float *b = box + 6 * (uint64_t)t;
        const bool overlap = !(qux < b[0] || qlx > b[3] || quy < b[1] || qly > b[4] ||
                               quz < b[2] || qlz > b[5]);
        if (!overlap)
            continue;
        const float v = closest_sqr(tri + 9 * (uint64_t)t, px, py, pz);
        if (v <= best)  // the reference keeps v unless v > best
            best = v;
    }
    out[i] = best;
}

hipError_t launch_tribox(const float *tri, uint64_t ntri, float R, GridGeom g, float *box,
                         uint32_t *count, hipStream_t s)
{
    if (ntri == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_tribox, dim3(blocks_for(ntri, 256)), dim3(256), 0, s, tri, ntri, R, g,
                       box, count);
    return hipGetLastError();
}

hipError_t launch_trifill(const float *box, uint64_t ntri, float R, GridGeom g, uint32_t *cursor,
                          uint32_t *list, hipStream_t s)
{
    if (ntri == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_trifill, dim3(blocks_for(ntri, 256)), dim3(256), 0, s, box, ntri, R, g,
                       cursor, list);
    return hipGetLastError();
}

hipError_t launch_ion2surf(const float *xyz, uint64_t n, const float *tri, const float *box,
                           GridGeom g, const uint32_t *cellstart, const uint32_t *list, float R,
                           float *out, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_ion2surf, dim3(blocks_for(n, 256)), dim3(256), 0, s, xyz, n, tri, box, g,
                       cellstart, list, R, out);
    return hipGetLastError();
}

} // namespace pp
