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
// distance from (px, py, pz) to the triangle with vertex 1 at (x1, y1, z1)
// and edges e0 = V2 - V1, e1 = V3 - V1 as the reference computes them.  The
// seven regions are resolved with selects instead of branches (lanes of a
// wavefront hit different regions), but each region's values come from the
// same operations in the same order as the reference, so the result is the
// same float:
//   region 0 (s+t < det, s >= 0, t >= 0): s, t times 1.f/det
//   regions 4 (d < 0) and 5:     s = CLAMP(-d/a), t = 0
//   regions 4 (d >= 0), 3, 2b:   s = 0, t = CLAMP(-e/c)
//   region 6b:                   s = CLAMP(-e/c), t = 0 (the reference's
//                                choice; the geometry would use -d/a)
//   region 2a:                   s = CLAMP(((c+e) - (b+d)) / denom), t = 1-s
//   regions 6a and 1:            s = CLAMP((((c+e) - b) - d) / denom), t = 1-s
// with denom = (a - 2b) + c and CLAMP(x) = MIN(MAX(x, 0), 1) as ternaries
// (src/PARAPROBE_Abinitio.h:47-49: a NaN clamps to 1).
__device__ inline float closest_sqr(float x1, float y1, float z1, float e0u, float e0v, float e0w,
                                    float e1u, float e1v, float e1w, float px, float py, float pz)
{
    const float v0u = x1 - px, v0v = y1 - py, v0w = z1 - pz;
    const float a = e0u * e0u + e0v * e0v + e0w * e0w;
    const float b = e0u * e1u + e0v * e1v + e0w * e1w;
    const float c = e1u * e1u + e1v * e1v + e1w * e1w;
    const float d = e0u * v0u + e0v * v0v + e0w * v0w;
    const float e = e1u * v0u + e1v * v0v + e1w * v0w;
    const float det = a * c - b * b;
    const float s = b * e - c * d;
    const float t = b * d - a * e;
    const bool inner = s + t < det;
    const bool sneg = s < 0.f, tneg = t < 0.f;
    const float tmp0 = b + d, tmp1 = c + e;
    const bool r0 = inner && !sneg && !tneg;
    const bool da = inner && ((sneg && tneg && d < 0.f) || (!sneg && tneg));
    const bool r6 = !inner && !sneg && tneg;
    const bool r6a = r6 && (a + d > b + e);
    const bool r2a = !inner && sneg && (tmp1 > tmp0);
    const bool ec = (inner && sneg && !(tneg && d < 0.f)) || (!inner && sneg && !r2a) ||
                    (r6 && !r6a);
    const float denom = a - 2.f * b + c;
    const float n1 = c + e - b - d;
    const float num = r0 ? 1.f : (da ? -d : (ec ? -e : (r2a ? tmp1 - tmp0 : n1)));
    const float den = r0 ? det : (da ? a : (ec ? c : denom));
    const float q = num / den;
    const float m = (q < 0.f) ? 0.f : q;
    const float cq = (m < 1.f) ? m : 1.f;
    // s-only: regions 4a, 5, 6b; t-only: 4b, 3, 2b; both from cq: 2a, 6a, 1
    const bool t_only = ec && !r6;
    const bool s_only = da || (ec && r6);
    float ss, tt;
    if (r0) {
        ss = s * q;
        tt = t * q;
    } else {
        ss = t_only ? 0.f : cq;
        tt = t_only ? cq : (s_only ? 0.f : 1.f - cq);
    }
    const float cx = x1 + ss * e0u + tt * e1u;
    const float cy = y1 + ss * e0v + tt * e1v;
    const float cz = z1 + ss * e0w + tt * e1w;
    const float dx = cx - px, dy = cy - py, dz = cz - pz;
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

// Ions into cell order: (x, y, z, original index as bits).
__global__ __launch_bounds__(256) void k_ionscatter(const float *__restrict__ xyz, uint64_t n,
                                                    const uint32_t *__restrict__ key,
                                                    const uint32_t *__restrict__ rank,
                                                    const uint32_t *__restrict__ cellstart,
                                                    float4 *__restrict__ ions)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    ions[cellstart[key[i]] + rank[i]] =
        make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], __uint_as_float((uint32_t)i));
}

// One workgroup per tile (up to 256 ions of one cell, one ion per thread):
// the cell's triangles pass through LDS in chunks of 256, and every lane
// keeps the smallest closest-point value not above R^2 over the triangles
// that pass the reference's overlap test against its float box [p - R,
// p + R].  A triangle is first tested against a lower bound of its value:
// the computed closest point lies in the triangle's box widened by delta
// (its barycentric coordinates stay in the simplex up to a few roundings;
// delta = 2^-16 max|coordinate| covers them with a wide margin), so the value
// is at least the squared distance from p to that widened box (from the box
// centre and half extents plus delta).  A bound above best * (1 + 2^-18)
// proves the value is above best, where the reference would not take it
// either; only lanes whose bound passes run the exact overlap test and the
// closest-point expression.  NaN values (only from overflowing arithmetic)
// are skipped.
__global__ __launch_bounds__(256) void k_surftile(const float4 *__restrict__ ions,
                                                  const SurfTile *__restrict__ tiles,
                                                  const uint32_t *__restrict__ tcellstart,
                                                  const uint32_t *__restrict__ list,
                                                  const float *__restrict__ tri,
                                                  const float *__restrict__ box, float R,
                                                  float delta, float *__restrict__ out)
{
    // per triangle of the chunk, 24 floats: centre (3), half extent + delta
    // (3), box (6), vertex 1 (3), edges e0 and e1 (6), padding (3)
    __shared__ float4 s_r[256][6];
    const SurfTile tl = tiles[blockIdx.x];
    const int tid = threadIdx.x;
    const bool act = tid < (int)tl.count;
    // inactive lanes sit at infinity: every bound fails for them
    const float4 p = act ? ions[tl.first + tid]
                         : make_float4(__builtin_inff(), __builtin_inff(), __builtin_inff(), 0.f);
    const float px = p.x, py = p.y, pz = p.z;
    const float RSQR = R * R;
    const float qlx = px - R, qly = py - R, qlz = pz - R;
    const float qux = px + R, quy = py + R, quz = pz + R;
    constexpr float kGrow = 1.0f + 0x1p-18f;
    float best = RSQR, bestg = RSQR * kGrow;
    const uint32_t ts = __builtin_amdgcn_readfirstlane(tcellstart[tl.cell]);
    const uint32_t te = __builtin_amdgcn_readfirstlane(tcellstart[tl.cell + 1]);
    for (uint32_t base = ts; base < te; base += 256) {
        const uint32_t m = __builtin_amdgcn_readfirstlane((te - base) < 256u ? (te - base) : 256u);
        __syncthreads();
        if ((uint32_t)tid < m) {
            const uint64_t t = list[base + tid];
            const float *bx = box + 6 * t;
            const float b0 = bx[0], b1 = bx[1], b2 = bx[2], b3 = bx[3], b4 = bx[4], b5 = bx[5];
            const float *v = tri + 9 * t;
            const float v0 = v[0], v1 = v[1], v2 = v[2], v3 = v[3], v4 = v[4], v5 = v[5];
            const float v6 = v[6], v7 = v[7], v8 = v[8];
            s_r[tid][0] = make_float4(0.5f * (b0 + b3), 0.5f * (b1 + b4), 0.5f * (b2 + b5),
                                      0.5f * (b3 - b0) + delta);
            s_r[tid][1] = make_float4(0.5f * (b4 - b1) + delta, 0.5f * (b5 - b2) + delta, b0, b1);
            s_r[tid][2] = make_float4(b2, b3, b4, b5);
            // vertex 1 and the two edges as the reference forms them
            s_r[tid][3] = make_float4(v0, v1, v2, v3 - v0);
            s_r[tid][4] = make_float4(v4 - v1, v5 - v2, v6 - v0, v7 - v1);
            s_r[tid][5] = make_float4(v8 - v2, 0.f, 0.f, 0.f);
        }
        __syncthreads();
        for (uint32_t j = 0; j < m; ++j) {
            const float4 r0 = s_r[j][0], r1 = s_r[j][1];
            const float gx = fmaxf(fabsf(px - r0.x) - r0.w, 0.f);
            const float gy = fmaxf(fabsf(py - r0.y) - r1.x, 0.f);
            const float gz = fmaxf(fabsf(pz - r0.z) - r1.y, 0.f);
            const bool pass = !(gx * gx + gy * gy + gz * gz > bestg);
            if (__builtin_amdgcn_ballot_w64(pass)) {
                const float4 r2 = s_r[j][2];
                // the reference's overlap test, evaluated without short cuts
                const bool out_of = (qux < r1.z) | (qlx > r2.y) | (quy < r1.w) | (qly > r2.z) |
                                    (quz < r2.x) | (qlz > r2.w);
                const bool need = pass & act & !out_of;
                if (__builtin_amdgcn_ballot_w64(need)) {
                    const float4 r3 = s_r[j][3], r4 = s_r[j][4], r5 = s_r[j][5];
                    const float v = closest_sqr(r3.x, r3.y, r3.z, r3.w, r4.x, r4.y, r4.z, r4.w,
                                                r5.x, px, py, pz);
                    if (need && v <= best) {
                        best = v;
                        bestg = v * kGrow;
                    }
                }
            }
        }
    }
    if (act)
        out[__float_as_uint(p.w)] = best;
}

__global__ __launch_bounds__(256) void k_fill(float *__restrict__ out, uint64_t n, float v)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = v;
}

hipError_t launch_fill(float *out, uint64_t n, float v, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(n, 256)), dim3(256), 0, s, out, n, v);
    return hipGetLastError();
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

hipError_t launch_ionscatter(const float *xyz, uint64_t n, const uint32_t *key,
                             const uint32_t *rank, const uint32_t *cellstart, float4 *ions,
                             hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_ionscatter, dim3(blocks_for(n, 256)), dim3(256), 0, s, xyz, n, key, rank,
                       cellstart, ions);
    return hipGetLastError();
}

hipError_t launch_surftile(const float4 *ions, const SurfTile *tiles, uint32_t ntiles,
                           const uint32_t *tcellstart, const uint32_t *list, const float *tri,
                           const float *box, float R, float delta, float *out, hipStream_t s)
{
    if (ntiles == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_surftile, dim3(ntiles), dim3(256), 0, s, ions, tiles, tcellstart, list,
                       tri, box, R, delta, out);
    return hipGetLastError();
}

} // namespace pp
