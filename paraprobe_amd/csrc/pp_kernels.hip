// pp_kernels.hip -- gfx950 kernels of the RDF / kNN spatial-statistics path.
//
// Replaces the reference's KD-tree range-ball queries
// (src/PARAPROBE_KDTree.cpp:423-655) and the OpenMP center loop of
// horderdist::compute_generic_spatstat3 (src/PARAPROBE_SolverHdl.cpp:3770-4176)
// with a uniform cell list and one persistent query kernel:
//   * cell list: bounding box, x-fastest linear cell keys computed in FP64
//     (times 32 plus the ion's sub-cell -- its x eighth of the cell, then its
//     y and z halves -- where those offsets stay small), a one-digit counting
//     sort (count with atomics, exclusive scan, scatter) into cell-sorted
//     float4 (x, y, z, type) plus the sorted ion2surf array, and each cell's
//     x-eighth starts;
//   * per run of jobs (pp_prep.hip): considered-center flags, stream
//     compaction one (job, center type) after the other, tiles of up to T
//     centers of one type in one cell;
//   * query (k_query, three modes): each workgroup pulls tiles from eight
//     per-XCD tile queues; the tile's stencil rows are cut to the bounding
//     box of its centers, their end cells to the x eighths that box's balls
//     reach; the four wavefronts split the concatenation of
//     the rows into 64-candidate chunks (a chunk may span rows), each lane
//     holds one candidate and loops over the T centers held in registers,
//     evaluating the reference's FP32
//     squared distance ((dx*dx + dy*dy) + dz*dz, no contraction) and binning
//     in-ball pairs with integer LDS atomics: mode 0 into pair counts per
//     (center type, env type, category) (RDF / RIPK); mode 1 into full-range
//     per-center histograms (16-bit pairs when a bound allows; CNTNB, and the
//     rerun of mode 2); mode 2 into both, the per-center histograms over each
//     env type's kNN range only.  After a tile one wavefront per center scans
//     its histograms and writes the category of every requested k-th
//     neighbour to a per-launch buffer (plain stores), which k_kreduce counts
//     per (task, k, category).
#include "pp_launch.h"

#pragma clang fp contract(off)

namespace pp {

// ---------------------------------------------------------------------------
// small helpers

__host__ __device__ inline uint32_t float_to_ord(float f)
{
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

float ord_to_float(uint32_t u)
{
    uint32_t v = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    float f;
    __builtin_memcpy(&f, &v, 4);
    return f;
}

static inline unsigned grid_for(uint64_t n, unsigned block)
{
    uint64_t g = (n + block - 1) / block;
    if (g > 0x7FFFFFFFull)
        g = 0x7FFFFFFFull;
    return (unsigned)(g ? g : 1);
}

// Inclusive prefix sum over the 64 lanes with DPP moves (no LDS round
// trips): shifts by 1, 2, 4, 8 inside each row of 16 lanes (a lane without a
// source adds 0), then row_bcast:15 adds row 0's total to row 1 and row 2's
// to row 3, and row_bcast:31 adds the total of rows 0-1 to rows 2 and 3.
__device__ inline uint32_t wave_incl_scan(uint32_t v, int lane)
{
    (void)lane;
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return v;
}

// Inclusive running maximum over the 64 lanes, with the same DPP moves as
// wave_incl_scan (a lane without a source takes 0, so that each step is one
// DPP-fused maximum)
__device__ inline uint32_t wave_max_scan(uint32_t v)
{
    v = __builtin_elementwise_max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false));
    v = __builtin_elementwise_max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false));
    v = __builtin_elementwise_max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false));
    v = __builtin_elementwise_max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false));
    v = __builtin_elementwise_max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false));
    v = __builtin_elementwise_max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false));
    return v;
}

// ---------------------------------------------------------------------------
// exclusive scan (reduce-then-scan, 2048 items per workgroup)

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

// in and out may alias (the block totals are scanned in place): each thread
// reads its items before the barrier and writes them after it
__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(const uint32_t *in, uint32_t *out,
                                                           uint32_t *sums, uint64_t n)
{
    __shared__ uint32_t s_w[kScanBlock / kWave];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t tot = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        tot += v[k];
    }
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    uint32_t incl = wave_incl_scan(tot, lane);
    if (lane == kWave - 1)
        s_w[wid] = incl;
    __syncthreads();
    uint32_t woff = 0;
    for (int w = 0; w < wid; ++w)
        woff += s_w[w];
    uint32_t run = woff + incl - tot;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n)
            out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == kScanBlock - 1)
        sums[blockIdx.x] = run;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_add(uint32_t *__restrict__ out,
                                                         const uint32_t *__restrict__ offs,
                                                         uint64_t n)
{
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    const uint32_t o = offs[blockIdx.x];
    for (int k = threadIdx.x; k < kScanTile; k += kScanBlock)
        if (base + k < n)
            out[base + k] += o;
}

__global__ void k_store_total(uint32_t *out, uint64_t n, const uint32_t *sums, uint32_t nblk)
{
    // out[n] = exclusive(out)[n-1] + in[n-1] == offs-added sum of the last tile
    if (threadIdx.x == 0 && blockIdx.x == 0)
        out[n] = sums[nblk];
}

size_t scan_workspace_words(uint64_t n)
{
    size_t words = 0;
    uint64_t m = n;
    do {
        uint64_t nb = (m + kScanTile - 1) / kScanTile;
        if (nb == 0)
            nb = 1;
        words += nb + 1 + 4;
        m = nb;
    } while (m > 1);
    return words + 16;
}

// sums buffer of level L holds nblk + 1 entries; its exclusive scan (level
// L + 1) gives every tile's offset and, at index nblk, the grand total.
static hipError_t scan_level(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *ws,
                             hipStream_t s)
{
    const uint64_t nb64 = (n + kScanTile - 1) / kScanTile;
    const uint32_t nblk = (uint32_t)(nb64 ? nb64 : 1);
    uint32_t *sums = ws;           // nblk + 1
    uint32_t *rest = ws + nblk + 4;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nblk), dim3(kScanBlock), 0, s, in, out, sums, n);
    if (nblk == 1) {
        // single tile: the tile total is the grand total
        hipLaunchKernelGGL(k_store_total, dim3(1), dim3(1), 0, s, out, n, sums, 0u);
        return hipGetLastError();
    }
    // scan the tile totals in place (with a total at index nblk)
    hipError_t e = scan_level(sums, sums, nblk, rest, s);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_scan_add, dim3(nblk), dim3(kScanBlock), 0, s, out, sums, n);
    hipLaunchKernelGGL(k_store_total, dim3(1), dim3(1), 0, s, out, n, sums, nblk);
    return hipGetLastError();
}

hipError_t scan_exclusive(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *ws,
                          hipStream_t s)
{
    if (n == 0) {
        return hipMemsetAsync(out, 0, sizeof(uint32_t), s);
    }
    return scan_level(in, out, n, ws, s);
}

// ---------------------------------------------------------------------------
// bounding box

__global__ __launch_bounds__(256) void k_bbox(const float *__restrict__ xyz, uint64_t n,
                                              uint32_t *__restrict__ out6)
{
    uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            uint32_t o = float_to_ord(xyz[3 * i + a]);
            mn[a] = o < mn[a] ? o : mn[a];
            mx[a] = o > mx[a] ? o : mx[a];
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        for (int d = kWave / 2; d > 0; d >>= 1) {
            uint32_t t0 = __shfl_xor(mn[a], d, kWave), t1 = __shfl_xor(mx[a], d, kWave);
            mn[a] = t0 < mn[a] ? t0 : mn[a];
            mx[a] = t1 > mx[a] ? t1 : mx[a];
        }
    }
    // one atomic per value and workgroup (atomics on one address serialise)
    __shared__ uint32_t s_mm[kWaves][6];
    const int wid = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0)
        for (int a = 0; a < 3; ++a) {
            s_mm[wid][a] = mn[a];
            s_mm[wid][3 + a] = mx[a];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        uint32_t v = s_mm[0][a];
        for (int w = 1; w < kWaves; ++w)
            v = a < 3 ? (s_mm[w][a] < v ? s_mm[w][a] : v) : (s_mm[w][a] > v ? s_mm[w][a] : v);
        if (a < 3)
            atomicMin(&out6[a], v);
        else
            atomicMax(&out6[a], v);
    }
}

hipError_t launch_bbox(const float *xyz, uint64_t n, uint32_t *out6, hipStream_t s)
{
    uint32_t init[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    hipError_t e = hipMemcpyAsync(out6, init, sizeof(init), hipMemcpyHostToDevice, s);
    if (e != hipSuccess)
        return e;
    unsigned g = grid_for(n, 256);
    if (g > 1024)
        g = 1024;
    hipLaunchKernelGGL(k_bbox, dim3(g), dim3(256), 0, s, xyz, n, out6);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// cell list

__device__ inline int cell_coord(float v, double lo, double inv, int dim)
{
    // FP64: the computed cell differs from the exact floor((v - lo) / w) only
    // for points within ~1e-15 relative of a cell face, far inside the
    // stencil margin.  The result is monotone in v, clamped to the grid.
    double t = ((double)v - lo) * inv;
    int c = (int)floor(t);
    c = c < 0 ? 0 : c;
    return c >= dim ? dim - 1 : c;
}

// The half of its cell an ion lies in along one axis (0 lower, 1 upper),
// from the same FP64 coordinate as cell_coord.
__device__ inline uint32_t cell_half(float v, double lo, double inv, int c)
{
    const double t = ((double)v - lo) * inv - (double)c;
    return t >= 0.5 ? 1u : 0u;
}

// The part of its cell an ion lies in along x, of 2^bits equal parts, from the
// same FP64 coordinate as cell_coord (monotone in v; clamped).
__device__ inline uint32_t cell_part(float v, double lo, double inv, int c, int bits)
{
    const double t = (((double)v - lo) * inv - (double)c) * (double)(1 << bits);
    const double f = floor(t);
    const double top = (double)((1 << bits) - 1);
    return (uint32_t)(f < 0.0 ? 0.0 : (f > top ? top : f));
}

// key = cell (x fastest), or with sub = 3 or 5 the cell times 2^sub plus the
// sub-cell of the ion: its x part of the cell (sub - 2 bits: halves or
// eighths) major, then its y and z halves.  A sort by key orders each cell's
// ions by x part, so the stencil rows can leave out the parts of their end
// cells beyond the balls, and within a part by quadrant, so that a tile of
// consecutive centers spans a small box (the cell offsets are every
// 2^sub-th key offset).
__global__ __launch_bounds__(256) void k_cellkeys(const float *__restrict__ xyz, uint64_t n,
                                                  GridGeom g, uint32_t *__restrict__ count,
                                                  uint32_t *__restrict__ key,
                                                  uint32_t *__restrict__ rank, int sub)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    const int cx = cell_coord(x, g.lo[0], g.inv_w, g.nx);
    const int cy = cell_coord(y, g.lo[1], g.inv_wyz, g.ny);
    const int cz = cell_coord(z, g.lo[2], g.inv_wyz, g.nz);
    uint32_t k = ((uint32_t)cz * (uint32_t)g.ny + (uint32_t)cy) * (uint32_t)g.nx + (uint32_t)cx;
    if (sub)
        k = (k << sub) | (cell_part(x, g.lo[0], g.inv_w, cx, sub - 2) << 2) |
            (cell_half(y, g.lo[1], g.inv_wyz, cy) << 1) | cell_half(z, g.lo[2], g.inv_wyz, cz);
    key[i] = k;
    rank[i] = atomicAdd(&count[k], 1u);
}

__global__ __launch_bounds__(256) void k_stride(const uint32_t *__restrict__ in,
                                                uint32_t *__restrict__ out, uint64_t nout,
                                                int shift)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nout)
        out[i] = in[i << shift];
}

hipError_t launch_stride(const uint32_t *in, uint32_t *out, uint64_t nout, int shift,
                         hipStream_t s)
{
    if (nout == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_stride, dim3(grid_for(nout, 256)), dim3(256), 0, s, in, out, nout, shift);
    return hipGetLastError();
}

// Each cell's x-part starts (XSplit records) from the sub-cell offsets.
__global__ __launch_bounds__(256) void k_xsplit(const uint32_t *__restrict__ start,
                                                uint64_t nc, int sub, uint2 *__restrict__ out)
{
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc)
        return;
    const uint64_t k = c << sub;
    const uint32_t base = start[k], cnt = start[k + (1u << sub)] - base;
    uint32_t lo = 0, hi = 0;
    if (cnt > 255u) {
        hi = 1u << 24;
    } else {
        const int parts = 1 << (sub - 2);
        for (int p = 1; p < parts; ++p) {
            const uint32_t off = start[k + ((uint64_t)p << 2)] - base;
            const int b = p - 1;
            if (b < 4)
                lo |= off << (8 * b);
            else
                hi |= off << (8 * (b - 4));
        }
    }
    out[c] = make_uint2(lo, hi);
}

hipError_t launch_xsplit(const uint32_t *start, uint64_t ncells, int sub, uint2 *out,
                         hipStream_t s)
{
    if (ncells == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_xsplit, dim3(grid_for(ncells, 256)), dim3(256), 0, s, start, ncells, sub,
                       out);
    return hipGetLastError();
}

hipError_t launch_cellkeys(const float *xyz, uint64_t n, GridGeom g, uint32_t *count,
                           uint32_t *key, uint32_t *rank, hipStream_t s, int sub)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_cellkeys, dim3(grid_for(n, 256)), dim3(256), 0, s, xyz, n, g, count,
                       key, rank, sub);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_scatter(const float *__restrict__ xyz,
                                                 const uint8_t *__restrict__ type,
                                                 const float *__restrict__ s2, uint64_t n,
                                                 const uint32_t *__restrict__ key,
                                                 const uint32_t *__restrict__ rank,
                                                 const uint32_t *__restrict__ cellstart,
                                                 float4 *__restrict__ pos,
                                                 float *__restrict__ s2sorted, int w_s2)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t d = cellstart[key[i]] + rank[i];
    pos[d] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2],
                         w_s2 ? s2[i] : __uint_as_float((uint32_t)type[i]));
    s2sorted[d] = s2[i];
}

hipError_t launch_scatter(const float *xyz, const uint8_t *type, const float *s2, uint64_t n,
                          const uint32_t *key, const uint32_t *rank, const uint32_t *cellstart,
                          float4 *pos, float *s2sorted, hipStream_t s, int w_s2)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_scatter, dim3(grid_for(n, 256)), dim3(256), 0, s, xyz, type, s2, n, key,
                       rank, cellstart, pos, s2sorted, w_s2);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_cellmin(const uint32_t *__restrict__ cellstart,
                                                 const float *__restrict__ s2, uint64_t nc,
                                                 float *__restrict__ out)
{
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc)
        return;
    float m = __builtin_inff();
    for (uint32_t i = cellstart[c]; i < cellstart[c + 1]; ++i) {
        const float v = s2[i];
        m = v >= m ? m : (v == v ? v : -__builtin_inff());
    }
    out[c] = m;
}

hipError_t launch_cellmin(const uint32_t *cellstart, const float *s2sorted, uint64_t ncells,
                          float *out, hipStream_t s)
{
    if (ncells == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_cellmin, dim3(grid_for(ncells, 256)), dim3(256), 0, s, cellstart,
                       s2sorted, ncells, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// typed candidate lists

__global__ __launch_bounds__(256) void k_typeflags(const uint8_t *__restrict__ type, uint64_t n,
                                                   const uint32_t *__restrict__ mask8,
                                                   uint32_t *__restrict__ flags)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t t = type[i];
    flags[i] = (mask8[t >> 5] >> (t & 31)) & 1u;
}

hipError_t launch_typeflags(const uint8_t *type, uint64_t n, const uint32_t *mask8, uint32_t *flags,
                            hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_typeflags, dim3(grid_for(n, 256)), dim3(256), 0, s, type, n, mask8, flags);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_gather(const float *__restrict__ xyz,
                                                const uint8_t *__restrict__ type,
                                                const float *__restrict__ s2,
                                                const uint32_t *__restrict__ flags,
                                                const uint32_t *__restrict__ pos, uint64_t n,
                                                float *__restrict__ oxyz,
                                                uint8_t *__restrict__ otype,
                                                float *__restrict__ os2)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flags[i])
        return;
    const uint32_t d = pos[i];
    oxyz[3 * (uint64_t)d] = xyz[3 * i];
    oxyz[3 * (uint64_t)d + 1] = xyz[3 * i + 1];
    oxyz[3 * (uint64_t)d + 2] = xyz[3 * i + 2];
    otype[d] = type[i];
    os2[d] = s2[i];
}

hipError_t launch_gather(const float *xyz, const uint8_t *type, const float *s2,
                         const uint32_t *flags, const uint32_t *pos, uint64_t n, float *oxyz,
                         uint8_t *otype, float *os2, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_gather, dim3(grid_for(n, 256)), dim3(256), 0, s, xyz, type, s2, flags, pos,
                       n, oxyz, otype, os2);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_typecount(const uint8_t *__restrict__ type, uint64_t n,
                                                   uint32_t *__restrict__ counts)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(&h[type[i]], 1u);
    __syncthreads();
    if (h[threadIdx.x])
        atomicAdd(&counts[threadIdx.x], h[threadIdx.x]);
}

// Occupancy of a coarse grid that tiles the cloud's box (cells of
// 1 / inv[a] along axis a): cell c is marked when an ion lies in it (several
// ions store the same 1, no atomics needed).
__global__ __launch_bounds__(256) void k_occupancy(const float *__restrict__ xyz, uint64_t n,
                                                   OccGrid g, uint8_t *__restrict__ occ)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        int c[3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
            c[a] = cell_coord(xyz[3 * i + a], g.lo[a], g.inv[a], g.d[a]);
        occ[((uint64_t)c[2] * g.d[1] + c[1]) * g.d[0] + c[0]] = 1u;
    }
}

hipError_t launch_occupancy(const float *xyz, uint64_t n, OccGrid g, uint8_t *occ, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    unsigned grid = grid_for(n, 256);
    if (grid > 8192)
        grid = 8192;
    hipLaunchKernelGGL(k_occupancy, dim3(grid), dim3(256), 0, s, xyz, n, g, occ);
    return hipGetLastError();
}

hipError_t launch_typecount(const uint8_t *type, uint64_t n, uint32_t *counts, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    unsigned g = grid_for(n, 256);
    if (g > 1024)
        g = 1024;
    hipLaunchKernelGGL(k_typecount, dim3(g), dim3(256), 0, s, type, n, counts);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// query kernel

struct Lay {
    int edges, envoff, trgslot, kinfo, ckoff, kidx, win, rows, rowpre, cen, meta, P, H, total;
};

__host__ __device__ inline int al4(int x) { return (x + 3) & ~3; }

// LDS layout (uint32 words) of the query kernel.  mode: 0 pair counts per
// center type only, 1 full-range per-center histograms, 2 kNN-range
// per-center histograms of maxHK words.
__host__ __device__ inline Lay make_layout(int T, int c16, int mode, int maxC, int maxE,
                                           int maxTrg, int maxKmax, int maxRows, int maxHK,
                                           int ntypes)
{
    Lay L;
    int o = 0;
    L.edges = o;  // the hybrid mode reads the (rarely needed) edges from global memory
    o += mode == 2 ? 0 : al4(maxC);
    L.envoff = o;  // per type id below ntypes
    o += al4(ntypes);
    L.trgslot = o;  // unused: tiles carry their center type
    L.kinfo = o;
    o += mode == 2 ? 2 * al4(ntypes) : 0;
    L.ckoff = o;
    o += mode == 2 ? kMaxEnv : 0;
    L.kidx = o;  // maxKmax: the largest kmax of the jobs whose k are not 1..kmax
    o += mode && maxKmax ? al4(maxKmax + 2) : 0;
    L.win = o;  // per wavefront: the k window of the k-th category extraction (uint16)
    o += mode ? kWaves * kWave / 2 : 0;
    L.rows = o;  // per row its first and end candidate
    o += al4(2 * maxRows);
    L.rowpre = o;  // per row its first index in the rows' concatenation
    o += al4(maxRows + 1);
    L.cen = o;
    o += 4 * T;
    L.meta = o;
    o += al4(2 * T);
    L.P = o;
    o += al4(maxE * maxC);  // pair counts of one center type (tiles hold one type)
    L.H = o;
    const int hc = c16 ? (T + 1) / 2 : T;
    o += mode == 1 ? al4(hc * maxE * maxC) : mode == 2 ? al4(hc * maxHK) : 0;
    L.total = o;
    return L;
}

size_t query_lds_bytes(int T, int c16, int mode, int maxC, int maxE, int maxTrg, int maxKmax,
                       int maxRows, int maxHK, int ntypes)
{
    return (size_t)make_layout(T, c16, mode, maxC, maxE, maxTrg, maxKmax, maxRows, maxHK, ntypes)
               .total *
           4u;
}

struct QArgs {
    const JobDev *jobs;
    const uint32_t *tile_off;
    const Tile *const *tiles;
    uint32_t *queue;
    uint32_t first_tile, end_tile;
    uint32_t grab;
    int njobs;
    int maxC, maxE, maxTrg, maxKmax, maxRows, maxHK;
    int ntypes;  // type ids of the launch's candidates are below this
};

__device__ inline int find_job(const uint32_t *__restrict__ off, int njobs, uint32_t g)
{
    int lo = 0, hi = njobs;  // off[lo] <= g < off[hi]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (off[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__device__ inline uint32_t rfl(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline float rflf(float v) { return __uint_as_float(rfl(__float_as_uint(v))); }
// a wave-uniform pointer into scalar registers (fields of a job record are
// read by vector loads, which leave even uniform values in vector registers)
template <typename P>
__device__ inline P *rflp(P *p)
{
    const uint64_t u = (uint64_t)p;
    return (P *)(((uint64_t)rfl((uint32_t)(u >> 32)) << 32) | (uint64_t)rfl((uint32_t)u));
}

// XCC_ID hardware register (accelerator complex die of this wave)
__device__ inline uint32_t xcc_id()
{
    return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 0xFu;
}

typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const fvec4 gfvec4;

// a compile-time switch passed to a generic lambda
template <bool B>
struct CFlag {
    static constexpr bool value = B;
};

typedef __attribute__((address_space(1))) uint16_t gu16;
typedef __attribute__((address_space(1))) const uint32_t gcu32;

// device-scope atomic add through the global address space: a flat atomic
// also counts against the LDS wait counter, so the LDS reads that follow it
// would wait for its completion
template <typename U>
__device__ inline void global_add(U *p, U v)
{
    typedef __attribute__((address_space(1))) U gU;
    __hip_atomic_fetch_add((gU *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// 16-byte load through the global address space (global_load_dwordx4
// rather than a flat load)
__device__ inline float4 load_g4(const gfvec4 *p, uint32_t i)
{
    const fvec4 v = p[i];
    return make_float4(v.x, v.y, v.z, v.w);
}

// The next tile of one workgroup (thread 0).  The launch's tiles form blocks
// of a.grab consecutive tiles; block b belongs to XCD b % 8, whose counter
// hands its blocks out in order, so all XCDs sweep the same spatial front (the
// ions of neighbouring tiles stay shared in the last-level cache) while each
// counter sits on its own line.  A workgroup whose XCD has no blocks left
// takes blocks of the next XCD.  Returns end_tile when every block is gone.
__device__ inline uint32_t next_tile(const QArgs &a, uint32_t myx, int &qx, uint32_t &gnext,
                                     uint32_t &left)
{
    if (left > 0) {
        --left;
        return gnext++;
    }
    const uint32_t span = a.end_tile - a.first_tile;
    const uint32_t nblk = (span + a.grab - 1) / a.grab;
    while (qx < kXcds) {
        const uint32_t x = (myx + (uint32_t)qx) & (uint32_t)(kXcds - 1);
        if (x < nblk) {
            const uint32_t k = atomicAdd(&a.queue[x * kQueueStride], 1u);
            const uint64_t b = (uint64_t)k * kXcds + x;
            if (b < nblk) {
                const uint32_t t0 = (uint32_t)b * a.grab;
                const uint32_t n = span - t0 < a.grab ? span - t0 : a.grab;
                left = n - 1;
                gnext = a.first_tile + t0 + 1;
                return a.first_tile + t0;
            }
        }
        ++qx;
    }
    return a.end_tile;
}

// Category of an in-range squared distance (d2 <= fl(R * R)).  The index
// coordinate y = (d - start) tmp + 1 has category floor(y) wherever y lies at
// least max_offset from every integer, and fmaf(sqrt(d2), isc, ioff1) is
// within y_err of it (amb = max_offset + y_err).  The shifted coordinate
// ya = fmaf(sqrt(d2), isc, ioffa), ioffa = fl(1 - start tmp + amb), is within
// y_err of y + amb; where its fraction is at least 2 amb, i.e. ya lies in
// [m + 2 amb, m + 1), y lies in [m + max_offset, m + 1 - max_offset) and the
// category is m = floor(ya).  When any active lane is nearer an edge (rare),
// the wave settles every lane with the exact edge table from the nearest
// ideal edge of the unshifted coordinate.  No LDS read on the common path.
template <typename EdgePtr>
__device__ inline int category_of(float d2, float isc, float ioff1, float ioffa, float amb2,
                                  float Cm1f, EdgePtr s_edges)
{
    const float s = __builtin_amdgcn_sqrtf(d2);
    const float ya = __builtin_fmaf(s, isc, ioffa);
    const bool near = __builtin_amdgcn_fractf(ya) < amb2;
    // floor for ya >= 0, 0 for ya < 0 (v_cvt_u32_f32 saturates: y < 1 is
    // the lowest category); ya < y + 0.5 + y_err < ncat for every in-ball
    // distance (d <= R = end puts y at most at ncat - 1).  Taken before the
    // test, so the common path runs straight through.
    int cat = (int)__float2uint_rz(ya);
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(near) != 0, 0)) {
        const float y = __builtin_fmaf(s, isc, ioff1);
        const int q = (int)__builtin_amdgcn_fmed3f(y + 0.5f, 1.0f, Cm1f);
        cat = q - (__float_as_uint(d2) < s_edges[q] ? 1 : 0);
    }
    return cat;
}

// The persistent query kernel.  MODE 0: pair counts per (center type, env
// type, category) only (RDF / RIPK).  MODE 1: full-range per-center
// histograms, from which RDF, 1NN, MKNN and CNTNB are extracted per tile.
// MODE 2: pair counts as in MODE 0 plus per-center histograms over the kNN
// range only (see JobDev::HK); centers whose kNN range holds too few
// neighbours are flagged for a MODE 1 rerun.
// Resident waves per SIMD asked of the compiler for 8-center tiles: 7 in the
// pair-count mode (72 VGPRs, no spills), 6 in the hybrid mode (7 fits in 72
// VGPRs with spills in the per-tile code: 0.3 % faster on the 20M-ion job,
// 13 % slower on the 100M-ion one, whose tiles are cheaper).
template <int T, bool C16, int MODE>
__global__ __launch_bounds__(kBlock, (T == 8 && MODE == 0) ? 7 : (T == 8 && MODE == 2) ? 6 : 1) void k_query(QArgs a)
{
    constexpr bool PC = MODE == 1;
    constexpr bool HY = MODE == 2;
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_g, s_ctot;
    const Lay L = make_layout(T, C16, MODE, a.maxC, a.maxE, a.maxTrg, a.maxKmax, a.maxRows,
                              a.maxHK, a.ntypes);
    uint32_t *s_edges = lds + L.edges;
    int32_t *s_envoff = (int32_t *)(lds + L.envoff);
    int2 *s_kinfo = (int2 *)(lds + L.kinfo);
    int32_t *s_ckoff = (int32_t *)(lds + L.ckoff);
    int32_t *s_kidx = (int32_t *)(lds + L.kidx);
    uint16_t *s_win = (uint16_t *)(lds + L.win);
    uint32_t *s_rows = lds + L.rows;
    uint32_t *s_rowpre = lds + L.rowpre;
    float4 *s_cen = (float4 *)(lds + L.cen);
    uint32_t *s_meta = lds + L.meta;  // [2c] tc | es << 8, [2c+1] unused
    uint32_t *s_P = lds + L.P;
    uint32_t *s_H = lds + L.H;

    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    int cur = -1;
    // tile queues (one counter per XCD); thread 0 keeps the grab state
    const uint32_t myx = xcc_id() & (uint32_t)(kXcds - 1);
    int qx = 0;
    uint32_t gnext = 0, left = 0;
    unsigned long long inball = 0;   // wave-uniform
    unsigned long long cand = 0;     // block-uniform
    unsigned long long pending = 0;  // pairs accumulated in s_P since its last flush

    for (;;) {
        __syncthreads();
        if (tid == 0)
            s_g = next_tile(a, myx, qx, gnext, left);
        __syncthreads();
        const uint32_t g = rfl(s_g);
        const bool done = g >= a.end_tile;
        const int job = done ? -1 : (int)rfl((uint32_t)find_job(a.tile_off, a.njobs, g));
        // the tile: cell key, first center, center count, target-type slot
        uint32_t t_key = 0, t_first = 0, t_count = 0;
        int ts = 0;
        if (!done) {
            const Tile *tp = a.tiles[job] + (g - a.tile_off[job]);
            t_key = rfl(tp->key);
            t_first = rfl(tp->first);
            t_count = rfl((uint32_t)tp->count);
            ts = (int)rfl(tp->pad);
        }
        // s_P holds the pair counts of one (job, center type)
        const int pkey = done ? -1 : job * kMaxTrg + ts;

        if (pkey != cur || (!done && pending > 0x7FFFFFFFull)) {
            if (cur >= 0) {
                // flush the accumulated pair counts and statistics
                const JobDev &P0 = a.jobs[cur / kMaxTrg];
                const int n = P0.E * P0.C;
                if (P0.families & (FAM_RDF | FAM_RIPK)) {
                    unsigned long long *dst = P0.P + (size_t)(cur % kMaxTrg) * n;
                    for (int i = tid; i < n; i += kBlock) {
                        const uint32_t v = s_P[i];
                        if (v)
                            global_add(&dst[i], (unsigned long long)v);
                    }
                }
                if (lane == 0 && inball)
                    atomicAdd(&P0.stats[0], inball);
                if (tid == 0 && cand)
                    atomicAdd(&P0.stats[1], cand);
                inball = 0;
                cand = 0;
                pending = 0;
                __syncthreads();
            }
            if (done)
                break;
            const JobDev &J = a.jobs[job];
            if (cur < 0 || job != cur / kMaxTrg) {
                if (!HY)
                    for (int i = tid; i < J.C; i += kBlock)
                        s_edges[i] = J.edges[i];
                for (int i = tid; i < a.ntypes; i += kBlock)
                    s_envoff[i] = J.envoff[i];
                if (HY) {
                    for (int i = tid; i < 2 * a.ntypes; i += kBlock)
                        ((int32_t *)s_kinfo)[i] = J.kinfo[i];
                    for (int i = tid; i < J.E; i += kBlock)
                        s_ckoff[i] = J.ckoff[i];
                }
                if ((PC || HY) && !J.k_dense)
                    for (int i = tid; i <= J.kmax; i += kBlock)
                        s_kidx[i] = J.kidx_of_k ? J.kidx_of_k[i] : -1;
            }
            const int n = J.E * J.C;
            for (int i = tid; i < n; i += kBlock)
                s_P[i] = 0u;
            cur = pkey;
            __syncthreads();
        }
        if (done)
            break;

        const JobDev &J = a.jobs[job];
        const int cnt = (int)t_count;
        const uint32_t key = t_key;
        const int C = J.C, E = J.E, EC = E * C;
        const int HK = HY ? J.HK : 0;
        const int nrows = J.nrows;
        // symmetric pair counting for this tile's center type: its own-type
        // candidates along srows[0, s_ns) of list S, the others along
        // srows[s_ns, nsrows) of list O
        const bool sym = !PC && ts == J.sym_slot;
        const int nrt = sym ? J.nsrows : nrows;
        const Row *rtab = sym ? J.srows : J.rows;
        const int cx = (int)(key % (uint32_t)J.nx);
        const int cy = (int)((key / (uint32_t)J.nx) % (uint32_t)J.ny);
        const int cz = (int)(key / ((uint32_t)J.nx * (uint32_t)J.ny));

        if (tid < T) {
            // absent centers of a partial tile: NaN coordinates compare false
            float px = __builtin_nanf(""), py = 0.f, pz = 0.f;
            uint32_t es = 0xFFu;
            if (tid < cnt) {
                const float4 *cp = J.cpos + J.centers[t_first + tid];
                px = cp->x;
                py = cp->y;
                pz = cp->z;
                const int eo = s_envoff[__float_as_uint(cp->w) & 0xFFu];
                // env slot = eo / C exactly (eo is a multiple of C below 2^24)
                if (eo >= 0)
                    es = (uint32_t)__builtin_rintf((float)eo * __builtin_amdgcn_rcpf((float)C));
            }
            s_meta[2 * tid] = (uint32_t)ts | (es << 8);
            s_cen[tid] = make_float4(px, py, pz, 0.f);
        }
        __syncthreads();
        for (int r = tid; r < nrt; r += kBlock) {
            const Row rw = rtab[r];
            // the row's candidate list
            const bool inS = sym && r < J.s_ns;
            const uint32_t *cst = inS ? J.cellstart_s : (sym ? J.cellstart_o : J.cellstart);
            const uint2 *xsp = inS ? J.xsplit_s : (sym ? J.xsplit_o : J.xsplit);
            const int xb = inS ? J.xbits_s : (sym ? J.xbits_o : J.xbits);
            uint32_t rs = 0, re = 0;
            const int y = cy + rw.dy, z = cz + rw.dz;
            if (y >= 0 && y < J.ny && z >= 0 && z < J.nz) {
                int x0 = cx - rw.dxm < 0 ? 0 : cx - rw.dxm;
                int x1 = cx + rw.dxm >= J.nx ? J.nx - 1 : cx + rw.dxm;
                // the parts of the own row before and after the own cell
                if (rw.pad & kRowBefore)
                    x1 = x1 < cx - 1 ? x1 : cx - 1;
                if (rw.pad & kRowAfter)
                    x0 = x0 > cx + 1 ? x0 : cx + 1;
                // the stencil row cut to the tile's centers: their bounding
                // box's distance in y and z from the row's cells bounds the x
                // extent of their balls (exact cell faces lo + c * w in FP64,
                // reach = R with a margin far above the FP32 rounding of the
                // in-ball test)
                float b0x = s_cen[0].x, b1x = b0x, b0y = s_cen[0].y, b1y = b0y;
                float b0z = s_cen[0].z, b1z = b0z;
                for (int c = 1; c < cnt; ++c) {
                    const float4 q = s_cen[c];
                    b0x = fminf(b0x, q.x);
                    b1x = fmaxf(b1x, q.x);
                    b0y = fminf(b0y, q.y);
                    b1y = fmaxf(b1y, q.y);
                    b0z = fminf(b0z, q.z);
                    b1z = fmaxf(b1z, q.z);
                }
                const double ylo = J.glo[1] + y * J.gwyz, zlo = J.glo[2] + z * J.gwyz;
                const double hy = fmax(0.0, fmax(ylo - (double)b1y, (double)b0y - (ylo + J.gwyz)));
                const double hz = fmax(0.0, fmax(zlo - (double)b1z, (double)b0z - (zlo + J.gwyz)));
                const double h2 = hy * hy + hz * hz;
                // the balls' x extent in cell units, [tlo, thi], for radius
                // sqrt(lim); x0 .. x1 cut to it
                double tlo = 0.0, thi = 0.0;
                auto cut = [&](double lim) {
                    if (h2 > lim) {
                        x1 = x0 - 1;
                        return;
                    }
                    const double ex = sqrt(lim - h2);
                    tlo = ((double)b0x - ex - J.glo[0]) * J.ginv;
                    thi = ((double)b1x + ex - J.glo[0]) * J.ginv;
                    const double a = floor(tlo), b = floor(thi);
                    if (a > (double)x0)
                        x0 = (int)a;
                    if (b < (double)x1)
                        x1 = (int)b;
                };
                cut(J.reach * J.reach);
                const uint32_t kr = ((uint32_t)z * (uint32_t)J.ny + (uint32_t)y) * (uint32_t)J.nx;
                if (inS && r >= 1 && r < J.s_nb1 && x0 <= x1) {
                    // a row before the own cell: the pairs with centers there
                    // are counted from those centers' tiles, so a cell whose
                    // own-type ions are all centers is needed only within the
                    // kNN reach.  The row keeps the cells holding a non-center
                    // (fa .. fb) and the kNN reach: their hull, trimmed at its
                    // ends as the full-reach row (as the kNN reach when no
                    // cell holds a non-center)
                    int fa = x1 + 1, fb = x0 - 1;
                    for (int x = x0; x <= x1; ++x)
                        if (!(J.cellmin_s[kr + (uint32_t)x] >= J.R2)) {
                            fa = fa > x1 ? x : fa;
                            fb = x;
                        }
                    if (fa > x0 || fb < x1) {
                        const double rtlo = tlo, rthi = thi;
                        if (J.kreach > 0.0)
                            cut(J.kreach * J.kreach);
                        else
                            x1 = x0 - 1;
                        if (fa <= fb) {
                            const bool kempty = x0 > x1;
                            x0 = kempty ? fa : (x0 < fa ? x0 : fa);
                            x1 = kempty ? fb : (x1 > fb ? x1 : fb);
                            tlo = rtlo;
                            thi = rthi;
                        }
                    }
                }
                if (x0 <= x1) {
                    const uint32_t k0 = kr + (uint32_t)x0;
                    const uint32_t k1 = k0 + (uint32_t)(x1 - x0);
                    rs = cst[k0];
                    re = cst[k1 + 1u];
                    if (xb) {
                        // the end cells' x parts beyond [tlo, thi]: an ion of
                        // cell x0 with x >= the balls' lower bound has
                        // t >= tlo, so its part is at least qa (the same
                        // monotone FP64 expression as the build); likewise at
                        // most qb in cell x1
                        const int P = 1 << xb;
                        if (floor(tlo) == (double)x0) {
                            const double f = floor((tlo - (double)x0) * (double)P);
                            const int qa = f < 0.0 ? 0 : (f > (double)(P - 1) ? P - 1 : (int)f);
                            const uint2 rec = xsp[k0];
                            if (qa > 0 && !xsplit_whole(rec))
                                rs += xsplit_at(rec, qa);
                        }
                        if (floor(thi) == (double)x1) {
                            const double f = floor((thi - (double)x1) * (double)P);
                            const int qb = f < 0.0 ? 0 : (f > (double)(P - 1) ? P - 1 : (int)f);
                            const uint2 rec = xsp[k1];
                            if (qb < P - 1 && !xsplit_whole(rec))
                                re = cst[k1] + xsplit_at(rec, qb + 1);
                        }
                        if (re < rs)
                            re = rs;
                    }
                }
            }
            s_rows[2 * r] = rs;
            s_rows[2 * r + 1] = re;
        }
        if (PC || HY) {
            const int nh = (C16 ? (cnt + 1) / 2 : cnt) * (PC ? EC : HK);
            for (int i = tid; i < nh; i += kBlock)
                s_H[i] = 0u;
        }
        __syncthreads();
        if (wid == 0) {
            // candidate prefix over the rows: the tile's candidates are the
            // concatenation of its rows, cut into 64-candidate chunks
            uint32_t carry = 0;
            for (int r0 = 0; r0 < nrt; r0 += kWave) {
                const int r = r0 + lane;
                uint32_t len = 0;
                if (r < nrt)
                    len = s_rows[2 * r + 1] - s_rows[2 * r];
                const uint32_t inc = wave_incl_scan(len, lane);
                if (r < nrt)
                    s_rowpre[r] = carry + inc - len;
                carry += __shfl(inc, kWave - 1, kWave);
            }
            if (lane == 0) {
                s_rowpre[nrt] = carry;
                s_ctot = carry;
            }
        }
        __syncthreads();
        {
            const unsigned long long pairs =
                (unsigned long long)rfl(s_ctot) * (unsigned long long)cnt;
            pending += pairs;
            if (tid == 0)
                cand += pairs;
        }

        // ---------------- main candidate loop ----------------
        {
            // the T centers in registers (no LDS read, and so no wait behind
            // the histogram atomics, in the common path); absent centers of a
            // partial tile have NaN coordinates, so their squared distances
            // compare false
            float cxr[T], cyr[T], czr[T];
#pragma unroll
            for (int c = 0; c < T; ++c) {
                const float4 p = s_cen[c];
                cxr[c] = p.x;
                cyr[c] = p.y;
                czr[c] = p.z;
            }
            const float R2 = J.R2;
            const float isc = J.idx_scale, ioff = J.idx_off;
            const float ioff1 = J.idx_off1, ioffa = J.idx_off1a, amb2 = J.amb2;
            const float Cm1f = (float)(C - 1);
            const bool rdf = (J.families & (FAM_RDF | FAM_RIPK)) != 0;
            const bool count_in = J.count_in != 0;
            // the hybrid mode's exact edge table, read through the global
            // address space (a flat load's wait would also wait for the
            // LDS atomics in flight)
            const gcu32 *gedges = rflp((const gcu32 *)J.edges);
            uint32_t *s_acc = PC ? s_H : s_P;
            const uint32_t ctot = rfl(s_ctot);
            unsigned long long my_in = 0;
            // symmetric tiles: the own type's pair-count row and kNN-range
            // entries (uniform), and the candidate indices where the cells
            // before the own cell (sb1) and after it (sb2) start in list S
            int woff_t = 0, kof_t = 0;
            float kthr_t = -1.0f;
            uint32_t sb1 = 0, sb2 = 0;
            if (sym) {
                woff_t = (int)rfl((uint32_t)s_envoff[J.sym_type]);
                if (HY) {
                    const int2 ki = s_kinfo[J.sym_type];
                    kof_t = (int)rfl((uint32_t)ki.x);
                    kthr_t = __int_as_float((int)rfl((uint32_t)ki.y));
                }
                sb1 = rfl(s_rowpre[1]);
                sb2 = rfl(s_rowpre[J.s_nb1]);
            }
            // One pass over rows [rlo, rhi) of list pos, whose candidates
            // are clo .. chi - 1 of the tile's concatenation.  Each
            // wavefront takes a contiguous block of the pass's chunks; chunk
            // q holds candidates clo + q * 64 .. clo + q * 64 + 63, so a chunk
            // may span rows and only the pass's last chunk is partly empty.
            // Every lane keeps its own cursor: row r, index j, row end re
            // (r == rhi: past the last row).  Two chunks alternate, so that
            // the load of the next is in flight while one is evaluated
            // against the T centers.  SYMF: the own-type pass of a symmetric
            // tile.
            auto sweep = [&](auto SYMF, const gfvec4 *__restrict__ pos, int rlo, int rhi,
                             uint32_t clo, uint32_t chi) {
            constexpr bool SY = !PC && decltype(SYMF)::value;
            const int nch = (int)((chi - clo + kWave - 1) / kWave);
            const int qa = (int)rfl((uint32_t)(wid * nch / kWaves));
            const int qb = (int)rfl((uint32_t)((wid + 1) * nch / kWaves));
            int r = rhi;
            uint32_t j = 0, re = 0xFFFFFFFFu;
            if (qa < qb) {
                // the row of candidate s0: the first h with rowpre[h] > s0,
                // less one (that row is non-empty)
                const uint32_t s0 = clo + (uint32_t)qa * kWave + (uint32_t)lane;
                if (s0 < chi) {
                    int lo = rlo, hi = rhi;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (s_rowpre[mid] <= s0)
                            lo = mid + 1;
                        else
                            hi = mid;
                    }
                    r = lo - 1;
                    j = s_rows[2 * r] + (s0 - s_rowpre[r]);
                    re = s_rows[2 * r + 1];
                }
            }
            // the chunk at the cursor (lanes past the last row load ion 0)
            auto fetch = [&](bool &ok) {
                ok = r < rhi;
                return load_g4(pos, j);
            };
            // the next chunk: 64 candidates on, across row ends (empty rows
            // are passed over)
            auto advance = [&]() {
                j += kWave;
                while (j >= re) {
                    if (r + 1 >= rhi) {
                        r = rhi;
                        j = 0;
                        re = 0xFFFFFFFFu;
                        break;
                    }
                    const uint32_t ov = j - re;
                    ++r;
                    j = s_rows[2 * r] + ov;
                    re = s_rows[2 * r + 1];
                }
            };
            auto body = [&](float4 p, const bool ok, int q) {
                // keep every use of this chunk's data below this point, so the
                // wait for its load is not hoisted into the prefetch of the next
                asm volatile("" : "+v"(p.x), "+v"(p.y), "+v"(p.z), "+v"(p.w));
                const float px = ok ? p.x : __builtin_nanf("");
                // (the own-type pass of a symmetric tile knows the type, and
                // its w holds ion2surf)
                const uint32_t ptype = SY ? 0u : __float_as_uint(p.w) & 0xFFu;
                // (a lane past the row end has a NaN x, so every distance
                // test fails and its type's table entries go unused)
                const int woff = SY ? 0 : s_envoff[ptype];
                const bool envok = woff >= 0;
                // the tile's center count as an opaque scalar, so that each
                // "c < ncen" below is one scalar compare and branch here (the
                // compiler keeps hoisted compares as lane masks, which cost
                // two vector instructions per center and chunk to branch on)
                int ncen = cnt;
                asm volatile("" : "+s"(ncen));
                // likewise the in-ball counting switch (kept as a lane mask,
                // it costs two vector instructions per chunk)
                int cin = (int)rfl(count_in ? 1u : 0u);
                asm volatile("" : "+s"(cin));
                // this candidate's pair-count row (only dereferenced when envok)
                uint32_t *acc_row = (PC ? s_H : s_P) + woff;
                if (PC) {
                    // 1) all T squared distances (the reference's FP32 expression)
                    float d2[T];
    #pragma unroll
                    for (int c = 0; c < T; ++c) {
                        const float dx = cxr[c] - px;
                        const float dy = cyr[c] - p.y;
                        const float dz = czr[c] - p.z;
                        d2[c] = dx * dx + dy * dy + dz * dz;
                    }
                    // 2) nearest ideal edge per pair and its exact d2 threshold; the
                    //    index is clamped to [1, C-1] (NaN -> 0), so every lane
                    //    reads inside the table and the LDS reads overlap
                    int qi[T];
                    uint32_t ed[T];
    #pragma unroll
                    for (int c = 0; c < T; ++c) {
                        const float sq = __builtin_amdgcn_sqrtf(d2[c]);
                        qi[c] = (int)__builtin_amdgcn_fmed3f(__builtin_fmaf(sq, isc, ioff), 1.0f, Cm1f);
                    }
    #pragma unroll
                    for (int c = 0; c < T; ++c)
                        ed[c] = s_edges[qi[c]];
                    // 3) in-ball pairs: count all types, bin env types
    #pragma unroll
                    for (int c = 0; c < T; ++c) {
                        const bool in = d2[c] <= R2;
                        if (__builtin_expect(cin != 0, 0))
                            my_in += (unsigned long long)__builtin_popcountll(
                                __builtin_amdgcn_ballot_w64(in));
                        if (in && envok) {
                            const int cat = qi[c] - (__float_as_uint(d2[c]) < ed[c] ? 1 : 0);
                            const uint32_t inc = (C16 && (c & 1)) ? 65536u : 1u;
                            __hip_atomic_fetch_add(&s_acc[(C16 ? (c >> 1) : c) * EC + woff + cat],
                                                   inc, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                    }
                } else if (!HY && SY) {
                    // direct mode, own-type pass of a symmetric tile (list S:
                    // the tile's type, ion2surf in w): weight 2 for a center
                    // in a cell after the own cell, none for a center in a
                    // cell before it (that lane is left out), else 1
                    (void)envok;
                    const uint32_t sq = (uint32_t)q * kWave + (uint32_t)lane;
                    const bool ctr = p.w >= R2;  // the center rule of the job's center lists
                    const bool before = sq - sb1 < sb2 - sb1;
                    const bool after = sq >= sb2;
                    const uint32_t wgt = (ctr && after) ? 2u : 1u;
                    const float pxs = (ctr && before) ? __builtin_nanf("") : px;
                    uint32_t *arow = s_P + woff_t;
#pragma unroll
                    for (int c = 0; c < T; ++c) {
                        if (c == 0 || c < ncen) {  // uniform: skips a partial tile's absent centers
                            const float dx = cxr[c] - pxs;
                            const float dy = cyr[c] - p.y;
                            const float dz = czr[c] - p.z;
                            const float d2 = dx * dx + dy * dy + dz * dz;
                            const bool in = d2 <= R2;
                            if (__builtin_expect(cin != 0, 0))
                                my_in += (unsigned long long)__builtin_popcountll(
                                             __builtin_amdgcn_ballot_w64(in)) +
                                         (unsigned long long)__builtin_popcountll(
                                             __builtin_amdgcn_ballot_w64(in && wgt == 2u));
                            if (in) {
                                const int cat = category_of(d2, isc, ioff1, ioffa, amb2, Cm1f, s_edges);
                                __hip_atomic_fetch_add(arow + cat, wgt, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                            }
                        }
                    }
                } else if (!HY) {
                    // direct mode (pair counts per center type): one center
                    // at a time keeps the register count, and so the number
                    // of resident wavefronts, higher
#pragma unroll
                    for (int c = 0; c < T; ++c) {
                        if (c == 0 || c < ncen) {  // uniform: skips a partial tile's absent centers
                            const float dx = cxr[c] - px;
                            const float dy = cyr[c] - p.y;
                            const float dz = czr[c] - p.z;
                            const float d2 = dx * dx + dy * dy + dz * dz;
                            const bool in = d2 <= R2;
                            if (__builtin_expect(cin != 0, 0))
                                my_in += (unsigned long long)__builtin_popcountll(
                                    __builtin_amdgcn_ballot_w64(in));
                            if (in && envok) {
                                const int cat = category_of(d2, isc, ioff1, ioffa, amb2, Cm1f, s_edges);
                                __hip_atomic_fetch_add(acc_row + cat, 1u,
                                                       __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                            }
                        }
                    }
                } else if (SY) {
                    // hybrid mode, own-type pass of a symmetric tile (list
                    // S: every candidate of the tile's type, its ion2surf
                    // in w): pair-count weight 2 for a center in a cell
                    // after the own cell (its own count too: same distance,
                    // same category, same row), 0 for a center in a cell
                    // before it (counted from that center's tile), else 1;
                    // the kNN range is read for every candidate
                    (void)ptype;
                    (void)envok;
                    const uint32_t sq = (uint32_t)q * kWave + (uint32_t)lane;
                    const bool ctr = p.w >= R2;  // the center rule of the job's center lists
                    const bool before = sq - sb1 < sb2 - sb1;
                    const bool after = sq >= sb2;
                    const uint32_t wgt = (ctr && after) ? 2u : ((ctr && before) ? 0u : 1u);
                    const bool rdf_w = rdf && wgt != 0u;
                    const float bthr = rdf_w ? R2 : kthr_t;
                    uint32_t *arow = s_P + woff_t;
                    uint32_t *krow = s_H + kof_t;
#pragma unroll
                    for (int c = 0; c < T; ++c) {
                        if (c == 0 || c < ncen) {  // uniform: skips a partial tile's absent centers
                            const float dx = cxr[c] - px;
                            const float dy = cyr[c] - p.y;
                            const float dz = czr[c] - p.z;
                            const float d2 = dx * dx + dy * dy + dz * dz;
                            if (__builtin_expect(cin != 0, 0)) {
                                const bool in = d2 <= R2;
                                my_in += (unsigned long long)__builtin_popcountll(
                                             __builtin_amdgcn_ballot_w64(in && wgt != 0u)) +
                                         (unsigned long long)__builtin_popcountll(
                                             __builtin_amdgcn_ballot_w64(in && wgt == 2u));
                            }
                            if (d2 <= bthr) {
                                const int cat = category_of(d2, isc, ioff1, ioffa, amb2, Cm1f, gedges);
                                if (rdf_w)
                                    __hip_atomic_fetch_add(arow + cat, wgt, __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                                if (d2 <= kthr_t)
                                    __hip_atomic_fetch_add(
                                        krow + (C16 ? (c >> 1) : c) * HK + cat,
                                        (C16 && (c & 1)) ? 65536u : 1u, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_WORKGROUP);
                            }
                        }
                    }
                } else {
                    // hybrid mode: the direct-mode pair counts, and the pairs
                    // within the kNN range of their type also into the
                    // center's own histogram
                    const int2 ki = s_kinfo[ptype];
                    const int kof = ki.x;
                    const float kthr = __int_as_float(ki.y);  // -1: no kNN for this type
                    uint32_t *krow = s_H + kof;  // only dereferenced when d2 <= kthr
                    const bool rdf_env = rdf && envok;
                    const float bthr = rdf_env ? R2 : kthr;
#pragma unroll
                    for (int c = 0; c < T; ++c) {
                        if (c == 0 || c < ncen) {  // uniform: skips a partial tile's absent centers
                            const float dx = cxr[c] - px;
                            const float dy = cyr[c] - p.y;
                            const float dz = czr[c] - p.z;
                            const float d2 = dx * dx + dy * dy + dz * dz;
                            if (__builtin_expect(cin != 0, 0))
                                my_in += (unsigned long long)__builtin_popcountll(
                                    __builtin_amdgcn_ballot_w64(d2 <= R2));
                            if (d2 <= bthr) {
                                const int cat = category_of(d2, isc, ioff1, ioffa, amb2, Cm1f, gedges);
                                if (rdf_env)
                                    __hip_atomic_fetch_add(acc_row + cat, 1u,
                                                           __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                                if (d2 <= kthr)
                                    __hip_atomic_fetch_add(
                                        krow + (C16 ? (c >> 1) : c) * HK + cat,
                                        (C16 && (c & 1)) ? 65536u : 1u, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_WORKGROUP);
                            }
                        }
                    }
                }
            };
            // the loads are unconditional (past the block's last chunk the
            // cursor stays and its chunk is loaded again, unused), so that no
            // branch joins paths with different loads in flight, which would
            // make the wait for one chunk wait for the other too
            bool oka = false, okb = false;
            float4 pa = make_float4(0.f, 0.f, 0.f, 0.f), pb = pa;
            if (qa < qb)
                pa = fetch(oka);
            for (int q = qa; q < qb; q += 2) {
                if (q + 1 < qb)
                    advance();
                pb = fetch(okb);
                body(pa, oka, q);
                if (q + 2 < qb)
                    advance();
                pa = fetch(oka);
                if (q + 1 < qb)
                    body(pb, okb, q + 1);
            }
            };
            if (sym) {
                const uint32_t bs = rfl(s_rowpre[J.s_ns]);
                const int ns = (int)rfl((uint32_t)J.s_ns);
                sweep(CFlag<true>(), rflp((const gfvec4 *)J.pos_s), 0, ns, 0u, bs);
                if (ns < nrt)
                    sweep(CFlag<false>(), rflp((const gfvec4 *)J.pos_o), ns, nrt, bs, ctot);
            } else {
                sweep(CFlag<false>(), rflp((const gfvec4 *)J.pos), 0, nrows, 0u, ctot);
            }
            inball += my_in;
        }
        __syncthreads();

        const bool want_k = (J.families & FAM_MKNN) && J.nk > 0;
        const bool want_1 = (J.families & FAM_1NN) != 0;
        const bool want_cn = (J.families & FAM_CNTNB) != 0;
        if (!PC) {
            // remove the self pair (d = 0) each center added to its own type
            if (tid < cnt) {
                const uint32_t m = s_meta[2 * tid];
                const uint32_t es = (m >> 8) & 0xFFu;
                if (es != 0xFFu)
                    __hip_atomic_fetch_add(&s_P[es * C + J.cat0], 0xFFFFFFFFu,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            if (!HY || !(want_k || want_1))
                continue;
        }

        // ---------------- per-center extraction ----------------
        const int cat0 = J.cat0;
        // each center's own d = 0 pair leaves its histogram of its own type
        // first (in a 16-bit half the count holds that pair, so nothing
        // borrows from the other half)
        if (tid < cnt) {
            const uint32_t es = (s_meta[2 * tid] >> 8) & 0xFFu;
            if (es != 0xFFu) {
                const int h0 = (C16 ? (tid >> 1) : tid) * (PC ? EC : HK) +
                               (PC ? (int)es * C : s_ckoff[es]) + cat0;
                __hip_atomic_fetch_add(&s_H[h0], (C16 && (tid & 1)) ? 0xFFFF0000u : 0xFFFFFFFFu,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        // the tasks of the tile's center type (every center of a tile has
        // it): per task l its id, env set and kNN-range categories, read once
        // here rather than by each center's wave in series, into the stencil
        // rows' words (free until the next tile) as far as they hold them
        const int Lt = (int)J.trg_ntask[ts];
        // (the rows' starts and prefix hold L.cen - L.rows words)
        const int Ls = Lt < (L.cen - L.rows) / 3 ? Lt : (L.cen - L.rows) / 3;
        uint32_t *s_task = s_rows;  // (the main loop's closing barrier freed them)
        for (int l = tid; l < Ls; l += kBlock) {
            const uint32_t tk = J.trg_tasks[ts * J.lmax + l];
            s_task[3 * l] = tk;
            s_task[3 * l + 1] = J.task_env[tk];
            s_task[3 * l + 2] = HY ? J.task_ck[tk] : (uint32_t)C;
        }
        __syncthreads();
        if (PC && (J.families & (FAM_RDF | FAM_RIPK))) {
            // pair counts per (center type, env type, category); each thread
            // owns distinct (env, category) entries, so no atomics are needed
            for (int idx = tid; idx < EC; idx += kBlock) {
                for (int c = 0; c < cnt; ++c) {
                    const uint32_t w = s_H[(C16 ? (c >> 1) : c) * EC + idx];
                    s_P[idx] += C16 ? ((c & 1) ? (w >> 16) : (w & 0xFFFFu)) : w;
                }
            }
        }
        if (want_k || want_1 || want_cn) {
            const int kmax = want_k ? J.kmax : 0;
            const bool kdense = J.k_dense != 0;
            const uint32_t need = want_k ? (uint32_t)kmax : 1u;
            const int lmax = J.lmax;
            // this tile's block of k-th neighbour rows in the chunk buffer
            const uint64_t tl = (uint64_t)(g - a.tile_off[job] - J.tile_base);
            const uint64_t TR = (uint64_t)T * (uint64_t)lmax;
            // stored through the global address space: a flat store counts
            // against the LDS wait counter as well, so every later LDS wait
            // of the extraction would wait for the store to complete
            gu16 *kb = J.KB ? (gu16 *)(J.KB + tl * (uint64_t)J.npad * TR) : nullptr;
            // read once: a load in the loop would put a wait for every store
            // in flight before the next store
            const uint16_t sentinel = (uint16_t)rfl((uint32_t)J.cat_sentinel);
            const uint32_t p1nn = rfl((uint32_t)J.nk) * (uint32_t)TR;  // the 1NN plane
            // the 1NN switch as an opaque scalar, so that its per-category
            // test is skipped, not masked, when the job has no 1NN
            int w1 = (int)rfl(want_1 ? 1u : 0u);
            asm volatile("" : "+s"(w1));
            for (int c = wid; c < cnt; c += kWaves) {
                const int hoff = (C16 ? (c >> 1) : c) * (PC ? EC : HK);
                const int sh = (C16 && (c & 1)) ? 16 : 0;
                // count of env slot e, category cat in this center's histogram
                // (its own d = 0 pair already removed)
                auto hv = [&](int e, int cat) -> uint32_t {
                    const uint32_t w =
                        s_H[hoff + (PC ? e * C : (int)rfl((uint32_t)s_ckoff[e])) + cat];
                    return C16 ? ((w >> sh) & 0xFFFFu) : w;
                };
                const int L = Lt;
                // a task whose kNN range is short of the full range must hold
                // its k neighbours there, else the center is rerun over the
                // full range (its rows here are then marked empty)
                bool short_range = false;
                for (int l = 0; l < L; ++l) {
                    const bool st = l < Ls;
                    const int tk = (int)rfl(st ? s_task[3 * l] : J.trg_tasks[ts * lmax + l]);
                    const int row = c * lmax + l;
                    const uint32_t envset = rfl(st ? s_task[3 * l + 1] : J.task_env[tk]);
                    const uint32_t ckw =
                        rfl(st ? s_task[3 * l + 2] : (HY ? J.task_ck[tk] : (uint32_t)C));
                    const int ck = HY ? (int)(ckw & 0x7FFFFFFFu) : C;
                    if ((want_k || want_1) && kb) {
                        // categories of the k-th nearest neighbours: a prefix
                        // scan over the center's category counts (two
                        // categories per lane) gives each category the first k
                        // it holds;
                        // the MKNN orders are then settled 64 at a time (lane =
                        // k): each category marks its first k in the wave's
                        // window row, and a running maximum over the window
                        // carries every category to the k after it, up to the
                        // next mark.  A window is written once every category
                        // that starts in it has been counted.
                        uint32_t carry = 0;
                        uint32_t kw = 0;  // the window holds k = kw + 1 .. kw + 64
                        // (window entries and wcur hold category + 1, 0 for
                        // none, so that the running maximum starts at 0)
                        uint32_t wcur = 0;  // category of k = kw, plus 1
                        uint16_t *wrow = s_win + wid * kWave;
                        if (want_k)
                            wrow[lane] = 0u;
                        // writes the window's orders (k <= lim: the scanned
                        // category, above: the sentinel) and moves it on
                        auto flush = [&](uint32_t lim) {
                            uint32_t m = wrow[lane];
                            wrow[lane] = 0u;
                            m = __builtin_elementwise_max(wave_max_scan(m), wcur);
                            wcur = (uint32_t)__builtin_amdgcn_readlane((int)m, kWave - 1);
                            const uint32_t k = kw + 1u + (uint32_t)lane;
                            if (k <= (uint32_t)kmax) {
                                const int kid = kdense ? (int)k - 1 : s_kidx[k];
                                if (kid >= 0)
                                    kb[(uint32_t)kid * (uint32_t)TR + (uint32_t)row] =
                                        k <= lim ? (uint16_t)(m - 1u) : sentinel;
                            }
                            kw += (uint32_t)kWave;
                        };
                        // a task of one neighbour type (the usual case) reads
                        // one histogram row, located once per task
                        const bool one_env = __builtin_popcount(envset) == 1;
                        const int e1 = one_env ? __builtin_ctz(envset) : 0;
                        const int off1 =
                            hoff + (PC ? e1 * C : (int)rfl((uint32_t)s_ckoff[e1]));
                        // 128 categories at a time, two per lane (ca, ca + 1):
                        // one scan of the pair sums gives both prefixes; the
                        // next 128 are read while these are settled
                        auto load2 = [&](int c0, uint32_t &a, uint32_t &b) {
                            a = 0u;
                            b = 0u;
                            if (one_env) {
                                if (c0 < ck) {
                                    const uint32_t w = s_H[off1 + c0];
                                    a = C16 ? ((w >> sh) & 0xFFFFu) : w;
                                }
                                if (c0 + 1 < ck) {
                                    const uint32_t w = s_H[off1 + c0 + 1];
                                    b = C16 ? ((w >> sh) & 0xFFFFu) : w;
                                }
                            } else {
                                for (int e = 0; e < E; ++e)
                                    if ((envset >> e) & 1u) {
                                        if (c0 < ck)
                                            a += hv(e, c0);
                                        if (c0 + 1 < ck)
                                            b += hv(e, c0 + 1);
                                    }
                            }
                        };
                        uint32_t va, vb;
                        load2(2 * lane, va, vb);
                        for (int cb = 0; cb < ck; cb += 2 * kWave) {
                            const int ca = cb + 2 * lane;
                            uint32_t na = 0u, nb = 0u;
                            if (cb + 2 * kWave < ck)
                                load2(ca + 2 * kWave, na, nb);
                            const uint32_t incl = wave_incl_scan(va + vb, lane);
                            const uint32_t cumb = carry + incl, cuma = cumb - vb;
                            const uint32_t preva = cuma - va;  // prevb = cuma
                            if (w1 != 0 && preva < 1u && cumb >= 1u)
                                kb[p1nn + (uint32_t)row] =
                                    (uint16_t)(cuma >= 1u ? ca : ca + 1);
                            const uint32_t cnew =
                                carry + (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
                            if (want_k) {
                                // each category's first k
                                const uint32_t fa = preva + 1u, fb = cuma + 1u;
                                while (kw < (uint32_t)kmax) {
                                    if (va > 0u && fa > kw && fa <= kw + (uint32_t)kWave)
                                        wrow[fa - kw - 1u] = (uint16_t)(ca + 1);
                                    if (vb > 0u && fb > kw && fb <= kw + (uint32_t)kWave)
                                        wrow[fb - kw - 1u] = (uint16_t)(ca + 2);
                                    const uint32_t wend = kw + (uint32_t)kWave < (uint32_t)kmax
                                                              ? kw + (uint32_t)kWave
                                                              : (uint32_t)kmax;
                                    if (cnew < wend)
                                        break;  // later chunks may still start in the window
                                    flush(wend);
                                }
                            }
                            carry = cnew;
                            if (carry >= need)
                                break;
                            va = na;
                            vb = nb;
                        }
                        if (HY && carry < need && (ckw >> 31)) {
                            short_range = true;
                            break;
                        }
                        // fewer than k env neighbours: R + EPSILON, i.e. the
                        // sentinel category (S.cpp:4093-4095, 4035-4036)
                        if (want_1 && carry == 0u && lane == 0)
                            kb[p1nn + (uint32_t)row] = sentinel;
                        if (want_k)
                            while (kw < (uint32_t)kmax)
                                flush(carry);
                    }
                    if (PC && want_cn) {
                        for (int e = 0; e < E; ++e) {
                            if (!((envset >> e) & 1u))
                                continue;
                            uint32_t tot = 0;
                            for (int cat = lane; cat < C; cat += kWave)
                                tot += hv(e, cat);
                            for (int d = kWave / 2; d > 0; d >>= 1)
                                tot += __shfl_xor(tot, d, kWave);
                            if (lane == 0) {
                                if ((int)tot < J.cn_cap)
                                    atomicAdd(&J.CN[(size_t)tk * J.cn_cap + tot], 1u);
                                else
                                    atomicMax((unsigned int *)&J.stats[2], 1u);
                            }
                        }
                    }
                }
                if (HY && short_range) {
                    // the rerun settles this center: every plane of its rows
                    // (some already written) is emptied for the reduction
                    if (lane == 0)
                        J.unres[J.centers[t_first + c]] = 1u;
                    if (kb) {
                        const int np = J.nplanes;
                        for (int i = lane; i < np * L; i += kWave) {
                            const int pl = i / L, l = i - pl * L;
                            kb[(uint32_t)pl * (uint32_t)TR + (uint32_t)(c * lmax + l)] = kKbEmpty;
                        }
                    }
                }
            }
        }
    }
}

template <int T, bool C16, int MODE>
static hipError_t launch_query_t(const QueryLaunch &q, hipStream_t s)
{
    const size_t lds = query_lds_bytes(T, C16, MODE, q.maxC, q.maxE, q.maxTrg, q.maxKmax,
                                       q.maxRows, q.maxHK, q.ntypes);
    QArgs a;
    a.jobs = q.jobs;
    a.tile_off = q.tile_off;
    a.tiles = q.tiles;
    a.queue = q.queue;
    a.first_tile = q.first_tile;
    a.end_tile = q.end_tile;
    a.grab = q.grab > 0 ? q.grab : 1u;
    a.njobs = q.njobs;
    a.maxC = q.maxC;
    a.maxE = q.maxE;
    a.maxTrg = q.maxTrg;
    a.maxKmax = q.maxKmax;
    a.ntypes = q.ntypes;
    a.maxRows = q.maxRows;
    a.maxHK = q.maxHK;
    hipError_t e = hipFuncSetAttribute((const void *)k_query<T, C16, MODE>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL((k_query<T, C16, MODE>), dim3(q.grid), dim3(kBlock), lds, s, a);
    return hipGetLastError();
}

template <int T, bool C16, int MODE>
static hipError_t occ_t(size_t lds, int *nb)
{
    hipError_t e = hipFuncSetAttribute((const void *)k_query<T, C16, MODE>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess)
        return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, k_query<T, C16, MODE>, kBlock, lds);
}

#define PP_DISPATCH_T(FN, C16, MODE, ...)                                    \
    do {                                                                     \
        if (T == 4) return FN<4, C16, MODE>(__VA_ARGS__);                    \
        if (T == 8) return FN<8, C16, MODE>(__VA_ARGS__);                    \
        if (T == 16) return FN<16, C16, MODE>(__VA_ARGS__);                  \
        if (T == 32) return FN<32, C16, MODE>(__VA_ARGS__);                  \
    } while (0)

#define PP_DISPATCH(FN, ...)                                                 \
    do {                                                                     \
        if (mode == 0)                                                       \
            PP_DISPATCH_T(FN, true, 0, __VA_ARGS__);                         \
        else if (mode == 1 && c16)                                           \
            PP_DISPATCH_T(FN, true, 1, __VA_ARGS__);                         \
        else if (mode == 1)                                                  \
            PP_DISPATCH_T(FN, false, 1, __VA_ARGS__);                        \
        else if (mode == 2 && c16)                                           \
            PP_DISPATCH_T(FN, true, 2, __VA_ARGS__);                         \
        else if (mode == 2)                                                  \
            PP_DISPATCH_T(FN, false, 2, __VA_ARGS__);                        \
        return hipErrorInvalidValue;                                         \
    } while (0)

hipError_t query_max_blocks_per_cu(int T, int c16, int mode, size_t lds, int *nb)
{
    PP_DISPATCH(occ_t, lds, nb);
}

// clears the tile queue counters (a kernel, not a memset node: the query
// launches may be captured into a HIP graph)
__global__ __launch_bounds__(64) void k_queue_clear(uint32_t *__restrict__ q)
{
    for (int i = threadIdx.x; i < kQueueWords; i += 64)
        q[i] = 0u;
}

hipError_t launch_query(const QueryLaunch &q, hipStream_t s)
{
    const int T = q.T, c16 = q.c16, mode = q.mode;
    hipLaunchKernelGGL(k_queue_clear, dim3(1), dim3(64), 0, s, q.queue);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    if (q.end_tile <= q.first_tile)
        return hipSuccess;
    PP_DISPATCH(launch_query_t, q, s);
}

// ---------------------------------------------------------------------------
// k-th neighbour histograms of one query chunk

// One work unit: a group of kb_group consecutive planes (MKNN orders, or the
// 1NN) of a block of tiles_per_unit tiles of one job.  A tile's rows of the
// group are one contiguous run (planes are padded to a multiple of 8 per
// tile, so the runs start on cache lines); the unit reads them two rows per
// thread and counts the categories per (task, plane, category) in 16-bit LDS
// counters, then adds them to the job's K / N1 with one global atomic per
// non-zero count.
constexpr uint32_t kRedStage = 1024;  // tiles staged in LDS at a time

__global__ __launch_bounds__(256) void k_kreduce(const JobDev *__restrict__ jobs, int njobs,
                                                 const uint32_t *__restrict__ unit_off, int T,
                                                 uint32_t tpu, int lds_words)
{
    extern __shared__ uint32_t h[];
    const uint32_t u = blockIdx.x;
    const int j = find_job(unit_off, njobs, u);
    const JobDev &J = jobs[j];
    const uint32_t lu = u - unit_off[j];
    const uint32_t nblk = (J.kb_tiles + tpu - 1) / tpu;
    const int G = J.kb_group;
    const int grp = (int)(lu / nblk);
    const uint32_t t0 = (lu - (uint32_t)grp * nblk) * tpu;
    const uint32_t t1 = t0 + tpu < J.kb_tiles ? t0 + tpu : J.kb_tiles;
    const int p0 = grp * G;
    const int np = J.nplanes, C = J.C;
    const uint32_t lmax = (uint32_t)J.lmax;
    const uint32_t TR = (uint32_t)T * lmax;  // even: T >= 4
    const int words16 = J.ntasks * G * C;
    // the task of row l of each center type, after the counters in LDS
    // (-1: the type has fewer tasks)
    const int nw = (words16 + 1) / 2, ntl_task = J.ntrg * (int)lmax;
    const bool use_lds = nw + ntl_task <= lds_words;
    int32_t *s_tk = (int32_t *)h + (use_lds ? nw : 0);
    for (int i = threadIdx.x; i < (use_lds ? nw : 0); i += blockDim.x)
        h[i] = 0u;
    const Tile *tiles = J.tiles + J.tile_base + t0;
    const uint32_t ntl = t1 - t0;
    const uint32_t EP = (uint32_t)G * TR / 2;  // row pairs per tile in the group
    const size_t tstride = (size_t)J.npad * TR;
    // read through the global address space (a flat load would also count
    // against the LDS wait counter, and the LDS reads between the loads would
    // wait for them)
    const gcu32 *kbw = (const gcu32 *)(J.KB + (size_t)t0 * tstride + (size_t)p0 * TR);
    const size_t twords = tstride / 2;
    // the unit's tiles go through LDS kRedStage at a time, each as its center
    // count, center type slot and that slot's task count, so that a row pair
    // needs one global load (its two categories) besides its tasks (count
    // <= 32, slot < 32 and task count <= 64 fit 8 bits each)
    __shared__ uint32_t s_ti[kRedStage];
    // one category of row r = c * lmax + l of plane pl of a tile (ti: its
    // staged info)
    // (LDS: the counters in LDS or, when they do not fit, global atomics,
    // a compile-time choice inside the sweep so that its waits stay exact)
    auto count = [&](auto LDS, uint32_t ti, uint32_t pl, uint32_t c, uint32_t l, uint32_t cat) {
        constexpr bool use_lds = decltype(LDS)::value;
        const int tks = use_lds ? s_tk[((ti >> 8) & 0xFFu) * lmax + l] : 0;
        if (c >= (ti & 0xFFu) || l >= (ti >> 16) || cat == kKbEmpty)
            return;
        const uint32_t tk =
            use_lds ? (uint32_t)tks : J.trg_tasks[((ti >> 8) & 0xFFu) * lmax + l];
        if (use_lds) {
            const uint32_t i16 = (tk * (uint32_t)G + pl) * (uint32_t)C + cat;
            atomicAdd(&h[i16 >> 1], 1u << ((i16 & 1u) << 4));
        } else {
            const int P = p0 + (int)pl;
            uint32_t *dst = P == J.nk ? J.N1 + (size_t)tk * C
                                      : J.K + ((size_t)tk * J.nk + (size_t)P) * C;
            global_add(&dst[cat], 1u);
        }
    };
    // row pair e (rows 2e and 2e + 1 of the group's run: plane pl, rows r and
    // r + 1 of it, TR being even) of staged tile st
    auto pair = [&](auto LDS, uint32_t st, uint32_t v, uint32_t pl, uint32_t c0, uint32_t l0,
                    uint32_t c1, uint32_t l1) {
        const uint32_t ti = s_ti[st];
        count(LDS, ti, pl, c0, l0, v & 0xFFFFu);
        count(LDS, ti, pl, c1, l1, v >> 16);
    };
    // thread -> (tile of the step, row pair) when a tile's pairs fit the
    // block: the divisions once, then the block steps over `per` tiles at a
    // time, four tiles' loads in flight per thread
    const bool fits = EP <= blockDim.x;
    const uint32_t per = fits ? blockDim.x / EP : 1u;
    const uint32_t dt = fits ? threadIdx.x / EP : 0u;
    const uint32_t e0 = fits ? threadIdx.x - dt * EP : 0u;
    const uint32_t pl0 = 2u * e0 / TR, r0 = 2u * e0 - pl0 * TR;
    const uint32_t c00 = r0 / lmax, l00 = r0 - c00 * lmax;
    const uint32_t c01 = (r0 + 1u) / lmax, l01 = r0 + 1u - c01 * lmax;
    const bool live0 = p0 + (int)pl0 < np;  // not a padding plane
    if (use_lds)
        for (int i = threadIdx.x; i < ntl_task; i += blockDim.x) {
            const int slot = i / (int)lmax, l = i - slot * (int)lmax;
            s_tk[i] = l < (int)J.trg_ntask[slot] ? (int32_t)J.trg_tasks[i] : -1;
        }
    auto sweep = [&](auto LDS) {
    for (uint32_t s0 = 0; s0 < ntl; s0 += kRedStage) {
        const uint32_t sn = ntl - s0 < kRedStage ? ntl - s0 : kRedStage;
        __syncthreads();  // the previous stage's rows are done with s_ti
        for (uint32_t i = threadIdx.x; i < sn; i += blockDim.x) {
            const Tile t = tiles[s0 + i];
            s_ti[i] = (uint32_t)t.count | (t.pad << 8) | (J.trg_ntask[t.pad] << 16);
        }
        __syncthreads();
        if (fits) {
            if (dt < per && live0)
                for (uint32_t st = dt; st < sn; st += 4u * per) {
                    uint32_t v[4];
                    // unconditional loads (past the stage's end: its last
                    // tile again, not counted), so that the four stay in flight
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const uint32_t sq = st + (uint32_t)q * per;
                        v[q] = kbw[(size_t)(s0 + (sq < sn ? sq : sn - 1u)) * twords + e0];
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const uint32_t sq = st + (uint32_t)q * per;
                        if (sq < sn)
                            pair(LDS, sq, v[q], pl0, c00, l00, c01, l01);
                    }
                }
        } else {
            for (uint32_t st = 0; st < sn; ++st)
                for (uint32_t e = threadIdx.x; e < EP; e += blockDim.x) {
                    const uint32_t pl = 2u * e / TR, r = 2u * e - pl * TR;
                    if (p0 + (int)pl >= np)
                        continue;
                    const uint32_t c0 = r / lmax, c1 = (r + 1u) / lmax;
                    pair(LDS, st, kbw[(size_t)(s0 + st) * twords + e], pl, c0, r - c0 * lmax,
                         c1, r + 1u - c1 * lmax);
                }
        }
    }
    };
    if (use_lds)
        sweep(CFlag<true>());
    else
        sweep(CFlag<false>());
    if (use_lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < (words16 + 1) / 2; i += blockDim.x) {
            const uint32_t w = h[i];
            if (!w)
                continue;
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                const uint32_t cnt = (w >> (16 * half)) & 0xFFFFu;
                const int i16 = 2 * i + half;
                if (!cnt || i16 >= words16)
                    continue;
                const int tk = i16 / (G * C);
                const int rem = i16 - tk * G * C;
                const int pl = rem / C, cat = rem - pl * C;
                const int P = p0 + pl;
                uint32_t *dst = P == J.nk ? J.N1 + (size_t)tk * C
                                          : J.K + ((size_t)tk * J.nk + (size_t)P) * C;
                global_add(&dst[cat], cnt);
            }
        }
    }
}

hipError_t launch_kreduce(const JobDev *jobs, int njobs, const uint32_t *unit_off, uint32_t units,
                          int T, uint32_t tiles_per_unit, int lds_words, hipStream_t s)
{
    if (units == 0)
        return hipSuccess;
    const size_t lds = (size_t)lds_words * 4u;
    hipError_t e = hipFuncSetAttribute((const void *)k_kreduce,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_kreduce, dim3(units), dim3(256), lds, s, jobs, njobs, unit_off, T,
                       tiles_per_unit, lds_words);
    return hipGetLastError();
}

} // namespace pp
