// pp_prep.hip -- the center lists and query tiles of a whole run of jobs in a
// fixed number of launches.
//
// A run (one job, a batch, a serving batch, or the full-range rerun of the
// short centers) is a sequence of segments, one per (job, center type): the
// considered centers of segment q are the ions of its job's cell-sorted list
// whose type has slot q.slot and whose ion2surf is at least fl32(R^2)
// (src/PARAPROBE_SolverHdl.cpp:3846-3860), or the ions a rerun selection
// flags.  All segments' ion ranges are laid end to end (the ion space, ion0
// onwards for segment q) and so are their cell ranges (the cell space,
// cell0 onwards): one flag pass, one exclusive scan and one compaction give
// the run's center list (segment after segment, each in cell order); one
// count pass, one scan and one write pass give its tiles -- up to T centers
// of one segment in one cell, first center as an index into the run's center
// list -- and each job's largest stencil candidate count.
#include "pp_launch.h"

namespace pp {

namespace {

constexpr int kPrepBlock = 256;

// The segment holding element e of a space whose segment starts are
// start(q): thread 0 of the block finds the block's first element's segment
// by bisection, every thread then steps forward (elements of one block span
// few segments).
template <typename Start>
__device__ inline int seg_of(uint64_t e, uint64_t block_first, int nseg, Start start,
                             int *s_first)
{
    if (threadIdx.x == 0) {
        int lo = 0, hi = nseg;  // start(lo) <= block_first < start(hi)
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (start(mid) <= block_first)
                lo = mid;
            else
                hi = mid;
        }
        *s_first = lo;
    }
    __syncthreads();
    int q = *s_first;
    while (q + 1 < nseg && start(q + 1) <= e)
        ++q;
    return q;
}

__global__ __launch_bounds__(kPrepBlock) void k_pflags(const PrepSeg *__restrict__ segs, int nseg,
                                                        uint64_t E, uint32_t *__restrict__ flags)
{
    __shared__ int s_first;
    const uint64_t b0 = (uint64_t)blockIdx.x * kPrepBlock;
    const uint64_t e = b0 + threadIdx.x;
    const int q = seg_of(e < E ? e : E - 1, b0, nseg, [&](int k) { return segs[k].ion0; },
                         &s_first);
    if (e >= E)
        return;
    const PrepSeg S = segs[q];  // by value: kept in registers, not re-read
    const uint64_t i = e - S.ion0;
    const uint32_t t = __float_as_uint(S.pos[i].w) & 0xFFu;
    const int ts = S.trgslot[t];
    const bool sel = S.src ? S.src[i] != 0u : (ts >= 0 && S.s2[i] >= S.R2);
    flags[e] = (sel && ts == S.slot) ? 1u : 0u;
}

// centers[cpos[e]] = the job-local sorted ion index of every flagged element
__global__ __launch_bounds__(kPrepBlock) void k_pcompact(const PrepSeg *__restrict__ segs,
                                                          int nseg, uint64_t E,
                                                          const uint32_t *__restrict__ flags,
                                                          const uint32_t *__restrict__ cpos,
                                                          uint32_t *__restrict__ centers)
{
    __shared__ int s_first;
    const uint64_t b0 = (uint64_t)blockIdx.x * kPrepBlock;
    const uint64_t e = b0 + threadIdx.x;
    const int q = seg_of(e < E ? e : E - 1, b0, nseg, [&](int k) { return segs[k].ion0; },
                         &s_first);
    if (e < E && flags[e])
        centers[cpos[e]] = (uint32_t)(e - segs[q].ion0);
}

__global__ __launch_bounds__(kPrepBlock) void k_ptilecount(const PrepSeg *__restrict__ segs,
                                                            int nseg, uint64_t Ec,
                                                            const uint32_t *__restrict__ cpos,
                                                            uint32_t *__restrict__ ntile)
{
    __shared__ int s_first;
    const uint64_t b0 = (uint64_t)blockIdx.x * kPrepBlock;
    const uint64_t e = b0 + threadIdx.x;
    const int q = seg_of(e < Ec ? e : Ec - 1, b0, nseg, [&](int k) { return segs[k].cell0; },
                         &s_first);
    if (e >= Ec)
        return;
    const PrepSeg S = segs[q];
    const uint64_t c = e - S.cell0;
    const uint32_t lo = cpos[S.ion0 + S.cellstart[c]], hi = cpos[S.ion0 + S.cellstart[c + 1]];
    ntile[e] = (hi - lo + (uint32_t)S.T - 1u) / (uint32_t)S.T;
}

// stencil row rw around cell (cx, cy, cz): its candidate range, empty when
// the row lies outside the grid
__device__ inline uint32_t row_count(const Row &rw, int cx, int cy, int cz, const GridGeom &g,
                                     const uint32_t *__restrict__ cellstart)
{
    const int y = cy + rw.dy, z = cz + rw.dz;
    if (y < 0 || y >= g.ny || z < 0 || z >= g.nz)
        return 0u;
    const int x0 = cx - rw.dxm < 0 ? 0 : cx - rw.dxm;
    const int x1 = cx + rw.dxm >= g.nx ? g.nx - 1 : cx + rw.dxm;
    const uint64_t k0 = ((uint64_t)z * g.ny + y) * g.nx + x0;
    return cellstart[k0 + (uint32_t)(x1 - x0) + 1u] - cellstart[k0];
}

// The tiles of every cell with centers, and each job's largest stencil
// candidate count (a bound on any center's neighbour count).
__global__ __launch_bounds__(kPrepBlock) void k_ptilewrite(const PrepSeg *__restrict__ segs,
                                                            int nseg, uint64_t Ec,
                                                            const uint32_t *__restrict__ cpos,
                                                            const uint32_t *__restrict__ toff,
                                                            Tile *__restrict__ tiles,
                                                            PrepJobStat *__restrict__ js)
{
    __shared__ int s_first;
    const uint64_t b0 = (uint64_t)blockIdx.x * kPrepBlock;
    const uint64_t e = b0 + threadIdx.x;
    const int q = seg_of(e < Ec ? e : Ec - 1, b0, nseg, [&](int k) { return segs[k].cell0; },
                         &s_first);
    // every lane stays to the wave's reduction below (no early return)
    uint32_t tot = 0;
    int job = 0;
    if (e < Ec) {
        const PrepSeg S = segs[q];  // by value: kept in registers, not re-read
        job = S.job;
        const uint64_t c = e - S.cell0;
        const uint32_t lo = cpos[S.ion0 + S.cellstart[c]],
                       hi = cpos[S.ion0 + S.cellstart[c + 1]];
        if (hi > lo) {
            const uint32_t T = (uint32_t)S.T, cnt = hi - lo, t0 = toff[e];
            for (uint32_t m = 0; m * T < cnt; ++m) {
                Tile t;
                t.key = (uint32_t)c;
                t.first = lo + m * T;
                const uint32_t left = cnt - m * T;
                t.count = (uint16_t)(left < T ? left : T);
                t.job = (uint16_t)S.job;
                t.pad = (uint32_t)S.slot;
                tiles[t0 + m] = t;
            }
            const int cx = (int)(c % (uint64_t)S.g.nx);
            const int cy = (int)((c / (uint64_t)S.g.nx) % (uint64_t)S.g.ny);
            const int cz = (int)(c / ((uint64_t)S.g.nx * (uint64_t)S.g.ny));
            for (int r = 0; r < S.nrows; ++r)
                tot += row_count(S.rows[r], cx, cy, cz, S.g, S.cellstart);
        }
    }
    // one atomic per wavefront and job: the lanes of a wave whose cells all
    // belong to one job (the usual case) take their maximum first (a single
    // counter line takes atomics one at a time)
    const int j0 = __shfl(job, 0, 64);
    const bool uniform = __all(job == j0) != 0;
    if (uniform) {
        for (int d = 32; d > 0; d >>= 1) {
            const uint32_t o = (uint32_t)__shfl_xor((int)tot, d, 64);
            tot = o > tot ? o : tot;
        }
        if ((threadIdx.x & 63) == 0 && tot > 0)
            atomicMax(&js[j0].maxcand, tot);
    } else if (tot > 0) {
        atomicMax(&js[job].maxcand, tot);
    }
}

// Per job: its first center and tile in the run's lists and their counts
// (a job's segments are consecutive).
__global__ void k_pjobstat(const PrepSeg *__restrict__ segs, const int32_t *__restrict__ jseg,
                           int njobs, const uint32_t *__restrict__ cpos,
                           const uint32_t *__restrict__ toff, PrepJobStat *__restrict__ js)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= njobs)
        return;
    const int a = jseg[j], b = jseg[j + 1];
    if (a == b) {
        js[j].center0 = js[j].centers = js[j].tile0 = js[j].tiles = 0u;
        return;
    }
    const PrepSeg &F = segs[a], &L = segs[b - 1];
    js[j].center0 = cpos[F.ion0];
    js[j].centers = cpos[L.ion0 + L.n] - cpos[F.ion0];
    js[j].tile0 = toff[F.cell0];
    js[j].tiles = toff[L.cell0 + L.nc] - toff[F.cell0];
}

__global__ void k_pclear(PrepJobStat *__restrict__ js, int njobs)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < njobs)
        js[j].maxcand = 0u;
}

unsigned grid_of(uint64_t n)
{
    uint64_t g = (n + kPrepBlock - 1) / kPrepBlock;
    return (unsigned)(g ? g : 1);
}

} // namespace

size_t prep_workspace_words(uint64_t E, uint64_t Ec)
{
    return scan_workspace_words(E > Ec ? E : Ec);
}

hipError_t launch_prep(const PrepLaunch &p, hipStream_t s)
{
    if (p.nseg <= 0 || p.njobs <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_pclear, dim3((p.njobs + 255) / 256), dim3(256), 0, s, p.js, p.njobs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    if (p.E > 0) {
        hipLaunchKernelGGL(k_pflags, dim3(grid_of(p.E)), dim3(kPrepBlock), 0, s, p.segs, p.nseg,
                           p.E, p.flags);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
    }
    if ((e = scan_exclusive(p.flags, p.cpos, p.E, p.ws, s)) != hipSuccess)
        return e;
    if (p.E > 0) {
        hipLaunchKernelGGL(k_pcompact, dim3(grid_of(p.E)), dim3(kPrepBlock), 0, s, p.segs, p.nseg,
                           p.E, p.flags, p.cpos, p.centers);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
    }
    if (p.Ec > 0) {
        hipLaunchKernelGGL(k_ptilecount, dim3(grid_of(p.Ec)), dim3(kPrepBlock), 0, s, p.segs,
                           p.nseg, p.Ec, p.cpos, p.ntile);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
    }
    if ((e = scan_exclusive(p.ntile, p.toff, p.Ec, p.ws, s)) != hipSuccess)
        return e;
    if (p.Ec > 0) {
        hipLaunchKernelGGL(k_ptilewrite, dim3(grid_of(p.Ec)), dim3(kPrepBlock), 0, s, p.segs,
                           p.nseg, p.Ec, p.cpos, p.toff, p.tiles, p.js);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
    }
    hipLaunchKernelGGL(k_pjobstat, dim3((p.njobs + 255) / 256), dim3(256), 0, s, p.segs, p.jseg,
                       p.njobs, p.cpos, p.toff, p.js);
    return hipGetLastError();
}

} // namespace pp
