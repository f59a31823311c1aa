// pp_npcorr.hip -- gfx950 kernel of the n-point correlation family (NPCORR,
// AnalysisSpatDistrType digit 5) of horderdist::compute_generic_spatstat3.
//
// Replaces the NPCORR branch of the center loop (src/PARAPROBE_SolverHdl.cpp:
// 3899-3960) and npc3d::add (src/PARAPROBE_NPointCorrelations3D.cpp:165-207):
// for every considered center (ion2surf >= fl32(R*R)) and every task whose
// central types hold its type, when more than K ions of any type lie within
// R (self excluded), the env-type neighbours within R are ordered by
// d = sqrtf(fl((u*u + v*v) + w*w)), (u, v, w) = neighbour - center; the
// K-th (0-based) one's difference vector is binned into a cube of
// (2h + 1)^3 cells of width `width` centred on the center, each component
// rounded half away from zero in double ((int)(t + 0.5) / (int)(t - 0.5) of
// t = fl(u * fl(1 / width))), when d <= maxlen.  The reference picks the K-th
// with nth_element over its KD-tree order, so equal distances leave the pick
// to that order; here ties are broken by the smaller (u, v, w),
// lexicographically, numerically.
//
// One wavefront per center over the 5 x 5 x 5 cells of a cell list at least
// R/2 wide.  For each task the wave first counts the env neighbours per
// distance bucket (kNpcBuckets buckets of width R / kNpcBuckets over
// d = sqrtf(d2), in LDS), finds the bucket that holds the K-th one, then
// gathers that bucket's neighbours (16 bytes each: d bits, then u, v, w as
// order-preserving integers) and ranks them; buckets are monotone in d, so
// every neighbour of a lower bucket precedes them.  When the bucket holds
// more than kNpcCap neighbours (shells of a perfect lattice) the wave falls
// back to a buffer of candidates that is sorted (bitonic) and cut to the
// K + 1 smallest whenever it fills up.  K must be below kNpcCap - 64.
#include "pp_launch.h"

#pragma clang fp contract(off)

namespace pp {

namespace {

__device__ inline uint32_t ordf(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ inline float unordf(uint32_t o)
{
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

__device__ inline bool key_less(const uint4 &a, const uint4 &b)
{
    if (a.x != b.x)
        return a.x < b.x;
    if (a.y != b.y)
        return a.y < b.y;
    if (a.z != b.z)
        return a.z < b.z;
    return a.w < b.w;
}

__device__ inline int cell_of(float v, double lo, double inv, int dim)
{
    int c = (int)floor(((double)v - lo) * inv);
    c = c < 0 ? 0 : c;
    return c >= dim ? dim - 1 : c;
}

// Bitonic sort of buf[0, kNpcCap) by key_less, one wavefront; entries at or
// beyond cnt are set to the largest key first so that they sort last.
__device__ void wave_sort(uint4 *buf, uint32_t cnt, int lane)
{
    for (uint32_t i = cnt + lane; i < (uint32_t)kNpcCap; i += kWave)
        buf[i] = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
    __builtin_amdgcn_wave_barrier();
    for (uint32_t k = 2; k <= (uint32_t)kNpcCap; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = lane; i < (uint32_t)kNpcCap; i += kWave) {
                const uint32_t p = i ^ j;
                if (p > i) {
                    const uint4 a = buf[i], b = buf[p];
                    const bool up = (i & k) == 0;
                    if (up ? key_less(b, a) : key_less(a, b)) {
                        buf[i] = b;
                        buf[p] = a;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
}

} // namespace

__global__ __launch_bounds__(256) void k_npcorr(NpcArgs a)
{
    __shared__ uint4 s_buf[kWaves][kNpcCap];
    __shared__ uint32_t s_hist[kWaves][kNpcBuckets];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    uint4 *buf = s_buf[wid];
    uint32_t *hist = s_hist[wid];
    const uint64_t ci = (uint64_t)blockIdx.x * kWaves + wid;
    if (ci >= a.n)
        return;
    const float4 me = a.pos[ci];
    const uint32_t mt = __float_as_uint(me.w);
    // considered: ion2surf >= fl32(R * R) (S.cpp:3846-3860)
    if (!(a.s2[ci] >= a.R2))
        return;
    bool any = false;
    for (int t = 0; t < a.ntasks; ++t)
        any |= ((a.trg[t * 8 + (mt >> 5)] >> (mt & 31)) & 1u) != 0;
    if (!any)
        return;
    const int cx = cell_of(me.x, a.g.lo[0], a.g.inv_w, a.g.nx);
    const int cy = cell_of(me.y, a.g.lo[1], a.g.inv_w, a.g.ny);
    const int cz = cell_of(me.z, a.g.lo[2], a.g.inv_w, a.g.nz);
    const int x0 = cx > 1 ? cx - 2 : 0, x1 = cx + 2 < a.g.nx ? cx + 2 : a.g.nx - 1;
    // visits every candidate of the stencil: f(j, p) per lane, j < e valid
    auto sweep = [&](auto &&f) {
        for (int dz = -2; dz <= 2; ++dz) {
            const int z = cz + dz;
            if (z < 0 || z >= a.g.nz)
                continue;
            for (int dy = -2; dy <= 2; ++dy) {
                const int y = cy + dy;
                if (y < 0 || y >= a.g.ny)
                    continue;
                const uint64_t row = ((uint64_t)z * a.g.ny + y) * a.g.nx;
                const uint32_t b = a.cellstart[row + x0], e = a.cellstart[row + x1 + 1];
                for (uint32_t j0 = b; j0 < e; j0 += kWave)
                    f(j0 + lane, e);
            }
        }
    };
    // all neighbours within R, any type, self excluded by index
    uint32_t ntot = 0;
    sweep([&](uint32_t j, uint32_t e) {
        if (j < e) {
            const float4 p = a.pos[j];
            const float dxx = me.x - p.x, dyy = me.y - p.y, dzz = me.z - p.z;
            const float d2 = dxx * dxx + dyy * dyy + dzz * dzz;
            ntot += (d2 <= a.R2 && j != ci) ? 1u : 0u;
        }
    });
    for (int o = kWave / 2; o > 0; o >>= 1)
        ntot += __shfl_xor(ntot, o, kWave);
    const uint32_t K = a.K;
    if (ntot <= K)
        return;
    // the env neighbour's key, when candidate j is one
    auto env_key = [&](uint32_t j, const uint32_t *env, uint4 &key, int &bucket) {
        const float4 p = a.pos[j];
        const uint32_t pt = __float_as_uint(p.w);
        const float dxx = me.x - p.x, dyy = me.y - p.y, dzz = me.z - p.z;
        const float d2 = dxx * dxx + dyy * dyy + dzz * dzz;
        if (!(d2 <= a.R2) || j == ci || !((env[pt >> 5] >> (pt & 31)) & 1u))
            return false;
        // the difference vector neighbour - center and its length by the
        // reference's expression (sqrt of a float)
        const float u = p.x - me.x, v = p.y - me.y, w = p.z - me.z;
        const float dd = __builtin_sqrtf(u * u + v * v + w * w);
        key = make_uint4(__float_as_uint(dd), ordf(u), ordf(v), ordf(w));
        const int bk = (int)(dd * a.inv_bucket);
        bucket = bk < kNpcBuckets - 1 ? bk : kNpcBuckets - 1;
        return true;
    };
    for (int t = 0; t < a.ntasks; ++t) {
        if (!((a.trg[t * 8 + (mt >> 5)] >> (mt & 31)) & 1u))
            continue;
        const uint32_t *env = a.env + t * 8;
        // pass 1: env neighbours per distance bucket
        for (int b = lane; b < kNpcBuckets; b += kWave)
            hist[b] = 0;
        __builtin_amdgcn_wave_barrier();
        uint32_t nenv = 0;
        sweep([&](uint32_t j, uint32_t e) {
            uint4 key;
            int bk;
            if (j < e && env_key(j, env, key, bk)) {
                __hip_atomic_fetch_add(&hist[bk], 1u, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                ++nenv;
            }
        });
        for (int o = kWave / 2; o > 0; o >>= 1)
            nenv += __shfl_xor(nenv, o, kWave);
        __builtin_amdgcn_wave_barrier();
        if (nenv <= K)
            continue;
        // the bucket of the K-th: lane l sums buckets [l * per, (l + 1) * per)
        constexpr int per = kNpcBuckets / kWave;
        uint32_t mine = 0;
        for (int q = 0; q < per; ++q)
            mine += hist[lane * per + q];
        uint32_t incl = mine;
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o, kWave);
            if (lane >= o)
                incl += v;
        }
        const uint32_t excl = incl - mine;
        // the lane whose range holds the K-th finds the bucket
        int bsel = -1;
        uint32_t below = 0;
        if (excl <= K && K < incl) {
            uint32_t c = excl;
            for (int q = 0; q < per; ++q) {
                const uint32_t h = hist[lane * per + q];
                if (K < c + h) {
                    bsel = lane * per + q;
                    below = c;
                    break;
                }
                c += h;
            }
        }
        const uint64_t own = __builtin_amdgcn_ballot_w64(bsel >= 0);
        if (own == 0)  // cannot happen: more than K env neighbours were counted
            continue;
        const int src = (int)__builtin_ctzll(own);
        bsel = __shfl(bsel, src, kWave);
        below = __shfl(below, src, kWave);
        const uint32_t inb = hist[bsel];
        uint4 pick;
        if (inb <= (uint32_t)kNpcCap) {
            // pass 2: the selected bucket's neighbours, ranked
            uint32_t cnt = 0;
            sweep([&](uint32_t j, uint32_t e) {
                uint4 key;
                int bk = -1;
                const bool take = j < e && env_key(j, env, key, bk) && bk == bsel;
                const uint64_t m = __builtin_amdgcn_ballot_w64(take);
                if (take)
                    buf[cnt + (uint32_t)__builtin_popcountll(m & ((1ull << lane) - 1ull))] = key;
                cnt += (uint32_t)__builtin_popcountll(m);
            });
            __builtin_amdgcn_wave_barrier();
            const uint32_t want = K - below;
            int found = -1;
            for (uint32_t q = lane; q < cnt; q += kWave) {
                const uint4 kq = buf[q];
                uint32_t rank = 0;
                for (uint32_t r = 0; r < cnt; ++r) {
                    const uint4 kr = buf[r];
                    rank += (key_less(kr, kq) || (r < q && !key_less(kq, kr))) ? 1u : 0u;
                }
                if (rank == want)
                    found = (int)q;
            }
            const uint64_t fm = __builtin_amdgcn_ballot_w64(found >= 0);
            if (fm == 0)  // cannot happen: the ranks are a permutation of [0, cnt)
                continue;
            const int q = __shfl(found, (int)__builtin_ctzll(fm), kWave);
            pick = buf[q];
        } else {
            // a crowded bucket: keep the K + 1 smallest in the buffer
            uint32_t cnt = 0;
            sweep([&](uint32_t j, uint32_t e) {
                uint4 key = make_uint4(0, 0, 0, 0);
                int bk;
                const bool take = j < e && env_key(j, env, key, bk);
                const uint64_t m = __builtin_amdgcn_ballot_w64(take);
                const uint32_t add = (uint32_t)__builtin_popcountll(m);
                if (add == 0)
                    return;
                if (cnt + add > (uint32_t)kNpcCap) {
                    wave_sort(buf, cnt, lane);
                    cnt = cnt < K + 1 ? cnt : K + 1;
                }
                if (take)
                    buf[cnt + (uint32_t)__builtin_popcountll(m & ((1ull << lane) - 1ull))] = key;
                cnt += add;
                __builtin_amdgcn_wave_barrier();
            });
            wave_sort(buf, cnt, lane);
            pick = buf[K];
        }
        if (lane == 0 && __uint_as_float(pick.x) <= a.maxlen) {
            // npc3d::add: t = u * (1 / width) in float, rounded half away
            // from zero through a double, plus the centre offset
            int ix[3];
            const uint32_t comp[3] = {pick.y, pick.z, pick.w};
            for (int q = 0; q < 3; ++q) {
                const float th = unordf(comp[q]) * a.inv_width;
                const int disc = (th >= 0.f) ? (int)((double)th + 0.5) : (int)((double)th - 0.5);
                ix[q] = disc + a.half;
            }
            const uint64_t bin =
                (uint64_t)ix[0] + (uint64_t)ix[1] * a.edge + (uint64_t)ix[2] * a.edge * a.edge;
            atomicAdd(&a.counts[(uint64_t)t * a.nxyz + bin], 1u);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

hipError_t launch_npcorr(const NpcArgs &a, hipStream_t s)
{
    if (a.n == 0)
        return hipSuccess;
    const uint64_t blocks = (a.n + kWaves - 1) / kWaves;
    if (blocks > 0x7FFFFFFFull)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_npcorr, dim3((unsigned)blocks), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

} // namespace pp
