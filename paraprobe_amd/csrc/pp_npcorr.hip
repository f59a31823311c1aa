// pp_npcorr.hip -- gfx950 kernel of the n-point correlation family (NPCORR,
// AnalysisSpatDistrType digit 5) of horderdist::compute_generic_spatstat3.
//
// Replaces the NPCORR branch of the center loop (src/PARAPROBE_SolverHdl.cpp:
// 3899-3960) and npc3d::add (src/PARAPROBE_NPointCorrelations3D.cpp:165-207):
// for every considered center (ion2surf >= fl32(R*R)) and every task whose
// central types hold its type, when more than K ions of the task's
// neighbour (env) types lie within R (self excluded; more than K of any type
// then hold as well), those neighbours are ordered by
// d = sqrtf(fl((u*u + v*v) + w*w)), (u, v, w) = neighbour - center; the
// K-th (0-based) one's difference vector is binned into a cube of
// (2h + 1)^3 cells of width `width` centred on the center, each component
// rounded half away from zero in double ((int)(t + 0.5) / (int)(t - 0.5) of
// t = fl(u * fl(1 / width))), when d <= maxlen.  The reference picks the K-th
// with nth_element over its KD-tree order, so equal distances leave the pick
// to that order; here equal d are ordered by the squared length d2 (of which
// d is a monotone function), then by the smaller (u, v, w),
// lexicographically, numerically.
//
// One wavefront per center.  Each task reads its own candidate list: the
// ions of its neighbour types, cell-sorted on a grid at least R/2 wide, so
// that the 5 x 5 x 5 cells around a center hold its ball; the stencil rows
// are cut to the part a sphere of the needed radius can reach.  A first sweep
// counts the candidates per distance bucket (kNpcBuckets buckets, uniform
// in d2) within a radius r1 that holds the K-th at the list's mean
// density with a margin, or within R when r1 holds too few; the bucket that
// holds the K-th
// is found by a scan, its candidates are gathered in one more sweep, limited
// to the bucket's spherical shell (16 bytes each: d2 bits, then u, v, w as
// order-preserving integers), and ranked; buckets are monotone in the key, so every
// candidate of a lower bucket precedes them.  When the bucket holds more
// than kNpcCap candidates (shells of a perfect lattice), the wave falls back
// to a buffer of candidates that is sorted (bitonic) and cut to the K + 1
// smallest whenever it fills up.  K must be below kNpcCap - 64.
#include "pp_launch.h"

#pragma clang fp contract(off)

namespace pp {

namespace {

typedef float nf4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const nf4 gnf4;
typedef __attribute__((address_space(1))) const uint32_t gcu32;

// 16-byte record load through the global address space: a flat load also
// counts against the LDS wait counter, so the bucket counters' LDS traffic
// would wait for the next step's records in flight
__device__ inline float4 rec_at(const gnf4 *p, uint32_t i)
{
    const nf4 v = p[i];
    return make_float4(v.x, v.y, v.z, v.w);
}

__device__ inline uint32_t ordf(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ inline float unordf(uint32_t o)
{
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

// Inclusive prefix sum over the 64 lanes with DPP moves: shifts by 1, 2, 4, 8
// inside each row of 16 lanes (a lane without a source adds 0), then
// row_bcast:15 and row_bcast:31 carry the row totals forward.
__device__ inline uint32_t wave_scan(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return v;
}

// a value of one lane, the lane index being the same in the whole wave
__device__ inline uint32_t lane_value(uint32_t v, int l)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
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

struct Rows {
    uint32_t b1, e1, b2, e2;  // up to two candidate index ranges of one stencil row
};

__global__ __launch_bounds__(256) void k_npcorr(NpcArgs a)
{
    __shared__ uint4 s_buf[kWaves][kNpcCap];
    __shared__ uint32_t s_hist[kWaves][kNpcBuckets + 1];
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
    const int cy = cell_of(me.y, a.g.lo[1], a.g.inv_w, a.g.ny);
    const int cz = cell_of(me.z, a.g.lo[2], a.g.inv_w, a.g.nz);
    // The candidates of a spherical shell ri <= d <= ro around the center in
    // a list with cell offsets cs: lane r < 25 takes the stencil row (dy, dz)
    // = (r % 5 - 2, r / 5 - 2) and keeps the x-cells where the shell can meet
    // the row -- one run, or two when the inner sphere cuts the row -- as
    // candidate index ranges.  The bounds are exact in double on the cell
    // boundaries lo + c * w (the same expression that bins the ions) and
    // widened by 1e-4 w, far more than the rounding of the FP32 distance
    // tests that decide membership.
    const double px = me.x, py = me.y, pz = me.z, cw = a.g.w, margin = 1e-4 * cw;
    auto rows_for = [&](const gcu32 *cs, double ri, double ro) {
        Rows rw{0u, 0u, 0u, 0u};
        const int y = cy + lane % 5 - 2, z = cz + lane / 5 - 2;
        if (lane >= 25 || y < 0 || y >= a.g.ny || z < 0 || z >= a.g.nz)
            return rw;
        const double ylo = a.g.lo[1] + y * cw, zlo = a.g.lo[2] + z * cw;
        const double hy0 = fmax(0.0, fmax(ylo - py, py - (ylo + cw)));
        const double hz0 = fmax(0.0, fmax(zlo - pz, pz - (zlo + cw)));
        const double hy1 = fmax(py - ylo, ylo + cw - py), hz1 = fmax(pz - zlo, zlo + cw - pz);
        const double hmin2 = hy0 * hy0 + hz0 * hz0, hmax2 = hy1 * hy1 + hz1 * hz1;
        ro += margin;
        ri = ri > margin ? ri - margin : 0.0;
        if (hmin2 > ro * ro)
            return rw;
        const double xo = sqrt(ro * ro - hmin2);
        const double xi2 = ri * ri - hmax2;
        const double xi = xi2 > 0.0 ? sqrt(xi2) : 0.0;
        auto cell = [&](double v) {
            const double t = floor((v - a.g.lo[0]) * a.g.inv_w);
            return t < 0.0 ? 0 : (t >= (double)a.g.nx ? a.g.nx - 1 : (int)t);
        };
        const int la = cell(px - xo), lb = cell(px - xi), ra = cell(px + xi), rb = cell(px + xo);
        const uint64_t row = ((uint64_t)z * a.g.ny + y) * a.g.nx;
        if (xi2 <= 0.0 || lb + 1 >= ra) {
            rw.b1 = cs[row + la];
            rw.e1 = cs[row + rb + 1];
        } else {
            rw.b1 = cs[row + la];
            rw.e1 = cs[row + lb + 1];
            rw.b2 = cs[row + ra];
            rw.e2 = cs[row + rb + 1];
        }
        return rw;
    };
    // Visits every candidate of the ranges in list p: f(q, valid) per lane
    // with the candidate's record q (only meaningful when valid).  The ranges
    // go to lanes (lane l < 25: the first range of row l, lane 25 + l: its
    // second) with an exclusive prefix of their lengths.  Long ranges are
    // walked one by one, 64 candidates a step; short ones (sparse type lists)
    // as one concatenation, each lane finding its range by a binary search
    // over the prefix.  The records of the next step are loaded while f runs
    // on the current ones; the loads are unconditional (the last step loads
    // its own records again), so that no branch joins paths with different
    // loads in flight.
    auto sweep = [&](const Rows &rw, const gnf4 *p, auto &&f) {
        const int src = lane >= 25 ? lane - 25 : lane;
        const uint32_t sb2 = __shfl(rw.b2, src, kWave), se2 = __shfl(rw.e2, src, kWave);
        const uint32_t b = lane < 25 ? rw.b1 : (lane < 50 ? sb2 : 0u);
        const uint32_t len = lane < 25 ? rw.e1 - rw.b1 : (lane < 50 ? se2 - sb2 : 0u);
        const uint32_t incl = wave_scan(len);
        const uint32_t excl = incl - len;
        const uint32_t total = lane_value(incl, kWave - 1);
        if (total == 0)
            return;
        uint64_t nonempty = __builtin_amdgcn_ballot_w64(len > 0);
        const uint32_t nne = (uint32_t)__builtin_popcountll(nonempty);
        if (total >= 32u * nne) {
            // cursor: step start cj in the range ending at ce (wave-uniform);
            // the ranges left are the set bits of `nonempty`
            uint32_t cj = 0, ce = 0;
            auto next_range = [&]() {
                if (!nonempty)
                    return false;
                const int nr = (int)__builtin_ctzll(nonempty);
                nonempty &= nonempty - 1;
                cj = __builtin_amdgcn_readlane(b, nr);
                ce = cj + __builtin_amdgcn_readlane(len, nr);
                return true;
            };
            bool has = next_range();
            bool cok = cj + lane < ce;
            float4 cur = rec_at(p, cok ? cj + lane : cj);
            while (has) {
                const float4 q = cur;
                const bool ok = cok;
                const uint32_t last = cj;
                cj += kWave;
                if (cj >= ce)
                    has = next_range();
                cok = has && cj + lane < ce;
                cur = rec_at(p, cok ? cj + lane : (has ? cj : last));
                f(q, ok);
            }
        } else {
            auto at = [&](uint32_t sl) {
                int k = 0;
                for (int st = kWave / 2; st > 0; st >>= 1) {
                    const uint32_t v = __shfl(excl, k + st, kWave);
                    k = v <= sl ? k + st : k;
                }
                return __shfl(b, k, kWave) + (sl - __shfl(excl, k, kWave));
            };
            const uint32_t first = lane_value(b, (int)__builtin_ctzll(nonempty));
            // (the searches run on every lane: a lane shuffle reads nothing
            // from lanes that are switched off)
            bool cok = (uint32_t)lane < total;
            const uint32_t j0 = at((uint32_t)lane);
            float4 cur = rec_at(p, cok ? j0 : first);
            for (uint32_t s0 = 0; s0 < total; s0 += kWave) {
                const float4 q = cur;
                const bool ok = cok;
                const uint32_t sn = s0 + kWave + (uint32_t)lane;
                cok = sn < total;
                const uint32_t jn = at(sn);
                cur = rec_at(p, cok ? jn : first);
                f(q, ok);
            }
        }
    };
    // a neighbour within R of the center (the center itself included when
    // its type is listed): its difference vector and squared length (the
    // same value as fl((u*u + v*v) + w*w), the squares of negated terms)
    auto neighbour = [&](const float4 &q, float &u, float &v, float &w, float &d2) {
        const float dxx = me.x - q.x, dyy = me.y - q.y, dzz = me.z - q.z;
        d2 = dxx * dxx + dyy * dyy + dzz * dzz;
        // the difference vector neighbour - center
        u = q.x - me.x;
        v = q.y - me.y;
        w = q.z - me.z;
        return d2 <= a.R2;
    };
    // its key: d2 bits, then u, v, w order-preserving (the order by d =
    // sqrtf(d2), a monotone function, with equal d ordered by d2 first)
    auto key_of = [&](float u, float v, float w, float d2) {
        return make_uint4(__float_as_uint(d2), ordf(u), ordf(v), ordf(w));
    };
    // its bucket (int)(d2 * inv_b2), monotone in the key, capped at bmax:
    // the last bucket when the buckets span R^2 (full), else the slot past
    // them, which collects what lies beyond and is never selected
    float inv_b2 = a.inv_bucket;
    bool full = true;
    uint32_t bmax = kNpcBuckets - 1;
    auto bucket_of = [&](float d2) {
        const uint32_t bk = __float2uint_rz(d2 * inv_b2);
        return bk < bmax ? bk : bmax;
    };
    const double reach = (double)a.maxlen * (1.0 + 0x1p-12);
    for (int t = 0; t < a.ntasks; ++t) {
        if (!((a.trg[t * 8 + (mt >> 5)] >> (mt & 31)) & 1u))
            continue;
        const gnf4 *ep = (const gnf4 *)a.epos[t];
        const gcu32 *ecs = (const gcu32 *)a.ecs[t];
        if (!ecs)  // no ion of the task's neighbour types
            continue;
        // the center is among its own candidates when its type is listed:
        // its key (d = 0, u = v = w = +0) is the smallest, so the K-th
        // neighbour is the (K + 1)-th candidate (a coincident ion has the
        // same key, so which of them stands for the center is immaterial)
        const bool self = ((a.env[t * 8 + (mt >> 5)] >> (mt & 31)) & 1u) != 0;
        const uint32_t Ke = a.K + (self ? 1u : 0u);
        // count the candidates per distance bucket: first within the radius
        // r1 that holds the K-th at the list's mean density with a margin
        // (buckets of r1^2 / kNpcBuckets in d2), and only when fewer than
        // K + 1 lie there, within R (buckets of R^2 / kNpcBuckets).
        // Candidates beyond r1 have larger keys than every one within it.
        const float r1 = a.r1[t];
        full = !(r1 < a.maxlen);
        inv_b2 = full ? a.inv_bucket : (float)kNpcBuckets / (r1 * r1);
        bmax = full ? kNpcBuckets - 1 : kNpcBuckets;
        // lane l sums buckets [l * per, (l + 1) * per): mine, and the
        // inclusive scan of those sums over the lanes; nenv = candidates
        // in the buckets
        constexpr int per = kNpcBuckets / kWave;
        uint32_t mine = 0, incl = 0, nenv = 0;
        for (;;) {
            for (int q = lane; q <= kNpcBuckets; q += kWave)
                hist[q] = 0;
            __builtin_amdgcn_wave_barrier();
            const Rows ball = rows_for(ecs, 0.0, full ? reach : (double)r1 * (1.0 + 0x1p-12));
            sweep(ball, ep, [&](const float4 &q, bool ok) {
                float u, v, w, d2;
                if (ok && neighbour(q, u, v, w, d2))
                    __hip_atomic_fetch_add(&hist[bucket_of(d2)], 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
            });
            __builtin_amdgcn_wave_barrier();
            mine = 0;
            for (int q = 0; q < per; ++q)
                mine += hist[lane * per + q];
            incl = wave_scan(mine);
            nenv = lane_value(incl, kWave - 1);
            if (nenv > Ke || full)
                break;
            full = true;
            inv_b2 = a.inv_bucket;
            bmax = kNpcBuckets - 1;
        }
        // more than K neighbours of the task's types (S.cpp:3940-3951; more
        // than K of any type then holds as well)
        if (nenv <= Ke)
            continue;
        uint4 pick;
        int bsel = -1;
        uint32_t below = 0, inb = 0;
        {
            // the bucket of the Ke-th
            const uint32_t excl = incl - mine;
            if (excl <= Ke && Ke < incl) {
                uint32_t c = excl;
                for (int q = 0; q < per; ++q) {
                    const uint32_t h = hist[lane * per + q];
                    if (Ke < c + h) {
                        bsel = lane * per + q;
                        below = c;
                        break;
                    }
                    c += h;
                }
            }
            const uint64_t own = __builtin_amdgcn_ballot_w64(bsel >= 0);
            if (own == 0)  // cannot happen: more than Ke candidates were counted
                continue;
            const int src = (int)__builtin_ctzll(own);
            bsel = (int)lane_value((uint32_t)bsel, src);
            below = lane_value(below, src);
            inb = hist[bsel];
        }
        if (inb <= (uint32_t)kNpcCap) {
            // the selected bucket's candidates, ranked; they lie in the shell
            // of the bucket's distances
            const double bw = 1.0 / (double)inv_b2;
            const Rows shell = rows_for(ecs, sqrt(bsel * bw) * (1.0 - 0x1p-12),
                                        (bsel == kNpcBuckets - 1 && full)
                                            ? reach
                                            : sqrt((bsel + 1) * bw) * (1.0 + 0x1p-12));
            uint32_t cnt = 0;
            sweep(shell, ep, [&](const float4 &q, bool ok) {
                float u = 0.f, v = 0.f, w = 0.f, d2 = 0.f;
                bool take = ok && neighbour(q, u, v, w, d2) && bucket_of(d2) == (uint32_t)bsel;
                const uint4 key = key_of(u, v, w, d2);
                const uint64_t m = __builtin_amdgcn_ballot_w64(take);
                if (take)
                    buf[cnt + (uint32_t)__builtin_popcountll(m & ((1ull << lane) - 1ull))] = key;
                cnt += (uint32_t)__builtin_popcountll(m);
            });
            __builtin_amdgcn_wave_barrier();
            const uint32_t want = Ke - below;
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
            const int q = (int)lane_value((uint32_t)found, (int)__builtin_ctzll(fm));
            pick = buf[q];
        } else {
            // a crowded bucket: keep the K + 1 smallest in the buffer; one
            // zero key (the center's own) is left out when the center's type
            // is listed
            const uint32_t K = a.K;
            bool dropped = !self;
            uint32_t cnt = 0;
            const Rows ball = rows_for(ecs, 0.0, reach);
            sweep(ball, ep, [&](const float4 &q, bool ok) {
                float u = 0.f, v = 0.f, w = 0.f, d2 = 0.f;
                bool take = ok && neighbour(q, u, v, w, d2);
                const uint4 key = key_of(u, v, w, d2);
                if (!dropped) {
                    const uint64_t zm = __builtin_amdgcn_ballot_w64(
                        take && key.x == 0u && key.y == 0x80000000u && key.z == 0x80000000u &&
                        key.w == 0x80000000u);
                    if (zm) {
                        if (lane == (int)__builtin_ctzll(zm))
                            take = false;
                        dropped = true;
                    }
                }
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
        // binned when d = sqrtf(d2) <= maxlen (npc3d::add)
        if (lane == 0 && __builtin_sqrtf(__uint_as_float(pick.x)) <= a.maxlen) {
            // npc3d::add: t = u * (1 / width) in float, rounded half away
            // from zero through a double, plus the centre offset
            int ix[3];
            const uint32_t comp[3] = {pick.y, pick.z, pick.w};
            for (int q = 0; q < 3; ++q) {
                const float th = unordf(comp[q]) * a.inv_width;
                const int disc = (th >= 0.f) ? (int)((double)th + 0.5) : (int)((double)th - 0.5);
                ix[q] = disc + a.half;
            }
            const uint64_t bin = (uint64_t)ix[0] + (uint64_t)ix[1] * a.edge +
                                 (uint64_t)ix[2] * a.edge * a.edge;
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
