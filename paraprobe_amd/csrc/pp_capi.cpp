// pp_capi.cpp -- host runtime behind include/pp_spatstat.h.
//
// Owns the device, the resident clouds and their cell lists, turns a pp_job
// (the settings and tasks of horderdist::compute_generic_spatstat3,
// src/PARAPROBE_SolverHdl.cpp:3685-4296) into device tables, runs the cell
// list, center, tile and query kernels, and composes the reference's
// histograms from the device counters.  Work runs on "slots" (a HIP stream
// with its own scratch memory): the context's main slot for the synchronous
// entry points, one slot per worker of a serving loop.
#include "../../include/pp_spatstat.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "pp_binning.h"
#include "pp_launch.h"

using namespace pp;

namespace {

constexpr double kEps = 1.0e-6;  // src/PARAPROBE_Numerics.h:45

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct Carver {
    char *base;
    size_t off = 0;
    explicit Carver(void *b) : base((char *)b) {}
    template <typename T>
    T *take(size_t count)
    {
        off = (off + 255) & ~(size_t)255;
        T *r = (T *)(base ? base + off : nullptr);
        off += count * sizeof(T);
        return r;
    }
};

struct CellList {
    GridGeom g{};
    float w = 0;                    // requested width (the grid may widen it)
    uint32_t *cellstart = nullptr;  // ncells + 1
    float4 *pos = nullptr;          // n
    float *s2 = nullptr;            // n
    uint2 *xsplit = nullptr;        // ncells x-part starts (xbits > 0)
    int xbits = 0;                  // x parts per cell: 2^xbits (0: none)
    int ay = 1;                     // requested y/z divisions of the x width
    // a list of one type for symmetric pair counting: pos[].w holds the
    // ion's ion2surf, and cellmin the least ion2surf per cell
    int w_s2 = 0;
    float *cellmin = nullptr;
    void *mem = nullptr;
    size_t bytes = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    DevBuf scratch;  // cell-list temporaries, job arenas
    DevBuf jobbuf;   // JobDev array + tile pointer table + tile offsets + queue
    DevBuf cnbuf;    // CNTNB counters
    DevBuf kbuf;     // k-th neighbour categories of one launch chunk
    hipEvent_t ev[6] = {};
    // event pairs around the query-kernel launches of the current call
    std::vector<hipEvent_t> qev;
    size_t nq = 0;
    // a slot whose buffers are carved from a fixed device arena (the
    // no-allocation entry point): ensure() never allocates there
    char *fixed = nullptr;
    size_t fixed_cap = 0, fixed_off = 0;
    const DevBuf *fixed_last = nullptr;
    // instantiated graphs of the launch groups of the slot's last runs, by
    // position in the run: the next run's capture updates them in place
    // (same kernels and chunk count) instead of instantiating anew
    std::vector<hipGraphExec_t> gexec;
};

// room the k-th neighbour buffer of a fixed-arena slot leaves for the launch
// descriptors of its group
constexpr size_t kFixedJobbufReserve = size_t(1) << 20;

// a query-kernel shape: centers per tile and cell widths
struct Shape {
    int T = 0;   // 0: choose
    int r = 0;   // 0: choose
    int ay = 0;  // y/z cells per x cell width; 0: choose
};

} // namespace

struct pp_ctx {
    int device = 0;
    int cu_count = 0;
    size_t lds_per_block = 0;
    std::mutex mu;  // err, tuning state
    std::string err;
    int tune_T = 0, tune_r = 0;
    int tune_ay = 0;  // y/z cells per x cell width (0: rule)
    double knn_factor = 4.0;  // kNN range: expected neighbours / k (0: full range)
    int sym_pairs = 1;        // symmetric pair counting (plan_symmetric) allowed
    int graph_launch = 1;     // batch launch groups captured into a HIP graph
    std::map<long, Shape> tuned;  // autotune bucket -> (T, cells per Rmax, y/z per x)
    Slot main;
    Slot noalloc;  // pp_spatstat_run_noalloc: buffers in the fixed device scratch
};

struct pp_cloud {
    pp_ctx *ctx = nullptr;
    uint64_t n = 0;
    float *xyz = nullptr;     // n x 3
    uint8_t *type = nullptr;  // n
    float *s2 = nullptr;      // n
    void *raw = nullptr;
    size_t raw_bytes = 0;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    bool resident = false;       // cell lists fixed at upload
    std::vector<CellList> lists; // resident widths, or the cached per-job list
    std::mutex mu;               // a non-resident cloud serves one call at a time
    uint64_t type_count[256] = {};
    // typed candidate lists: the ions of a type set, on the grid of one of
    // the lists above (so the center tiles of that list index them directly)
    struct SubList {
        uint32_t mask[8];
        int w_s2;
        const CellList *base;
        CellList cl;
    };
    std::list<SubList> subs;
    std::mutex sub_mu;
    double occupied = 0;  // nm^3 of the coarse occupancy grid's cells that hold ions
    double density() const
    {
        // ions per nm^3 inside the tip: over the cells that hold ions when
        // measured at upload, else the bounding box taken as two thirds full
        // (cylinders pi/4, frusta ~0.6)
        if (occupied > 0)
            return (double)n / occupied;
        double v = 1.0;
        for (int a = 0; a < 3; ++a)
            v *= std::max(1e-3, (double)hi[a] - (double)lo[a]);
        return (double)n / v * 1.5;
    }
};

namespace {

pp_status fail(pp_ctx *c, pp_status st, const std::string &msg)
{
    if (c) {
        std::lock_guard<std::mutex> lk(c->mu);
        c->err = msg;
    }
    return st;
}

#define PP_HIP(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail((ctx), _e == hipErrorOutOfMemory ? PP_ERR_OOM_DEVICE : PP_ERR_HIP,   \
                        std::string(#expr) + ": " + hipGetErrorString(_e));                  \
    } while (0)

pp_status ensure(pp_ctx *c, Slot &s, DevBuf &b, size_t bytes)
{
    if (b.cap >= bytes && b.p)
        return PP_OK;
    if (s.fixed) {
        // carve from the arena: the buffer carved last grows in place, any
        // other takes a new region (the arena is reset per call)
        const size_t start = (&b == s.fixed_last && b.p)
                                 ? (size_t)((char *)b.p - s.fixed)
                                 : ((s.fixed_off + 255) & ~(size_t)255);
        if (start + bytes > s.fixed_cap) {
            char msg[192];
            std::snprintf(msg, sizeof(msg),
                          "the job needs more than the %zu MiB of the no-allocation scratch "
                          "(at least %zu MiB)",
                          s.fixed_cap >> 20, (start + bytes + (size_t(1) << 20) - 1) >> 20);
            return fail(c, PP_ERR_UNSUPPORTED, msg);
        }
        b.p = s.fixed + start;
        b.cap = bytes;
        s.fixed_off = start + bytes;
        s.fixed_last = &b;
        return PP_OK;
    }
    if (b.p) {
        PP_HIP(c, hipStreamSynchronize(s.stream));
        PP_HIP(c, hipFree(b.p));
        b.p = nullptr;
        b.cap = 0;
    }
    size_t want = bytes + bytes / 8 + (1u << 20);
    PP_HIP(c, hipMalloc(&b.p, want));
    b.cap = want;
    return PP_OK;
}

pp_status slot_init(pp_ctx *c, Slot &s)
{
    PP_HIP(c, hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    for (auto &e : s.ev)
        PP_HIP(c, hipEventCreate(&e));
    return PP_OK;
}

void slot_free(Slot &s)
{
    if (s.stream)
        (void)hipStreamSynchronize(s.stream);
    for (DevBuf *b : {&s.scratch, &s.jobbuf, &s.cnbuf, &s.kbuf})
        if (b->p) {
            if (!s.fixed)
                (void)hipFree(b->p);
            b->p = nullptr;
            b->cap = 0;
        }
    for (auto &e : s.ev)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
    for (hipEvent_t e : s.qev)
        (void)hipEventDestroy(e);
    s.qev.clear();
    s.nq = 0;
    for (hipGraphExec_t e : s.gexec)
        if (e)
            (void)hipGraphExecDestroy(e);
    s.gexec.clear();
    if (s.stream) {
        (void)hipStreamDestroy(s.stream);
        s.stream = nullptr;
    }
}

inline bool has_type(const uint8_t *mask32, int t) { return (mask32[t >> 3] >> (t & 7)) & 1; }

// The host-side view of one job.
struct JobPlan {
    HistGeometry hg;
    EdgeTable et;
    int C = 0;
    float R = 0, R2 = 0;
    int E = 0, ntrg = 0;
    int32_t envoff[256];
    int32_t trgslot[256];
    int env_type[kMaxEnv];
    int trg_type[kMaxTrg];
    int ntasks = 0;
    std::vector<uint32_t> task_trg, task_env;
    uint32_t families = 0;
    bool percenter = false;
    std::vector<int64_t> kdesc;     // as given (k - 1, descending)
    std::vector<int32_t> kidx_of_k; // kmax + 1
    std::vector<int32_t> kth_to_u;  // kdesc entry -> unique index
    int kmax = 0, nku = 0;
    int cat0 = 0, cat_sentinel = 0;
    int r = 0;          // cells per Rmax of this job's cell list (along x)
    int ay = 1;         // y and z cells per x cell width
    // query mode: 0 pair counts only, 1 full-range per-center histograms,
    // 2 hybrid (kNN-range per-center histograms + rerun of short centers)
    int mode = 0;
    int HK = 0, trunc_any = 0;
    std::vector<int32_t> kinfo;      // [256][2]: kNN-range offset, threshold bits
    std::vector<int32_t> ckoff;      // [kMaxEnv]
    std::vector<uint32_t> task_ck;   // [ntasks]
    int lmax = 1;                    // most tasks sharing one center type
    std::vector<uint32_t> trg_ntask; // [ntrg]
    std::vector<uint32_t> trg_tasks; // [ntrg][lmax]
    // uint16 planes of k-th neighbour categories per center and task: the
    // MKNN orders, then the 1NN
    int nplanes() const
    {
        return ((families & PP_FAM_MKNN) ? nku : 0) + ((families & PP_FAM_1NN) ? 1 : 0);
    }
    // k -> order index entries the query kernel keeps in LDS: none when the
    // orders are k = 1 .. kmax (the index is k - 1)
    int kidx_len() const { return kmax > 0 && nku == kmax ? 0 : kmax; }
    // planes per tile in the buffer: a multiple of 8, so that tile blocks and
    // groups of planes start on cache lines
    int npad() const { return (nplanes() + 7) & ~7; }
    // planes per reduction work unit: as many (8, 4, 2, 1) as the unit's
    // 16-bit LDS counters allow; 8 when even one plane's do not fit (the
    // unit then adds straight to global memory)
    int kb_group() const
    {
        const int extra = ntrg * lmax;  // the reduction's task table
        int g = 8;
        while (g > 1 && (ntasks * g * C + 1) / 2 + extra > kReduceLdsWords)
            g /= 2;
        return (ntasks * g * C + 1) / 2 + extra > kReduceLdsWords ? 8 : g;
    }
};

pp_status plan_job(pp_ctx *c, const pp_job *j, JobPlan &p)
{
    if (!j)
        return fail(c, PP_ERR_INVALID_ARG, "job is NULL");
    if (j->families & PP_FAM_NPCORR)
        return fail(c, PP_ERR_UNSUPPORTED, "NPCORR runs through pp_npcorr_run");
    const uint32_t fams = j->families & (PP_FAM_RDF | PP_FAM_1NN | PP_FAM_RIPK | PP_FAM_MKNN |
                                         PP_FAM_CNTNB);
    if (fams == 0)
        return fail(c, PP_ERR_INVALID_ARG, "No spatial statistics task to do");
    // Settings::checkUserInput, src/PARAPROBE_Settings.cpp:871-885
    if (!(j->rmin >= 0.f))
        return fail(c, PP_ERR_INVALID_ARG, "SpatStatRadiusMin must be positive or zero");
    if (!(j->rincr >= kEps))
        return fail(c, PP_ERR_INVALID_ARG, "SpatStatRadiusIncr must be positive");
    if (!(j->rmax >= kEps))
        return fail(c, PP_ERR_INVALID_ARG, "SpatStatRadiusMax must be positive");
    if (!(j->rmin < j->rmax))
        return fail(c, PP_ERR_INVALID_ARG, "SpatStatRadiusMin is larger than SpatStatRadiusMax");
    if (j->ntasks < 0 || j->ntasks > kMaxTasks)
        return fail(c, PP_ERR_INVALID_ARG, "ntasks out of range (0..64)");
    if (j->ntasks > 0 && (!j->trg_mask || !j->env_mask))
        return fail(c, PP_ERR_INVALID_ARG, "task masks are NULL");
    p.hg = HistGeometry(j->rmin, j->rincr, j->rmax);
    std::string why;
    if (!p.et.build(p.hg, &why))
        return fail(c, PP_ERR_UNSUPPORTED, why);
    if (!p.et.exact_up_to(p.hg, j->rmax))
        return fail(c, PP_ERR_UNSUPPORTED,
                    "bins too fine for the float index coordinate up to Rmax");
    p.C = (int)p.hg.ncat();
    p.R = j->rmax;
    p.R2 = p.R * p.R;
    p.families = fams;
    p.ntasks = j->ntasks;
    // type slots
    for (int t = 0; t < 256; ++t) {
        p.envoff[t] = -1;
        p.trgslot[t] = -1;
    }
    for (int t = 0; t < 256; ++t) {
        bool isenv = false, istrg = false;
        for (int k = 0; k < j->ntasks; ++k) {
            isenv |= has_type(j->env_mask + 32 * k, t);
            istrg |= has_type(j->trg_mask + 32 * k, t);
        }
        if (isenv) {
            if (p.E >= kMaxEnv)
                return fail(c, PP_ERR_UNSUPPORTED, "more than 32 distinct neighbour types");
            p.env_type[p.E] = t;
            p.envoff[t] = p.E * p.C;
            ++p.E;
        }
        if (istrg) {
            if (p.ntrg >= kMaxTrg)
                return fail(c, PP_ERR_UNSUPPORTED, "more than 32 distinct central types");
            p.trg_type[p.ntrg] = t;
            p.trgslot[t] = p.ntrg;
            ++p.ntrg;
        }
    }
    p.task_trg.assign(p.ntasks, 0u);
    p.task_env.assign(p.ntasks, 0u);
    for (int k = 0; k < p.ntasks; ++k) {
        for (int s = 0; s < p.ntrg; ++s)
            if (has_type(j->trg_mask + 32 * k, p.trg_type[s]))
                p.task_trg[k] |= 1u << s;
        for (int s = 0; s < p.E; ++s)
            if (has_type(j->env_mask + 32 * k, p.env_type[s]))
                p.task_env[k] |= 1u << s;
    }
    // MKNN k list (S.cpp:3739-3767)
    p.kmax = 0;
    p.nku = 0;
    if (fams & PP_FAM_MKNN) {
        if (j->nk < 0 || (j->nk > 0 && !j->kminus1_desc))
            return fail(c, PP_ERR_INVALID_ARG, "bad MKNN k list");
        p.kdesc.assign(j->kminus1_desc, j->kminus1_desc + j->nk);
        for (int64_t v : p.kdesc) {
            if (v < 0)
                return fail(c, PP_ERR_INVALID_ARG, "Instructing a kth order value smaller than 1");
            if (v + 1 > kMaxKmax)
                return fail(c, PP_ERR_UNSUPPORTED, "k larger than 4096");
            p.kmax = std::max(p.kmax, (int)v + 1);
        }
        p.kidx_of_k.assign(p.kmax + 1, -1);
        std::vector<int> uk;
        for (int64_t v : p.kdesc)
            uk.push_back((int)v + 1);
        std::sort(uk.begin(), uk.end());
        uk.erase(std::unique(uk.begin(), uk.end()), uk.end());
        for (size_t u = 0; u < uk.size(); ++u)
            p.kidx_of_k[uk[u]] = (int32_t)u;
        p.nku = (int)uk.size();
        p.kth_to_u.resize(p.kdesc.size());
        for (size_t q = 0; q < p.kdesc.size(); ++q)
            p.kth_to_u[q] = p.kidx_of_k[p.kdesc[q] + 1];
    }
    p.percenter = (fams & (PP_FAM_1NN | PP_FAM_CNTNB)) || ((fams & PP_FAM_MKNN) && p.nku > 0);
    p.trg_ntask.assign(std::max(p.ntrg, 1), 0u);
    p.lmax = 1;
    for (int s = 0; s < p.ntrg; ++s) {
        for (int k = 0; k < p.ntasks; ++k)
            p.trg_ntask[s] += (p.task_trg[k] >> s) & 1u;
        p.lmax = std::max(p.lmax, (int)p.trg_ntask[s]);
    }
    p.trg_tasks.assign((size_t)std::max(p.ntrg, 1) * p.lmax, 0u);
    for (int s = 0; s < p.ntrg; ++s) {
        int l = 0;
        for (int k = 0; k < p.ntasks; ++k)
            if ((p.task_trg[k] >> s) & 1u)
                p.trg_tasks[(size_t)s * p.lmax + l++] = (uint32_t)k;
    }
    // k-th neighbour categories travel as uint16, kKbEmpty reserved
    if (p.nplanes() > 0 && p.C >= (int)kKbEmpty)
        return fail(c, PP_ERR_UNSUPPORTED, "kNN histograms of more than 65533 bins");
    p.cat0 = (int)p.hg.category(0.0f);
    p.cat_sentinel = (int)p.hg.category((float)((double)p.R + kEps));
    return PP_OK;
}

// Stencil rows for cell width w and radius R: cell offset (dx, dy, dz) is
// visited iff the closest points of the two cells are within R, with a
// relative margin of 2^-20 on R^2 that covers the FP32 rounding of the
// in-ball test (at most ~6e-7 relative).
std::vector<Row> make_rows(double w, double wyz, double R)
{
    const double lim = R * R * (1.0 + std::ldexp(1.0, -20));
    const int dmax = (int)std::floor(std::sqrt(lim) / wyz) + 1;
    std::vector<Row> rows;
    for (int dz = -dmax; dz <= dmax; ++dz)
        for (int dy = -dmax; dy <= dmax; ++dy) {
            const double my = std::max(0, std::abs(dy) - 1), mz = std::max(0, std::abs(dz) - 1);
            const double rem = lim - (my * my + mz * mz) * wyz * wyz;
            if (rem < 0)
                continue;
            const int mx = (int)std::floor(std::sqrt(rem) / w);
            Row r;
            r.dy = (int16_t)dy;
            r.dz = (int16_t)dz;
            r.dxm = (int16_t)(mx + 1);
            r.pad = 0;
            rows.push_back(r);
        }
    return rows;
}


// Hybrid mode for kNN jobs: per neighbour type e a range rho_e in which the
// k-th neighbour is expected with margin (factor * k neighbours expected at
// the cloud's mean density), as a number of complete categories ck_e.  Pairs
// within it also enter the center's own histogram; RDF pairs go straight to
// the per-type counts.  Without CNTNB only; falls back to full-range
// per-center histograms when the ranges save little.
void plan_hybrid(const pp_cloud *cl, JobPlan &p, double factor)
{
    p.mode = p.percenter ? 1 : 0;
    p.HK = 0;
    p.trunc_any = 0;
    p.kinfo.assign(512, 0);
    const float mone = -1.0f;
    int32_t mone_bits;
    std::memcpy(&mone_bits, &mone, 4);
    for (int t = 0; t < 256; ++t) {
        p.kinfo[2 * t] = -1;
        p.kinfo[2 * t + 1] = mone_bits;
    }
    p.ckoff.assign(kMaxEnv, 0);
    p.task_ck.assign(std::max(p.ntasks, 1), 0u);
    const bool mknn = (p.families & PP_FAM_MKNN) && p.nku > 0;
    const bool knn = mknn || (p.families & PP_FAM_1NN);
    if (!p.percenter || !knn || (p.families & PP_FAM_CNTNB) || !(factor > 0) || cl->n == 0 ||
        p.E == 0)
        return;
    const double need = mknn ? (double)p.kmax : 1.0;
    // neighbours expected in the range: factor x max(k, (k + 6 sqrt(k) + 10) / 4)
    // -- at the default factor 4 at least k + 6 sqrt(k) + 10, so that for small
    // k the count's spread (about the square root of its mean) rarely leaves a
    // center short of k and re-run over the full range
    const double expect = factor * std::max(need, (need + 6.0 * std::sqrt(need) + 10.0) / 4.0);
    const double dens = cl->density();
    const int catR = (int)p.hg.category_d2(p.R2);
    std::vector<double> rho(p.E, 0.0);
    for (int tk = 0; tk < p.ntasks; ++tk) {
        if (!p.task_trg[tk])
            continue;
        double dS = 0;
        for (int e = 0; e < p.E; ++e)
            if ((p.task_env[tk] >> e) & 1u)
                dS += dens * (double)cl->type_count[p.env_type[e]] / (double)cl->n;
        const double rS = dS > 0 ? std::cbrt(expect / (4.18879020478639 * dS)) : (double)p.R;
        for (int e = 0; e < p.E; ++e)
            if ((p.task_env[tk] >> e) & 1u)
                rho[e] = std::max(rho[e], rS);
    }
    std::vector<int> ck(p.E);
    std::vector<char> tr(p.E, 0);
    int HK = 0;
    for (int e = 0; e < p.E; ++e) {
        ck[e] = catR + 1;
        if (rho[e] > 0 && rho[e] < (double)p.R) {
            const float r2 = (float)(rho[e] * rho[e]);
            const int cc = (int)p.hg.category_d2(r2);
            if (cc >= 2 && cc <= catR) {
                ck[e] = cc;
                tr[e] = 1;
            }
        }
        HK += ck[e];
    }
    if ((double)HK > 0.75 * (double)p.E * (double)p.C)
        return;
    int off = 0;
    for (int e = 0; e < p.E; ++e) {
        p.ckoff[e] = off;
        const int t = p.env_type[e];
        p.kinfo[2 * t] = off;
        // truncated: d2 <= the float below edge ck_e, i.e. category < ck_e
        const uint32_t thr = tr[e] ? p.et.edges[ck[e]] - 1u : [&] {
            uint32_t b;
            std::memcpy(&b, &p.R2, 4);
            return b;
        }();
        p.kinfo[2 * t + 1] = (int32_t)thr;
        off += ck[e];
    }
    for (int tk = 0; tk < p.ntasks; ++tk) {
        int m = catR + 1;
        bool t = false;
        for (int e = 0; e < p.E; ++e)
            if ((p.task_env[tk] >> e) & 1u) {
                m = std::min(m, ck[e]);
                t = t || tr[e];
            }
        p.task_ck[tk] = (uint32_t)m | (t ? 0x80000000u : 0u);
        p.trunc_any |= t ? 1 : 0;
    }
    p.HK = HK;
    p.mode = 2;
}

// Cell list of n ions (device arrays xyz, type, s2) on grid g: key and rank
// per ion, exclusive scan of the cell counts, scatter into cell-sorted float4.
pp_status build_on_grid(pp_ctx *c, Slot &s, const float *xyz, const uint8_t *type,
                        const float *s2, uint64_t n, const GridGeom &g, float w, CellList &L,
                        size_t extra_scratch_bytes = 0, int w_s2 = 0)
{
    const uint64_t nc = g.ncells();
    // each cell's ions by sub-cell: x eighths (or halves), then y and z
    // halves (tiles of consecutive centers span a small box, and the stencil
    // rows leave out the x parts of their end cells beyond the balls),
    // unless the sub-cell offsets would be large next to the cloud
    auto fits = [&](int b) {
        return (nc << b) <= std::max<uint64_t>(4 * n, (uint64_t)1 << 22) &&
               (nc << b) < 0xFFFFFFF0ull;
    };
    const int sub = fits(5) ? 5 : (fits(3) ? 3 : 0);
    Carver m0(nullptr);
    m0.take<uint32_t>(nc + 1);
    m0.take<float4>(n);
    m0.take<float>(n);
    if (sub)
        m0.take<uint2>(nc);
    if (w_s2)
        m0.take<float>(nc);
    const size_t bytes = m0.off + 256;
    if (L.mem && L.bytes < bytes) {
        PP_HIP(c, hipStreamSynchronize(s.stream));
        PP_HIP(c, hipFree(L.mem));
        L.mem = nullptr;
        L.bytes = 0;
    }
    if (!L.mem) {
        PP_HIP(c, hipMalloc(&L.mem, bytes));
        L.bytes = bytes;
    }
    Carver m(L.mem);
    L.cellstart = m.take<uint32_t>(nc + 1);
    L.pos = m.take<float4>(n);
    L.s2 = m.take<float>(n);
    L.xsplit = sub ? m.take<uint2>(nc) : nullptr;
    L.xbits = sub ? sub - 2 : 0;
    L.w_s2 = w_s2;
    L.cellmin = w_s2 ? m.take<float>(nc) : nullptr;
    L.g = g;
    L.w = w;
    const uint64_t nk = nc << sub;
    // temporaries after the first extra_scratch_bytes of the scratch arena
    Carver t0(nullptr);
    t0.off = extra_scratch_bytes;
    t0.take<uint32_t>(nk);
    t0.take<uint32_t>(n);
    t0.take<uint32_t>(n);
    t0.take<uint32_t>(scan_workspace_words(nk));
    if (sub)
        t0.take<uint32_t>(nk + 1);
    pp_status st = ensure(c, s, s.scratch, t0.off + 256);
    if (st != PP_OK)
        return st;
    Carver t(s.scratch.p);
    t.off = extra_scratch_bytes;
    uint32_t *count = t.take<uint32_t>(nk);
    uint32_t *key = t.take<uint32_t>(n);
    uint32_t *rank = t.take<uint32_t>(n);
    uint32_t *ws = t.take<uint32_t>(scan_workspace_words(nk));
    uint32_t *start = sub ? t.take<uint32_t>(nk + 1) : L.cellstart;
    PP_HIP(c, hipMemsetAsync(count, 0, nk * sizeof(uint32_t), s.stream));
    PP_HIP(c, launch_cellkeys(xyz, n, g, count, key, rank, s.stream, sub));
    PP_HIP(c, scan_exclusive(count, start, nk, ws, s.stream));
    PP_HIP(c, launch_scatter(xyz, type, s2, n, key, rank, start, L.pos, L.s2, s.stream, w_s2));
    if (sub) {
        PP_HIP(c, launch_stride(start, L.cellstart, nc + 1, sub, s.stream));
        PP_HIP(c, launch_xsplit(start, nc, sub, L.xsplit, s.stream));
    }
    if (w_s2)
        PP_HIP(c, launch_cellmin(L.cellstart, L.s2, nc, L.cellmin, s.stream));
    return PP_OK;
}

pp_status build_celllist(pp_ctx *c, Slot &s, pp_cloud *cl, float w, CellList &L, int ay = 1)
{
    const uint64_t n = cl->n;
    GridGeom g{};
    // grow the cell when the grid would be huge compared with the cloud
    double ww = w;
    for (;;) {
        uint64_t tot = 1;
        int d[3];
        for (int a = 0; a < 3; ++a) {
            d[a] = (int)std::floor(((double)cl->hi[a] - (double)cl->lo[a]) / (a ? ww / ay : ww)) +
                   1;
            tot *= (uint64_t)d[a];
        }
        if (tot <= std::max<uint64_t>(8 * n + 1024, 1u << 20) && tot < 0xFFFFFFF0ull) {
            g.nx = d[0];
            g.ny = d[1];
            g.nz = d[2];
            break;
        }
        ww *= 1.25;
    }
    g.w = ww;
    g.inv_w = 1.0 / ww;
    g.wyz = ww / ay;
    g.inv_wyz = 1.0 / g.wyz;
    for (int a = 0; a < 3; ++a)
        g.lo[a] = cl->lo[a];
    const pp_status st = build_on_grid(c, s, cl->xyz, cl->type, cl->s2, n, g, w, L);
    L.ay = ay;
    return st;
}

// The ions of the types in `mask`, cell-sorted on the grid of `base`, into
// out (its own allocation, out.mem).
pp_status build_typed_list(pp_ctx *c, Slot &s, const pp_cloud *cl, const CellList *base,
                           const uint32_t mask[8], CellList &out, int w_s2 = 0)
{
    const uint64_t n = cl->n;
    Carver t0(nullptr);
    t0.take<uint32_t>(8);
    t0.take<uint32_t>(n);
    t0.take<uint32_t>(n + 1);
    t0.take<uint32_t>(scan_workspace_words(n));
    const size_t head = t0.off;
    t0.take<uint32_t>(4);
    pp_status st = ensure(c, s, s.scratch, t0.off + 256);
    if (st != PP_OK)
        return st;
    Carver t(s.scratch.p);
    uint32_t *dmask = t.take<uint32_t>(8);
    uint32_t *flags = t.take<uint32_t>(n);
    uint32_t *pos = t.take<uint32_t>(n + 1);
    uint32_t *ws = t.take<uint32_t>(scan_workspace_words(n));
    PP_HIP(c, hipMemcpyAsync(dmask, mask, 32, hipMemcpyHostToDevice, s.stream));
    PP_HIP(c, launch_typeflags(cl->type, n, dmask, flags, s.stream));
    PP_HIP(c, scan_exclusive(flags, pos, n, ws, s.stream));
    uint32_t m = 0;
    PP_HIP(c, hipMemcpyAsync(&m, pos + n, 4, hipMemcpyDeviceToHost, s.stream));
    PP_HIP(c, hipStreamSynchronize(s.stream));
    // gathered sub-cloud: kept in its own allocation until the list is built
    Carver g0(nullptr);
    g0.take<float>(3 * (size_t)m);
    g0.take<uint8_t>(m);
    g0.take<float>(m);
    void *tmp = nullptr;
    PP_HIP(c, hipMalloc(&tmp, g0.off + 256));
    Carver gm(tmp);
    float *sx = gm.take<float>(3 * (size_t)m);
    uint8_t *sty = gm.take<uint8_t>(m);
    float *ss2 = gm.take<float>(m);
    hipError_t e = launch_gather(cl->xyz, cl->type, cl->s2, flags, pos, n, sx, sty, ss2, s.stream);
    if (e != hipSuccess) {
        (void)hipFree(tmp);
        return fail(c, PP_ERR_HIP, std::string("gather: ") + hipGetErrorString(e));
    }
    st = build_on_grid(c, s, sx, sty, ss2, m, base->g, base->w, out, head, w_s2);
    (void)hipStreamSynchronize(s.stream);
    (void)hipFree(tmp);
    if (st != PP_OK && out.mem) {
        (void)hipFree(out.mem);
        out.mem = nullptr;
    }
    return st;
}

// The typed candidate list of `mask` on the grid of `base` (built once per
// cloud, base, type set and w_s2; the caller holds no lock).
pp_status get_sublist(pp_ctx *c, Slot &s, pp_cloud *cl, const CellList *base,
                      const uint32_t mask[8], const CellList *&out, int w_s2 = 0)
{
    std::lock_guard<std::mutex> lk(cl->sub_mu);
    for (const auto &sb : cl->subs)
        if (sb.base == base && sb.w_s2 == w_s2 && std::memcmp(sb.mask, mask, 32) == 0) {
            out = &sb.cl;
            return PP_OK;
        }
    if (s.fixed)
        return PP_OK;  // no new list without allocating: out keeps the full list
    cl->subs.emplace_back();
    pp_cloud::SubList &sb = cl->subs.back();
    std::memcpy(sb.mask, mask, 32);
    sb.w_s2 = w_s2;
    sb.base = base;
    const pp_status st = build_typed_list(c, s, cl, base, mask, sb.cl, w_s2);
    if (st != PP_OK) {
        cl->subs.pop_back();
        return st;
    }
    out = &sb.cl;
    return PP_OK;
}

// Device-side state of one prepared job.
struct JobRun {
    JobPlan plan;
    const pp_cloud *cloud = nullptr;
    const CellList *cl = nullptr;    // the full list: centers and tiles
    const CellList *cand = nullptr;  // candidates: cl or a typed list on its grid
    std::vector<Row> rows;
    // symmetric pair counting (plan_symmetric): the center type slot and
    // its type, list S (that type alone, ion2surf in w), list O (the other
    // candidates, or none), the rows (own cell, cells before it, cells
    // after it on S, then every row on O) and the own type's kNN reach
    int sym_slot = -1, sym_type = -1;
    const CellList *symS = nullptr, *symO = nullptr;
    std::vector<Row> srows;
    int s_nb1 = 0, s_ns = 0;
    double kreach = -1.0;
    size_t ncells = 0;
    // device pointers: the hybrid mode's rerun marks (per sorted ion) and the
    // tables, carved from the slot's scratch arena; centers and tiles point
    // into the lists of the run (prep_run)
    uint32_t *flags = nullptr, *tables = nullptr;
    const uint32_t *centers = nullptr;
    Tile *tiles = nullptr;
    int T_req = 8;  // centers per tile of this job's tiles
    unsigned long long *P = nullptr, *stats = nullptr;
    uint32_t *K = nullptr, *N1 = nullptr, *CN = nullptr;
    int cn_cap = 0;
    // table offsets (words) inside `tables`
    size_t o_edges = 0, o_envoff = 0, o_trgslot = 0, o_ttrg = 0, o_tenv = 0, o_kidx = 0,
           o_tntask = 0, o_ttasks = 0, o_kinfo = 0, o_ckoff = 0, o_taskck = 0, o_rows = 0,
           o_srows = 0, tables_words = 0;
    // read back after prep: centers, tiles, largest stencil candidate count
    uint32_t h_counts[4] = {0, 0, 0, 0};
    uint32_t considered = 0, rerun = 0;   // centers of the job, of its full-range rerun
    // in-ball pairs counted by the kernel; otherwise every candidate is a
    // neighbour type of the RDF, whose pair counts sum to the same total
    bool count_in = true;
    int T_used = 0, c16_used = 0;
    std::vector<uint32_t> h_tables;       // host staging of the device tables
    JobDev dev{};
};

// Word offsets of a job's device tables (the size in tables_words).
void layout_tables(JobRun &r)
{
    const JobPlan &p = r.plan;
    size_t o = 0;
    r.o_edges = o;
    o += p.C;
    r.o_envoff = o;
    o += 256;
    r.o_trgslot = o;
    o += 256;
    r.o_ttrg = o;
    o += std::max(p.ntasks, 1);
    r.o_tenv = o;
    o += std::max(p.ntasks, 1);
    r.o_kidx = o;
    o += p.kmax + 1;
    r.o_tntask = o;
    o += p.trg_ntask.size();
    r.o_ttasks = o;
    o += p.trg_tasks.size();
    r.o_kinfo = (o + 1) & ~(size_t)1;
    o = r.o_kinfo + 512;
    r.o_ckoff = o;
    o += kMaxEnv;
    r.o_taskck = o;
    o += std::max(p.ntasks, 1);
    r.o_rows = (o + 1) & ~(size_t)1;
    o = r.o_rows + 2 * r.rows.size();
    r.o_srows = o;
    o += 2 * r.srows.size();
    r.tables_words = (o + 63) & ~(size_t)63;
}

// A job's counters (cleared together with every other job's of the run).
void layout_outputs(JobRun &r, Carver &m)
{
    const JobPlan &p = r.plan;
    r.P = m.take<unsigned long long>((size_t)std::max(p.ntrg * p.E * p.C, 1));
    r.stats = m.take<unsigned long long>(8);  // [4..7]: the full-range rerun's (unused)
    r.K = m.take<uint32_t>((size_t)std::max(p.ntasks * p.nku * p.C, 1));
    r.N1 = m.take<uint32_t>((size_t)std::max(p.ntasks * p.C, 1));
    // CN is carved later: its size follows from the stencil bound
}

// The lists one preparation of a run builds (sized for every job of the
// run; a rerun of some of them fits too).
struct PrepArena {
    uint64_t E = 0, Ec = 0, nion = 0;
    int nseg = 0, njobs = 0;
    uint32_t *flags = nullptr, *cpos = nullptr, *centers = nullptr, *ntile = nullptr,
             *toff = nullptr, *ws = nullptr;
    Tile *tiles = nullptr;
    PrepSeg *segs = nullptr;
    int32_t *jseg = nullptr;
    PrepJobStat *js = nullptr;
};

void size_prep(PrepArena &a, const std::vector<JobRun> &runs)
{
    a.E = a.Ec = a.nion = 0;
    a.nseg = 0;
    a.njobs = (int)runs.size();
    for (const JobRun &r : runs) {
        const uint64_t k = (uint64_t)r.plan.ntrg;
        a.E += (uint64_t)r.cloud->n * k;
        a.Ec += (uint64_t)r.ncells * k;
        a.nion += r.cloud->n;
        a.nseg += r.plan.ntrg;
    }
}

void layout_prep(PrepArena &a, Carver &m)
{
    a.flags = m.take<uint32_t>(a.E + 1);
    a.cpos = m.take<uint32_t>(a.E + 1);
    a.centers = m.take<uint32_t>(a.nion + 1);
    a.ntile = m.take<uint32_t>(a.Ec + 1);
    a.toff = m.take<uint32_t>(a.Ec + 1);
    a.ws = m.take<uint32_t>(prep_workspace_words(a.E, a.Ec));
    a.tiles = m.take<Tile>(a.nion + 1);  // at most one tile per center
    a.segs = m.take<PrepSeg>((size_t)std::max(a.nseg, 1));
    a.jseg = m.take<int32_t>((size_t)a.njobs + 1);
    a.js = m.take<PrepJobStat>((size_t)std::max(a.njobs, 1));
}

// Bytes of one job's share of a run's scratch (admission control).
size_t prep_bytes(uint64_t n, uint64_t nc, int ntrg)
{
    const uint64_t k = (uint64_t)std::max(ntrg, 1);
    return (size_t)(n * k * 8 + n * 4 + n * 16 + nc * k * 8 +
                    4 * prep_workspace_words(n * k, nc * k) + 4096);
}

// The host image of a job's device tables.
void stage_tables(const JobRun &r, uint32_t *h)
{
    const JobPlan &p = r.plan;
    std::memcpy(&h[r.o_edges], p.et.edges.data(), p.C * 4);
    std::memcpy(&h[r.o_envoff], p.envoff, 256 * 4);
    std::memcpy(&h[r.o_trgslot], p.trgslot, 256 * 4);
    for (int k = 0; k < p.ntasks; ++k) {
        h[r.o_ttrg + k] = p.task_trg[k];
        h[r.o_tenv + k] = p.task_env[k];
    }
    for (int k = 0; k <= p.kmax; ++k)
        h[r.o_kidx + k] = (uint32_t)(p.kmax > 0 ? p.kidx_of_k[k] : -1);
    std::memcpy(&h[r.o_tntask], p.trg_ntask.data(), p.trg_ntask.size() * 4);
    std::memcpy(&h[r.o_ttasks], p.trg_tasks.data(), p.trg_tasks.size() * 4);
    if (!p.kinfo.empty())
        std::memcpy(&h[r.o_kinfo], p.kinfo.data(), 512 * 4);
    if (!p.ckoff.empty())
        std::memcpy(&h[r.o_ckoff], p.ckoff.data(), kMaxEnv * 4);
    if (!p.task_ck.empty())
        std::memcpy(&h[r.o_taskck], p.task_ck.data(), p.task_ck.size() * 4);
    if (!r.rows.empty())
        std::memcpy(&h[r.o_rows], r.rows.data(), r.rows.size() * sizeof(Row));
    if (!r.srows.empty())
        std::memcpy(&h[r.o_srows], r.srows.data(), r.srows.size() * sizeof(Row));
}

// Center lists and tiles of the jobs `ids` in one pass (pp_prep.hip): per
// job one segment per center type (every tile holds centers of one type),
// selected by the considered rule, or, for the full-range rerun, by the
// job's rerun marks.  Reads back each job's center and tile counts and its
// stencil candidate bound; the lists stay valid until the next prep_run.
pp_status prep_run(pp_ctx *c, Slot &sl, std::vector<JobRun> &runs, const std::vector<int> &ids,
                   PrepArena &a, bool rerun)
{
    const int nj = (int)ids.size();
    if (nj == 0)
        return PP_OK;
    std::vector<PrepSeg> segs;
    std::vector<int32_t> jseg(nj + 1, 0);
    uint64_t E = 0, Ec = 0;
    for (int q = 0; q < nj; ++q) {
        JobRun &r = runs[ids[q]];
        const JobPlan &p = r.plan;
        jseg[q] = (int32_t)segs.size();
        for (int slot = 0; slot < p.ntrg; ++slot) {
            PrepSeg g{};
            g.pos = r.cl->pos;
            g.s2 = r.cl->s2;
            g.src = rerun ? r.flags : nullptr;
            g.trgslot = (const int32_t *)(r.tables + r.o_trgslot);
            g.cellstart = r.cl->cellstart;
            g.rows = (const Row *)(r.tables + r.o_rows);
            g.g = r.cl->g;
            g.ion0 = E;
            g.cell0 = Ec;
            g.n = (uint32_t)r.cloud->n;
            g.nc = (uint32_t)r.ncells;
            g.nrows = (int)r.rows.size();
            g.R2 = p.R2;
            g.slot = slot;
            g.T = r.T_req;
            g.job = q;
            segs.push_back(g);
            E += r.cloud->n;
            Ec += r.ncells;
        }
    }
    jseg[nj] = (int32_t)segs.size();
    if (E > a.E || Ec > a.Ec || (int)segs.size() > a.nseg || nj > a.njobs)
        return fail(c, PP_ERR_HIP, "internal: preparation larger than its arena");
    hipStream_t s = sl.stream;
    if (!segs.empty())
        PP_HIP(c, hipMemcpyAsync(a.segs, segs.data(), segs.size() * sizeof(PrepSeg),
                                 hipMemcpyHostToDevice, s));
    PP_HIP(c, hipMemcpyAsync(a.jseg, jseg.data(), jseg.size() * 4, hipMemcpyHostToDevice, s));
    PrepLaunch L;
    L.segs = a.segs;
    L.nseg = (int)segs.size();
    L.jseg = a.jseg;
    L.njobs = nj;
    L.E = E;
    L.Ec = Ec;
    L.flags = a.flags;
    L.cpos = a.cpos;
    L.centers = a.centers;
    L.ntile = a.ntile;
    L.toff = a.toff;
    L.ws = a.ws;
    L.tiles = a.tiles;
    L.js = a.js;
    std::vector<PrepJobStat> js(nj);
    if (L.nseg > 0) {
        PP_HIP(c, launch_prep(L, s));
        PP_HIP(c, hipMemcpyAsync(js.data(), a.js, nj * sizeof(PrepJobStat), hipMemcpyDeviceToHost,
                                 s));
    }
    PP_HIP(c, hipStreamSynchronize(s));  // segs and jseg are host memory of this frame
    for (int q = 0; q < nj; ++q) {
        JobRun &r = runs[ids[q]];
        r.h_counts[0] = js[q].centers;
        r.h_counts[1] = js[q].tiles;
        r.h_counts[2] = js[q].maxcand;
        r.centers = a.centers;
        r.tiles = a.tiles + js[q].tile0;
    }
    return PP_OK;
}

void fill_jobdev(JobRun &r)
{
    const JobPlan &p = r.plan;
    JobDev &d = r.dev;
    d.pos = r.cand->pos;
    d.cellstart = r.cand->cellstart;
    d.xsplit = r.cand->xsplit;
    d.xbits = r.cand->xbits;
    d.cpos = r.cl->pos;
    d.nx = r.cl->g.nx;
    d.ny = r.cl->g.ny;
    d.nz = r.cl->g.nz;
    for (int a = 0; a < 3; ++a)
        d.glo[a] = r.cl->g.lo[a];
    d.gw = r.cl->g.w;
    d.ginv = r.cl->g.inv_w;
    d.gwyz = r.cl->g.wyz;
    d.reach = std::sqrt((double)p.R2) * (1.0 + std::ldexp(1.0, -12)) + 1e-4 * r.cl->g.w;
    d.nrows = (int)r.rows.size();
    d.rows = (const Row *)(r.tables + r.o_rows);
    d.centers = r.centers;
    d.R2 = p.R2;
    d.C = p.C;
    d.edges = r.tables + r.o_edges;
    d.idx_scale = p.et.idx_scale;
    d.idx_off = p.et.idx_off;
    d.idx_off1 = p.et.idx_off1;
    {
        const float amb = p.et.amb_margin(p.hg, p.R);  // at most 0.5
        d.idx_off1a = (float)(1.0 - p.hg.start * p.hg.tmp + (double)amb);
        d.amb2 = 2.0f * amb;
    }
    d.cat0 = p.cat0;
    d.cat_sentinel = p.cat_sentinel;
    d.E = p.E;
    d.ntrg = p.ntrg;
    d.envoff = (const int32_t *)(r.tables + r.o_envoff);
    d.trgslot = (const int32_t *)(r.tables + r.o_trgslot);
    d.ntasks = p.ntasks;
    d.task_trg = r.tables + r.o_ttrg;
    d.task_env = r.tables + r.o_tenv;
    d.families = p.families;
    d.kmax = p.kmax;
    d.nk = p.nku;
    d.kidx_of_k = (const int32_t *)(r.tables + r.o_kidx);
    d.k_dense = p.kmax > 0 && p.nku == p.kmax ? 1 : 0;
    d.P = r.P;
    d.K = r.K;
    d.N1 = r.N1;
    d.CN = r.CN;
    d.cn_cap = r.cn_cap;
    d.percenter = p.percenter ? 1 : 0;
    d.stats = r.stats;
    d.lmax = p.lmax;
    d.trg_ntask = r.tables + r.o_tntask;
    d.trg_tasks = r.tables + r.o_ttasks;
    d.nplanes = p.nplanes();
    d.npad = p.npad();
    d.kb_group = p.kb_group();
    d.KB = nullptr;  // set per launch chunk
    d.tiles = r.tiles;
    d.tile_base = 0;
    d.kb_tiles = 0;
    d.HK = p.HK;
    d.trunc_any = p.trunc_any;
    d.kinfo = (const int32_t *)(r.tables + r.o_kinfo);
    d.ckoff = (const int32_t *)(r.tables + r.o_ckoff);
    d.task_ck = r.tables + r.o_taskck;
    d.unres = r.flags;  // reused once the center list is built
    d.count_in = r.count_in ? 1 : 0;
    d.sym_slot = r.sym_slot;
    d.sym_type = r.sym_type;
    d.nsrows = (int)r.srows.size();
    d.s_nb1 = r.s_nb1;
    d.s_ns = r.s_ns;
    d.srows = (const Row *)(r.tables + r.o_srows);
    d.kreach = r.kreach;
    d.pos_s = r.symS ? r.symS->pos : nullptr;
    d.cellstart_s = r.symS ? r.symS->cellstart : nullptr;
    d.xsplit_s = r.symS ? r.symS->xsplit : nullptr;
    d.xbits_s = r.symS ? r.symS->xbits : 0;
    d.cellmin_s = r.symS ? r.symS->cellmin : nullptr;
    d.pos_o = r.symO ? r.symO->pos : nullptr;
    d.cellstart_o = r.symO ? r.symO->cellstart : nullptr;
    d.xsplit_o = r.symO ? r.symO->xsplit : nullptr;
    d.xbits_o = r.symO ? r.symO->xbits : 0;
}

// One more than the largest type id among a cloud's ions (the query kernel's
// per-type tables hold that many entries).
int type_span(const pp_cloud *cl)
{
    int t = 255;
    while (t > 0 && cl->type_count[t] == 0)
        --t;
    return t + 1;
}

// Stencil rows a job's launch in kernel mode `mode` lays out in LDS.
int query_rows(const JobRun &r, int mode)
{
    return (int)(mode != 1 && r.sym_slot >= 0 ? r.srows.size() : r.rows.size());
}

// Symmetric pair counting for a job with pair counts (RDF, RIPK) in the
// pair-count or the hybrid mode: the center type t with the most ions among the job's center types
// that are also neighbour types, when t holds at least half of the job's
// candidates.  A pair of two centers of t lands in the same pair-count row
// (center type t, neighbour type t) and, its squared distance being the same
// float either way, in the same category from either side, so the tile of
// the center whose cell comes first in cell order counts it for both (the
// later tile reads it only for its kNN range).  t's tiles then read t's ions
// from a list of t alone (ion2surf in w, the least ion2surf per cell) along
// the own cell, the cells before it and the cells after it, and the other
// candidates from a list without t along every row.  Kept off when the
// longer row table would cost the kernel resident workgroups.
pp_status plan_symmetric(pp_ctx *c, Slot &sl, pp_cloud *cl, JobRun &r)
{
    r.sym_slot = -1;
    r.srows.clear();
    const JobPlan &p = r.plan;
    if ((p.mode != 0 && p.mode != 2) || !(p.families & (PP_FAM_RDF | PP_FAM_RIPK)) ||
        cl->n == 0 || sl.fixed)
        return PP_OK;
    // the candidates' types: every type of the cloud on the full list, the
    // neighbour types on a typed list
    uint32_t cmask[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint64_t ncand = 0;
    for (int t = 0; t < 256; ++t)
        if (cl->type_count[t] > 0 && (r.cand == r.cl || p.envoff[t] >= 0)) {
            cmask[t >> 5] |= 1u << (t & 31);
            ncand += cl->type_count[t];
        }
    int best = -1, bslot = -1;
    for (int s = 0; s < p.ntrg; ++s) {
        const int t = p.trg_type[s];
        if (p.envoff[t] >= 0 && (best < 0 || cl->type_count[t] > cl->type_count[best])) {
            best = t;
            bslot = s;
        }
    }
    if (best < 0 || 2 * cl->type_count[best] < ncand)
        return PP_OK;
    // the row table: own cell, the cells before it (rows of lower z, or of
    // the same z and lower y, and the own row's part below x), the cells
    // after it, then every row on the other candidates
    std::vector<Row> before, after;
    for (const Row &w : r.rows) {
        if (w.dy == 0 && w.dz == 0) {
            Row b = w, a = w;
            b.pad = kRowBefore;
            a.pad = kRowAfter;
            before.push_back(b);
            after.push_back(a);
        } else if (w.dz < 0 || (w.dz == 0 && w.dy < 0)) {
            before.push_back(w);
        } else {
            after.push_back(w);
        }
    }
    uint32_t omask[8];
    bool others = false;
    for (int k = 0; k < 8; ++k) {
        omask[k] = cmask[k] & ~(k == (best >> 5) ? 1u << (best & 31) : 0u);
        others = others || omask[k] != 0;
    }
    const size_t nsr = 1 + before.size() + after.size() + (others ? r.rows.size() : 0);
    if ((int)nsr > kMaxRows)
        return PP_OK;
    {
        // (resident workgroups as the register counts allow: 7 in the
        // pair-count mode, 6 in the hybrid mode)
        const size_t gran = 512, lds_cu = 160 * 1024, cap = p.mode == 2 ? 6 : 7;
        auto wgs = [&](int nrows) {
            const size_t l = query_lds_bytes(r.T_req, 1, p.mode, std::max(p.C, 1),
                                             std::max(p.E, 1), std::max(p.ntrg, 1),
                                             p.kidx_len(), nrows, std::max(p.HK, 1),
                                             type_span(cl));
            return std::min<size_t>(cap, lds_cu / ((l + gran - 1) / gran * gran));
        };
        if (wgs((int)nsr) < wgs((int)r.rows.size()))
            return PP_OK;
    }
    uint32_t smask[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    smask[best >> 5] = 1u << (best & 31);
    const CellList *S = nullptr, *O = nullptr;
    pp_status st = get_sublist(c, sl, cl, r.cl, smask, S, 1);
    if (st != PP_OK)
        return st;
    if (S == r.cl || !S->w_s2)
        return PP_OK;
    if (others) {
        st = get_sublist(c, sl, cl, r.cl, omask, O, 0);
        if (st != PP_OK)
            return st;
        if (O == r.cl)
            return PP_OK;
    }
    r.srows.push_back(Row{0, 0, 0, 0});
    r.srows.insert(r.srows.end(), before.begin(), before.end());
    r.srows.insert(r.srows.end(), after.begin(), after.end());
    r.s_nb1 = 1 + (int)before.size();
    r.s_ns = r.s_nb1 + (int)after.size();
    if (others)
        r.srows.insert(r.srows.end(), r.rows.begin(), r.rows.end());
    r.symS = S;
    r.symO = O;
    r.sym_slot = bslot;
    r.sym_type = best;
    // t's kNN range (d2 <= its threshold), widened like the in-ball reach
    r.kreach = -1.0;
    if (!p.kinfo.empty() && p.kinfo[2 * best] >= 0) {
        float kthr;
        std::memcpy(&kthr, &p.kinfo[2 * best + 1], 4);
        if (kthr >= 0.0f)
            r.kreach = std::sqrt((double)kthr) * (1.0 + std::ldexp(1.0, -12)) + 1e-4 * r.cl->g.w;
    }
    return PP_OK;
}

void compose_outputs(const JobRun &r, const std::vector<unsigned long long> &P,
                     const std::vector<uint32_t> &K, const std::vector<uint32_t> &N1,
                     const std::vector<uint32_t> &CN, const unsigned long long *stats,
                     pp_hist_out *o, bool *overflow)
{
    const JobPlan &p = r.plan;
    const int C = p.C;
    o->nbins = p.hg.nbins;
    o->centers_considered = r.h_counts[0];
    o->ions_discarded = (int64_t)r.cloud->n - (int64_t)r.h_counts[0];
    // the kernel's ballot count includes each considered center's self pair
    // (d = 0, always in its own cell); the reference excludes it by index
    o->pairs_inball = (int64_t)stats[0] - (int64_t)r.h_counts[0];
    if (!r.count_in) {
        o->pairs_inball = 0;
        for (unsigned long long v : P)
            o->pairs_inball += (int64_t)v;
    }
    o->pairs_candidate = (int64_t)stats[1];
    *overflow = stats[2] != 0;
    if (p.families & (PP_FAM_RDF | PP_FAM_RIPK)) {
        std::vector<int64_t> acc((size_t)p.ntasks * C, 0);
        for (int t = 0; t < p.ntasks; ++t)
            for (int tc = 0; tc < p.ntrg; ++tc) {
                if (!((p.task_trg[t] >> tc) & 1u))
                    continue;
                for (int e = 0; e < p.E; ++e) {
                    if (!((p.task_env[t] >> e) & 1u))
                        continue;
                    const unsigned long long *src = &P[((size_t)tc * p.E + e) * C];
                    for (int k = 0; k < C; ++k)
                        acc[(size_t)t * C + k] += (int64_t)src[k];
                }
            }
        if ((p.families & PP_FAM_RDF) && o->rdf)
            std::memcpy(o->rdf, acc.data(), acc.size() * 8);
        if ((p.families & PP_FAM_RIPK) && o->ripk)
            std::memcpy(o->ripk, acc.data(), acc.size() * 8);
    }
    if ((p.families & PP_FAM_1NN) && o->nn1)
        for (size_t q = 0; q < (size_t)p.ntasks * C; ++q)
            o->nn1[q] = N1[q];
    if ((p.families & PP_FAM_MKNN) && o->mknn) {
        const size_t nk = p.kdesc.size();
        for (int t = 0; t < p.ntasks; ++t)
            for (size_t kth = 0; kth < nk; ++kth) {
                const int u = p.kth_to_u[kth];
                for (int k = 0; k < C; ++k)
                    o->mknn[((size_t)t * nk + kth) * C + k] =
                        K[((size_t)t * p.nku + u) * C + k];
            }
    }
    if (p.families & PP_FAM_CNTNB) {
        int64_t len = 0;
        for (int t = 0; t < p.ntasks; ++t)
            for (int64_t b = r.cn_cap - 1; b >= 0; --b)
                if (CN[(size_t)t * r.cn_cap + b]) {
                    len = std::max(len, b + 1);
                    break;
                }
        o->cntnb_len = len;
        if (len > o->cntnb_cap)
            *overflow = true;
        if (o->cntnb) {
            for (int t = 0; t < p.ntasks; ++t)
                for (int64_t b = 0; b < o->cntnb_cap; ++b)
                    o->cntnb[(size_t)t * o->cntnb_cap + b] =
                        b < r.cn_cap ? (int64_t)CN[(size_t)t * r.cn_cap + b] : 0;
        }
    }
}


// ---------------------------------------------------------------------------
// kernel shape: centers per tile and cells per Rmax

// Density rule: cells of Rmax / r, r the nearest integer to the width that
// holds ~50 candidates, capped at 4 (finer cells cut fewer candidates per
// in-ball pair than their per-tile and per-row work costs; the autotuner also
// tries 5).
// Ions per nm^3 among a job's candidates: the neighbour types alone when they
// are a minority (the typed candidate list), else every ion.
double cand_density(const pp_cloud *cl, const JobPlan &p)
{
    if (cl->n == 0)
        return 0.0;
    double envn = 0;
    for (int e = 0; e < p.E; ++e)
        envn += (double)cl->type_count[p.env_type[e]];
    const double frac = envn / (double)cl->n;
    return cl->density() * (frac < 0.8 ? frac : 1.0);
}


int rule_r(double density, double R)
{
    // with the stencil rows trimmed to the x parts of their end cells, larger
    // cells lose little to whole cells and save per-row work (measured: 20M
    // ions at 27 ions/nm^3 and R = 5 fastest at R/4; 100M ions at 22 ions/nm^3
    // and R = 3 fastest at R/2)
    const int r = (int)std::lround(R * std::cbrt(std::max(density, 1e-9) / 50.0));
    return std::min(4, std::max(1, r));
}

long tune_bucket(double density, const JobPlan &p)
{
    const double nball = density * 4.18879 * (double)p.R * p.R * p.R;
    const long b = (long)std::lround(2.0 * std::log2(std::max(nball, 1.0)));
    return b * 4 + (p.percenter ? 1 : 0) + ((p.families & (PP_FAM_RDF | PP_FAM_RIPK)) ? 2 : 0);
}

// Chooses the cell list a job runs on: a resident cloud offers fixed widths
// (pick the one closest to Rmax / r in log scale), a non-resident one builds
// (or reuses) the width Rmax / r.
pp_status pick_celllist(pp_ctx *c, Slot &s, pp_cloud *cl, const JobPlan &p,
                        const CellList *&out)
{
    const double want = (double)p.R / p.r * (1.0 + std::ldexp(1.0, -10));
    if (cl->resident) {
        double best = 1e300;
        for (const CellList &L : cl->lists) {
            const double d = std::fabs(std::log(L.g.w / want));
            if (d < best) {
                best = d;
                out = &L;
            }
        }
        return out ? PP_OK : fail(c, PP_ERR_INVALID_ARG, "resident cloud without a cell list");
    }
    if (s.fixed) {
        if (!cl->lists.empty() && cl->lists[0].mem && cl->lists[0].w == (float)want &&
            cl->lists[0].ay == p.ay) {
            out = &cl->lists[0];
            return PP_OK;
        }
        return fail(c, PP_ERR_INVALID_ARG,
                    "the no-allocation entry point needs a resident cloud (uploaded with a "
                    "cell width) or one whose cell list this job's width already built");
    }
    if (cl->lists.empty())
        cl->lists.emplace_back();
    CellList &L = cl->lists[0];
    if (!(L.mem && L.w == (float)want && L.ay == p.ay)) {
        {
            std::lock_guard<std::mutex> lk(cl->sub_mu);
            for (auto &sb : cl->subs)
                if (sb.cl.mem)
                    (void)hipFree(sb.cl.mem);
            cl->subs.clear();
        }
        pp_status st = build_celllist(c, s, cl, (float)want, L, p.ay);
        if (st != PP_OK)
            return st;
    }
    out = &L;
    return PP_OK;
}

// ---------------------------------------------------------------------------
// running jobs

// Runs that share the device through several streams (the serving loop's
// slots) take turns for their query launches, in the order of their sequence
// numbers (the order their batches were formed in): each waits, on the host,
// for its turn and, on the device, for the query launches of the run before
// it, so that a run's preparation overlaps the query of the run before while
// the queries themselves keep arrival order (a small batch whose preparation
// ends first does not overtake a larger one formed before it).
struct LaunchOrder {
    std::mutex mu;
    std::condition_variable cv;
    hipEvent_t ev = nullptr;
    bool armed = false;
    uint64_t turn = 0;  // sequence number of the run whose query launches go next
};

// One run's place in a LaunchOrder.  take() blocks until it is this run's
// turn and holds the order's lock; release() passes the turn on.  The
// destructor passes on a turn never taken too (a run that failed before its
// launches), so that later runs never wait for it.
class OrderTurn {
  public:
    OrderTurn(LaunchOrder *o, uint64_t seq) : o_(o), seq_(seq) {}
    OrderTurn(const OrderTurn &) = delete;
    OrderTurn &operator=(const OrderTurn &) = delete;
    ~OrderTurn() { release(); }
    void take()
    {
        if (!o_ || taken_)
            return;
        lk_ = std::unique_lock<std::mutex>(o_->mu);
        o_->cv.wait(lk_, [&] { return o_->turn == seq_; });
        taken_ = true;
    }
    void release()
    {
        if (!o_ || released_)
            return;
        take();
        ++o_->turn;
        released_ = true;
        lk_.unlock();
        o_->cv.notify_all();
    }

  private:
    LaunchOrder *o_;
    uint64_t seq_;
    std::unique_lock<std::mutex> lk_;
    bool taken_ = false, released_ = false;
};

pp_status run_jobs(pp_ctx *c, Slot &sl, int njobs, const pp_cloud *const *clouds,
                   const pp_job *jobs, pp_hist_out *outs, Shape force = Shape(),
                   LaunchOrder *order = nullptr, uint64_t seq = 0);

constexpr size_t kKbChunkBytes = size_t(2) << 30;  // k-th neighbour buffer of one launch chunk
constexpr int kNoAllocTimedLaunches = 64;          // query launches the fixed-arena slot times
constexpr int kReduceRows = 32768;                  // rows per reduction work unit

// One query launch over the global tile range [g0, g1).
struct KChunk {
    uint32_t g0 = 0, g1 = 0;
    std::vector<uint32_t> base, ntl;  // per job: first local tile and tile count in the chunk
    size_t bytes = 0;                 // k-th neighbour buffer bytes
};

// Bytes of one tile's block of k-th neighbour categories (planes padded to a
// multiple of 8).
size_t kb_tile_bytes(const JobPlan &p, int T)
{
    const int np = p.npad();
    if (np == 0)
        return 0;
    const size_t TR = (size_t)T * (size_t)p.lmax;
    return TR * 2 * (size_t)np;
}

// Splits the launch's tiles (job after job) into chunks whose k-th neighbour
// blocks fit `budget` bytes (at least one tile per chunk).
void plan_chunks(const std::vector<const JobPlan *> &plans, const std::vector<uint32_t> &tile_off,
                 int T, size_t budget, std::vector<KChunk> &out)
{
    const int njobs = (int)plans.size();
    const size_t slack = 512;  // alignment of a job's array in the chunk buffer
    out.clear();
    KChunk cur;
    cur.base.assign(njobs, 0u);
    cur.ntl.assign(njobs, 0u);
    auto close = [&](uint32_t g1) {
        cur.g1 = g1;
        out.push_back(cur);
        cur.g0 = g1;
        cur.bytes = 0;
        std::fill(cur.base.begin(), cur.base.end(), 0u);
        std::fill(cur.ntl.begin(), cur.ntl.end(), 0u);
    };
    for (int i = 0; i < njobs; ++i) {
        const uint32_t n = tile_off[i + 1] - tile_off[i];
        const size_t tb = kb_tile_bytes(*plans[i], T);
        uint32_t done = 0;
        while (done < n) {
            uint32_t take = n - done;
            if (tb) {
                const size_t avail = budget > cur.bytes + slack ? budget - cur.bytes - slack : 0;
                uint64_t fit = avail / tb;
                if (fit == 0 && cur.bytes == 0)
                    fit = 1;
                if (fit == 0) {
                    close(tile_off[i] + done);
                    continue;
                }
                take = (uint32_t)std::min<uint64_t>(fit, take);
                cur.bytes += (size_t)take * tb + slack;
            }
            if (cur.ntl[i] == 0)
                cur.base[i] = done;
            cur.ntl[i] += take;
            done += take;
            if (done < n)
                close(tile_off[i] + done);
        }
    }
    if (out.empty() || cur.g0 < tile_off[njobs])
        close(tile_off[njobs]);
}

// Host images of device tables handed to asynchronous copies; they must
// outlive the launches that read them.
struct Keep {
    std::deque<std::vector<JobDev>> devs;
    std::deque<std::vector<uint32_t>> words;
    std::deque<std::vector<const Tile *>> tiles;
    bool graph = false;               // capture each group's launches into a graph
    hipStream_t stream = nullptr;     // the stream the graphs were launched on
    size_t ngraphs = 0;               // groups launched as graphs so far in this run
    Keep() = default;
    Keep(const Keep &) = delete;
    Keep &operator=(const Keep &) = delete;
    ~Keep()
    {
        // the launched graphs finish before the slot's next run updates
        // their executables
        if (ngraphs)
            (void)hipStreamSynchronize(stream);
    }
};

// One query launch group: the jobs `ids` (one kernel mode, tiles of T
// centers) with their device descriptors `base`, as launch chunks, each
// followed by the reduction of its k-th neighbour categories.  The group's
// LDS layout must fit (query_lds_bytes).  With keep.graph the group's kernel
// launches (queue clears, query chunks, reductions; the descriptors are
// copied before) are captured into one HIP graph, which is launched once.
pp_status launch_group(pp_ctx *c, Slot &sl, std::vector<JobRun> &runs, const std::vector<int> &ids,
                       const std::vector<JobDev> &base, int mode, int &T, Keep &keep)
{
    const int nj = (int)ids.size();
    if (nj == 0)
        return PP_OK;
    int maxC = 1, maxE = 1, maxTrg = 1, maxK = 0, maxRows = 1, maxHK = 1, ntypes = 1;
    uint32_t maxcand = 0;
    for (int q = 0; q < nj; ++q) {
        const JobRun &r = runs[ids[q]];
        const JobPlan &p = r.plan;
        maxC = std::max(maxC, p.C);
        maxE = std::max(maxE, p.E);
        maxTrg = std::max(maxTrg, p.ntrg);
        maxK = std::max(maxK, p.kidx_len());
        maxRows = std::max(maxRows, query_rows(r, mode));
        if (mode == 2)
            maxHK = std::max(maxHK, p.HK);
        ntypes = std::max(ntypes, type_span(r.cloud));
        maxcand = std::max(maxcand, r.h_counts[2]);
    }
    const int c16 = maxcand < 65535u ? 1 : 0;
    const size_t lds =
        query_lds_bytes(T, c16, mode, maxC, maxE, maxTrg, maxK, maxRows, maxHK, ntypes);
    if (lds > c->lds_per_block)
        return fail(c, PP_ERR_UNSUPPORTED, "histograms too large for LDS (bins x types x tile)");
    std::vector<uint32_t> tile_off(nj + 1, 0);
    std::vector<const JobPlan *> plans(nj);
    uint32_t total = 0;
    for (int q = 0; q < nj; ++q) {
        tile_off[q] = total;
        total += runs[ids[q]].h_counts[1];
        plans[q] = &runs[ids[q]].plan;
    }
    tile_off[nj] = total;
    // launch chunks: the k-th neighbour categories of a chunk's tiles must
    // fit the chunk buffer
    std::vector<KChunk> chunks;
    size_t kb_budget = kKbChunkBytes;
    if (sl.fixed) {
        // what the fixed arena has left after the descriptors' reserve (or
        // the k-th neighbour buffer already carved, when that is more)
        const size_t used = (sl.fixed_off + 255) & ~(size_t)255;
        const size_t room =
            sl.fixed_cap > used + kFixedJobbufReserve ? sl.fixed_cap - used - kFixedJobbufReserve : 0;
        kb_budget = std::min(kb_budget, std::max(room, sl.kbuf.cap > 256 ? sl.kbuf.cap - 256 : 0));
    }
    plan_chunks(plans, tile_off, T, kb_budget, chunks);
    const size_t nch = chunks.size();
    size_t kb_need = 0;
    for (const KChunk &ch : chunks)
        kb_need = std::max(kb_need, ch.bytes);
    if (kb_need) {
        pp_status st = ensure(c, sl, sl.kbuf, kb_need + 256);
        if (st != PP_OK)
            return st;
    }
    Carver jb0(nullptr);
    jb0.take<JobDev>((size_t)nj * nch);
    jb0.take<const Tile *>(nj);
    jb0.take<uint32_t>(nj + 1);
    jb0.take<uint32_t>((size_t)(nj + 1) * nch);
    jb0.take<uint32_t>(kQueueWords);
    pp_status st = ensure(c, sl, sl.jobbuf, jb0.off + 256);
    if (st != PP_OK)
        return st;
    Carver jb(sl.jobbuf.p);
    JobDev *d_jobs = jb.take<JobDev>((size_t)nj * nch);
    const Tile **d_tiles = jb.take<const Tile *>(nj);
    uint32_t *d_off = jb.take<uint32_t>(nj + 1);
    uint32_t *d_uoff = jb.take<uint32_t>((size_t)(nj + 1) * nch);
    uint32_t *d_queue = jb.take<uint32_t>(kQueueWords);
    keep.devs.emplace_back((size_t)nj * nch);
    std::vector<JobDev> &hj = keep.devs.back();
    keep.words.emplace_back((size_t)(nj + 1) * nch, 0u);
    std::vector<uint32_t> &huoff = keep.words.back();
    keep.words.push_back(tile_off);
    const std::vector<uint32_t> &htoff = keep.words.back();
    keep.tiles.emplace_back(nj);
    std::vector<const Tile *> &ht = keep.tiles.back();
    uint32_t tpu_max = 1;
    std::vector<uint32_t> tpus(nch);  // tiles per reduction unit, per chunk
    int red_lds = 0;
    {
        int maxTR = 1;
        for (int q = 0; q < nj; ++q)
            if (plans[q]->nplanes() > 0) {
                maxTR = std::max(maxTR, T * plans[q]->lmax);
                const int w = (plans[q]->ntasks * plans[q]->kb_group() * plans[q]->C + 1) / 2 +
                              plans[q]->ntrg * plans[q]->lmax;
                if (w <= kReduceLdsWords)
                    red_lds = std::max(red_lds, w);
            }
        // a unit adds at most tpu * T * lmax rows to one 16-bit counter
        tpu_max = (uint32_t)std::max(1, std::min(kReduceRows, 65535) / maxTR);
    }
    for (size_t k = 0; k < nch; ++k) {
        const KChunk &ch = chunks[k];
        // units of up to tpu_max tiles, but at least ~4 per CU when the
        // chunk's groups of planes hold fewer tiles than that (small jobs)
        uint64_t gt = 0;
        for (int q = 0; q < nj; ++q) {
            const int np = plans[q]->nplanes(), G = plans[q]->kb_group();
            if (np > 0)
                gt += (uint64_t)((np + G - 1) / G) * ch.ntl[q];
        }
        const uint64_t want = gt / (4u * (uint64_t)std::max(c->cu_count, 1));
        const uint32_t tpu = tpus[k] =
            (uint32_t)std::max<uint64_t>(std::min<uint64_t>(tpu_max, 32u),
                                         std::min<uint64_t>(tpu_max, want));
        Carver kc(sl.kbuf.p);
        uint32_t units = 0;
        for (int q = 0; q < nj; ++q) {
            JobDev d = base[q];
            const int np = plans[q]->nplanes();
            huoff[k * (nj + 1) + q] = units;
            if (np > 0 && ch.ntl[q] > 0) {
                const size_t TR = (size_t)T * plans[q]->lmax;
                const int G = plans[q]->kb_group();
                d.KB = kc.take<uint16_t>((size_t)ch.ntl[q] * plans[q]->npad() * TR);
                d.tile_base = ch.base[q];
                d.kb_tiles = ch.ntl[q];
                units += (uint32_t)((np + G - 1) / G) * ((ch.ntl[q] + tpu - 1) / tpu);
            }
            hj[k * nj + q] = d;
        }
        huoff[k * (nj + 1) + nj] = units;
    }
    for (int q = 0; q < nj; ++q)
        ht[q] = runs[ids[q]].tiles;
    PP_HIP(c, hipMemcpyAsync(d_jobs, hj.data(), sizeof(JobDev) * hj.size(), hipMemcpyHostToDevice,
                             sl.stream));
    PP_HIP(c, hipMemcpyAsync(d_tiles, ht.data(), sizeof(Tile *) * nj, hipMemcpyHostToDevice,
                             sl.stream));
    PP_HIP(c, hipMemcpyAsync(d_off, htoff.data(), 4 * (nj + 1), hipMemcpyHostToDevice, sl.stream));
    PP_HIP(c, hipMemcpyAsync(d_uoff, huoff.data(), 4 * huoff.size(), hipMemcpyHostToDevice,
                             sl.stream));
    int nb = 0;
    PP_HIP(c, query_max_blocks_per_cu(T, c16, mode, lds, &nb));
    if (nb < 1)
        nb = 1;
    if (std::getenv("PP_TRACE_LAUNCH"))
        std::fprintf(stderr,
                     "pp: query mode=%d T=%d c16=%d lds=%zu blocks/CU=%d jobs=%d tiles=%u chunks=%zu "
                     "maxC=%d maxE=%d maxTrg=%d kmax=%d rows=%d HK=%d\n",
                     mode, T, c16, lds, nb, nj, total, nch, maxC, maxE, maxTrg, maxK, maxRows,
                     maxHK);
    const bool capture = keep.graph && nch > 0;  // (no empty graphs)
    if (capture)
        PP_HIP(c, hipStreamBeginCapture(sl.stream, hipStreamCaptureModeRelaxed));
    hipError_t le = hipSuccess;
    for (size_t k = 0; k < nch && le == hipSuccess; ++k) {
        const KChunk &ch = chunks[k];
        QueryLaunch ql{};
        ql.jobs = d_jobs + k * nj;
        ql.njobs = nj;
        ql.tile_off = d_off;
        ql.tiles = d_tiles;
        ql.first_tile = ch.g0;
        ql.end_tile = ch.g1;
        ql.queue = d_queue;
        ql.T = T;
        ql.c16 = c16;
        ql.mode = mode;
        ql.maxC = maxC;
        ql.maxE = maxE;
        ql.maxTrg = maxTrg;
        ql.maxKmax = maxK;
        ql.ntypes = ntypes;
        ql.maxRows = maxRows;
        ql.maxHK = maxHK;
        const uint32_t want = (uint32_t)c->cu_count * (uint32_t)nb;
        ql.grid = (int)std::max<uint32_t>(
            1u, std::min<uint32_t>(want, std::max<uint32_t>(ch.g1 - ch.g0, 1u)));
        // several tiles per queue atomic when tiles are many and cheap (the
        // stencil bounds a tile's pair evaluations by maxcand * T); blocks of
        // tiles cost the XCDs their shared front where tiles are dearer
        // (measured: 100M ions at Rmax 3 nm, 426.5 ms per job with blocks of
        // 4 tiles, 414.7 ms with single tiles)
        ql.grab = (uint64_t)maxcand * (uint64_t)T < (1u << 15)
                      ? std::max<uint32_t>(1u, std::min<uint32_t>(
                                                   4u, (ch.g1 - ch.g0) / ((uint32_t)ql.grid * 16u)))
                      : 1u;
        const bool timed = !capture && (!sl.fixed || sl.qev.size() >= 2 * (sl.nq + 1));
        if (timed && le == hipSuccess) {
            // time the query launch alone (not its reduction); a fixed-arena
            // slot uses the events made when it was set up
            while (sl.qev.size() < 2 * (sl.nq + 1)) {
                hipEvent_t e = nullptr;
                le = hipEventCreate(&e);
                if (le != hipSuccess)
                    break;
                sl.qev.push_back(e);
            }
            if (le == hipSuccess)
                le = hipEventRecord(sl.qev[2 * sl.nq], sl.stream);
        }
        if (le == hipSuccess)
            le = launch_query(ql, sl.stream);
        if (timed && le == hipSuccess) {
            le = hipEventRecord(sl.qev[2 * sl.nq + 1], sl.stream);
            ++sl.nq;
        }
        const uint32_t units = huoff[k * (nj + 1) + nj];
        if (le == hipSuccess && units > 0)
            le = launch_kreduce(d_jobs + k * nj, nj, d_uoff + k * (nj + 1), units, T, tpus[k],
                                red_lds, sl.stream);
    }
    if (capture) {
        // the capture always ends, so the stream is usable after an error
        hipGraph_t gr = nullptr;
        const hipError_t ee = hipStreamEndCapture(sl.stream, &gr);
        if (le == hipSuccess)
            le = ee;
        if (le == hipSuccess) {
            // this group's executable from the slot's last run, updated to
            // the new capture when the kernels and chunk count match
            const size_t pos = keep.ngraphs;
            if (sl.gexec.size() <= pos)
                sl.gexec.resize(pos + 1, nullptr);
            hipGraphExec_t &ex = sl.gexec[pos];
            if (ex) {
                hipGraphNode_t bad = nullptr;
                hipGraphExecUpdateResult res = hipGraphExecUpdateError;
                if (hipGraphExecUpdate(ex, gr, &bad, &res) != hipSuccess ||
                    res != hipGraphExecUpdateSuccess) {
                    (void)hipGetLastError();
                    (void)hipGraphExecDestroy(ex);
                    ex = nullptr;
                }
            }
            if (!ex)
                le = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
            if (le == hipSuccess) {
                keep.stream = sl.stream;
                ++keep.ngraphs;
                le = hipGraphLaunch(ex, sl.stream);
            } else {
                ex = nullptr;
            }
        }
        if (gr)
            (void)hipGraphDestroy(gr);
    }
    PP_HIP(c, le);
    for (int q = 0; q < nj; ++q) {
        runs[ids[q]].T_used = T;
        runs[ids[q]].c16_used = c16;
    }
    return PP_OK;
}

// Times the candidate shapes on this job and keeps the fastest for its bucket.
pp_status autotune(pp_ctx *c, Slot &sl, const pp_cloud *cl, const pp_job *job, const JobPlan &p,
                   long bucket)
{
    pp_hist_out scratch;
    double best = 1e300;
    Shape win;
    win.T = 8;
    win.r = rule_r(cand_density(cl, p), p.R);
    win.ay = 1;
    // cubic cells of Rmax/1 .. Rmax/5, and (for the lists jobs build for
    // themselves) cells of Rmax by Rmax/2 and Rmax/2 by Rmax/4; each shape is
    // timed twice (the first run of a width builds its list), the faster run
    // counts
    std::vector<Shape> cands;
    for (int T : {8, 16}) {
        for (int r = 1; r <= 5; ++r)
            cands.push_back(Shape{T, r, 1});
        for (int r = 1; r <= 2 && !cl->resident; ++r)
            cands.push_back(Shape{T, r, 2});
    }
    for (const Shape &f : cands)
        for (int rep = 0; rep < 2; ++rep) {
            std::memset(&scratch, 0, sizeof(scratch));
            pp_status st = run_jobs(c, sl, 1, &cl, job, &scratch, f);
            if (st != PP_OK)
                break;
            if (scratch.ms_query < best) {
                best = scratch.ms_query;
                win = f;
            }
        }
    std::lock_guard<std::mutex> lk(c->mu);
    c->tuned[bucket] = win;
    return PP_OK;
}

pp_status run_jobs(pp_ctx *c, Slot &sl, int njobs, const pp_cloud *const *clouds,
                   const pp_job *jobs, pp_hist_out *outs, Shape force, LaunchOrder *order,
                   uint64_t seq)
{
    auto t_start = std::chrono::steady_clock::now();
    OrderTurn turn(order, seq);
    if (!outs)
        return fail(c, PP_ERR_INVALID_ARG, "outputs are NULL");
    sl.nq = 0;
    std::vector<JobRun> runs(njobs);
    int tuneT, tuneR, tuneAY, graph_launch;
    double knn_factor;
    int sym_pairs;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        tuneT = c->tune_T;
        tuneR = c->tune_r;
        tuneAY = c->tune_ay;
        knn_factor = c->knn_factor;
        sym_pairs = c->sym_pairs;
        graph_launch = c->graph_launch;
    }
    // the jobs' plans (histogram geometry, exact edge tables, task tables)
    // are independent host work: on several threads for larger runs; a job
    // whose plan fails is planned again below, in order, for its message
    std::vector<char> planned(njobs, 0);
    if (njobs >= 16) {
        const int nt = std::max(1, std::min(8, (int)std::thread::hardware_concurrency()));
        std::atomic<int> next{0};
        auto work = [&]() {
            for (int i = next++; i < njobs; i = next++)
                planned[i] = plan_job(nullptr, &jobs[i], runs[i].plan) == PP_OK ? 1 : 0;
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nt; ++t) {
            try {
                pool.emplace_back(work);
            } catch (...) {
                break;  // no more threads: the calling thread plans the rest
            }
        }
        work();
        for (std::thread &t : pool)
            t.join();
    }
    for (int i = 0; i < njobs; ++i) {
        if (!clouds[i])
            return fail(c, PP_ERR_INVALID_ARG, "cloud is NULL");
        pp_status st = planned[i] ? PP_OK : plan_job(c, &jobs[i], runs[i].plan);
        if (st != PP_OK)
            return st;
        runs[i].cloud = clouds[i];
        JobPlan &p = runs[i].plan;
        int jT = 8, jAY = 0;
        p.r = rule_r(cand_density(clouds[i], p), p.R);
        bool by_rule = true;  // the cell shape is the density rule's
        if (force.T > 0 || force.r > 0) {
            by_rule = force.r <= 0;
            if (force.T > 0)
                jT = force.T;
            if (force.r > 0)
                p.r = force.r;
            jAY = force.ay;
        } else if (tuneT < 0 && !sl.fixed) {
            by_rule = false;
            const long b = tune_bucket(cand_density(clouds[i], p), p);
            bool have;
            {
                std::lock_guard<std::mutex> lk(c->mu);
                have = c->tuned.count(b) != 0;
            }
            if (!have) {
                st = autotune(c, sl, clouds[i], &jobs[i], p, b);
                if (st != PP_OK)
                    return st;
            }
            std::lock_guard<std::mutex> lk(c->mu);
            const Shape &w = c->tuned[b];
            jT = w.T;
            p.r = w.r;
            jAY = w.ay;
        } else {
            if (tuneT > 0)
                jT = tuneT;
            if (tuneR > 0) {
                p.r = tuneR;
                by_rule = false;
            }
        }
        runs[i].T_req = jT;
        p.ay = jAY > 0 ? jAY : tuneAY > 0 ? tuneAY : 1;
        // the rule's cells of Rmax/4 and Rmax/2 are taken twice as long
        // along x in the lists jobs build for themselves (Rmax/2 by Rmax/4,
        // Rmax by Rmax/2): the same stencil rows, each of half as many cells
        if (by_rule && tuneAY <= 0 && jAY <= 0 && (p.r == 4 || p.r == 2) &&
            !clouds[i]->resident) {
            p.r /= 2;
            p.ay = 2;
        }
        plan_hybrid(clouds[i], p, knn_factor);
    }
    PP_HIP(c, hipSetDevice(c->device));
    // a non-resident cloud's cached cell list is rebuilt per width: hold it
    std::vector<std::unique_lock<std::mutex>> locks;
    {
        std::set<pp_cloud *> own;
        for (int i = 0; i < njobs; ++i)
            if (!clouds[i]->resident)
                own.insert(const_cast<pp_cloud *>(clouds[i]));
        for (pp_cloud *cl : own)
            locks.emplace_back(cl->mu);
    }
    PP_HIP(c, hipEventRecord(sl.ev[0], sl.stream));
    for (int i = 0; i < njobs; ++i) {
        pp_status st = pick_celllist(c, sl, const_cast<pp_cloud *>(clouds[i]), runs[i].plan,
                                     runs[i].cl);
        if (st != PP_OK)
            return st;
        runs[i].cand = runs[i].cl;
        {
            // only the neighbour types are candidates when they are a minority
            const JobPlan &p = runs[i].plan;
            const pp_cloud *cl = clouds[i];
            uint64_t inenv = 0;
            uint32_t mask[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int e = 0; e < p.E; ++e) {
                inenv += cl->type_count[p.env_type[e]];
                mask[p.env_type[e] >> 5] |= 1u << (p.env_type[e] & 31);
            }
            if (cl->n > 0 && (double)inenv < 0.8 * (double)cl->n) {
                st = get_sublist(c, sl, const_cast<pp_cloud *>(cl), runs[i].cl, mask,
                                 runs[i].cand);
                if (st != PP_OK)
                    return st;
            }
        }
        {
            // the RDF pair counts give the in-ball total when every ion type
            // of the cloud is a neighbour type of the job
            const JobPlan &p = runs[i].plan;
            bool all_env = runs[i].cand == runs[i].cl;
            for (int t = 0; t < 256 && all_env; ++t)
                if (clouds[i]->type_count[t] > 0 && p.envoff[t] < 0)
                    all_env = false;
            runs[i].count_in = !(all_env && (p.families & (PP_FAM_RDF | PP_FAM_RIPK)));
        }
        runs[i].rows = make_rows(runs[i].cl->g.w, runs[i].cl->g.wyz, runs[i].plan.R);
        if ((int)runs[i].rows.size() > kMaxRows)
            return fail(c, PP_ERR_UNSUPPORTED, "stencil too large (cell width far below Rmax)");
        runs[i].ncells = runs[i].cl->g.ncells();
        if (sym_pairs) {
            st = plan_symmetric(c, sl, const_cast<pp_cloud *>(clouds[i]), runs[i]);
            if (st != PP_OK)
                return st;
        }
    }
    struct Back {
        std::vector<unsigned long long> P;
        std::vector<uint32_t> K, N1, CN;
        unsigned long long stats[4];
    };
    std::vector<Back> back(njobs);
    PP_HIP(c, hipEventRecord(sl.ev[1], sl.stream));
    // centers per tile: the job's request, halved while its LDS layout with
    // 32-bit per-center counters does not fit
    for (int i = 0; i < njobs; ++i) {
        JobRun &r = runs[i];
        const JobPlan &p = r.plan;
        auto need = [&](int t) {
            return query_lds_bytes(t, 0, p.mode, std::max(p.C, 1), std::max(p.E, 1),
                                   std::max(p.ntrg, 1), p.kidx_len(), query_rows(r, p.mode),
                                   std::max(p.HK, 1), type_span(r.cloud));
        };
        while (need(r.T_req) > c->lds_per_block && r.T_req > 8)
            r.T_req /= 2;
        if (need(r.T_req) > c->lds_per_block)
            return fail(c, PP_ERR_UNSUPPORTED, "histograms too large for LDS (bins x types x tile)");
    }
    // scratch (after the cell lists: a build may have used it): the counters
    // of every job together (one clear, one read-back), the tables of every
    // job together (one upload), each job's rerun marks, the run's lists
    PrepArena pa;
    size_prep(pa, runs);
    size_t tables_words = 0;
    for (int i = 0; i < njobs; ++i) {
        layout_tables(runs[i]);
        tables_words += runs[i].tables_words;
    }
    auto carve = [&](Carver &m, char *&out0, char *&out1) {
        out0 = m.base ? m.base + ((m.off + 255) & ~(size_t)255) : nullptr;
        for (int i = 0; i < njobs; ++i)
            layout_outputs(runs[i], m);
        out1 = m.base ? m.base + m.off : nullptr;
        uint32_t *tab = m.take<uint32_t>(tables_words);
        for (int i = 0; i < njobs; ++i) {
            runs[i].tables = tab;
            tab = tab ? tab + runs[i].tables_words : nullptr;
        }
        for (int i = 0; i < njobs; ++i)
            runs[i].flags = m.take<uint32_t>(runs[i].cloud->n + 1);
        layout_prep(pa, m);
    };
    char *out0 = nullptr, *out1 = nullptr;
    Carver m0(nullptr);
    carve(m0, out0, out1);
    pp_status st = ensure(c, sl, sl.scratch, m0.off + 256);
    if (st != PP_OK)
        return st;
    Carver m(sl.scratch.p);
    carve(m, out0, out1);
    const size_t out_bytes = (size_t)(out1 - out0);
    {
        std::vector<uint32_t> h(tables_words, 0u);
        size_t o = 0;
        for (int i = 0; i < njobs; ++i) {
            stage_tables(runs[i], h.data() + o);
            o += runs[i].tables_words;
        }
        if (tables_words)
            PP_HIP(c, hipMemcpyAsync(runs[0].tables, h.data(), tables_words * 4,
                                     hipMemcpyHostToDevice, sl.stream));
        PP_HIP(c, hipMemsetAsync(out0, 0, out_bytes, sl.stream));
        std::vector<int> all(njobs);
        for (int i = 0; i < njobs; ++i)
            all[i] = i;
        st = prep_run(c, sl, runs, all, pa, false);  // synchronises: h may go
        if (st != PP_OK)
            return st;
    }
    PP_HIP(c, hipEventRecord(sl.ev[2], sl.stream));
    size_t cn_words = 0;
    {
        // CNTNB counters: no neighbour count exceeds the largest stencil candidate count
        for (int i = 0; i < njobs; ++i) {
            JobRun &r = runs[i];
            r.cn_cap = (r.plan.families & PP_FAM_CNTNB) ? (int)r.h_counts[2] + 1 : 1;
            cn_words += ((size_t)std::max(r.plan.ntasks, 1) * (size_t)r.cn_cap + 63) & ~(size_t)63;
        }
        st = ensure(c, sl, sl.cnbuf, cn_words * 4 + 256);
        if (st != PP_OK)
            return st;
        PP_HIP(c, hipMemsetAsync(sl.cnbuf.p, 0, cn_words * 4, sl.stream));
        size_t o = 0;
        for (int i = 0; i < njobs; ++i) {
            JobRun &r = runs[i];
            r.CN = (uint32_t *)sl.cnbuf.p + o;
            o += ((size_t)std::max(r.plan.ntasks, 1) * (size_t)r.cn_cap + 63) & ~(size_t)63;
            fill_jobdev(r);
            r.considered = r.h_counts[0];
        }
    }
    // launch groups of one kernel mode, tile size and counter width (pair
    // counts, hybrid, full-range per center); a group whose common LDS
    // layout does not fit runs job by job.  Then the full-range rerun of the
    // centers whose kNN range fell short.
    auto lds_of = [&](const std::vector<int> &ids, int mode, int T, int c16) {
        int maxC = 1, maxE = 1, maxTrg = 1, maxK = 0, maxRows = 1, maxHK = 1, ntypes = 1;
        for (int i : ids) {
            const JobPlan &p = runs[i].plan;
            maxC = std::max(maxC, p.C);
            maxE = std::max(maxE, p.E);
            maxTrg = std::max(maxTrg, p.ntrg);
            maxK = std::max(maxK, p.kidx_len());
            maxRows = std::max(maxRows, query_rows(runs[i], mode));
            if (mode == 2)
                maxHK = std::max(maxHK, p.HK);
            ntypes = std::max(ntypes, type_span(runs[i].cloud));
        }
        return query_lds_bytes(T, c16, mode, maxC, maxE, maxTrg, maxK, maxRows, maxHK, ntypes);
    };
    struct Group {
        int mode, T, c16;
        std::vector<int> ids;
        std::vector<JobDev> devs;
    };
    auto form_groups = [&](const std::vector<int> &ids, const std::vector<JobDev> &devs,
                           const std::vector<int> &modes) {
        std::map<std::array<int, 3>, Group> by;
        for (size_t k = 0; k < ids.size(); ++k) {
            const JobRun &r = runs[ids[k]];
            const int c16 = r.h_counts[2] < 65535u ? 1 : 0;
            Group &g = by[{modes[k], r.T_req, c16}];
            g.mode = modes[k];
            g.T = r.T_req;
            g.c16 = c16;
            g.ids.push_back(ids[k]);
            g.devs.push_back(devs[k]);
        }
        std::vector<Group> out;
        for (int mode : {0, 2, 1})
            for (auto &kv : by) {
                Group &g = kv.second;
                if (g.mode != mode)
                    continue;
                if (g.ids.size() > 1 && lds_of(g.ids, g.mode, g.T, g.c16) > c->lds_per_block) {
                    for (size_t k = 0; k < g.ids.size(); ++k)
                        out.push_back(Group{g.mode, g.T, g.c16, {g.ids[k]}, {g.devs[k]}});
                } else {
                    out.push_back(std::move(g));
                }
            }
        return out;
    };
    std::vector<int> hy;  // hybrid jobs with a kNN range short of Rmax: they may rerun
    {
        std::vector<int> ids, modes;
        std::vector<JobDev> devs;
        // the rerun marks of the jobs that may rerun, cleared by one memset
        // over their span (the jobs' marks are carved one after another)
        char *mk0 = nullptr, *mk1 = nullptr;
        for (int i = 0; i < njobs; ++i) {
            JobRun &r = runs[i];
            ids.push_back(i);
            devs.push_back(r.dev);
            modes.push_back(r.plan.mode);
            if (r.plan.mode == 2 && r.plan.trunc_any) {
                char *b0 = (char *)r.flags, *b1 = (char *)(r.flags + r.cloud->n);
                mk0 = mk0 && mk0 < b0 ? mk0 : b0;
                mk1 = mk1 && mk1 > b1 ? mk1 : b1;
                hy.push_back(i);
            }
        }
        if (mk0 && mk1 > mk0)
            PP_HIP(c, hipMemsetAsync(mk0, 0, (size_t)(mk1 - mk0), sl.stream));
        if (order) {
            turn.take();
            if (order->armed)
                PP_HIP(c, hipStreamWaitEvent(sl.stream, order->ev, 0));
        }
        Keep keep;
        // graphs for runs of several jobs only: for a single job the graph's
        // instantiation (~0.1 ms) costs more than its launches
        keep.graph = graph_launch != 0 && njobs > 1;
        PP_HIP(c, hipEventRecord(sl.ev[3], sl.stream));
        for (Group &g : form_groups(ids, devs, modes)) {
            int T = g.T;
            st = launch_group(c, sl, runs, g.ids, g.devs, g.mode, T, keep);
            if (st != PP_OK)
                return st;
        }
        if (!hy.empty()) {
            // the rerun's tiles: the flagged centers, with the largest tile
            // that fits the full-range layout
            for (int i : hy) {
                JobRun &r = runs[i];
                const JobPlan &p = r.plan;
                while (r.T_req > 8 &&
                       query_lds_bytes(r.T_req, 0, 1, std::max(p.C, 1), std::max(p.E, 1),
                                       std::max(p.ntrg, 1), p.kidx_len(), (int)r.rows.size(),
                                       1, type_span(r.cloud)) > c->lds_per_block)
                    r.T_req /= 2;
            }
            st = prep_run(c, sl, runs, hy, pa, true);
            if (st != PP_OK)
                return st;
            std::vector<int> ids1, modes1;
            std::vector<JobDev> devs1;
            for (int i : hy) {
                JobRun &r = runs[i];
                r.rerun = r.h_counts[0];
                r.h_counts[0] = r.considered;
                if (r.h_counts[1] == 0)
                    continue;
                JobDev d = r.dev;
                d.families &= FAM_MKNN | FAM_1NN;
                d.percenter = 1;
                d.stats = r.stats + 4;
                d.centers = r.centers;
                d.tiles = r.tiles;
                ids1.push_back(i);
                devs1.push_back(d);
                modes1.push_back(1);
            }
            std::vector<int> tmain(njobs);
            for (int i : ids1)
                tmain[i] = runs[i].T_used;
            for (Group &g : form_groups(ids1, devs1, modes1)) {
                int T = g.T;
                st = launch_group(c, sl, runs, g.ids, g.devs, g.mode, T, keep);
                if (st != PP_OK)
                    return st;
            }
            for (int i : ids1)
                runs[i].T_used = tmain[i];  // the job's main launch
        }
        PP_HIP(c, hipEventRecord(sl.ev[4], sl.stream));
        // read back (the counters of every job in one copy)
        std::vector<char> hout(out_bytes);
        std::vector<uint32_t> hcn(cn_words);
        if (out_bytes)
            PP_HIP(c, hipMemcpyAsync(hout.data(), out0, out_bytes, hipMemcpyDeviceToHost,
                                     sl.stream));
        if (cn_words)
            PP_HIP(c, hipMemcpyAsync(hcn.data(), sl.cnbuf.p, cn_words * 4, hipMemcpyDeviceToHost,
                                     sl.stream));
        if (order) {
            // the next run's queries follow this run's read-back too: a copy
            // that needs compute units would otherwise wait behind the next
            // run's persistent query kernel, which holds them all, and this
            // run would complete only after the next one
            PP_HIP(c, hipEventRecord(order->ev, sl.stream));
            order->armed = true;
            turn.release();
        }
        PP_HIP(c, hipStreamSynchronize(sl.stream));
        for (int i = 0; i < njobs; ++i) {
            JobRun &r = runs[i];
            const JobPlan &p = r.plan;
            Back &b = back[i];
            auto host = [&](const void *dev) { return hout.data() + ((const char *)dev - out0); };
            b.P.resize((size_t)std::max(p.ntrg * p.E * p.C, 1));
            std::memcpy(b.P.data(), host(r.P), b.P.size() * 8);
            std::memcpy(b.stats, host(r.stats), sizeof(b.stats));
            b.K.resize((size_t)std::max(p.ntasks * p.nku * p.C, 1));
            std::memcpy(b.K.data(), host(r.K), b.K.size() * 4);
            b.N1.resize((size_t)std::max(p.ntasks * p.C, 1));
            std::memcpy(b.N1.data(), host(r.N1), b.N1.size() * 4);
            if (p.families & PP_FAM_CNTNB) {
                const uint32_t *src = hcn.data() + (r.CN - (uint32_t *)sl.cnbuf.p);
                b.CN.assign(src, src + (size_t)p.ntasks * r.cn_cap);
            }
        }
    }
    float ms_idx = 0, ms_prep = 0, ms_q = 0;
    PP_HIP(c, hipEventElapsedTime(&ms_idx, sl.ev[0], sl.ev[1]));
    PP_HIP(c, hipEventElapsedTime(&ms_prep, sl.ev[1], sl.ev[2]));
    PP_HIP(c, hipEventElapsedTime(&ms_q, sl.ev[3], sl.ev[4]));
    double ms_kern = 0;
    for (size_t q = 0; q < sl.nq; ++q) {
        float ms = 0;
        PP_HIP(c, hipEventElapsedTime(&ms, sl.qev[2 * q], sl.qev[2 * q + 1]));
        ms_kern += ms;
    }
    const double ms_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
            .count();
    bool any_overflow = false;
    for (int i = 0; i < njobs; ++i) {
        bool overflow = false;
        compose_outputs(runs[i], back[i].P, back[i].K, back[i].N1, back[i].CN, back[i].stats,
                        &outs[i], &overflow);
        if (runs[i].cand != runs[i].cl)
            outs[i].pairs_inball = -1;  // only neighbour types were scanned
        outs[i].ms_index = ms_idx;
        outs[i].ms_prep = ms_prep;
        outs[i].ms_query = ms_q;
        outs[i].ms_total = ms_total;
        outs[i].tile_centers = runs[i].T_used;
        outs[i].cells_per_rmax = runs[i].cloud->resident ? 0 : runs[i].plan.r;
        outs[i].counters16 = runs[i].c16_used;
        outs[i].percenter = runs[i].plan.mode;
        outs[i].centers_rerun = runs[i].rerun;
        outs[i].ms_kernel = ms_kern;
        any_overflow |= overflow;
    }
    if (any_overflow)
        return fail(c, PP_ERR_OVERFLOW, "a CNTNB neighbour count exceeded cntnb_cap");
    return PP_OK;
}

// Device bytes of one job's arena (admission control of the serving loop).
size_t job_footprint(const pp_cloud *cl, const JobPlan &p)
{
    JobRun r;
    r.plan = p;
    r.cloud = cl;
    size_t nc = 1;
    double w = (double)p.R / std::max(p.r, 1), wyz = w / p.ay;
    if (cl->resident && !cl->lists.empty())
        for (const CellList &L : cl->lists)
            nc = std::max<size_t>(nc, L.g.ncells()), w = L.g.w, wyz = L.g.wyz;
    else {
        double v = 1.0;
        for (int a = 0; a < 3; ++a)
            v *= std::floor(((double)cl->hi[a] - (double)cl->lo[a]) / (a ? w / p.ay : w)) + 1.0;
        nc = (size_t)v;
    }
    r.ncells = nc;
    r.rows = make_rows(w, wyz, p.R);
    layout_tables(r);
    Carver m(nullptr);
    layout_outputs(r, m);
    m.take<uint32_t>(r.tables_words);
    m.take<uint32_t>(cl->n + 1);  // rerun marks
    m.off += prep_bytes(cl->n, nc, p.ntrg);
    size_t cn = (p.families & PP_FAM_CNTNB) ? (size_t)p.ntasks * 65536 * 4 : 0;
    // k-th neighbour chunk buffer: at most one tile per considered center
    const size_t kb = std::min(kKbChunkBytes, (size_t)cl->n * kb_tile_bytes(p, 8) + 1024);
    return m.off + cn + (p.nplanes() ? kb : 0) + 4096;
}

pp_status upload_common(pp_ctx *c, const float *xyz, const uint8_t *type, const float *s2,
                        uint64_t n, float cell_width, hipMemcpyKind kind, pp_cloud **out)
{
    if (!c || !out)
        return fail(c, PP_ERR_INVALID_ARG, "NULL argument");
    if (n > 0 && (!xyz || !type || !s2))
        return fail(c, PP_ERR_INVALID_ARG, "NULL input arrays");
    if (n >= 0xFFFFFFF0ull)
        return fail(c, PP_ERR_INVALID_ARG, "more than 2^32 - 16 ions");
    if (cell_width < 0 || std::isnan(cell_width))
        return fail(c, PP_ERR_INVALID_ARG, "cell_width must be >= 0");
    PP_HIP(c, hipSetDevice(c->device));
    Slot &sl = c->main;
    std::unique_ptr<pp_cloud> cl(new pp_cloud());
    cl->ctx = c;
    cl->n = n;
    Carver m0(nullptr);
    m0.take<float>(3 * n);
    m0.take<uint8_t>(n);
    m0.take<float>(n);
    cl->raw_bytes = m0.off + 256;
    PP_HIP(c, hipMalloc(&cl->raw, cl->raw_bytes));
    auto cleanup = [&](pp_status st) {
        (void)hipStreamSynchronize(sl.stream);
        if (cl->raw)
            (void)hipFree(cl->raw);
        for (CellList &L : cl->lists)
            if (L.mem)
                (void)hipFree(L.mem);
        return st;
    };
    Carver m(cl->raw);
    cl->xyz = m.take<float>(3 * n);
    cl->type = m.take<uint8_t>(n);
    cl->s2 = m.take<float>(n);
    if (n > 0) {
        if (hipMemcpyAsync(cl->xyz, xyz, 12 * n, kind, sl.stream) != hipSuccess ||
            hipMemcpyAsync(cl->type, type, n, kind, sl.stream) != hipSuccess ||
            hipMemcpyAsync(cl->s2, s2, 4 * n, kind, sl.stream) != hipSuccess)
            return cleanup(fail(c, PP_ERR_HIP, "input copy failed"));
        if (ensure(c, sl, sl.scratch, 64 + 1024) != PP_OK)
            return cleanup(PP_ERR_OOM_DEVICE);
        uint32_t *bb = (uint32_t *)sl.scratch.p;
        uint32_t h[6];
        if (launch_bbox(cl->xyz, n, bb, sl.stream) != hipSuccess ||
            hipMemcpyAsync(h, bb, sizeof(h), hipMemcpyDeviceToHost, sl.stream) != hipSuccess ||
            hipStreamSynchronize(sl.stream) != hipSuccess)
            return cleanup(fail(c, PP_ERR_HIP, "bounding box failed"));
        uint32_t *tc = bb + 8;
        uint32_t htc[256];
        if (hipMemsetAsync(tc, 0, 256 * 4, sl.stream) != hipSuccess ||
            launch_typecount(cl->type, n, tc, sl.stream) != hipSuccess ||
            hipMemcpyAsync(htc, tc, sizeof(htc), hipMemcpyDeviceToHost, sl.stream) != hipSuccess ||
            hipStreamSynchronize(sl.stream) != hipSuccess)
            return cleanup(fail(c, PP_ERR_HIP, "type count failed"));
        for (int t = 0; t < 256; ++t)
            cl->type_count[t] = htc[t];
        for (int a = 0; a < 3; ++a) {
            cl->lo[a] = ord_to_float(h[a]);
            cl->hi[a] = ord_to_float(h[3 + a]);
            if (!std::isfinite(cl->lo[a]) || !std::isfinite(cl->hi[a]))
                return cleanup(fail(c, PP_ERR_INVALID_ARG, "non-finite coordinates"));
        }
        // the volume the ions occupy, on a grid of cubic cells over the box
        // (about 2^18, and at least 8 ions per cell on average, so that a
        // cell inside the tip is never empty by chance): the density rule of
        // the cell width needs the tip's density, not the box's
        {
            double ext[3], vbox = 1.0;
            for (int a = 0; a < 3; ++a) {
                ext[a] = std::max(1e-3, (double)cl->hi[a] - (double)cl->lo[a]);
                vbox *= ext[a];
            }
            const double target = std::min(262144.0, std::max(1.0, (double)n / 8.0));
            const double w = std::cbrt(vbox / target);
            OccGrid og{};
            uint64_t ncell = 1;
            double vcell = 1.0;
            for (int a = 0; a < 3; ++a) {
                // cells of about w that tile the box's extent exactly (a thin
                // box gets one layer as thick as the box)
                og.d[a] = (int)std::min(4096.0, std::max(1.0, std::round(ext[a] / w)));
                og.lo[a] = cl->lo[a];
                og.inv[a] = (double)og.d[a] / ext[a];
                vcell *= ext[a] / (double)og.d[a];
                ncell *= (uint64_t)og.d[a];
            }
            if (ensure(c, sl, sl.scratch, ncell + 1024) != PP_OK)
                return cleanup(PP_ERR_OOM_DEVICE);
            std::vector<uint8_t> hocc(ncell);
            uint8_t *occ = (uint8_t *)sl.scratch.p;
            if (hipMemsetAsync(occ, 0, ncell, sl.stream) != hipSuccess ||
                launch_occupancy(cl->xyz, n, og, occ, sl.stream) != hipSuccess ||
                hipMemcpyAsync(hocc.data(), occ, ncell, hipMemcpyDeviceToHost, sl.stream) !=
                    hipSuccess ||
                hipStreamSynchronize(sl.stream) != hipSuccess)
                return cleanup(fail(c, PP_ERR_HIP, "occupancy failed"));
            uint64_t k = 0;
            for (uint8_t v : hocc)
                k += v;
            cl->occupied = (double)k * vcell;
        }
    }
    if (cell_width > 0) {
        cl->lists.emplace_back();
        pp_status st = build_celllist(c, sl, cl.get(), cell_width, cl->lists.back());
        if (st != PP_OK)
            return cleanup(st);
        cl->resident = true;
        if (hipStreamSynchronize(sl.stream) != hipSuccess)
            return cleanup(fail(c, PP_ERR_HIP, "cell list build failed"));
    }
    *out = cl.release();
    return PP_OK;
}

} // namespace

// ---------------------------------------------------------------------------
// serving loop

namespace {

struct Request {
    uint64_t ticket = 0;
    const pp_cloud *cloud = nullptr;
    pp_job job{};
    std::vector<uint8_t> trg, env;
    std::vector<int64_t> kd;
    pp_hist_out *out = nullptr;
    size_t footprint = 0;
    double work = 0;  // estimated candidate pair evaluations
    std::chrono::steady_clock::time_point arrival;
};

// Candidate pair evaluations a job is expected to make: centers of its
// target types times the candidates of a stencil at the cloud's mean
// density (neighbour types only when they are a minority), per-center
// histogram modes weighted up.
double job_work(const pp_cloud *cl, const JobPlan &p)
{
    if (cl->n == 0)
        return 0.0;
    double centers = 0, envn = 0;
    for (int s = 0; s < p.ntrg; ++s)
        centers += (double)cl->type_count[p.trg_type[s]];
    for (int e = 0; e < p.E; ++e)
        envn += (double)cl->type_count[p.env_type[e]];
    const double frac = envn < 0.8 * (double)cl->n ? envn / (double)cl->n : 1.0;
    const double R = p.R;
    const double cand = cl->density() * frac * 4.18879020478639 * R * R * R * 2.5;
    return centers * cand * (p.percenter ? 1.5 : 1.0);
}

// Default estimated work that one serving batch may gather (about 3 ms of
// query kernel): small requests share launches, while the requests of a batch,
// which all complete when its last one does, wait little for each other.
constexpr double kBatchWork = 5.0e9;

} // namespace

struct pp_server {
    pp_ctx *ctx = nullptr;
    pp_serve_opts opts{};
    size_t budget = 0;
    std::vector<std::unique_ptr<Slot>> slots;
    std::vector<std::thread> workers;
    LaunchOrder order;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request> pending;
    size_t used = 0;
    bool stop = false;
    uint64_t next_ticket = 1, submitted = 0, completed = 0;
    uint64_t next_seq = 0;  // sequence number of the next batch formed
    std::vector<std::pair<uint64_t, int32_t>> done;

    // Moves the next batch out of `pending` (called with mu held): the
    // waiting requests in arrival order, as many as fit the remaining budget,
    // max_batch_jobs and about opts.batch_work of estimated work (a larger request
    // runs alone).  No request overtakes an earlier one, so the tail latency
    // stays that of a first-come first-served loop.  Returns false when the
    // oldest waiting request does not fit the budget yet.
    bool pick(std::vector<Request> &batch, size_t &fp)
    {
        double work = 0;
        while ((int)batch.size() < opts.max_batch_jobs && !pending.empty()) {
            Request &r = pending.front();
            if (used + r.footprint > budget)
                break;
            if (!batch.empty() && work + r.work > opts.batch_work)
                break;
            used += r.footprint;
            fp += r.footprint;
            work += r.work;
            batch.push_back(std::move(r));
            pending.pop_front();
        }
        return !batch.empty();
    }

    void worker(int i)
    {
        (void)hipSetDevice(ctx->device);
        Slot &sl = *slots[i];
        for (;;) {
            std::vector<Request> batch;
            size_t fp = 0;
            uint64_t seq = 0;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return (stop && pending.empty()) || pick(batch, fp); });
                if (batch.empty())
                    return;  // stop requested and nothing left
                seq = next_seq++;  // under mu: batches are numbered as formed
            }
            std::vector<const pp_cloud *> cls;
            std::vector<pp_job> jobs;
            std::vector<pp_hist_out> outs;
            for (Request &r : batch) {
                cls.push_back(r.cloud);
                jobs.push_back(r.job);
                outs.push_back(*r.out);
            }
            const pp_status st = run_jobs(ctx, sl, (int)batch.size(), cls.data(), jobs.data(),
                                          outs.data(), Shape(), &order, seq);
            for (size_t k = 0; k < batch.size(); ++k)
                *batch[k].out = outs[k];
            std::lock_guard<std::mutex> lk(mu);
            used -= fp;
            completed += batch.size();
            for (Request &r : batch)
                done.emplace_back(r.ticket, (int32_t)st);
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" {

pp_status pp_spatstat_create(int32_t device, pp_ctx **out)
{
    if (!out)
        return PP_ERR_INVALID_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return PP_ERR_NO_DEVICE;
    if (device < 0 || device >= ndev)
        return PP_ERR_INVALID_ARG;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess)
        return PP_ERR_HIP;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return PP_ERR_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess)
        return PP_ERR_HIP;
    std::unique_ptr<pp_ctx> c(new pp_ctx());
    c->device = device;
    c->cu_count = prop.multiProcessorCount;
    c->lds_per_block = prop.sharedMemPerBlock;
    if (slot_init(c.get(), c->main) != PP_OK || slot_init(c.get(), c->noalloc) != PP_OK) {
        slot_free(c->main);
        slot_free(c->noalloc);
        return PP_ERR_HIP;
    }
    // event pairs for the no-allocation entry point's query launches, made now
    for (int k = 0; k < 2 * kNoAllocTimedLaunches; ++k) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) {
            slot_free(c->main);
            slot_free(c->noalloc);
            return PP_ERR_HIP;
        }
        c->noalloc.qev.push_back(e);
    }
    *out = c.release();
    return PP_OK;
}

pp_status pp_spatstat_destroy(pp_ctx *c)
{
    if (!c)
        return PP_ERR_INVALID_ARG;
    (void)hipSetDevice(c->device);
    slot_free(c->main);
    slot_free(c->noalloc);
    delete c;
    return PP_OK;
}

const char *pp_last_error_string(const pp_ctx *c) { return c ? c->err.c_str() : "no context"; }

pp_status pp_hist_supported(float rmin, float rincr, float rmax)
{
    if (!(rmin >= 0.f) || !(rincr >= kEps) || !(rmax >= kEps) || !(rmin < rmax))
        return PP_ERR_INVALID_ARG;
    HistGeometry hg(rmin, rincr, rmax);
    EdgeTable et;
    if (!et.build(hg, nullptr) || !et.exact_up_to(hg, rmax))
        return PP_ERR_UNSUPPORTED;
    return PP_OK;
}

pp_status pp_hist_geometry(float rmin, float rincr, float rmax, int64_t *nbins)
{
    if (!nbins)
        return PP_ERR_INVALID_ARG;
    HistGeometry g(rmin, rincr, rmax);
    *nbins = g.nbins;
    return g.valid ? PP_OK : PP_ERR_INVALID_ARG;
}

pp_status pp_cloud_upload(pp_ctx *c, const float *xyz, const uint8_t *type,
                          const float *ion2surf_sqr, uint64_t n, float cell_width, pp_cloud **out)
{
    return upload_common(c, xyz, type, ion2surf_sqr, n, cell_width, hipMemcpyHostToDevice, out);
}

pp_status pp_cloud_from_device(pp_ctx *c, const float *d_xyz, const uint8_t *d_type,
                               const float *d_s2, uint64_t n, float cell_width, pp_cloud **out)
{
    return upload_common(c, d_xyz, d_type, d_s2, n, cell_width, hipMemcpyDeviceToDevice, out);
}

pp_status pp_cloud_add_cell_list(pp_ctx *c, pp_cloud *cl, float cell_width)
{
    if (!c || !cl || !(cell_width > 0))
        return fail(c, PP_ERR_INVALID_ARG, "bad cloud or cell width");
    if (!cl->resident)
        return fail(c, PP_ERR_INVALID_ARG, "cloud is not resident (upload it with a cell width)");
    PP_HIP(c, hipSetDevice(c->device));
    CellList L;
    pp_status st = build_celllist(c, c->main, cl, cell_width, L);
    if (st != PP_OK) {
        if (L.mem)
            (void)hipFree(L.mem);
        return st;
    }
    PP_HIP(c, hipStreamSynchronize(c->main.stream));
    cl->lists.push_back(L);
    return PP_OK;
}

pp_status pp_cloud_free(pp_ctx *c, pp_cloud *cl)
{
    if (!c || !cl)
        return PP_ERR_INVALID_ARG;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    if (cl->raw)
        (void)hipFree(cl->raw);
    for (CellList &L : cl->lists)
        if (L.mem)
            (void)hipFree(L.mem);
    for (auto &sb : cl->subs)
        if (sb.cl.mem)
            (void)hipFree(sb.cl.mem);
    delete cl;
    return PP_OK;
}

pp_status pp_cloud_count(const pp_cloud *cl, uint64_t *n)
{
    if (!cl || !n)
        return PP_ERR_INVALID_ARG;
    *n = cl->n;
    return PP_OK;
}

pp_status pp_cloud_types(pp_ctx *c, const pp_cloud *cl, uint8_t *types)
{
    if (!c || !cl || (cl->n > 0 && !types))
        return fail(c, PP_ERR_INVALID_ARG, "NULL argument");
    if (cl->n == 0)
        return PP_OK;
    PP_HIP(c, hipSetDevice(c->device));
    PP_HIP(c, hipMemcpyAsync(types, cl->type, cl->n, hipMemcpyDeviceToHost, c->main.stream));
    PP_HIP(c, hipStreamSynchronize(c->main.stream));
    return PP_OK;
}

pp_status pp_cloud_relabel(pp_ctx *c, const pp_cloud *src, const uint8_t *types, pp_cloud **out)
{
    if (!c || !src || !out || (src->n > 0 && !types))
        return fail(c, PP_ERR_INVALID_ARG, "NULL argument");
    PP_HIP(c, hipSetDevice(c->device));
    const uint64_t n = src->n;
    void *d = nullptr;
    PP_HIP(c, hipMalloc(&d, n + 256));
    std::unique_ptr<void, void (*)(void *)> own(d, [](void *p) { (void)hipFree(p); });
    if (n > 0) {
        PP_HIP(c, hipMemcpyAsync(d, types, n, hipMemcpyHostToDevice, c->main.stream));
        PP_HIP(c, hipStreamSynchronize(c->main.stream));
    }
    const float w = src->resident && !src->lists.empty() ? src->lists[0].w : 0.f;
    pp_cloud *cl = nullptr;
    pp_status st = upload_common(c, src->xyz, (const uint8_t *)d, src->s2, n, w,
                                 hipMemcpyDeviceToDevice, &cl);
    if (st != PP_OK)
        return st;
    if (src->resident)
        for (size_t i = 1; i < src->lists.size(); ++i)
            if ((st = pp_cloud_add_cell_list(c, cl, src->lists[i].w)) != PP_OK) {
                pp_cloud_free(c, cl);
                return st;
            }
    *out = cl;
    return PP_OK;
}

pp_status pp_cloud_footprint(const pp_cloud *cl, uint64_t *bytes)
{
    if (!cl || !bytes)
        return PP_ERR_INVALID_ARG;
    size_t b = cl->raw_bytes;
    for (const CellList &L : cl->lists)
        b += L.bytes;
    for (const auto &sb : cl->subs)
        b += sb.cl.bytes;
    *bytes = b;
    return PP_OK;
}

pp_status pp_spatstat_run(pp_ctx *c, const pp_cloud *cl, const pp_job *job, pp_hist_out *out)
{
    if (!c || !cl || !job || !out)
        return fail(c, PP_ERR_INVALID_ARG, "NULL argument");
    return run_jobs(c, c->main, 1, &cl, job, out);
}

pp_status pp_spatstat_run_noalloc(pp_ctx *c, const pp_cloud *cl, const pp_job *job,
                                  pp_hist_out *out)
{
    if (!c || !cl || !job || !out)
        return fail(c, PP_ERR_INVALID_ARG, "NULL argument");
    PP_HIP(c, hipSetDevice(c->device));
    void *base = nullptr;
    size_t bytes = 0;
    PP_HIP(c, noalloc_scratch(c->device, &base, &bytes));
    // one device copy of the scratch: calls on one device take turns
    static std::mutex dev_mu[kMaxDevices];
    std::lock_guard<std::mutex> lk(dev_mu[c->device]);
    Slot &sl = c->noalloc;
    for (DevBuf *b : {&sl.scratch, &sl.jobbuf, &sl.cnbuf, &sl.kbuf})
        *b = DevBuf();
    sl.fixed = (char *)base;
    sl.fixed_cap = bytes;
    sl.fixed_off = 0;
    sl.fixed_last = nullptr;
    const pp_status st = run_jobs(c, sl, 1, &cl, job, out);
    for (DevBuf *b : {&sl.scratch, &sl.jobbuf, &sl.cnbuf, &sl.kbuf})
        *b = DevBuf();
    return st;
}

pp_status pp_batch_run(pp_ctx *c, int32_t njobs, const pp_cloud *const *clouds, const pp_job *jobs,
                       pp_hist_out *outs)
{
    if (!c || njobs < 0 || (njobs > 0 && (!clouds || !jobs || !outs)))
        return fail(c, PP_ERR_INVALID_ARG, "NULL argument");
    if (njobs > 65535)
        return fail(c, PP_ERR_INVALID_ARG, "at most 65535 jobs per batch");
    if (njobs == 0)
        return PP_OK;
    return run_jobs(c, c->main, njobs, clouds, jobs, outs);
}

pp_status pp_set_tuning(pp_ctx *c, int32_t centers_per_tile, int32_t cells_per_rmax)
{
    if (!c)
        return PP_ERR_INVALID_ARG;
    const bool autot = centers_per_tile == -1 && cells_per_rmax == -1;
    if (!autot && !(centers_per_tile == 0 || centers_per_tile == 4 || centers_per_tile == 8 ||
                    centers_per_tile == 16 || centers_per_tile == 32))
        return fail(c, PP_ERR_INVALID_ARG,
                    "centers_per_tile must be 0, 4, 8, 16 or 32 (-1: autotune)");
    if (!autot && (cells_per_rmax < 0 || cells_per_rmax > 8))
        return fail(c, PP_ERR_INVALID_ARG, "cells_per_rmax must be 0..8 (-1: autotune)");
    std::lock_guard<std::mutex> lk(c->mu);
    c->tune_T = centers_per_tile;
    c->tune_r = cells_per_rmax;
    return PP_OK;
}

pp_status pp_set_pair_symmetry(pp_ctx *c, int32_t on)
{
    if (!c)
        return PP_ERR_INVALID_ARG;
    if (on != 0 && on != 1)
        return fail(c, PP_ERR_INVALID_ARG, "on must be 0 or 1");
    std::lock_guard<std::mutex> lk(c->mu);
    c->sym_pairs = on;
    return PP_OK;
}

pp_status pp_set_cell_aspect(pp_ctx *c, int32_t yz_per_x)
{
    if (!c)
        return PP_ERR_INVALID_ARG;
    if (yz_per_x < 0 || yz_per_x > 4)
        return fail(c, PP_ERR_INVALID_ARG, "yz_per_x must be 0..4");
    std::lock_guard<std::mutex> lk(c->mu);
    c->tune_ay = yz_per_x;
    return PP_OK;
}

namespace {

// ion2surf of n ions against ntri triangles; dev: the three arrays are device
// memory, else host memory (staged through one device allocation).
pp_status ion2surf_run(pp_ctx *c, const float *xyz, uint64_t n, const float *tri, uint64_t ntri,
                       float R, float *out, bool dev)
{
    if (!c)
        return PP_ERR_INVALID_ARG;
    if (!(R > 0.f) || !std::isfinite(R))
        return fail(c, PP_ERR_INVALID_ARG, "R must be positive and finite");
    if ((n > 0 && (!xyz || !out)) || (ntri > 0 && !tri))
        return fail(c, PP_ERR_INVALID_ARG, "NULL input or output arrays");
    if (n >= 0xFFFFFFF0ull || ntri >= 0xFFFFFFF0ull)
        return fail(c, PP_ERR_INVALID_ARG, "more than 2^32 - 16 ions or triangles");
    if (n == 0)
        return PP_OK;
    PP_HIP(c, hipSetDevice(c->device));
    Slot &sl = c->main;
    hipStream_t s = sl.stream;
    const float r2 = R * R;
    if (ntri == 0 && !dev) {
        // no hull: every ion keeps fl32(R * R)
        std::fill(out, out + n, r2);
        return PP_OK;
    }
    Carver m0(nullptr);
    if (!dev) {
        m0.take<float>(3 * n);
        m0.take<float>(n);
        m0.take<float>(9 * ntri + 1);
    }
    m0.take<float>(6 * ntri + 1);
    m0.take<uint32_t>(8);
    void *mem = nullptr;
    PP_HIP(c, hipMalloc(&mem, m0.off + 256));
    std::unique_ptr<void, void (*)(void *)> own(mem, [](void *p) { (void)hipFree(p); });
    Carver m(mem);
    const float *d_xyz = xyz, *d_tri = tri;
    float *d_out = out;
    if (!dev) {
        float *x = m.take<float>(3 * n);
        d_out = m.take<float>(n);
        float *t = m.take<float>(9 * ntri + 1);
        PP_HIP(c, hipMemcpyAsync(x, xyz, 12 * n, hipMemcpyHostToDevice, s));
        PP_HIP(c, hipMemcpyAsync(t, tri, 36 * ntri, hipMemcpyHostToDevice, s));
        d_xyz = x;
        d_tri = t;
    }
    float *d_box = m.take<float>(6 * ntri + 1);
    uint32_t *d_bb = m.take<uint32_t>(8);
    if (ntri == 0) {
        PP_HIP(c, launch_fill(d_out, n, r2, s));
        PP_HIP(c, hipStreamSynchronize(s));
        return PP_OK;
    }
    // grid over the ions and the triangle vertices, cells of at least R
    uint32_t hb[12];
    PP_HIP(c, launch_bbox(d_xyz, n, d_bb, s));
    PP_HIP(c, hipMemcpyAsync(hb, d_bb, 24, hipMemcpyDeviceToHost, s));
    PP_HIP(c, hipStreamSynchronize(s));
    PP_HIP(c, launch_bbox(d_tri, 3 * ntri, d_bb, s));
    PP_HIP(c, hipMemcpyAsync(hb + 6, d_bb, 24, hipMemcpyDeviceToHost, s));
    PP_HIP(c, hipStreamSynchronize(s));
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(ord_to_float(hb[a]), ord_to_float(hb[6 + a]));
        hi[a] = std::max(ord_to_float(hb[3 + a]), ord_to_float(hb[9 + a]));
        if (!std::isfinite(lo[a]) || !std::isfinite(hi[a]))
            return fail(c, PP_ERR_INVALID_ARG, "non-finite coordinates");
    }
    // cells of at least R/2 holding a few hundred ions: a tile is one cell's
    // ions against the triangles whose R-inflated box touches the cell
    double w = 0.5 * (double)R, M = 0.0;
    {
        double vol = 1.0;
        for (int a = 0; a < 3; ++a) {
            vol *= std::max(hi[a] - lo[a], 1e-3);
            M = std::max(M, std::max(std::fabs(lo[a]), std::fabs(hi[a])));
        }
        w = std::max(w, std::cbrt(256.0 * vol / (double)n));
    }
    GridGeom g{};
    for (;;) {
        g.nx = (int)std::floor((hi[0] - lo[0]) / w) + 1;
        g.ny = (int)std::floor((hi[1] - lo[1]) / w) + 1;
        g.nz = (int)std::floor((hi[2] - lo[2]) / w) + 1;
        if (g.ncells() <= (1ull << 24))
            break;
        w *= 1.25;
    }
    g.w = w;
    g.inv_w = 1.0 / w;
    g.wyz = w;
    g.inv_wyz = g.inv_w;
    for (int a = 0; a < 3; ++a)
        g.lo[a] = lo[a];
    const uint64_t nc = g.ncells();
    Carver t0(nullptr);
    t0.take<uint32_t>(nc + 1);
    t0.take<uint32_t>(nc + 1);
    t0.take<uint32_t>(nc + 1);
    t0.take<uint32_t>(scan_workspace_words(std::max<uint64_t>(nc, n)));
    t0.take<uint32_t>(n);
    t0.take<uint32_t>(n);
    t0.take<float4>(n);
    void *tm = nullptr;
    PP_HIP(c, hipMalloc(&tm, t0.off + 256));
    std::unique_ptr<void, void (*)(void *)> own_t(tm, [](void *p) { (void)hipFree(p); });
    Carver t(tm);
    uint32_t *count = t.take<uint32_t>(nc + 1);
    uint32_t *tcs = t.take<uint32_t>(nc + 1);
    uint32_t *ics = t.take<uint32_t>(nc + 1);
    uint32_t *ws = t.take<uint32_t>(scan_workspace_words(std::max<uint64_t>(nc, n)));
    uint32_t *key = t.take<uint32_t>(n);
    uint32_t *rank = t.take<uint32_t>(n);
    float4 *ions = t.take<float4>(n);
    // ions in cell order
    PP_HIP(c, hipMemsetAsync(count, 0, 4 * (nc + 1), s));
    PP_HIP(c, launch_cellkeys(d_xyz, n, g, count, key, rank, s));
    PP_HIP(c, scan_exclusive(count, ics, nc, ws, s));
    PP_HIP(c, launch_ionscatter(d_xyz, n, key, rank, ics, ions, s));
    // triangle lists per cell
    PP_HIP(c, hipMemsetAsync(count, 0, 4 * (nc + 1), s));
    PP_HIP(c, launch_tribox(d_tri, ntri, R, g, d_box, count, s));
    std::vector<uint32_t> h_ics(nc + 1), h_tcs(nc + 1);
    // the per-cell triangle counts, summed in 64 bits: the 32-bit offsets of
    // the cell lists must not wrap
    PP_HIP(c, hipMemcpyAsync(h_tcs.data(), count, 4 * nc, hipMemcpyDeviceToHost, s));
    PP_HIP(c, hipStreamSynchronize(s));
    uint64_t total = 0;
    for (uint64_t k = 0; k < nc; ++k)
        total += h_tcs[k];
    if (total > 0xFFFFFFF0ull)
        return fail(c, PP_ERR_UNSUPPORTED,
                    "triangles too large for the surface grid (over 2^32 cell entries)");
    PP_HIP(c, scan_exclusive(count, tcs, nc, ws, s));
    PP_HIP(c, hipMemcpyAsync(h_ics.data(), ics, 4 * (nc + 1), hipMemcpyDeviceToHost, s));
    PP_HIP(c, hipMemcpyAsync(h_tcs.data(), tcs, 4 * (nc + 1), hipMemcpyDeviceToHost, s));
    PP_HIP(c, hipStreamSynchronize(s));
    const uint32_t entries = h_tcs[nc];
    std::vector<SurfTile> tiles;
    for (uint64_t k = 0; k < nc; ++k) {
        if (h_tcs[k + 1] == h_tcs[k])
            continue;
        for (uint32_t f = h_ics[k]; f < h_ics[k + 1]; f += 256)
            tiles.push_back({(uint32_t)k, f, std::min<uint32_t>(256, h_ics[k + 1] - f), 0});
    }
    void *lm = nullptr;
    PP_HIP(c, hipMalloc(&lm, 4 * (size_t)std::max<uint32_t>(entries, 1) +
                                 sizeof(SurfTile) * (tiles.size() + 1) + 512));
    std::unique_ptr<void, void (*)(void *)> own_l(lm, [](void *p) { (void)hipFree(p); });
    Carver l(lm);
    uint32_t *list = l.take<uint32_t>(std::max<uint32_t>(entries, 1));
    SurfTile *d_tiles = l.take<SurfTile>(tiles.size() + 1);
    PP_HIP(c, hipMemcpyAsync(count, tcs, 4 * (nc + 1), hipMemcpyDeviceToDevice, s));
    PP_HIP(c, launch_trifill(d_box, ntri, R, g, count, list, s));
    if (!tiles.empty())
        PP_HIP(c, hipMemcpyAsync(d_tiles, tiles.data(), sizeof(SurfTile) * tiles.size(),
                                 hipMemcpyHostToDevice, s));
    PP_HIP(c, launch_fill(d_out, n, r2, s));
    // delta bounds how far a computed closest point can stray outside the
    // triangle's box: a few roundings of coordinates up to M in magnitude
    const float delta = (float)(std::ldexp(M, -16));
    PP_HIP(c, launch_surftile(ions, d_tiles, (uint32_t)tiles.size(), tcs, list, d_tri, d_box, R,
                              delta, d_out, s));
    if (!dev)
        PP_HIP(c, hipMemcpyAsync(out, d_out, 4 * n, hipMemcpyDeviceToHost, s));
    PP_HIP(c, hipStreamSynchronize(s));
    return PP_OK;
}

} // namespace

pp_status pp_ion2surf(pp_ctx *c, const float *xyz, uint64_t n, const float *tri, uint64_t ntri,
                      float R, float *out)
{
    return ion2surf_run(c, xyz, n, tri, ntri, R, out, false);
}

pp_status pp_ion2surf_device(pp_ctx *c, const float *d_xyz, uint64_t n, const float *d_tri,
                             uint64_t ntri, float R, float *d_out)
{
    return ion2surf_run(c, d_xyz, n, d_tri, ntri, R, d_out, true);
}

namespace {

// POS (words = 4) or EPOS (words = 11) decoding and ranging; dev: pos and
// the outputs are device memory.
pp_status pos_decode_run(pp_ctx *c, const void *pos, uint64_t nbytes, const float *lo,
                         const float *hi, const uint8_t *rt, int32_t nranges, float *xyz,
                         float *mq, uint8_t *type, bool dev, int words = 4)
{
    if (!c)
        return PP_ERR_INVALID_ARG;
    const uint64_t rb = 4ull * (uint64_t)words;
    if (nbytes % rb != 0)
        return fail(c, PP_ERR_INVALID_ARG,
                    words == 4 ? "POS size is not an integer multiple of the 16-byte record"
                               : "EPOS size is not an integer multiple of the 44-byte record");
    const uint64_t n = nbytes / rb;
    if (n > 0xFFFFFFFDull)
        return fail(c, PP_ERR_INVALID_ARG, "more than 2^32 - 3 ions");
    if (nranges < 0 || nranges > kMaxRanges)
        return fail(c, PP_ERR_INVALID_ARG, "0 to 4096 ranges");
    if (nranges > 0 && (!lo || !hi || !rt))
        return fail(c, PP_ERR_INVALID_ARG, "NULL range arrays");
    if (n > 0 && (!pos || !xyz || !type))
        return fail(c, PP_ERR_INVALID_ARG, "NULL input or output arrays");
    if (dev && ((uintptr_t)pos & (words == 4 ? 15u : 3u)))
        return fail(c, PP_ERR_INVALID_ARG, "device record buffer misaligned");
    if (n == 0)
        return PP_OK;
    PP_HIP(c, hipSetDevice(c->device));
    hipStream_t s = c->main.stream;
    Carver m0(nullptr);
    m0.take<float>(2 * (size_t)nranges + 1);
    m0.take<uint8_t>((size_t)nranges + 1);
    if (!dev) {
        m0.take<uint32_t>((size_t)words * n);
        m0.take<float>(3 * n);
        m0.take<float>(n);
        m0.take<uint8_t>(n);
    }
    void *mem = nullptr;
    PP_HIP(c, hipMalloc(&mem, m0.off + 256));
    std::unique_ptr<void, void (*)(void *)> own(mem, [](void *p) { (void)hipFree(p); });
    Carver m(mem);
    float *d_rng = m.take<float>(2 * (size_t)nranges + 1);
    uint8_t *d_rt = m.take<uint8_t>((size_t)nranges + 1);
    if (nranges > 0) {
        PP_HIP(c, hipMemcpyAsync(d_rng, lo, 4 * (size_t)nranges, hipMemcpyHostToDevice, s));
        PP_HIP(c, hipMemcpyAsync(d_rng + nranges, hi, 4 * (size_t)nranges, hipMemcpyHostToDevice,
                                 s));
        PP_HIP(c, hipMemcpyAsync(d_rt, rt, (size_t)nranges, hipMemcpyHostToDevice, s));
    }
    const uint32_t *d_pos = (const uint32_t *)pos;
    float *d_xyz = xyz, *d_mq = mq;
    uint8_t *d_type = type;
    if (!dev) {
        uint32_t *p = m.take<uint32_t>((size_t)words * n);
        d_xyz = m.take<float>(3 * n);
        float *q = m.take<float>(n);
        d_type = m.take<uint8_t>(n);
        d_mq = mq ? q : nullptr;
        PP_HIP(c, hipMemcpyAsync(p, pos, nbytes, hipMemcpyHostToDevice, s));
        d_pos = p;
    }
    PP_HIP(c, launch_posdecode(d_pos, n, words, d_rng, d_rng + nranges, d_rt, nranges, d_xyz,
                               d_mq, d_type, s));
    if (!dev) {
        PP_HIP(c, hipMemcpyAsync(xyz, d_xyz, 12 * n, hipMemcpyDeviceToHost, s));
        if (mq)
            PP_HIP(c, hipMemcpyAsync(mq, d_mq, 4 * n, hipMemcpyDeviceToHost, s));
        PP_HIP(c, hipMemcpyAsync(type, d_type, n, hipMemcpyDeviceToHost, s));
    }
    PP_HIP(c, hipStreamSynchronize(s));
    return PP_OK;
}

} // namespace

pp_status pp_pos_decode(pp_ctx *c, const void *pos, uint64_t nbytes, const float *lo,
                        const float *hi, const uint8_t *rt, int32_t nranges, float *xyz,
                        float *mq, uint8_t *type)
{
    return pos_decode_run(c, pos, nbytes, lo, hi, rt, nranges, xyz, mq, type, false);
}

pp_status pp_pos_decode_device(pp_ctx *c, const void *d_pos, uint64_t nbytes, const float *lo,
                               const float *hi, const uint8_t *rt, int32_t nranges, float *d_xyz,
                               float *d_mq, uint8_t *d_type)
{
    return pos_decode_run(c, d_pos, nbytes, lo, hi, rt, nranges, d_xyz, d_mq, d_type, true);
}

namespace {

pp_status cloud_from_records(pp_ctx *c, const void *pos, uint64_t nbytes, int words,
                             const float *lo, const float *hi, const uint8_t *rt,
                             int32_t nranges, const float *tri, uint64_t ntri, float R,
                             float cell_width, pp_cloud **out)
{
    if (!c || !out || (nbytes > 0 && !pos) || (ntri > 0 && !tri))
        return fail(c, PP_ERR_INVALID_ARG, "NULL argument");
    if (nbytes % (4ull * words) != 0)
        return fail(c, PP_ERR_INVALID_ARG, "file size is not an integer multiple of the record");
    if (ntri >= 0xFFFFFFF0ull)
        return fail(c, PP_ERR_INVALID_ARG, "more than 2^32 - 16 triangles");
    const uint64_t n = nbytes / (4ull * words);
    PP_HIP(c, hipSetDevice(c->device));
    hipStream_t s = c->main.stream;
    Carver m0(nullptr);
    m0.take<uint32_t>((size_t)words * n + 4);
    m0.take<float>(3 * n + 1);
    m0.take<uint8_t>(n + 1);
    m0.take<float>(n + 1);
    m0.take<float>(9 * ntri + 1);
    void *mem = nullptr;
    PP_HIP(c, hipMalloc(&mem, m0.off + 256));
    std::unique_ptr<void, void (*)(void *)> own(mem, [](void *p) { (void)hipFree(p); });
    Carver m(mem);
    uint32_t *d_pos = m.take<uint32_t>((size_t)words * n + 4);
    float *d_xyz = m.take<float>(3 * n + 1);
    uint8_t *d_type = m.take<uint8_t>(n + 1);
    float *d_s2 = m.take<float>(n + 1);
    float *d_tri = m.take<float>(9 * ntri + 1);
    if (n > 0)
        PP_HIP(c, hipMemcpyAsync(d_pos, pos, nbytes, hipMemcpyHostToDevice, s));
    if (ntri > 0)
        PP_HIP(c, hipMemcpyAsync(d_tri, tri, 36 * ntri, hipMemcpyHostToDevice, s));
    pp_status st = pos_decode_run(c, d_pos, nbytes, lo, hi, rt, nranges, d_xyz, nullptr, d_type,
                                  true, words);
    if (st != PP_OK)
        return st;
    if ((st = ion2surf_run(c, d_xyz, n, d_tri, ntri, R, d_s2, true)) != PP_OK)
        return st;
    return upload_common(c, d_xyz, d_type, d_s2, n, cell_width, hipMemcpyDeviceToDevice, out);
}

} // namespace

pp_status pp_cloud_from_pos(pp_ctx *c, const void *pos, uint64_t nbytes, const float *lo,
                            const float *hi, const uint8_t *rt, int32_t nranges,
                            const float *tri, uint64_t ntri, float R, float cell_width,
                            pp_cloud **out)
{
    return cloud_from_records(c, pos, nbytes, 4, lo, hi, rt, nranges, tri, ntri, R, cell_width,
                              out);
}

pp_status pp_cloud_from_epos(pp_ctx *c, const void *epos, uint64_t nbytes, const float *lo,
                             const float *hi, const uint8_t *rt, int32_t nranges,
                             const float *tri, uint64_t ntri, float R, float cell_width,
                             pp_cloud **out)
{
    return cloud_from_records(c, epos, nbytes, 11, lo, hi, rt, nranges, tri, ntri, R, cell_width,
                              out);
}

pp_status pp_epos_decode(pp_ctx *c, const void *epos, uint64_t nbytes, const float *lo,
                         const float *hi, const uint8_t *rt, int32_t nranges, float *xyz,
                         float *mq, uint8_t *type)
{
    return pos_decode_run(c, epos, nbytes, lo, hi, rt, nranges, xyz, mq, type, false, 11);
}

int64_t pp_npcorr_geometry(float rmax, float rincr)
{
    if (!(rmax > 0.f) || !(rincr > 0.f) || !std::isfinite(rmax) || !std::isfinite(rincr))
        return -1;
    // npc3d::npc3d (NPointCorrelations3D.cpp:50-60): float ceil of the float
    // quotient, one guard layer, odd edge
    const float realhalf = std::ceil(rmax / rincr) + 1.f;
    if (!(realhalf < 1e6f))
        return -1;
    const int64_t half = (int64_t)realhalf;
    const int64_t edge = 1 + 2 * half;
    // MAXIMUM_NPC3D_BIN_RESOLUTION (src/PARAPROBE_Numerics.h:73; S.cpp:3706-3711)
    if (edge * edge * edge > 503ll * 503ll * 503ll)
        return -1;
    return edge;
}

pp_status pp_npcorr_run(pp_ctx *c, const pp_cloud *cl, int32_t ntasks, const uint8_t *trg_mask,
                        const uint8_t *env_mask, float rmax, float rincr, int64_t knn_order,
                        uint32_t *counts, int64_t *centers_binned)
{
    if (!c || !cl || ntasks < 0 || ntasks > kMaxTasks || (ntasks > 0 && (!trg_mask || !env_mask)))
        return fail(c, PP_ERR_INVALID_ARG, "bad NPCORR arguments");
    if (knn_order < 0 || knn_order >= kNpcCap - 64)
        return fail(c, PP_ERR_UNSUPPORTED, "NPCORR: SpatStatKNNOrder must be below 448");
    const int64_t edge = pp_npcorr_geometry(rmax, rincr);
    if (edge < 0)
        return fail(c, PP_ERR_UNSUPPORTED,
                    "Computation of 2-point statistics would require too large buffers");
    if (ntasks > 0 && !counts)
        return fail(c, PP_ERR_INVALID_ARG, "NULL counts");
    const uint64_t nxyz = (uint64_t)edge * edge * edge;
    if (centers_binned)
        *centers_binned = 0;
    if (ntasks == 0)
        return PP_OK;
    std::fill(counts, counts + (size_t)ntasks * nxyz, 0u);
    if (cl->n == 0)
        return PP_OK;
    PP_HIP(c, hipSetDevice(c->device));
    Slot &sl = c->main;
    hipStream_t s = sl.stream;
    // a cell list at least R/2 wide: the 5 x 5 x 5 cells around a center
    // hold its ball
    CellList L;
    pp_status st = build_celllist(c, sl, const_cast<pp_cloud *>(cl),
                                  0.5f * rmax * (1.0f + std::ldexp(1.0f, -10)), L);
    std::unique_ptr<void, void (*)(void *)> own_l(L.mem, [](void *p) {
        if (p)
            (void)hipFree(p);
    });
    if (st != PP_OK)
        return st;
    // each task's candidates: the ions of its neighbour types on the same
    // grid (one list per distinct type set)
    std::vector<CellList> envl;
    std::vector<int> lidx(ntasks);
    struct FreeLists {
        std::vector<CellList> &v;
        ~FreeLists()
        {
            for (CellList &E : v)
                if (E.mem)
                    (void)hipFree(E.mem);
        }
    } free_lists{envl};
    std::vector<std::array<uint32_t, 8>> masks;
    for (int32_t t = 0; t < ntasks; ++t) {
        std::array<uint32_t, 8> mk{};
        std::memcpy(mk.data(), env_mask + 32 * (size_t)t, 32);
        int k = 0;
        while (k < (int)masks.size() && masks[k] != mk)
            ++k;
        if (k == (int)masks.size()) {
            masks.push_back(mk);
            envl.emplace_back();
            st = build_typed_list(c, sl, cl, &L, mk.data(), envl.back());
            if (st != PP_OK)
                return st;
        }
        lidx[t] = k;
    }
    void *mem = nullptr;
    const size_t cbytes = 4 * (size_t)ntasks * nxyz;
    Carver m0(nullptr);
    m0.take<uint32_t>((size_t)ntasks * nxyz);
    m0.take<uint32_t>(8 * (size_t)ntasks);
    m0.take<uint32_t>(8 * (size_t)ntasks);
    m0.take<const float4 *>((size_t)ntasks);
    m0.take<const uint32_t *>((size_t)ntasks);
    m0.take<float>((size_t)ntasks);
    PP_HIP(c, hipMalloc(&mem, m0.off + 256));
    std::unique_ptr<void, void (*)(void *)> own(mem, [](void *p) { (void)hipFree(p); });
    Carver m(mem);
    uint32_t *d_counts = m.take<uint32_t>((size_t)ntasks * nxyz);
    uint32_t *d_trg = m.take<uint32_t>(8 * (size_t)ntasks);
    uint32_t *d_env = m.take<uint32_t>(8 * (size_t)ntasks);
    const float4 **d_epos = m.take<const float4 *>((size_t)ntasks);
    const uint32_t **d_ecs = m.take<const uint32_t *>((size_t)ntasks);
    float *d_r1 = m.take<float>((size_t)ntasks);
    std::vector<const float4 *> h_epos(ntasks);
    std::vector<const uint32_t *> h_ecs(ntasks);
    // the first counting radius: 1.4 times the radius that holds K + 2
    // neighbour-type ions at the cloud's mean density (R when that is not
    // well below R)
    std::vector<float> h_r1(ntasks);
    const double dens = cl->density();
    for (int32_t t = 0; t < ntasks; ++t) {
        h_epos[t] = envl[lidx[t]].pos;
        h_ecs[t] = envl[lidx[t]].cellstart;
        double m_env = 0;
        for (int ty = 0; ty < 256; ++ty)
            if ((env_mask[32 * (size_t)t + ty / 8] >> (ty % 8)) & 1u)
                m_env += (double)cl->type_count[ty];
        const double rho = dens * m_env / (double)cl->n;
        const double r = rho > 0 ? 1.4 * std::cbrt(3.0 * (double)(knn_order + 2) /
                                                   (4.0 * 3.14159265358979 * rho))
                                 : (double)rmax;
        h_r1[t] = r < 0.8 * (double)rmax ? (float)r : rmax;
    }
    PP_HIP(c, hipMemsetAsync(d_counts, 0, cbytes, s));
    PP_HIP(c, hipMemcpyAsync(d_trg, trg_mask, 32 * (size_t)ntasks, hipMemcpyHostToDevice, s));
    PP_HIP(c, hipMemcpyAsync(d_env, env_mask, 32 * (size_t)ntasks, hipMemcpyHostToDevice, s));
    PP_HIP(c, hipMemcpyAsync(d_epos, h_epos.data(), sizeof(void *) * (size_t)ntasks,
                             hipMemcpyHostToDevice, s));
    PP_HIP(c, hipMemcpyAsync(d_ecs, h_ecs.data(), sizeof(void *) * (size_t)ntasks,
                             hipMemcpyHostToDevice, s));
    PP_HIP(c, hipMemcpyAsync(d_r1, h_r1.data(), 4 * (size_t)ntasks, hipMemcpyHostToDevice, s));
    NpcArgs a{};
    a.pos = L.pos;
    a.s2 = L.s2;
    a.epos = d_epos;
    a.ecs = d_ecs;
    a.r1 = d_r1;
    a.g = L.g;
    a.n = cl->n;
    a.R2 = rmax * rmax;
    a.ntasks = ntasks;
    a.trg = d_trg;
    a.env = d_env;
    a.K = (uint32_t)knn_order;
    a.maxlen = rmax;
    a.inv_width = 1.f / rincr;
    a.inv_bucket = (float)kNpcBuckets / a.R2;
    a.half = (int)((edge - 1) / 2);
    a.edge = (uint64_t)edge;
    a.nxyz = nxyz;
    a.counts = d_counts;
    PP_HIP(c, launch_npcorr(a, s));
    PP_HIP(c, hipMemcpyAsync(counts, d_counts, cbytes, hipMemcpyDeviceToHost, s));
    PP_HIP(c, hipStreamSynchronize(s));
    if (centers_binned) {
        int64_t tot = 0;
        for (size_t i = 0; i < (size_t)ntasks * nxyz; ++i)
            tot += counts[i];
        *centers_binned = tot;
    }
    return PP_OK;
}

pp_status pp_set_graph_launch(pp_ctx *c, int32_t on)
{
    if (!c)
        return PP_ERR_INVALID_ARG;
    if (on != 0 && on != 1)
        return fail(c, PP_ERR_INVALID_ARG, "graph launch flag must be 0 or 1");
    std::lock_guard<std::mutex> lk(c->mu);
    c->graph_launch = on;
    return PP_OK;
}

pp_status pp_set_knn_range(pp_ctx *c, float factor)
{
    if (!c)
        return PP_ERR_INVALID_ARG;
    if (!(factor >= 0.f) || factor > 1e6f)
        return fail(c, PP_ERR_INVALID_ARG, "kNN range factor must be >= 0");
    std::lock_guard<std::mutex> lk(c->mu);
    c->knn_factor = factor;
    return PP_OK;
}

pp_status pp_serve_create(pp_ctx *c, const pp_serve_opts *opts, pp_server **out)
{
    if (!c || !opts || !out)
        return fail(c, PP_ERR_INVALID_ARG, "NULL argument");
    if (opts->slots < 1 || opts->slots > 8 || opts->max_batch_jobs < 1 ||
        opts->max_batch_jobs > 65535)
        return fail(c, PP_ERR_INVALID_ARG, "slots must be 1..8, max_batch_jobs 1..65535");
    if (!(opts->batch_work >= 0) || !std::isfinite(opts->batch_work))
        return fail(c, PP_ERR_INVALID_ARG, "batch_work must be finite and >= 0");
    PP_HIP(c, hipSetDevice(c->device));
    std::unique_ptr<pp_server> s(new pp_server());
    s->ctx = c;
    s->opts = *opts;
    if (s->opts.batch_work == 0)
        s->opts.batch_work = kBatchWork;
    s->budget = opts->hbm_budget_bytes;
    if (s->budget == 0) {
        size_t fr = 0, tot = 0;
        PP_HIP(c, hipMemGetInfo(&fr, &tot));
        s->budget = fr / 2;
    }
    for (int i = 0; i < opts->slots; ++i) {
        s->slots.emplace_back(new Slot());
        if (slot_init(c, *s->slots.back()) != PP_OK) {
            for (auto &sl : s->slots)
                slot_free(*sl);
            return PP_ERR_HIP;
        }
    }
    if (hipEventCreateWithFlags(&s->order.ev, hipEventDisableTiming) != hipSuccess) {
        for (auto &sl : s->slots)
            slot_free(*sl);
        return PP_ERR_HIP;
    }
    for (int i = 0; i < opts->slots; ++i)
        s->workers.emplace_back(&pp_server::worker, s.get(), i);
    *out = s.release();
    return PP_OK;
}

pp_status pp_serve_submit(pp_server *s, const pp_cloud *cl, const pp_job *job, pp_hist_out *out,
                          uint64_t *ticket)
{
    if (!s || !cl || !job || !out || !ticket)
        return PP_ERR_INVALID_ARG;
    if (!cl->resident)
        return fail(s->ctx, PP_ERR_INVALID_ARG, "serving needs resident clouds");
    Request r;
    JobPlan p;
    pp_status st = plan_job(s->ctx, job, p);
    if (st != PP_OK)
        return st;
    p.r = rule_r(cand_density(cl, p), p.R);
    r.cloud = cl;
    r.job = *job;
    const int nt = std::max(job->ntasks, 0);
    r.trg.assign(job->trg_mask, job->trg_mask + 32 * (size_t)nt);
    r.env.assign(job->env_mask, job->env_mask + 32 * (size_t)nt);
    if (job->nk > 0 && job->kminus1_desc)
        r.kd.assign(job->kminus1_desc, job->kminus1_desc + job->nk);
    r.job.trg_mask = r.trg.data();
    r.job.env_mask = r.env.data();
    r.job.kminus1_desc = r.kd.empty() ? nullptr : r.kd.data();
    r.out = out;
    r.footprint = job_footprint(cl, p);
    r.work = job_work(cl, p);
    r.arrival = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (r.footprint > s->budget)
        return fail(s->ctx, PP_ERR_ADMISSION_REJECTED, "request larger than the HBM budget");
    r.ticket = s->next_ticket++;
    *ticket = r.ticket;
    ++s->submitted;
    s->pending.push_back(std::move(r));
    s->cv_work.notify_one();
    return PP_OK;
}

pp_status pp_serve_poll(pp_server *s, uint64_t *tickets, int32_t *statuses, int32_t cap,
                        int32_t *n)
{
    if (!s || !n || (cap > 0 && (!tickets || !statuses)))
        return PP_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    const int32_t k = (int32_t)std::min<size_t>((size_t)std::max(cap, 0), s->done.size());
    for (int32_t i = 0; i < k; ++i) {
        tickets[i] = s->done[i].first;
        statuses[i] = s->done[i].second;
    }
    s->done.erase(s->done.begin(), s->done.begin() + k);
    *n = k;
    return PP_OK;
}

pp_status pp_serve_drain(pp_server *s)
{
    if (!s)
        return PP_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->completed == s->submitted; });
    return PP_OK;
}

pp_status pp_serve_destroy(pp_server *s)
{
    if (!s)
        return PP_ERR_INVALID_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers)
        t.join();
    (void)hipSetDevice(s->ctx->device);
    for (auto &sl : s->slots)
        slot_free(*sl);
    if (s->order.ev)
        (void)hipEventDestroy(s->order.ev);
    delete s;
    return PP_OK;
}

} // extern "C"
