// pp_capi.cpp -- host runtime behind include/pp_spatstat.h.
//
// Owns the device, the resident clouds and their cell lists, turns a pp_job
// (the settings and tasks of horderdist::compute_generic_spatstat3,
// src/PARAPROBE_SolverHdl.cpp:3685-4296) into device tables, runs the cell
// list, center, tile and query kernels, and composes the reference's
// histograms from the device counters.
#include "../../include/pp_spatstat.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
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
    float w = 0;
    uint32_t *cellstart = nullptr;  // ncells + 1
    float4 *pos = nullptr;          // n
    float *s2 = nullptr;            // n
    void *mem = nullptr;
    size_t bytes = 0;
};

} // namespace

struct pp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int cu_count = 0;
    size_t lds_per_block = 0;
    int tune_T = 0, tune_r = 0;
    DevBuf scratch;   // per-call scratch (cell-list temporaries, job arenas)
    DevBuf jobbuf;    // JobDev array + tile pointer table + tile offsets + queue
    DevBuf cnbuf;     // CNTNB counters
    hipEvent_t ev[6] = {};
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
    bool resident = false;    // cell list fixed at upload
    CellList cl;              // resident or cached per-job cell list
};

namespace {

pp_status fail(pp_ctx *c, pp_status st, const std::string &msg)
{
    if (c)
        c->err = msg;
    return st;
}

#define PP_HIP(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail((ctx), _e == hipErrorOutOfMemory ? PP_ERR_OOM_DEVICE : PP_ERR_HIP,   \
                        std::string(#expr) + ": " + hipGetErrorString(_e));                  \
    } while (0)

pp_status ensure(pp_ctx *c, DevBuf &b, size_t bytes)
{
    if (b.cap >= bytes && b.p)
        return PP_OK;
    if (b.p) {
        PP_HIP(c, hipStreamSynchronize(c->stream));
        PP_HIP(c, hipFree(b.p));
        b.p = nullptr;
        b.cap = 0;
    }
    size_t want = bytes + bytes / 8 + (1u << 20);
    PP_HIP(c, hipMalloc(&b.p, want));
    b.cap = want;
    return PP_OK;
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
};

pp_status plan_job(pp_ctx *c, const pp_job *j, JobPlan &p)
{
    if (!j)
        return fail(c, PP_ERR_INVALID_ARG, "job is NULL");
    if (j->families & PP_FAM_NPCORR)
        return fail(c, PP_ERR_UNSUPPORTED, "NPCORR (n-point correlation) is not provided");
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
    p.cat0 = (int)p.hg.category(0.0f);
    p.cat_sentinel = (int)p.hg.category((float)((double)p.R + kEps));
    return PP_OK;
}

// Stencil rows for cell width w and radius R: cell offset (dx, dy, dz) is
// visited iff the closest points of the two cells are within R, with a
// relative margin of 2^-20 on R^2 that covers the FP32 rounding of the
// in-ball test (at most ~6e-7 relative).
std::vector<Row> make_rows(double w, double R)
{
    const double lim = R * R * (1.0 + std::ldexp(1.0, -20));
    const int dmax = (int)std::floor(std::sqrt(lim) / w) + 1;
    std::vector<Row> rows;
    for (int dz = -dmax; dz <= dmax; ++dz)
        for (int dy = -dmax; dy <= dmax; ++dy) {
            const double my = std::max(0, std::abs(dy) - 1), mz = std::max(0, std::abs(dz) - 1);
            const double rem = lim - (my * my + mz * mz) * w * w;
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

pp_status build_celllist(pp_ctx *c, pp_cloud *cl, float w, CellList &L)
{
    const uint64_t n = cl->n;
    GridGeom g{};
    g.w = w;
    // grow the cell when the grid would be huge compared with the cloud
    double ww = w;
    for (;;) {
        uint64_t tot = 1;
        int d[3];
        for (int a = 0; a < 3; ++a) {
            d[a] = (int)std::floor(((double)cl->hi[a] - (double)cl->lo[a]) / ww) + 1;
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
    for (int a = 0; a < 3; ++a)
        g.lo[a] = cl->lo[a];
    const uint64_t nc = g.ncells();
    // persistent part
    Carver m0(nullptr);
    m0.take<uint32_t>(nc + 1);
    m0.take<float4>(n);
    m0.take<float>(n);
    const size_t bytes = m0.off + 256;
    if (L.mem && L.bytes < bytes) {
        PP_HIP(c, hipStreamSynchronize(c->stream));
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
    L.g = g;
    L.w = (float)ww;
    // temporaries
    Carver t0(nullptr);
    t0.take<uint32_t>(nc);
    t0.take<uint32_t>(n);
    t0.take<uint32_t>(n);
    t0.take<uint32_t>(scan_workspace_words(nc));
    pp_status st = ensure(c, c->scratch, t0.off + 256);
    if (st != PP_OK)
        return st;
    Carver t(c->scratch.p);
    uint32_t *count = t.take<uint32_t>(nc);
    uint32_t *key = t.take<uint32_t>(n);
    uint32_t *rank = t.take<uint32_t>(n);
    uint32_t *ws = t.take<uint32_t>(scan_workspace_words(nc));
    PP_HIP(c, hipMemsetAsync(count, 0, nc * sizeof(uint32_t), c->stream));
    PP_HIP(c, launch_cellkeys(cl->xyz, n, g, count, key, rank, c->stream));
    PP_HIP(c, scan_exclusive(count, L.cellstart, nc, ws, c->stream));
    PP_HIP(c, launch_scatter(cl->xyz, cl->type, cl->s2, n, key, rank, L.cellstart, L.pos, L.s2,
                             c->stream));
    return PP_OK;
}

// Device-side state of one prepared job.
struct JobRun {
    JobPlan plan;
    const pp_cloud *cloud = nullptr;
    const CellList *cl = nullptr;
    std::vector<Row> rows;
    size_t ncells = 0;
    // device pointers (carved from the scratch arena)
    uint32_t *flags = nullptr, *cpos = nullptr, *centers = nullptr, *ntile = nullptr,
             *toff = nullptr, *ws = nullptr, *tables = nullptr, *maxcand = nullptr;
    Tile *tiles = nullptr;
    unsigned long long *P = nullptr, *stats = nullptr;
    uint32_t *K = nullptr, *N1 = nullptr, *CN = nullptr;
    int cn_cap = 0;
    // table offsets (words) inside `tables`
    size_t o_edges = 0, o_envoff = 0, o_trgslot = 0, o_ttrg = 0, o_tenv = 0, o_kidx = 0,
           o_rows = 0, tables_words = 0;
    // read back after prep
    uint32_t h_counts[4] = {0, 0, 0, 0};  // considered, tiles, maxcand, pad
    JobDev dev{};
};

void layout_job(JobRun &r, Carver &m)
{
    const JobPlan &p = r.plan;
    const uint64_t n = r.cloud->n;
    const uint64_t nc = r.ncells;
    r.flags = m.take<uint32_t>(n);
    r.cpos = m.take<uint32_t>(n + 1);
    r.centers = m.take<uint32_t>(n);
    r.ntile = m.take<uint32_t>(nc);
    r.toff = m.take<uint32_t>(nc + 1);
    r.ws = m.take<uint32_t>(scan_workspace_words(std::max<uint64_t>(n, nc)));
    // tiles: at most one per considered center
    r.tiles = m.take<Tile>(n + 1);
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
    r.o_rows = (o + 1) & ~(size_t)1;
    o = r.o_rows + 2 * r.rows.size();
    r.tables_words = o;
    r.tables = m.take<uint32_t>(o);
    r.maxcand = m.take<uint32_t>(4);
    r.P = m.take<unsigned long long>((size_t)std::max(p.ntrg * p.E * p.C, 1));
    r.stats = m.take<unsigned long long>(4);
    r.K = m.take<uint32_t>((size_t)std::max(p.ntasks * p.nku * p.C, 1));
    r.N1 = m.take<uint32_t>((size_t)std::max(p.ntasks * p.C, 1));
    // CN is carved later: its size follows from the stencil bound
}

pp_status prep_tiles(pp_ctx *c, JobRun &r, int T);

pp_status prep_job(pp_ctx *c, JobRun &r, int T)
{
    const JobPlan &p = r.plan;
    const uint64_t n = r.cloud->n;
    hipStream_t s = c->stream;
    // tables
    std::vector<uint32_t> h(r.tables_words, 0u);
    std::memcpy(&h[r.o_edges], p.et.edges.data(), p.C * 4);
    std::memcpy(&h[r.o_envoff], p.envoff, 256 * 4);
    std::memcpy(&h[r.o_trgslot], p.trgslot, 256 * 4);
    for (int k = 0; k < p.ntasks; ++k) {
        h[r.o_ttrg + k] = p.task_trg[k];
        h[r.o_tenv + k] = p.task_env[k];
    }
    for (int k = 0; k <= p.kmax; ++k)
        h[r.o_kidx + k] = (uint32_t)(p.kmax > 0 ? p.kidx_of_k[k] : -1);
    if (!r.rows.empty())
        std::memcpy(&h[r.o_rows], r.rows.data(), r.rows.size() * sizeof(Row));
    PP_HIP(c, hipMemcpyAsync(r.tables, h.data(), h.size() * 4, hipMemcpyHostToDevice, s));
    PP_HIP(c, hipMemsetAsync(r.P, 0, sizeof(unsigned long long) * std::max(p.ntrg * p.E * p.C, 1), s));
    PP_HIP(c, hipMemsetAsync(r.stats, 0, sizeof(unsigned long long) * 4, s));
    PP_HIP(c, hipMemsetAsync(r.K, 0, 4 * (size_t)std::max(p.ntasks * p.nku * p.C, 1), s));
    PP_HIP(c, hipMemsetAsync(r.N1, 0, 4 * (size_t)std::max(p.ntasks * p.C, 1), s));
    // centers
    const int32_t *d_trgslot = (const int32_t *)(r.tables + r.o_trgslot);
    PP_HIP(c, launch_flags(r.cl->pos, r.cl->s2, n, d_trgslot, p.R2, r.flags, s));
    PP_HIP(c, scan_exclusive(r.flags, r.cpos, n, r.ws, s));
    PP_HIP(c, launch_compact(r.flags, r.cpos, n, r.centers, s));
    return prep_tiles(c, r, T);
}

pp_status prep_tiles(pp_ctx *c, JobRun &r, int T)
{
    const uint64_t n = r.cloud->n;
    const uint64_t nc = r.ncells;
    hipStream_t s = c->stream;
    PP_HIP(c, hipMemsetAsync(r.maxcand, 0, 16, s));
    PP_HIP(c, launch_tilecount(r.cl->cellstart, r.cpos, nc, T, r.ntile, s));
    PP_HIP(c, scan_exclusive(r.ntile, r.toff, nc, r.ws, s));
    PP_HIP(c, launch_tilewrite(r.cl->cellstart, r.cpos, r.toff, nc, T, 0, r.cl->g,
                               (const Row *)(r.tables + r.o_rows), (int)r.rows.size(), r.tiles,
                               r.maxcand, s));
    PP_HIP(c, hipMemcpyAsync(&r.h_counts[0], r.cpos + n, 4, hipMemcpyDeviceToHost, s));
    PP_HIP(c, hipMemcpyAsync(&r.h_counts[1], r.toff + nc, 4, hipMemcpyDeviceToHost, s));
    PP_HIP(c, hipMemcpyAsync(&r.h_counts[2], r.maxcand, 4, hipMemcpyDeviceToHost, s));
    return PP_OK;
}

void fill_jobdev(JobRun &r)
{
    const JobPlan &p = r.plan;
    JobDev &d = r.dev;
    d.pos = r.cl->pos;
    d.cellstart = r.cl->cellstart;
    d.nx = r.cl->g.nx;
    d.ny = r.cl->g.ny;
    d.nz = r.cl->g.nz;
    d.nrows = (int)r.rows.size();
    d.rows = (const Row *)(r.tables + r.o_rows);
    d.centers = r.centers;
    d.R2 = p.R2;
    d.C = p.C;
    d.edges = r.tables + r.o_edges;
    d.idx_scale = p.et.idx_scale;
    d.idx_off = p.et.idx_off;
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
    d.P = r.P;
    d.K = r.K;
    d.N1 = r.N1;
    d.CN = r.CN;
    d.cn_cap = r.cn_cap;
    d.percenter = p.percenter ? 1 : 0;
    d.stats = r.stats;
}

int choose_cells_per_rmax(pp_ctx *c)
{
    // finer cells cut the candidates per in-ball pair (6.3 at Rmax, 3.7 at
    // Rmax/2, 2.3 at Rmax/4) for a few more, shorter stencil rows; Rmax/4
    // measured fastest on the 20M-ion tip
    return c->tune_r > 0 ? c->tune_r : 4;
}

pp_status ensure_celllist(pp_ctx *c, pp_cloud *cl, const JobPlan &p, const CellList *&out)
{
    if (cl->resident) {
        out = &cl->cl;
        return PP_OK;
    }
    const int r = choose_cells_per_rmax(c);
    const float w = (float)((double)p.R / r * (1.0 + std::ldexp(1.0, -10)));
    if (!(cl->cl.mem && cl->cl.w == w)) {
        // a cell list at a different width is rebuilt (grid growth may widen it)
        pp_status st = build_celllist(c, cl, w, cl->cl);
        if (st != PP_OK)
            return st;
        cl->cl.w = w;
    }
    out = &cl->cl;
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

int choose_T(pp_ctx *c, const JobPlan &p)
{
    if (c->tune_T > 0)
        return c->tune_T;
    (void)p;
    return 8;  // measured fastest for both modes (fewer VGPRs, more resident waves)
}

pp_status run_jobs(pp_ctx *c, int njobs, const pp_cloud *const *clouds, const pp_job *jobs,
                   pp_hist_out *outs)
{
    auto t_start = std::chrono::steady_clock::now();
    std::vector<JobRun> runs(njobs);
    int T = 0;
    for (int i = 0; i < njobs; ++i) {
        if (!clouds[i])
            return fail(c, PP_ERR_INVALID_ARG, "cloud is NULL");
        if (!outs)
            return fail(c, PP_ERR_INVALID_ARG, "outputs are NULL");
        pp_status st = plan_job(c, &jobs[i], runs[i].plan);
        if (st != PP_OK)
            return st;
        runs[i].cloud = clouds[i];
        T = std::max(T, choose_T(c, runs[i].plan));
    }
    PP_HIP(c, hipSetDevice(c->device));
    PP_HIP(c, hipEventRecord(c->ev[0], c->stream));
    // cell lists (per-job ones are rebuilt only when the width changes)
    for (int i = 0; i < njobs; ++i) {
        pp_status st = ensure_celllist(c, const_cast<pp_cloud *>(clouds[i]), runs[i].plan, runs[i].cl);
        if (st != PP_OK)
            return st;
        runs[i].rows = make_rows(runs[i].cl->g.w, runs[i].plan.R);
        if ((int)runs[i].rows.size() > kMaxRows)
            return fail(c, PP_ERR_UNSUPPORTED, "stencil too large (cell width far below Rmax)");
        runs[i].ncells = runs[i].cl->g.ncells();
    }
    PP_HIP(c, hipEventRecord(c->ev[1], c->stream));
    // job arenas (after the cell lists: build_celllist may have used scratch)
    Carver m0(nullptr);
    for (int i = 0; i < njobs; ++i)
        layout_job(runs[i], m0);
    pp_status st = ensure(c, c->scratch, m0.off + 256);
    if (st != PP_OK)
        return st;
    Carver m(c->scratch.p);
    for (int i = 0; i < njobs; ++i)
        layout_job(runs[i], m);
    for (int i = 0; i < njobs; ++i) {
        st = prep_job(c, runs[i], T);
        if (st != PP_OK)
            return st;
    }
    PP_HIP(c, hipStreamSynchronize(c->stream));
    PP_HIP(c, hipEventRecord(c->ev[2], c->stream));
    // launch configuration
    int maxC = 1, maxE = 1, maxTrg = 1, maxK = 0, maxRows = 1;
    bool pc = false;
    uint32_t maxcand = 0, total = 0;
    std::vector<uint32_t> tile_off(njobs + 1, 0);
    for (int i = 0; i < njobs; ++i) {
        const JobPlan &p = runs[i].plan;
        maxC = std::max(maxC, p.C);
        maxE = std::max(maxE, p.E);
        maxTrg = std::max(maxTrg, p.ntrg);
        maxK = std::max(maxK, p.kmax);
        maxRows = std::max(maxRows, (int)runs[i].rows.size());
        pc |= p.percenter;
        maxcand = std::max(maxcand, runs[i].h_counts[2]);
        tile_off[i] = total;
        total += runs[i].h_counts[1];
    }
    tile_off[njobs] = total;
    int c16 = maxcand < 65535u ? 1 : 0;
    size_t lds = query_lds_bytes(T, c16, pc, maxC, maxE, maxTrg, maxK, maxRows);
    while (lds > c->lds_per_block && T > 8) {
        // smaller tiles: rebuild the tile lists
        T /= 2;
        for (int i = 0; i < njobs; ++i) {
            st = prep_tiles(c, runs[i], T);
            if (st != PP_OK)
                return st;
        }
        PP_HIP(c, hipStreamSynchronize(c->stream));
        total = 0;
        for (int i = 0; i < njobs; ++i) {
            tile_off[i] = total;
            total += runs[i].h_counts[1];
        }
        tile_off[njobs] = total;
        lds = query_lds_bytes(T, c16, pc, maxC, maxE, maxTrg, maxK, maxRows);
    }
    if (lds > c->lds_per_block)
        return fail(c, PP_ERR_UNSUPPORTED, "histograms too large for LDS (bins x types x tile)");
    // CNTNB counters: no neighbour count exceeds the largest stencil candidate count
    {
        size_t words = 0;
        for (int i = 0; i < njobs; ++i) {
            JobRun &r = runs[i];
            r.cn_cap = (r.plan.families & PP_FAM_CNTNB) ? (int)r.h_counts[2] + 1 : 1;
            words += ((size_t)std::max(r.plan.ntasks, 1) * (size_t)r.cn_cap + 63) & ~(size_t)63;
        }
        st = ensure(c, c->cnbuf, words * 4 + 256);
        if (st != PP_OK)
            return st;
        size_t o = 0;
        for (int i = 0; i < njobs; ++i) {
            JobRun &r = runs[i];
            r.CN = (uint32_t *)c->cnbuf.p + o;
            const size_t w = (size_t)std::max(r.plan.ntasks, 1) * (size_t)r.cn_cap;
            o += (w + 63) & ~(size_t)63;
            PP_HIP(c, hipMemsetAsync(r.CN, 0, w * 4, c->stream));
            fill_jobdev(r);
        }
    }
    // device job table: JobDev[njobs] | Tile*[njobs] | tile_off[njobs+1] | queue
    Carver jb0(nullptr);
    jb0.take<JobDev>(njobs);
    jb0.take<const Tile *>(njobs);
    jb0.take<uint32_t>(njobs + 1);
    jb0.take<uint32_t>(kXcds + 1);
    st = ensure(c, c->jobbuf, jb0.off + 256);
    if (st != PP_OK)
        return st;
    Carver jb(c->jobbuf.p);
    JobDev *d_jobs = jb.take<JobDev>(njobs);
    const Tile **d_tiles = jb.take<const Tile *>(njobs);
    uint32_t *d_off = jb.take<uint32_t>(njobs + 1);
    uint32_t *d_queue = jb.take<uint32_t>(kXcds + 1);
    std::vector<JobDev> hj(njobs);
    std::vector<const Tile *> ht(njobs);
    for (int i = 0; i < njobs; ++i) {
        hj[i] = runs[i].dev;
        ht[i] = runs[i].tiles;
    }
    PP_HIP(c, hipMemcpyAsync(d_jobs, hj.data(), sizeof(JobDev) * njobs, hipMemcpyHostToDevice, c->stream));
    PP_HIP(c, hipMemcpyAsync(d_tiles, ht.data(), sizeof(Tile *) * njobs, hipMemcpyHostToDevice, c->stream));
    PP_HIP(c, hipMemcpyAsync(d_off, tile_off.data(), 4 * (njobs + 1), hipMemcpyHostToDevice, c->stream));
    int nb = 0;
    PP_HIP(c, query_max_blocks_per_cu(T, c16, pc, lds, &nb));
    if (nb < 1)
        nb = 1;
    QueryLaunch q{};
    q.jobs = d_jobs;
    q.njobs = njobs;
    q.tile_off = d_off;
    q.tiles = d_tiles;
    q.total_tiles = total;
    q.queue = d_queue;
    q.T = T;
    q.c16 = c16;
    q.percenter = pc ? 1 : 0;
    q.maxC = maxC;
    q.maxE = maxE;
    q.maxTrg = maxTrg;
    q.maxKmax = maxK;
    q.maxRows = maxRows;
    const uint32_t want = (uint32_t)c->cu_count * (uint32_t)nb;
    q.grid = (int)std::max<uint32_t>(1u, std::min<uint32_t>(want, std::max<uint32_t>(total, 1u)));
    PP_HIP(c, hipEventRecord(c->ev[3], c->stream));
    PP_HIP(c, launch_query(q, c->stream));
    PP_HIP(c, hipEventRecord(c->ev[4], c->stream));
    // read back
    for (int i = 0; i < njobs; ++i) {
        JobRun &r = runs[i];
        const JobPlan &p = r.plan;
        std::vector<unsigned long long> P((size_t)std::max(p.ntrg * p.E * p.C, 1));
        std::vector<uint32_t> K((size_t)std::max(p.ntasks * p.nku * p.C, 1));
        std::vector<uint32_t> N1((size_t)std::max(p.ntasks * p.C, 1));
        std::vector<uint32_t> CN;
        unsigned long long stats[4];
        PP_HIP(c, hipMemcpyAsync(P.data(), r.P, P.size() * 8, hipMemcpyDeviceToHost, c->stream));
        PP_HIP(c, hipMemcpyAsync(stats, r.stats, sizeof(stats), hipMemcpyDeviceToHost, c->stream));
        if (p.families & PP_FAM_MKNN)
            PP_HIP(c, hipMemcpyAsync(K.data(), r.K, K.size() * 4, hipMemcpyDeviceToHost, c->stream));
        if (p.families & PP_FAM_1NN)
            PP_HIP(c, hipMemcpyAsync(N1.data(), r.N1, N1.size() * 4, hipMemcpyDeviceToHost, c->stream));
        if (p.families & PP_FAM_CNTNB) {
            // counts never exceed the largest stencil candidate count
            CN.assign((size_t)p.ntasks * r.cn_cap, 0u);
            PP_HIP(c, hipMemcpyAsync(CN.data(), r.CN, CN.size() * 4, hipMemcpyDeviceToHost,
                                     c->stream));
        }
        PP_HIP(c, hipStreamSynchronize(c->stream));
        bool overflow = false;
        compose_outputs(r, P, K, N1, CN, stats, &outs[i], &overflow);
        float ms_idx = 0, ms_prep = 0, ms_q = 0;
        PP_HIP(c, hipEventElapsedTime(&ms_idx, c->ev[0], c->ev[1]));
        PP_HIP(c, hipEventElapsedTime(&ms_prep, c->ev[1], c->ev[2]));
        PP_HIP(c, hipEventElapsedTime(&ms_q, c->ev[3], c->ev[4]));
        outs[i].ms_index = ms_idx;
        outs[i].ms_prep = ms_prep;
        outs[i].ms_query = ms_q;
        outs[i].tile_centers = T;
        outs[i].cells_per_rmax = r.cloud->resident ? 0 : choose_cells_per_rmax(c);
        outs[i].counters16 = c16;
        outs[i].percenter = pc ? 1 : 0;
        outs[i].ms_total = std::chrono::duration<double, std::milli>(
                               std::chrono::steady_clock::now() - t_start)
                               .count();
        if (overflow)
            return fail(c, PP_ERR_OVERFLOW, "a CNTNB neighbour count exceeded cntnb_cap");
    }
    return PP_OK;
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
    pp_cloud *cl = new pp_cloud();
    cl->ctx = c;
    cl->n = n;
    Carver m0(nullptr);
    m0.take<float>(3 * n);
    m0.take<uint8_t>(n);
    m0.take<float>(n);
    cl->raw_bytes = m0.off + 256;
    hipError_t e = hipMalloc(&cl->raw, cl->raw_bytes);
    if (e != hipSuccess) {
        delete cl;
        return fail(c, e == hipErrorOutOfMemory ? PP_ERR_OOM_DEVICE : PP_ERR_HIP,
                    std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    Carver m(cl->raw);
    cl->xyz = m.take<float>(3 * n);
    cl->type = m.take<uint8_t>(n);
    cl->s2 = m.take<float>(n);
    auto cleanup = [&](pp_status st) {
        if (cl->raw)
            (void)hipFree(cl->raw);
        if (cl->cl.mem)
            (void)hipFree(cl->cl.mem);
        delete cl;
        return st;
    };
    if (n > 0) {
        if (hipMemcpyAsync(cl->xyz, xyz, 12 * n, kind, c->stream) != hipSuccess ||
            hipMemcpyAsync(cl->type, type, n, kind, c->stream) != hipSuccess ||
            hipMemcpyAsync(cl->s2, s2, 4 * n, kind, c->stream) != hipSuccess)
            return cleanup(fail(c, PP_ERR_HIP, "input copy failed"));
        uint32_t *bb = nullptr;
        if (ensure(c, c->scratch, 64) != PP_OK)
            return cleanup(PP_ERR_OOM_DEVICE);
        bb = (uint32_t *)c->scratch.p;
        uint32_t h[6];
        if (launch_bbox(cl->xyz, n, bb, c->stream) != hipSuccess ||
            hipMemcpyAsync(h, bb, sizeof(h), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return cleanup(fail(c, PP_ERR_HIP, "bounding box failed"));
        for (int a = 0; a < 3; ++a) {
            cl->lo[a] = ord_to_float(h[a]);
            cl->hi[a] = ord_to_float(h[3 + a]);
            if (!std::isfinite(cl->lo[a]) || !std::isfinite(cl->hi[a]))
                return cleanup(fail(c, PP_ERR_INVALID_ARG, "non-finite coordinates"));
        }
    }
    if (cell_width > 0) {
        pp_status st = build_celllist(c, cl, cell_width, cl->cl);
        if (st != PP_OK)
            return cleanup(st);
        cl->resident = true;
        if (hipStreamSynchronize(c->stream) != hipSuccess)
            return cleanup(fail(c, PP_ERR_HIP, "cell list build failed"));
    }
    *out = cl;
    return PP_OK;
}

} // namespace

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
    pp_ctx *c = new pp_ctx();
    c->device = device;
    c->cu_count = prop.multiProcessorCount;
    c->lds_per_block = prop.sharedMemPerBlock;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return PP_ERR_HIP;
    }
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return PP_ERR_HIP;
        }
    *out = c;
    return PP_OK;
}

pp_status pp_spatstat_destroy(pp_ctx *c)
{
    if (!c)
        return PP_ERR_INVALID_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->scratch.p)
        (void)hipFree(c->scratch.p);
    if (c->jobbuf.p)
        (void)hipFree(c->jobbuf.p);
    if (c->cnbuf.p)
        (void)hipFree(c->cnbuf.p);
    for (auto &e : c->ev)
        if (e)
            (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return PP_OK;
}

const char *pp_last_error_string(const pp_ctx *c) { return c ? c->err.c_str() : "no context"; }

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

pp_status pp_cloud_free(pp_ctx *c, pp_cloud *cl)
{
    if (!c || !cl)
        return PP_ERR_INVALID_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (cl->raw)
        (void)hipFree(cl->raw);
    if (cl->cl.mem)
        (void)hipFree(cl->cl.mem);
    delete cl;
    return PP_OK;
}

pp_status pp_cloud_footprint(const pp_cloud *cl, uint64_t *bytes)
{
    if (!cl || !bytes)
        return PP_ERR_INVALID_ARG;
    *bytes = cl->raw_bytes + cl->cl.bytes;
    return PP_OK;
}

pp_status pp_spatstat_run(pp_ctx *c, const pp_cloud *cl, const pp_job *job, pp_hist_out *out)
{
    if (!c || !cl || !job || !out)
        return fail(c, PP_ERR_INVALID_ARG, "NULL argument");
    return run_jobs(c, 1, &cl, job, out);
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
    return run_jobs(c, njobs, clouds, jobs, outs);
}

pp_status pp_set_tuning(pp_ctx *c, int32_t centers_per_tile, int32_t cells_per_rmax)
{
    if (!c)
        return PP_ERR_INVALID_ARG;
    if (!(centers_per_tile == 0 || centers_per_tile == 8 || centers_per_tile == 16 ||
          centers_per_tile == 32))
        return fail(c, PP_ERR_INVALID_ARG, "centers_per_tile must be 0, 8, 16 or 32");
    if (cells_per_rmax < 0 || cells_per_rmax > 4)
        return fail(c, PP_ERR_INVALID_ARG, "cells_per_rmax must be 0..4");
    c->tune_T = centers_per_tile;
    c->tune_r = cells_per_rmax;
    return PP_OK;
}

} // extern "C"
