// pp_cpu_baseline.cpp -- the reference's CPU algorithm for the RDF / kNN path,
// restated for timing on the host cores.
//
// TEST / BENCH INFRASTRUCTURE ONLY: the cpu_baseline leg of bench.py and the
// tests load it; the product library never links or calls it.
//
// The structure follows the reference step by step, because the point of
// this file is to time that structure:
//   * decompositor::loadpartitioning (src/PARAPROBE_SolverHdl.cpp:1819-1934):
//     sort all z, split at the (t / nthreads) quantiles of
//     mathHdl::quantiles_nosort (src/PARAPROBE_Math.cpp:46-77), one z-slab per
//     thread, ions in the half-open [zmi, zmx) (S.cpp:1330-1468);
//   * threadmemory::build_local_kdtree + kd_tree::build + pack_p3dm1_dist
//     (S.cpp:1593-1663, src/PARAPROBE_KDTree.cpp:34-193): median splits with
//     nth_element cycling x, y, z, leaves of fewer than 256 + 1 points, ions
//     and ion2surf permuted into leaf order;
//   * kd_tree::range_rball_noclear_nosort / _external
//     (src/PARAPROBE_KDTree.cpp:423-655): box pruning with the float
//     cuboid::outside_proximity, self excluded by index in the own slab;
//   * horderdist::compute_generic_spatstat3 (S.cpp:3770-4176): the team
//     sweeps the regions one after the other with `omp for schedule(dynamic,1)
//     nowait` and a barrier per region; per considered center the neighbours
//     are materialised as (sqrt(d2), type), probing the other slabs when the
//     ball crosses the slab ((z - R) > zmi && (z + R) < zmx); the type
//     reorganiser (counting sort by type); RDF / RIPK add every env distance;
//     1NN the minimum or R + EPSILON; CNTNB one count per env type; MKNN
//     nth_element for every k in descending order or R + EPSILON;
//     thread-local double histograms reduced under `omp critical`.
// Output layout matches the C ABI: [lowest, bins..., highest] as int64.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

constexpr double kEpsilon = 1.0e-6;
constexpr size_t kLeaf = 256;
constexpr int kStack = 32;

struct P3 {
    float x, y, z;
    uint32_t m;
};
struct Nbor {
    float d;
    uint32_t m;
};
struct Box {
    float mn[3], mx[3];
    // squared float distance of p to the box, 0 inside (cuboid::outside_proximity)
    float outside(const P3 &p) const
    {
        const float c[3] = {p.x, p.y, p.z};
        float r = 0.f;
        for (int a = 0; a < 3; ++a) {
            if (c[a] < mn[a]) {
                const float t = mn[a] - c[a];
                r += t * t;
            } else if (c[a] > mx[a]) {
                const float t = c[a] - mx[a];
                r += t * t;
            }
        }
        return r;
    }
};
struct Node {
    size_t i0 = (size_t)-1, i1 = (size_t)-1, split = (size_t)-1;
    float pos = 3.402823466e38f;
    bool leaf() const { return split == (size_t)-1; }
};

struct Hist {
    double start = 0, width = 0, end = 0, tmp = 0;
    uint32_t nbins = 0;
    std::vector<double> cnts;
    double lowest = 0, highest = 0;
    Hist() = default;
    Hist(float s, float w, float e)
    {
        const double ds = s, dw = w, de = e;
        if (dw > kEpsilon) {
            const size_t nb = (size_t)((de - ds) / dw);
            start = ds;
            width = dw;
            end = de;
            nbins = (uint32_t)nb;
            tmp = 1.0 / (de - ds) * (double)nb;
            cnts.assign(nbins, 0.0);
        }
    }
    void add(float x)
    {
        const double xx = x;
        if (xx >= start && xx <= end) {
            const float rb = (float)((xx - start) * tmp);
            const unsigned b = (unsigned)std::floor(rb);
            if (b < nbins)
                cnts[b] += 1.0;
            else
                highest += 1.0;
        } else {
            if (xx < start)
                lowest += 1.0;
            if (xx > end)
                highest += 1.0;
        }
    }
    void merge(const Hist &o)
    {
        lowest += o.lowest;
        highest += o.highest;
        for (uint32_t b = 0; b < nbins; ++b)
            cnts[b] += o.cnts[b];
    }
};

struct KdTree {
    std::vector<Node> nodes;
    float mn[3] = {3.4e38f, 3.4e38f, 3.4e38f}, mx[3] = {-3.4e38f, -3.4e38f, -3.4e38f};

    void build(const std::vector<P3> &pts, std::vector<size_t> &perm)
    {
        const size_t count = pts.size();
        perm.resize(count);
        for (size_t i = 0; i < count; ++i)
            perm[i] = i;
        if (count == 0)
            return;
        struct Task {
            size_t first, last, node, dim;
        };
        Task st[kStack];
        int top = 0;
        nodes.push_back(Node());
        st[0] = {0, count - 1, 0, 0};
        auto coord = [&](size_t i, size_t d) {
            return d == 0 ? pts[i].x : (d == 1 ? pts[i].y : pts[i].z);
        };
        while (top >= 0) {
            const Task t = st[top];
            if (t.last - t.first < kLeaf) {
                Node &n = nodes[t.node];
                n.i0 = t.first;
                n.i1 = t.last;
                for (size_t k = t.first; k <= t.last; ++k) {
                    const P3 &p = pts[perm[k]];
                    const float c[3] = {p.x, p.y, p.z};
                    for (int a = 0; a < 3; ++a) {
                        mn[a] = std::min(mn[a], c[a]);
                        mx[a] = std::max(mx[a], c[a]);
                    }
                }
                --top;
                continue;
            }
            const size_t k = (t.first + t.last) / 2;
            const size_t d = t.dim;
            std::nth_element(perm.begin() + t.first, perm.begin() + k, perm.begin() + t.last + 1,
                             [&](size_t a, size_t b) { return coord(a, d) < coord(b, d); });
            const size_t i0 = nodes.size(), i1 = i0 + 1;
            nodes[t.node].pos = coord(perm[k], d);
            nodes[t.node].split = k;
            nodes[t.node].i0 = i0;
            nodes[t.node].i1 = i1;
            const size_t nd = (d + 1) % 3;
            st[top] = {t.first, k, i0, nd};
            nodes.push_back(Node());
            ++top;
            st[top] = {k + 1, t.last, i1, nd};
            nodes.push_back(Node());
        }
    }

    // range query: self excluded by index when self != SIZE_MAX
    void range(const P3 &target, size_t self, const std::vector<P3> &sorted, float r2,
               std::vector<Nbor> &out) const
    {
        visit(target, self, sorted, r2,
              [&](const P3 &q, float d) { out.push_back(Nbor{std::sqrt(d), q.m}); });
    }
    // the same query returning the neighbours themselves (the _p3dm1 variants)
    void range_p3(const P3 &target, size_t self, const std::vector<P3> &sorted, float r2,
                  std::vector<P3> &out) const
    {
        visit(target, self, sorted, r2, [&](const P3 &q, float) { out.push_back(q); });
    }
    template <class F>
    void visit(const P3 &target, size_t self, const std::vector<P3> &sorted, float r2,
               F &&emit) const
    {
        if (nodes.empty())
            return;
        struct TT {
            Box bx;
            size_t node, dim;
            float d;
        };
        TT st[kStack];
        Box root;
        for (int a = 0; a < 3; ++a) {
            root.mn[a] = mn[a];
            root.mx[a] = mx[a];
        }
        st[0] = {root, 0, 0, root.outside(target)};
        int top = 0;
        const float tc[3] = {target.x, target.y, target.z};
        while (top >= 0) {
            const TT t = st[top];
            if (t.d > r2) {
                --top;
                continue;
            }
            const Node &n = nodes[t.node];
            if (n.leaf()) {
                for (size_t i = n.i0; i <= n.i1; ++i) {
                    const P3 &q = sorted[i];
                    const float dx = target.x - q.x, dy = target.y - q.y, dz = target.z - q.z;
                    const float d = dx * dx + dy * dy + dz * dz;
                    if (d > r2)
                        continue;
                    if (i != self)
                        emit(q, d);
                }
                --top;
                continue;
            }
            const int a = (int)t.dim;
            Box lo = t.bx, hi = t.bx;
            lo.mx[a] = n.pos;
            hi.mn[a] = n.pos;
            TT nearT, farT;
            if (tc[a] <= n.pos) {
                nearT.bx = lo;
                nearT.node = n.i0;
                farT.bx = hi;
                farT.node = n.i1;
            } else {
                nearT.bx = hi;
                nearT.node = n.i1;
                farT.bx = lo;
                farT.node = n.i0;
            }
            const size_t nd = (t.dim + 1) % 3;
            nearT.d = 0.f;
            nearT.dim = nd;
            farT.d = farT.bx.outside(target);
            farT.dim = nd;
            st[top] = farT;
            ++top;
            st[top] = nearT;
        }
    }
};

struct Region {
    float zmi, zmx;
    std::vector<P3> ions;  // leaf order
    std::vector<float> s2; // leaf order
    KdTree tree;
};

inline bool has_type(const uint8_t *mask32, uint32_t t)
{
    return t < 256 && ((mask32[t >> 3] >> (t & 7)) & 1);
}

float lerpf(float v0, float v1, float t) { return (1.f - t) * v0 + t * v1; }

} // namespace

// decompositor::loadpartitioning (S.cpp:1819-1934) and build_local_kdtree
// (S.cpp:1593-1663): one z-slab per thread at the ion-count quantiles, each
// with its KD-tree, ions and ion2surf in leaf order.
static void build_regions(int64_t n, const float *xyz, const uint8_t *type,
                          const float *ion2surf, int nthreads, std::vector<Region> &reg)
{
    // ---- loadpartitioning: quantile z-slabs
    std::vector<float> zs((size_t)n);
    for (int64_t i = 0; i < n; ++i)
        zs[i] = xyz[3 * i + 2];
    std::sort(zs.begin(), zs.end());
    reg.assign(nthreads, Region());
    for (int t = 0; t < nthreads; ++t) {
        auto quant = [&](int tid) {
            const float q = (float)tid / (float)nthreads;
            const float poi = lerpf(-0.5f, (float)n - 0.5f, q);
            const size_t left = (size_t)std::max<int64_t>((int64_t)std::floor(poi), 0);
            const size_t right = (size_t)std::min<int64_t>((int64_t)std::ceil(poi), n - 1);
            return lerpf(zs[left], zs[right], poi - (float)left);
        };
        reg[t].zmi = (t == 0) ? (n ? zs.front() - 1.0e-4f : 0.f) : quant(t);
        reg[t].zmx = (t == nthreads - 1) ? (n ? zs.back() + 1.0e-4f : 0.f) : quant(t + 1);
    }
    // ---- read_localions + build_local_kdtree, one region per thread
#pragma omp parallel num_threads(nthreads)
    {
        const int t = omp_get_thread_num();
        Region &r = reg[t];
        std::vector<P3> loc;
        std::vector<float> ls2;
        for (int64_t i = 0; i < n; ++i) {
            const float z = xyz[3 * i + 2];
            if (z >= r.zmi && z < r.zmx) {
                loc.push_back(P3{xyz[3 * i], xyz[3 * i + 1], z, (uint32_t)type[i]});
                ls2.push_back(ion2surf[i]);
            }
        }
        std::vector<size_t> perm;
        r.tree.build(loc, perm);
        r.ions.reserve(loc.size());
        r.s2.reserve(loc.size());
        for (size_t k = 0; k < perm.size(); ++k) {
            r.ions.push_back(loc[perm[k]]);
            r.s2.push_back(ls2[perm[k]]);
        }
    }
}

extern "C" {

// Returns 0 on success.  timings[0] = index build seconds (slabs + KD-trees),
// timings[1] = compute seconds (the DescrStatsThreadparallelCompute span).
int pp_cpu_spatstat(int64_t n, const float *xyz, const uint8_t *type, const float *ion2surf,
                    int32_t ntasks, const uint8_t *trg_mask, const uint8_t *env_mask, float rmin,
                    float rincr, float rmax, uint32_t families, int32_t nk,
                    const int64_t *kminus1_desc, int32_t nthreads, int64_t *rdf, int64_t *nn1,
                    int64_t *ripk, int64_t *mknn, int64_t *cntnb, int64_t cntnb_cap,
                    int64_t *stats, double *timings)
{
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
    const bool doRDF = families & 1u, do1NN = families & 2u, doRIPK = families & 4u,
               doMKNN = (families & 8u) && nk > 0, doCNT = families & 32u;
    Hist proto(rmin, rincr, rmax);
    if (proto.nbins == 0)
        return 1;
    const int64_t C = (int64_t)proto.nbins + 2;
    const float R = rmax, RSQR = R * R;
    const double t0 = omp_get_wtime();
    std::vector<Region> reg;
    build_regions(n, xyz, type, ion2surf, nthreads, reg);
    const double t1 = omp_get_wtime();

    // ---- compute_generic_spatstat3
    uint8_t anytrg[32] = {0};
    for (int t = 0; t < ntasks; ++t)
        for (int b = 0; b < 32; ++b)
            anytrg[b] |= trg_mask[32 * t + b];
    const uint32_t maxtype = 256;
    std::vector<Hist> g_rdf(ntasks, proto), g_1nn(ntasks, proto), g_mknn;
    if (doMKNN)
        g_mknn.assign((size_t)ntasks * nk, proto);
    std::vector<std::vector<double>> g_cnt(ntasks);
    int64_t considered = 0, discarded = 0, inball = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : considered, discarded, inball)
    {
        std::vector<Hist> m_rdf(ntasks, proto), m_1nn(ntasks, proto), m_mknn;
        if (doMKNN)
            m_mknn.assign((size_t)ntasks * nk, proto);
        std::vector<std::vector<double>> m_cnt(ntasks);
        for (int thr = 0; thr < nthreads; ++thr) {
            const Region &cur = reg[thr];
#pragma omp for schedule(dynamic, 1) nowait
            for (size_t i = 0; i < cur.ions.size(); ++i) {
                const P3 me = cur.ions[i];
                if (!(cur.s2[i] >= RSQR && has_type(anytrg, me.m))) {
                    ++discarded;
                    continue;
                }
                ++considered;
                std::vector<Nbor> nb;
                cur.tree.range(me, i, cur.ions, RSQR, nb);
                if (!((me.z - R) > cur.zmi && (me.z + R) < cur.zmx)) {
                    for (int o = thr + 1; o < nthreads; ++o)
                        reg[o].tree.range(me, (size_t)-1, reg[o].ions, RSQR, nb);
                    for (int o = thr - 1; o > -1; --o)
                        reg[o].tree.range(me, (size_t)-1, reg[o].ions, RSQR, nb);
                }
                inball += (int64_t)nb.size();
                // type reorganiser: counting sort of the neighbours by type
                std::vector<size_t> cnt(maxtype, 0), s(maxtype, 0);
                for (const Nbor &q : nb)
                    cnt[q.m]++;
                for (uint32_t a = 1; a < maxtype; ++a)
                    s[a] = s[a - 1] + cnt[a - 1];
                std::vector<float> by(nb.size(), 3.402823466e38f);
                std::vector<size_t> fill(s);
                for (const Nbor &q : nb)
                    by[fill[q.m]++] = q.d;
                for (int t = 0; t < ntasks; ++t) {
                    if (!has_type(trg_mask + 32 * t, me.m))
                        continue;
                    const uint8_t *env = env_mask + 32 * t;
                    if (doRDF || doRIPK)
                        for (uint32_t e = 0; e < maxtype; ++e)
                            if (has_type(env, e))
                                for (size_t q = s[e]; q < s[e] + cnt[e]; ++q)
                                    m_rdf[t].add(by[q]);
                    if (do1NN) {
                        float closest = 3.402823466e38f;
                        bool found = false;
                        for (uint32_t e = 0; e < maxtype; ++e)
                            if (has_type(env, e))
                                for (size_t q = s[e]; q < s[e] + cnt[e]; ++q)
                                    if (!(by[q] > closest)) {
                                        closest = by[q];
                                        found = true;
                                    }
                        m_1nn[t].add(found ? closest : (float)((double)R + kEpsilon));
                    }
                    if (doCNT)
                        for (uint32_t e = 0; e < maxtype; ++e)
                            if (has_type(env, e)) {
                                auto &v = m_cnt[t];
                                if (cnt[e] >= v.size())
                                    v.resize(cnt[e] + 1, 0.0);
                                v[cnt[e]] += 1.0;
                            }
                    if (doMKNN) {
                        std::vector<float> tmp;
                        for (uint32_t e = 0; e < maxtype; ++e)
                            if (has_type(env, e))
                                for (size_t q = s[e]; q < s[e] + cnt[e]; ++q)
                                    tmp.push_back(by[q]);
                        for (int kth = 0; kth < nk; ++kth) {
                            const size_t k = (size_t)kminus1_desc[kth];
                            Hist &h = m_mknn[(size_t)t * nk + kth];
                            if (tmp.size() > k) {
                                std::nth_element(tmp.begin(), tmp.begin() + k, tmp.end());
                                h.add(tmp[k]);
                            } else {
                                h.add((float)((double)R + kEpsilon));
                            }
                        }
                    }
                }
            }
#pragma omp barrier
        }
#pragma omp critical
        {
            for (int t = 0; t < ntasks; ++t) {
                g_rdf[t].merge(m_rdf[t]);
                g_1nn[t].merge(m_1nn[t]);
                auto &g = g_cnt[t];
                if (m_cnt[t].size() > g.size())
                    g.resize(m_cnt[t].size(), 0.0);
                for (size_t b = 0; b < m_cnt[t].size(); ++b)
                    g[b] += m_cnt[t][b];
            }
            for (size_t q = 0; q < m_mknn.size(); ++q)
                g_mknn[q].merge(m_mknn[q]);
        }
    }
    const double t2 = omp_get_wtime();
    auto emit = [&](const Hist &h, int64_t *dst) {
        dst[0] = (int64_t)h.lowest;
        for (uint32_t b = 0; b < h.nbins; ++b)
            dst[1 + b] = (int64_t)h.cnts[b];
        dst[h.nbins + 1] = (int64_t)h.highest;
    };
    for (int t = 0; t < ntasks; ++t) {
        if (doRDF && rdf)
            emit(g_rdf[t], rdf + t * C);
        if (doRIPK && ripk)
            emit(g_rdf[t], ripk + t * C);
        if (do1NN && nn1)
            emit(g_1nn[t], nn1 + t * C);
        if (doMKNN && mknn)
            for (int kth = 0; kth < nk; ++kth)
                emit(g_mknn[(size_t)t * nk + kth], mknn + ((int64_t)t * nk + kth) * C);
        if (doCNT && cntnb)
            for (int64_t b = 0; b < cntnb_cap; ++b)
                cntnb[t * cntnb_cap + b] =
                    b < (int64_t)g_cnt[t].size() ? (int64_t)g_cnt[t][b] : 0;
    }
    stats[0] = considered;
    stats[1] = discarded;
    stats[2] = inball;
    timings[0] = t1 - t0;
    timings[1] = t2 - t1;
    return 0;
}

// ---------------------------------------------------------------------------
// Ion-to-surface squared distance, timed the reference's way
// (solver::surface_distancing2, src/PARAPROBE_SolverHdl.cpp:312-484): a
// bounding-volume tree over the triangles' vertex boxes, one leaf per
// triangle, queried per ion with the float box [p - R, p + R] by a stack walk
// (Tree::query, src/PARAPROBE_AABBTree.cpp:331-368), the candidates' closest-
// point values folded into a running best that starts at fl32(R * R), ions
// spread over the threads with schedule(dynamic).  The reference inserts the
// triangles one by one with surface-area costs (Tree::insertLeaf,
// AABBTree.cpp:383-510); this tree is built top-down by median splits, which
// queries at least as fast.  Returns the build time in timings[0] and the
// query time in timings[1].
int pp_cpu_ion2surf(int64_t n, const float *xyz, int64_t ntri, const float *tri, float R,
                    int32_t nthreads, int64_t stride, float *out, double *timings)
{
    if (n < 0 || ntri < 0 || !(R > 0.f) || stride < 1)
        return -1;
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
    const double t0 = omp_get_wtime();
    struct Node {
        float lo[3], hi[3];
        int32_t left, right;  // children, or (-1, triangle) for a leaf
    };
    std::vector<Node> nodes;
    std::vector<float> box(6 * (size_t)ntri);
    std::vector<int32_t> ids((size_t)ntri);
    for (int64_t t = 0; t < ntri; ++t) {
        const float *v = tri + 9 * t;
        for (int k = 0; k < 3; ++k) {
            box[6 * t + k] = std::fmin(std::fmin(v[k], v[3 + k]), v[6 + k]);
            box[6 * t + 3 + k] = std::fmax(std::fmax(v[k], v[3 + k]), v[6 + k]);
        }
        ids[t] = (int32_t)t;
    }
    nodes.reserve(2 * (size_t)ntri + 1);
    // iterative top-down build over ids[b, e)
    struct Job {
        int64_t b, e;
        int32_t node;
    };
    std::vector<Job> work;
    if (ntri > 0) {
        nodes.push_back(Node{});
        work.push_back({0, ntri, 0});
    }
    while (!work.empty()) {
        Job j = work.back();
        work.pop_back();
        Node nd;
        for (int k = 0; k < 3; ++k) {
            nd.lo[k] = INFINITY;
            nd.hi[k] = -INFINITY;
        }
        for (int64_t i = j.b; i < j.e; ++i)
            for (int k = 0; k < 3; ++k) {
                nd.lo[k] = std::fmin(nd.lo[k], box[6 * ids[i] + k]);
                nd.hi[k] = std::fmax(nd.hi[k], box[6 * ids[i] + 3 + k]);
            }
        if (j.e - j.b == 1) {
            nd.left = -1;
            nd.right = ids[j.b];
            nodes[j.node] = nd;
            continue;
        }
        int ax = 0;
        for (int k = 1; k < 3; ++k)
            if (nd.hi[k] - nd.lo[k] > nd.hi[ax] - nd.lo[ax])
                ax = k;
        const int64_t mid = (j.b + j.e) / 2;
        std::nth_element(ids.begin() + j.b, ids.begin() + mid, ids.begin() + j.e,
                         [&](int32_t a, int32_t b) {
                             return box[6 * a + ax] + box[6 * a + 3 + ax] <
                                    box[6 * b + ax] + box[6 * b + 3 + ax];
                         });
        nd.left = (int32_t)nodes.size();
        nd.right = nd.left + 1;
        nodes.push_back(Node{});
        nodes.push_back(Node{});
        nodes[j.node] = nd;
        work.push_back({j.b, mid, nd.left});
        work.push_back({mid, j.e, nd.right});
    }
    const double t1 = omp_get_wtime();
    const float RSQR = R * R;
    auto closest = [](const float *t, const float *p) {
        auto clamp01 = [](float x) {
            const float m = (x < 0.f) ? 0.f : x;
            return (m < 1.f) ? m : 1.f;
        };
        const float e0u = t[3] - t[0], e0v = t[4] - t[1], e0w = t[5] - t[2];
        const float e1u = t[6] - t[0], e1v = t[7] - t[1], e1w = t[8] - t[2];
        const float v0u = t[0] - p[0], v0v = t[1] - p[1], v0w = t[2] - p[2];
        const float a = e0u * e0u + e0v * e0v + e0w * e0w;
        const float b = e0u * e1u + e0v * e1v + e0w * e1w;
        const float c = e1u * e1u + e1v * e1v + e1w * e1w;
        const float d = e0u * v0u + e0v * v0v + e0w * v0w;
        const float e = e1u * v0u + e1v * v0v + e1w * v0w;
        const float det = a * c - b * b;
        float s = b * e - c * d, u = b * d - a * e;
        if (s + u < det) {
            if (s < 0.f) {
                if (u < 0.f && d < 0.f) {
                    s = clamp01(-d / a);
                    u = 0.f;
                } else {
                    s = 0.f;
                    u = clamp01(-e / c);
                }
            } else if (u < 0.f) {
                s = clamp01(-d / a);
                u = 0.f;
            } else {
                const float inv = 1.f / det;
                s *= inv;
                u *= inv;
            }
        } else if (s < 0.f) {
            if (c + e > b + d) {
                s = clamp01(((c + e) - (b + d)) / (a - 2.f * b + c));
                u = 1.f - s;
            } else {
                u = clamp01(-e / c);
                s = 0.f;
            }
        } else if (u < 0.f) {
            if (a + d > b + e) {
                s = clamp01((c + e - b - d) / (a - 2.f * b + c));
                u = 1.f - s;
            } else {
                s = clamp01(-e / c);
                u = 0.f;
            }
        } else {
            s = clamp01((c + e - b - d) / (a - 2.f * b + c));
            u = 1.f - s;
        }
        const float dx = (t[0] + s * e0u + u * e1u) - p[0];
        const float dy = (t[1] + s * e0v + u * e1v) - p[1];
        const float dz = (t[2] + s * e0w + u * e1w) - p[2];
        return dx * dx + dy * dy + dz * dz;
    };
    const int64_t m = (n + stride - 1) / stride;
#pragma omp parallel
    {
        std::vector<int32_t> stack;
        stack.reserve(256);
#pragma omp for schedule(dynamic, 64)
        for (int64_t q = 0; q < m; ++q) {
            const int64_t i = q * stride;
            const float *p = xyz + 3 * i;
            const float lo[3] = {p[0] - R, p[1] - R, p[2] - R};
            const float hi[3] = {p[0] + R, p[1] + R, p[2] + R};
            float best = RSQR;
            stack.clear();
            if (!nodes.empty())
                stack.push_back(0);
            while (!stack.empty()) {
                const Node &nd = nodes[stack.back()];
                stack.pop_back();
                if (hi[0] < nd.lo[0] || lo[0] > nd.hi[0] || hi[1] < nd.lo[1] || lo[1] > nd.hi[1] ||
                    hi[2] < nd.lo[2] || lo[2] > nd.hi[2])
                    continue;
                if (nd.left < 0) {
                    const float v = closest(tri + 9 * (int64_t)nd.right, p);
                    if (!(v > best))
                        best = v;
                } else {
                    stack.push_back(nd.left);
                    stack.push_back(nd.right);
                }
            }
            out[q] = best;
        }
    }
    const double t2 = omp_get_wtime();
    if (timings) {
        timings[0] = t1 - t0;
        timings[1] = t2 - t1;
    }
    return 0;
}

// NPCORR, the n-point correlation branch of the center loop (S.cpp:3899-3960)
// with npc3d::add (src/PARAPROBE_NPointCorrelations3D.cpp:165-207): the
// neighbours within R as positions (the _p3dm1 range queries, the other slabs
// probed when the ball crosses the own slab), per task the env ones as
// (difference vector, sqrt length), nth_element for the K-th, binned.  Equal
// lengths are ordered by the squared length, then the smaller (u, v, w), so
// that the result is
// deterministic (the reference leaves them to nth_element over its KD-tree
// order).  counts is [ntasks][edge^3]; timings as for pp_cpu_spatstat.
int pp_cpu_npcorr(int64_t n, const float *xyz, const uint8_t *type, const float *ion2surf,
                  int32_t ntasks, const uint8_t *trg_mask, const uint8_t *env_mask, float rmax,
                  float rincr, int64_t K, int32_t nthreads, uint32_t *counts, double *timings)
{
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
    const float realhalf = std::ceil(rmax / rincr) + 1.f;
    const int64_t half = (int64_t)realhalf, edge = 1 + 2 * half, nxyz = edge * edge * edge;
    const float R = rmax, RSQR = R * R, inv = 1.f / rincr;
    const double t0 = omp_get_wtime();
    std::vector<Region> reg;
    build_regions(n, xyz, type, ion2surf, nthreads, reg);
    const double t1 = omp_get_wtime();
    std::memset(counts, 0, sizeof(uint32_t) * (size_t)(ntasks * nxyz));
    struct NP {
        float u, v, w, d2, d;
    };
    auto less = [](const NP &a, const NP &b) {
        if (a.d != b.d)
            return a.d < b.d;
        if (a.d2 != b.d2)
            return a.d2 < b.d2;
        if (a.u != b.u)
            return a.u < b.u;
        if (a.v != b.v)
            return a.v < b.v;
        return a.w < b.w;
    };
#pragma omp parallel num_threads(nthreads)
    {
        std::vector<uint32_t> m_cnt((size_t)(ntasks * nxyz), 0u);
        std::vector<P3> nb;
        std::vector<NP> tmp;
        for (int thr = 0; thr < nthreads; ++thr) {
            const Region &cur = reg[thr];
#pragma omp for schedule(dynamic, 1) nowait
            for (size_t i = 0; i < cur.ions.size(); ++i) {
                const P3 me = cur.ions[i];
                if (!(cur.s2[i] >= RSQR))
                    continue;
                bool any = false;
                for (int t = 0; t < ntasks; ++t)
                    any |= has_type(trg_mask + 32 * t, me.m);
                if (!any)
                    continue;
                nb.clear();
                cur.tree.range_p3(me, i, cur.ions, RSQR, nb);
                if (!((me.z - R) > cur.zmi && (me.z + R) < cur.zmx)) {
                    for (int o = thr + 1; o < nthreads; ++o)
                        reg[o].tree.range_p3(me, (size_t)-1, reg[o].ions, RSQR, nb);
                    for (int o = thr - 1; o > -1; --o)
                        reg[o].tree.range_p3(me, (size_t)-1, reg[o].ions, RSQR, nb);
                }
                if (nb.size() <= (size_t)K)
                    continue;
                for (int t = 0; t < ntasks; ++t) {
                    if (!has_type(trg_mask + 32 * t, me.m))
                        continue;
                    tmp.clear();
                    for (const P3 &q : nb)
                        if (has_type(env_mask + 32 * t, q.m)) {
                            NP e;
                            e.u = q.x - me.x;
                            e.v = q.y - me.y;
                            e.w = q.z - me.z;
                            e.d2 = e.u * e.u + e.v * e.v + e.w * e.w;
                            e.d = std::sqrt(e.d2);
                            tmp.push_back(e);
                        }
                    if (tmp.size() <= (size_t)K)
                        continue;
                    std::nth_element(tmp.begin(), tmp.begin() + K, tmp.end(), less);
                    const NP k = tmp[(size_t)K];
                    if (!(k.d <= rmax))
                        continue;
                    const float c[3] = {k.u * inv, k.v * inv, k.w * inv};
                    int64_t ix[3];
                    for (int q = 0; q < 3; ++q) {
                        const int disc = (c[q] >= 0.f) ? (int)((double)c[q] + 0.5)
                                                       : (int)((double)c[q] - 0.5);
                        ix[q] = disc + half;
                    }
                    m_cnt[(size_t)(t * nxyz + ix[0] + ix[1] * edge + ix[2] * edge * edge)] += 1;
                }
            }
#pragma omp barrier
        }
#pragma omp critical
        for (size_t b = 0; b < m_cnt.size(); ++b)
            counts[b] += m_cnt[b];
    }
    const double t2 = omp_get_wtime();
    if (timings) {
        timings[0] = t1 - t0;
        timings[1] = t2 - t1;
    }
    return 0;
}

} // extern "C"
