// pp_launch.h -- host-side launchers for the gfx950 kernels of the spatial
// statistics path (implemented in pp_kernels.hip).
#pragma once
#include "pp_internal.h"

namespace pp {

struct GridGeom {
    double lo[3];
    double inv_w;
    double w;               // cell width along x
    double wyz, inv_wyz;    // cell width along y and z (w / an integer)
    int nx, ny, nz;
    uint64_t ncells() const { return (uint64_t)nx * ny * nz; }
};

// Exclusive scan of n uint32 values; out has n + 1 entries (out[n] = total).
// ws must hold scan_workspace_words(n) uint32.
size_t scan_workspace_words(uint64_t n);
hipError_t scan_exclusive(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *ws,
                          hipStream_t s);

// Ordered-uint min/max of an AoS xyz array: out[0..2] min, out[3..5] max
// (order-preserving encodings of float; decode with ord_to_float).
hipError_t launch_bbox(const float *xyz, uint64_t n, uint32_t *out6, hipStream_t s);
float ord_to_float(uint32_t u);

// Cell-list build: key/rank per ion (sub = 3 or 5: keys of sub-cells, count
// holds 1 << sub entries per cell: the x part of the cell (sub - 2 bits, the
// major sub-key), then the y and z halves), then scatter into cell-sorted
// float4 (x, y, z, type) and the sorted ion2surf array; launch_stride takes
// every (1 << shift)-th offset (the cell offsets of sub-cell keys);
// launch_xsplit writes each cell's x-part starts (XSplit).
hipError_t launch_cellkeys(const float *xyz, uint64_t n, GridGeom g, uint32_t *count,
                           uint32_t *key, uint32_t *rank, hipStream_t s, int sub = 0);
hipError_t launch_xsplit(const uint32_t *start, uint64_t ncells, int sub, uint2 *out,
                         hipStream_t s);
hipError_t launch_stride(const uint32_t *in, uint32_t *out, uint64_t nout, int shift,
                         hipStream_t s);
// w_s2: the float4's w holds the ion's ion2surf instead of its type (a list
// of one type for symmetric pair counting)
hipError_t launch_scatter(const float *xyz, const uint8_t *type, const float *s2, uint64_t n,
                          const uint32_t *key, const uint32_t *rank, const uint32_t *cellstart,
                          float4 *pos, float *s2sorted, hipStream_t s, int w_s2 = 0);
// the least ion2surf of each cell's ions (NaN as -inf; +inf for an empty cell)
hipError_t launch_cellmin(const uint32_t *cellstart, const float *s2sorted, uint64_t ncells,
                          float *out, hipStream_t s);

// Typed candidate lists: flag the ions whose type is in a 256-bit mask,
// and gather them (order kept) after an exclusive scan of the flags.
hipError_t launch_typeflags(const uint8_t *type, uint64_t n, const uint32_t *mask8, uint32_t *flags,
                            hipStream_t s);
hipError_t launch_gather(const float *xyz, const uint8_t *type, const float *s2,
                         const uint32_t *flags, const uint32_t *pos, uint64_t n, float *oxyz,
                         uint8_t *otype, float *os2, hipStream_t s);
// Per-type ion counts of a cloud (counts has 256 entries, zeroed by the caller).
hipError_t launch_typecount(const uint8_t *type, uint64_t n, uint32_t *counts, hipStream_t s);
// A coarse grid that tiles a box exactly: d[a] cells of 1 / inv[a] along axis a.
struct OccGrid {
    double lo[3], inv[3];
    int d[3];
};
// Marks occ[cell] = 1 for every cell of g holding an ion (occ cleared by the
// caller).
hipError_t launch_occupancy(const float *xyz, uint64_t n, OccGrid g, uint8_t *occ, hipStream_t s);

// Centers and tiles of a run of jobs in one pass (pp_prep.hip).  One segment
// per (job, center type slot), a job's segments consecutive; segment q's ions
// occupy [ion0, ion0 + n) of the run's ion space and its cells
// [cell0, cell0 + nc) of the cell space.
struct PrepSeg {
    const float4 *pos;         // the job's cell-sorted ions (x, y, z, type bits)
    const float *s2;           // their ion2surf
    const uint32_t *src;       // a selection per sorted ion (the rerun), or null
    const int32_t *trgslot;    // [256] target slot of a type id, or -1
    const uint32_t *cellstart; // nc + 1 cell offsets
    const Row *rows;           // the job's stencil rows
    GridGeom g;
    uint64_t ion0, cell0;
    uint32_t n, nc;
    int nrows;
    float R2;                  // fl32(R * R): considered iff ion2surf >= R2
    int slot, T, job;
};
// Per job: its centers [center0, center0 + centers) of the run's center list,
// its tiles [tile0, tile0 + tiles) of the run's tile list, and the largest
// stencil candidate count of a cell with centers.
struct PrepJobStat {
    uint32_t center0, centers, tile0, tiles, maxcand, pad[3];
};
struct PrepLaunch {
    const PrepSeg *segs;       // device [nseg]
    int nseg;
    const int32_t *jseg;       // device [njobs + 1]: job j's segments [jseg[j], jseg[j + 1])
    int njobs;
    uint64_t E, Ec;            // sizes of the ion and cell spaces
    uint32_t *flags;           // E
    uint32_t *cpos;            // E + 1: exclusive scan of flags (center positions)
    uint32_t *centers;         // the run's center list (job-local sorted ion indices)
    uint32_t *ntile;           // Ec
    uint32_t *toff;            // Ec + 1
    uint32_t *ws;              // prep_workspace_words(E, Ec)
    Tile *tiles;               // the run's tiles (Tile::first indexes centers, Tile::job the job)
    PrepJobStat *js;           // device [njobs]
};
size_t prep_workspace_words(uint64_t E, uint64_t Ec);
hipError_t launch_prep(const PrepLaunch &p, hipStream_t s);

// Query launch over all jobs of a launch.
struct QueryLaunch {
    const JobDev *jobs;        // device array
    int njobs;
    const uint32_t *tile_off;  // device [njobs + 1] prefix of per-job tile counts
    const Tile *const *tiles;  // device [njobs] per-job tile arrays
    uint32_t first_tile;       // global tile range [first_tile, end_tile) of this launch
    uint32_t end_tile;
    uint32_t *queue;           // device counters (kQueueWords), zeroed by the launcher
    uint32_t grab;             // tiles handed out per queue atomic
    int T;                     // centers per tile
    int c16;                   // 16-bit per-center counters
    int mode;                  // 0 pair counts, 1 full-range per-center, 2 hybrid
    // maxKmax: the largest kmax of the jobs whose k are not 1..kmax (0: none)
    int maxC, maxE, maxTrg, maxKmax, maxRows, maxHK;
    int ntypes;                // type ids of the candidates are below this
    int grid;                  // workgroups (persistent)
};
size_t query_lds_bytes(int T, int c16, int mode, int maxC, int maxE, int maxTrg, int maxKmax,
                       int maxRows, int maxHK, int ntypes);
hipError_t query_max_blocks_per_cu(int T, int c16, int mode, size_t lds, int *nb);
hipError_t launch_query(const QueryLaunch &q, hipStream_t s);

// Histograms of the k-th neighbour categories one query chunk wrote
// (JobDev::KB) into K and N1.  unit_off[njobs + 1] is the prefix of the
// per-job work units (groups of kb_group planes x blocks of tiles_per_unit
// tiles); lds_words is the largest (ntasks * kb_group * C + 1) / 2 + ntrg *
// lmax among the jobs that reduce through LDS (16-bit counters, two per word:
// a unit adds at most tiles_per_unit * T * lmax <= 65535 rows to one counter;
// then the task of each center type's rows).
hipError_t launch_kreduce(const JobDev *jobs, int njobs, const uint32_t *unit_off,
                          uint32_t units, int T, uint32_t tiles_per_unit, int lds_words,
                          hipStream_t s);
constexpr int kReduceLdsWords = 16384;  // LDS words of a reduction workgroup at most

// Ion-to-surface squared distances (pp_surfdist.hip): triangle boxes and
// per-cell triangle counts on grid g, the cell lists (cursor starts as a copy
// of the cell offsets), ions in cell order, and one workgroup per tile of
// ions; launch_fill sets n floats to v.
hipError_t launch_fill(float *out, uint64_t n, float v, hipStream_t s);
hipError_t launch_tribox(const float *tri, uint64_t ntri, float R, GridGeom g, float *box,
                         uint32_t *count, hipStream_t s);
hipError_t launch_trifill(const float *box, uint64_t ntri, float R, GridGeom g, uint32_t *cursor,
                          uint32_t *list, hipStream_t s);
// One tile = up to 256 ions of one cell (ions[first, first + count) in cell
// order) tested against that cell's triangles.
struct SurfTile {
    uint32_t cell, first, count, pad;
};
hipError_t launch_ionscatter(const float *xyz, uint64_t n, const uint32_t *key,
                             const uint32_t *rank, const uint32_t *cellstart, float4 *ions,
                             hipStream_t s);
hipError_t launch_surftile(const float4 *ions, const SurfTile *tiles, uint32_t ntiles,
                           const uint32_t *tcellstart, const uint32_t *list, const float *tri,
                           const float *box, float R, float delta, float *out, hipStream_t s);

// POS records (words = 4: 16 bytes each, 16-byte aligned) or EPOS records
// (words = 11: 44 bytes each, 4-byte aligned) -> xyz, mq (may be null) and
// ranged types (pp_io.hip); at most kMaxRanges ranges.
constexpr int kMaxRanges = 4096;
hipError_t launch_posdecode(const uint32_t *rec, uint64_t n, int words, const float *lo,
                            const float *hi, const uint8_t *rtype, int nranges, float *xyz,
                            float *mq, uint8_t *type, hipStream_t s);

// NPCORR (pp_npcorr.hip): one wavefront per center over a cell list at least
// R/2 wide; per-task counts of the K-th env neighbour's binned difference vector.
constexpr int kNpcCap = 512;      // LDS entries per wavefront; K < kNpcCap - 64
constexpr int kNpcBuckets = 256;  // distance buckets per task
struct NpcArgs {
    const float4 *pos;          // cell-sorted (x, y, z, type bits): the centers
    const float *s2;            // cell-sorted ion2surf
    GridGeom g;
    uint64_t n;
    float R2;                   // fl32(R * R)
    int ntasks;
    const uint32_t *trg, *env;  // [ntasks][8] 256-bit type masks
    const float4 *const *epos;  // [ntasks] the ions of the task's env types, cell-sorted on g
    const uint32_t *const *ecs; // [ntasks] their cell offsets (ncells + 1), or null: none
    const float *r1;            // [ntasks] first counting radius (>= maxlen: R at once)
    uint32_t K;                 // SpatStatKNNOrder (0-based)
    float maxlen;               // npc3d fvec_maxlen (SpatStatRadiusMax)
    float inv_width;            // fl32(1 / SpatStatRadiusIncr)
    float inv_bucket;           // kNpcBuckets / R2: distance bucket of d2 = d2 * inv_bucket
    int half;                   // halfedgelen
    uint64_t edge, nxyz;        // edgelen, edgelen^3
    uint32_t *counts;           // [ntasks][nxyz]
};
hipError_t launch_npcorr(const NpcArgs &a, hipStream_t s);

// The fixed scratch of the no-allocation entry point (pp_noalloc.hip): the
// device address of this device's copy (looked up once per device) and its
// size.
constexpr size_t kNoAllocBytes = size_t(256) << 20;
constexpr int kMaxDevices = 64;
hipError_t noalloc_scratch(int device, void **ptr, size_t *bytes);

} // namespace pp
