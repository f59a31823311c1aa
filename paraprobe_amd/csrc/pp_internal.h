// pp_internal.h -- structures shared by the host runtime and the gfx950 kernels
// of the RDF / kNN spatial-statistics path.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pp {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kWaves = 4;          // wavefronts per query workgroup
constexpr int kBlock = kWave * kWaves;
constexpr int kMaxRows = 512;      // stencil rows (dy, dz) per query
constexpr int kMaxEnv = 32;        // distinct neighbour (env) types per job
constexpr int kMaxTrg = 32;        // distinct center (target) types per job
constexpr int kMaxTasks = 64;      // type-pair tasks per job
constexpr int kMaxKmax = 4096;     // largest MKNN order
constexpr int kXcds = 8;           // MI355X accelerator complex dies
// Tile queues of a query launch: one counter per XCD, each on its own
// 256-byte line (atomics on one line serialise at ~11 ns each).
constexpr int kQueueStride = 64;
constexpr int kQueueWords = kXcds * kQueueStride;

// Where the x parts of a cell start in a cell list sorted by sub-cell key
// (x part major): byte p - 1 of the 8-byte record is the offset of part p
// (p = 1 .. 2^xbits - 1) from the cell's first ion; byte 7 = 1 marks a cell
// of more than 255 ions, whose ions are never trimmed.
__host__ __device__ inline uint32_t xsplit_at(uint2 rec, int p)
{
    const int b = p - 1;
    return ((b < 4 ? rec.x : rec.y) >> (8 * (b & 3))) & 0xFFu;
}
__host__ __device__ inline bool xsplit_whole(uint2 rec) { return (rec.y >> 24) != 0u; }
// k-th neighbour category of a row that the reduction skips (categories are
// below 65535: jobs with kNN and more bins are refused)
constexpr uint16_t kKbEmpty = 0xFFFFu;

// Stencil row: cells (cx-dxm .. cx+dxm, cy+dy, cz+dz) are contiguous in the
// cell-sorted ion array because the cell key is x-fastest.
struct Row {
    int16_t dy, dz, dxm, pad;
};

// Row::pad flags of the symmetric rows: the part of the own row before the
// own cell (x < cx), or after it (x > cx)
constexpr int16_t kRowBefore = 1;
constexpr int16_t kRowAfter = 2;

// One query tile: up to T considered centers of one type in one cell of one job.
struct Tile {
    uint32_t key;    // linear cell key
    uint32_t first;  // first entry in the job's center list
    uint16_t count;  // centers in this tile
    uint16_t job;    // unused (the launch's tile offsets locate the job)
    uint32_t pad;    // target-type slot of the tile's centers
};

enum : uint32_t {
    FAM_RDF = 1u << 0,
    FAM_1NN = 1u << 1,
    FAM_RIPK = 1u << 2,
    FAM_MKNN = 1u << 3,
    FAM_NPCORR = 1u << 4,
    FAM_CNTNB = 1u << 5,
};

// Everything a query kernel needs to know about one job; lives in device
// memory as an array indexed by Tile::job.
struct JobDev {
    // candidates: cell-sorted ions (x, y, z, type in the bits of w) and cell
    // offsets -- every ion, or only the job's neighbour types on the same grid
    const float4 *pos;
    const uint32_t *cellstart;
    // x-part starts of every cell of the candidate list (xbits = 0: the
    // list has no x parts and rows keep whole cells)
    const uint2 *xsplit;
    int xbits;
    const float4 *cpos;       // centers' cell-sorted ions (the full list)
    int nx, ny, nz;
    // the grid: cell c spans [glo + c * gw, glo + (c + 1) * gw) along x,
    // with gwyz for gw along y and z
    double glo[3], gw, ginv, gwyz;
    double reach;             // R widened by the rounding margins of the in-ball test
    int nrows;
    const Row *rows;
    const uint32_t *centers;  // considered centers, sorted-array indices
    float R2;                 // fl32(R * R)
    // binning: category = q - 1 + (bits(d2) >= edges[q]), q from the nearest
    // ideal edge; edges[] holds the exact smallest d2 bit pattern per category
    int C;                    // nbins + 2 categories (lowest, bins, highest)
    const uint32_t *edges;    // C entries, edges[0] unused
    float idx_scale, idx_off;
    // floor form: category = floor(fmaf(sqrt(d2), idx_scale, idx_off1))
    // wherever that argument lies at least amb from an integer; the kernel
    // tests the shifted coordinate fmaf(sqrt(d2), idx_scale, idx_off1a) with
    // idx_off1a = fl(1 - start tmp + amb): clear when its fraction is at
    // least amb2 = 2 amb, and then its floor is the category
    float idx_off1, idx_off1a, amb2;
    int cat0;                 // category of d = 0 (self pair)
    int cat_sentinel;         // category of float(R + 1e-6)
    // types
    int E;                    // env slots
    int ntrg;                 // target slots
    const int32_t *envoff;    // [256] env slot * C, or -1
    const int32_t *trgslot;   // [256] target slot, or -1
    // tasks
    int ntasks;
    const uint32_t *task_trg; // [ntasks] bit set over target slots
    const uint32_t *task_env; // [ntasks] bit set over env slots
    uint32_t families;
    int kmax, nk;
    const int32_t *kidx_of_k; // [kmax + 1] index into the descending k list, or -1
    int k_dense;              // every k of 1..kmax requested: kidx_of_k[k] = k - 1
    // outputs (device)
    unsigned long long *P;    // [ntrg][E][C] pair counts per center type / env type / category
    uint32_t *K;              // [ntasks][nk][C] MKNN
    uint32_t *N1;             // [ntasks][C] 1NN
    uint32_t *CN;             // [ntasks][cn_cap] neighbour-count histogram
    int cn_cap;
    int percenter;            // 1 when MKNN, 1NN or CNTNB need per-center histograms
    unsigned long long *stats;// [0] in-ball pairs (all types), [1] candidate evaluations
    int count_in;             // count in-ball pairs in stats[0]; 0 when every candidate is
                              // a neighbour type and the pair counts give the total
    // tasks of each center type: trg_tasks[slot * lmax + l], l < trg_ntask[slot]
    int lmax;
    const uint32_t *trg_ntask;
    const uint32_t *trg_tasks;
    // k-th neighbour categories of the tiles of one launch chunk, reduced into
    // K / N1 afterwards: per tile, npad (nplanes rounded up to a multiple of
    // 8) planes of T * lmax uint16 (row c * lmax + l: center c of the tile,
    // its center type's task trg_tasks[slot * lmax + l]; plane nk is the 1NN
    // when requested), so that the tiles' blocks and their groups of planes
    // start on cache lines.  Rows of absent centers (c >= the tile's count),
    // of missing tasks (l >= trg_ntask[slot]) and of the padding planes are
    // never written and the reduction skips them; rows of a center handed to
    // the full-range rerun hold kKbEmpty in every plane.  Tile tl of the
    // chunk is the job's tile tile_base + tl.
    uint16_t *KB;
    const Tile *tiles;        // the job's tiles (their slot and center count)
    int nplanes, npad;
    int kb_group;             // planes per reduction work unit (8, 4, 2 or 1)
    uint32_t tile_base;
    uint32_t kb_tiles;        // tiles of this job in the chunk
    // kNN-range histograms (hybrid mode): per center, env slot e holds the
    // categories [0, ck_e) at ckoff[e] (HK words in all); a pair enters them
    // iff d2 <= kthr of its type, i.e. iff its category is below ck_e.
    // kinfo[2 t] = ckoff of type t (or -1), kinfo[2 t + 1] = kthr bits.
    // task_ck[tk] = categories complete for task tk, bit 31 set when that is
    // short of the full range: a center with fewer than k neighbours there
    // is flagged in unres (per sorted ion) and re-run over the full range.
    int HK;
    int trunc_any;
    const int32_t *kinfo;
    const int32_t *ckoff;
    const uint32_t *task_ck;
    uint32_t *unres;
    // Symmetric pair counting (pair-count and hybrid modes): the tiles of center type slot
    // sym_slot (-1: none) read the candidates of their own type from list S,
    // whose float4 w holds the ion's ion2surf (every ion of S has that type),
    // along srows[0, s_ns): the own cell (srows[0]), then the cells before it
    // in cell order (srows[1, s_nb1)), then the cells after it.  A pair with a
    // center in a cell after the own cell is counted for both centers (the
    // same distance, so the same category, and the same pair-count row);
    // with a center in a cell before it, it is left to that center's tile
    // (only its kNN range is read there).  A row before the own cell whose
    // own-type ions are all centers (cellmin_s: the least ion2surf per cell)
    // is cut to the kNN reach (left out without kNN).  The other candidates come from list O along
    // srows[s_ns, nsrows), counted as in any other tile.
    int sym_slot, sym_type;
    int nsrows, s_nb1, s_ns;
    const Row *srows;
    double kreach;            // the own type's kNN range widened as reach (< 0: none)
    const float4 *pos_s;
    const uint32_t *cellstart_s;
    const uint2 *xsplit_s;
    int xbits_s;
    const float *cellmin_s;
    const float4 *pos_o;
    const uint32_t *cellstart_o;
    const uint2 *xsplit_o;
    int xbits_o;
};

} // namespace pp
