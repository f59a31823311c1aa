/*
 * pp_spatstat.h -- C ABI of the MI355X (gfx950) RDF / kNN spatial-statistics path.
 *
 * Drop-in replacement for the compute step of the reference's spatial
 * statistics: the KD-tree range-ball queries and the OpenMP center loop of
 *   horderdist::compute_generic_spatstat3()      src/PARAPROBE_SolverHdl.cpp:3685-4296
 * fed by
 *   solver::init_spatialindexing()               src/PARAPROBE_SolverHdl.cpp:687-731
 * and called from
 *   solver::characterize_spatstat()              src/PARAPROBE_SolverHdl.cpp:745-776.
 * The reference function takes no arguments and reads globals
 * (Settings::SpatStat*, spatstat_tasks, the per-thread KD-trees); here every
 * input is explicit, every host buffer is owned by the caller, the library owns
 * device buffers behind opaque handles, and no exception crosses the ABI.
 *
 * Histogram layout (every family): one int64 histogram is
 *   [cnts_lowest, cnts[0 .. nbins-1], cnts_highest]       (nbins + 2 entries)
 * matching histogram::cnts_lowest / cnts / cnts_highest
 * (src/PARAPROBE_Histogram.h:168-170; the reference keeps doubles, which are
 * exact for these integer counts).  Family arrays are task-major in task parse
 * order; MKNN is [task][kth][nbins + 2] with kth indexing the k list sorted in
 * descending order (S.cpp:3759, 4086); CNTNB is [task][cntnb_cap] indexed by
 * the neighbour count (discrhistogram, src/PARAPROBE_DiscreteHistogram.h:83-97).
 */
#ifndef PP_SPATSTAT_H
#define PP_SPATSTAT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pp_status {
    PP_OK = 0,
    PP_ERR_INVALID_ARG = 1,        /* bad settings or arguments (cf. Settings::checkUserInput) */
    PP_ERR_OOM_DEVICE = 2,         /* device allocation failed (cf. bad_alloc catches) */
    PP_ERR_HIP = 3,                /* HIP runtime error */
    PP_ERR_ADMISSION_REJECTED = 4, /* serving: the job can never fit the HBM budget */
    PP_ERR_NO_DEVICE = 5,          /* no gfx950 device */
    PP_ERR_UNSUPPORTED = 6,        /* e.g. NPCORR, or bins finer than float resolution */
    PP_ERR_OVERFLOW = 7,           /* a CNTNB count exceeded cntnb_cap */
    PP_ERR_NOT_READY = 8           /* serving: ticket still in flight */
} pp_status;

/* Family bits: digit d of the reference's AnalysisSpatDistrType maps to
 * bit (d - 1) (src/PARAPROBE_Settings.cpp:347-393). */
#define PP_FAM_RDF (1u << 0)
#define PP_FAM_1NN (1u << 1)
#define PP_FAM_RIPK (1u << 2)
#define PP_FAM_MKNN (1u << 3)
#define PP_FAM_NPCORR (1u << 4) /* through pp_npcorr_run, not pp_spatstat_run */
#define PP_FAM_CNTNB (1u << 5)

typedef struct pp_ctx pp_ctx;
typedef struct pp_cloud pp_cloud;
typedef struct pp_server pp_server;

/* One spatial-statistics job: the settings compute_generic_spatstat3 reads
 * (Settings::SpatStatRadiusMin/Incr/Max, SpatStatDo*, DescrStatMKNNCode) and
 * the parsed tasks (TypeSpecDescrStat trgcandidates/envcandidates,
 * src/PARAPROBE_PeriodicTable.h:100-127) as 256-bit type masks. */
typedef struct pp_job {
    int32_t ntasks;
    const uint8_t *trg_mask;     /* ntasks x 32 bytes; bit t: type id t is a central type */
    const uint8_t *env_mask;     /* ntasks x 32 bytes; bit t: type id t is a neighbour type */
    float rmin, rincr, rmax;     /* SpatStatRadiusMin / Incr / Max (float settings) */
    uint32_t families;           /* PP_FAM_* */
    int32_t nk;                  /* MKNN: number of k values */
    const int64_t *kminus1_desc; /* MKNN: k - 1 values in descending order (S.cpp:3739-3760) */
} pp_job;

/* Caller-allocated outputs.  Any family pointer may be NULL when that family
 * is not requested. */
typedef struct pp_hist_out {
    int64_t *rdf;    /* [ntasks][nbins + 2] */
    int64_t *nn1;    /* [ntasks][nbins + 2] */
    int64_t *ripk;   /* [ntasks][nbins + 2] */
    int64_t *mknn;   /* [ntasks][nk][nbins + 2] */
    int64_t *cntnb;  /* [ntasks][cntnb_cap] */
    int64_t cntnb_cap;
    /* out: */
    int64_t nbins;
    int64_t cntnb_len;          /* largest count seen + 1 (discrhistogram size) */
    int64_t centers_considered; /* MyIonsCurrRegionConsider summed over regions */
    int64_t ions_discarded;     /* MyIonsCurrRegionDiscard summed over regions */
    int64_t pairs_inball;       /* sum over considered centers of |neighbours within R|, all types */
    int64_t pairs_candidate;    /* candidate pair-distance evaluations made by the GPU */
    double ms_index;            /* cell-list build (0 when resident) */
    double ms_prep;             /* center flags, compaction, tiles */
    double ms_query;            /* query kernel */
    double ms_total;            /* whole call, host wall clock */
    int32_t tile_centers;       /* query-kernel configuration used: centers per tile, */
    int32_t cells_per_rmax;     /*   cell width along x as a fraction of Rmax (0: resident list), */
    int32_t counters16;         /*   16-bit per-center counters, */
    int32_t percenter;          /*   per-center histograms: 0 none, 1 full range, 2 kNN range
                                 *   only (with an exact full-range rerun of the centers
                                 *   whose k-th neighbour lies beyond it) */
    int64_t centers_rerun;      /* centers of that full-range rerun */
    double ms_kernel;           /* the query kernel launches alone (HIP events around each;
                                 * 0 when they ran inside a HIP graph) */
} pp_hist_out;

/* Context bound to one device; one per host thread (calls on a context are
 * serialised by the caller). */
pp_status pp_spatstat_create(int32_t device, pp_ctx **out);
pp_status pp_spatstat_destroy(pp_ctx *ctx);
const char *pp_last_error_string(const pp_ctx *ctx);

/* Histogram geometry of histogram(rmin, rincr, rmax)
 * (src/PARAPROBE_Histogram.h:54-89), for sizing outputs. */
pp_status pp_hist_geometry(float rmin, float rincr, float rmax, int64_t *nbins);

/* Whether the GPU path bins histogram(rmin, rincr, rmax) exactly (PP_OK), or
 * refuses it (PP_ERR_UNSUPPORTED: bins finer than the float index coordinate
 * resolves; PP_ERR_INVALID_ARG: settings Settings::checkUserInput rejects).
 * Host only; a job with such settings fails the same way. */
pp_status pp_hist_supported(float rmin, float rincr, float rmax);

/* Uploads a point cloud (ions: x,y,z AoS float, uint8 type id, squared
 * distance to the tip surface -- the reference's ion2surf -- as float).
 * cell_width > 0 builds a resident cell list of that width now (serving /
 * batching: every later job with any Rmax reuses it); cell_width == 0 defers
 * the cell list to each job (width chosen from that job's Rmax).
 * Replaces decompositor::loadpartitioning + threadmemory::build_local_kdtree
 * (S.cpp:1819-1934, 1593-1663). */
pp_status pp_cloud_upload(pp_ctx *ctx, const float *xyz, const uint8_t *type,
                          const float *ion2surf_sqr, uint64_t n, float cell_width,
                          pp_cloud **out);
/* Same, from device pointers already resident in HBM (no host copy). */
pp_status pp_cloud_from_device(pp_ctx *ctx, const float *d_xyz, const uint8_t *d_type,
                               const float *d_ion2surf_sqr, uint64_t n, float cell_width,
                               pp_cloud **out);
pp_status pp_cloud_free(pp_ctx *ctx, pp_cloud *cloud);
/* Number of ions of a cloud (the sum over the z-slabs the reference forms in
 * rndlabeler::learn_original_types_global, S.cpp:2856-2859). */
pp_status pp_cloud_count(const pp_cloud *cloud, uint64_t *n);
/* Copies the cloud's ion types (n bytes, upload order) to host memory:
 * rndlabeler::learn_original_types_global (S.cpp:2850-2880). */
pp_status pp_cloud_types(pp_ctx *ctx, const pp_cloud *cloud, uint8_t *types);
/* A new cloud with src's positions and ion2surf (copied on the device) and
 * the given host types (n bytes, src's upload order), resident with src's
 * cell widths when src is resident: the relabelled tip of
 * rndlabeler::apply_shuffled_types_global (S.cpp:2907-2942).  Free it with
 * pp_cloud_free (rndlabeler::reset_original_types_global, S.cpp:2951-2983). */
pp_status pp_cloud_relabel(pp_ctx *ctx, const pp_cloud *src, const uint8_t *types,
                           pp_cloud **out);
/* Adds one more resident cell list of the given width to a resident cloud;
 * each job then uses the resident width that suits its Rmax best. */
pp_status pp_cloud_add_cell_list(pp_ctx *ctx, pp_cloud *cloud, float cell_width);
/* Device bytes a cloud holds. */
pp_status pp_cloud_footprint(const pp_cloud *cloud, uint64_t *bytes);

/* Runs one job synchronously: the compute step of compute_generic_spatstat3
 * (S.cpp:3702-4176).  Histograms are written to *out. */
pp_status pp_spatstat_run(pp_ctx *ctx, const pp_cloud *cloud, const pp_job *job, pp_hist_out *out);

/* pp_spatstat_run without device allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Every temporary of the job (center lists, tiles, tables, counters, the k-th
 * neighbour categories) is carved from a fixed device scratch of 256 MiB that
 * the library holds per device until the process ends; a job that needs more
 * (roughly 32 bytes per ion plus the k-th neighbour categories) returns
 * PP_ERR_UNSUPPORTED.  The cloud must be resident (uploaded with a cell
 * width), or have the cell list of this job's width from an earlier
 * pp_spatstat_run; a job whose neighbour types would get a typed candidate
 * list uses the full list instead when that typed list does not exist yet
 * (same results).  Results equal pp_spatstat_run's.  Calls on one device take
 * turns on the scratch (a lock inside the library); each call clears what it
 * needs cleared, so nothing carries over between calls. */
pp_status pp_spatstat_run_noalloc(pp_ctx *ctx, const pp_cloud *cloud, const pp_job *job,
                                  pp_hist_out *out);

/* Runs njobs jobs (each on its own cloud) in one persistent launch of the
 * query kernel; outs[i] receives job i.  Results equal njobs separate
 * pp_spatstat_run calls. */
pp_status pp_batch_run(pp_ctx *ctx, int32_t njobs, const pp_cloud *const *clouds,
                       const pp_job *jobs, pp_hist_out *outs);

/* Query-kernel tuning: centers per tile (4, 8, 16 or 32) and the cell width as a
 * fraction of Rmax (cells per Rmax: 1..8).  0 = built-in choice (density
 * rule); -1 for both = autotune: the first job of each (density x Rmax^3,
 * families) bucket times the candidate shapes (8 or 16 centers per tile;
 * cubic cells of Rmax/1 .. Rmax/5 and, for the lists jobs build for
 * themselves, cells of Rmax by Rmax/2 and Rmax/2 by Rmax/4), the faster of
 * two runs each, and the winner is kept for the bucket. */
pp_status pp_set_tuning(pp_ctx *ctx, int32_t centers_per_tile, int32_t cells_per_rmax);

/* Cell shape of the cell lists jobs build for themselves (clouds uploaded
 * without a cell width): cells of Rmax / cells_per_rmax along x and
 * yz_per_x times finer along y and z (1..4).  0 = built-in choice: cubic
 * cells, except where the density rule picks cells of Rmax/4 or Rmax/2,
 * which it builds twice as long along x (Rmax/2 by Rmax/4, Rmax by
 * Rmax/2).  Resident
 * clouds keep their cubic cells.  Results do not depend on it. */
pp_status pp_set_cell_aspect(pp_ctx *ctx, int32_t yz_per_x);

/* Symmetric pair counting (1 = allowed, the default; 0 = off).  In a job with
 * RDF or RIPK (alone, or with MKNN / 1NN over kNN ranges), the center type
 * with the most ions, when it is also a neighbour type holding at least half
 * of the candidates, counts each pair of two of its centers once for both:
 * the same squared distance from either side gives the same category.  A
 * tuning switch of this library (no counterpart in the reference); results
 * do not depend on it. */
pp_status pp_set_pair_symmetry(pp_ctx *ctx, int32_t on);

/* Squared distance of every ion to a triangulated tip surface, the ion2surf
 * input of pp_cloud_upload: replaces solver::surface_distancing2
 * (src/PARAPROBE_SolverHdl.cpp:312-484).  Each ion gets the smallest
 * closestPointOnTriangle value (src/PARAPROBE_Math.cpp:88-171) over the
 * triangles whose vertex box overlaps its box [p - R, p + R], or fl32(R * R)
 * when that is smaller or there is none; pass R = max(SpatStatRadiusMax,
 * SurfaceCellsCarvingRadius).  xyz: n x 3 floats, triangles: ntri x 9 floats
 * (three vertices), out: n floats; all host memory.  The expression is the
 * reference's, quirks included (its region-6 far branch clamps -e/c where
 * the geometry asks for -d/a).  A NaN distance (only from arithmetic that
 * overflows) is skipped; the reference's result then depends on its R-tree's
 * candidate order. */
pp_status pp_ion2surf(pp_ctx *ctx, const float *xyz, uint64_t n, const float *triangles,
                      uint64_t ntri, float R, float *ion2surf_sqr);

/* pp_ion2surf on device memory (the same device as ctx): d_xyz, d_triangles
 * and d_out stay on the GPU, so the result can feed pp_cloud_from_device. */
pp_status pp_ion2surf_device(pp_ctx *ctx, const float *d_xyz, uint64_t n,
                             const float *d_triangles, uint64_t ntri, float R, float *d_out);

/* NPCORR, the n-point correlation family (AnalysisSpatDistrType digit 5;
 * S.cpp:3899-3960, npc3d, src/PARAPROBE_NPointCorrelations3D.cpp:50-207).
 * pp_npcorr_geometry gives the edge length of its cubic histogram of
 * npc3d(rmax, rincr): 1 + 2 (ceilf(rmax / rincr) + 1), or -1 when edge^3
 * exceeds 503^3 (the reference then skips the statistics).
 * pp_npcorr_run counts, per task, the difference vector (neighbour - center)
 * of the knn_order-th (0-based, SpatStatKNNOrder) nearest env-type neighbour
 * of every considered center of a central type that has more than knn_order
 * neighbours of any type within rmax, binned to the nearest cell (halves away
 * from zero) when its length is <= rmax; counts is [ntasks][edge^3] uint32
 * (npc3d::cnts), x fastest.  Equal distances are ordered by the squared
 * distance, then the smaller (u, v, w) (the reference leaves them to
 * nth_element over its KD-tree order).  knn_order must be below 448.  *centers_binned (may be NULL)
 * receives the number of counts added. */
int64_t pp_npcorr_geometry(float rmax, float rincr);
pp_status pp_npcorr_run(pp_ctx *ctx, const pp_cloud *cloud, int32_t ntasks,
                        const uint8_t *trg_mask, const uint8_t *env_mask, float rmax,
                        float rincr, int64_t knn_order, uint32_t *counts,
                        int64_t *centers_binned);

/* POS records to positions and ranged ion types on the GPU: replaces the
 * record conversion of solverHdl::load_pos_sequentially
 * (S.cpp:6797-6930, src/PARAPROBE_EPOSEndianness.h:45-84: 16-byte records of
 * four big-endian floats x, y, z, mass-to-charge) and solverHdl::
 * identify_ions / PeriodicTable::mq2ionname (S.cpp:7340-7438,
 * src/PARAPROBE_PeriodicTable.cpp:65-91).  nbytes must be a multiple of 16.
 * Ranges are tested in the given order (the reference's: type id ascending,
 * Unknown's dummy range [0, 0.001] first); an ion takes the type of the first
 * closed range [lo, hi] holding its mass-to-charge (a NaN matches the first
 * range), else type 0.  nranges = 0 (no range file) gives type 0 everywhere.
 * Outputs: xyz (3 floats per ion), mq (may be NULL), type (1 byte per ion);
 * all host memory here. */
pp_status pp_pos_decode(pp_ctx *ctx, const void *pos, uint64_t nbytes, const float *range_lo,
                        const float *range_hi, const uint8_t *range_type, int32_t nranges,
                        float *xyz, float *mq, uint8_t *type);
/* Same with pos, xyz, mq and type in device memory (pos 16-byte aligned);
 * the ranges stay host arrays.  The outputs can feed pp_ion2surf_device and
 * pp_cloud_from_device directly. */
pp_status pp_pos_decode_device(pp_ctx *ctx, const void *d_pos, uint64_t nbytes,
                               const float *range_lo, const float *range_hi,
                               const uint8_t *range_type, int32_t nranges, float *d_xyz,
                               float *d_mq, uint8_t *d_type);

/* The same for EPOS files (solverHdl::load_epos_sequentially, S.cpp:
 * 6933-7066; epos(epos_be_2_stl_le), src/PARAPROBE_EPOSEndianness.h:86-160:
 * 44-byte records of nine big-endian floats x, y, z, mass-to-charge, tof,
 * vdc, vpu, detx, dety and two big-endian ints); x, y, z and the
 * mass-to-charge are decoded.  nbytes must be a multiple of 44. */
pp_status pp_epos_decode(pp_ctx *ctx, const void *epos, uint64_t nbytes, const float *range_lo,
                         const float *range_hi, const uint8_t *range_type, int32_t nranges,
                         float *xyz, float *mq, uint8_t *type);

/* POS bytes + ranges + triangle hull -> a cloud, entirely on the device:
 * pp_pos_decode, then pp_ion2surf with R (an empty hull gives every ion
 * fl32(R * R), as solver::surface_distancing2 does without a hull,
 * S.cpp:333-358), then the cloud as pp_cloud_upload would make it.  pos and
 * triangles are host memory. */
pp_status pp_cloud_from_pos(pp_ctx *ctx, const void *pos, uint64_t nbytes, const float *range_lo,
                            const float *range_hi, const uint8_t *range_type, int32_t nranges,
                            const float *triangles, uint64_t ntri, float R, float cell_width,
                            pp_cloud **out);
/* The same from EPOS bytes (44-byte records). */
pp_status pp_cloud_from_epos(pp_ctx *ctx, const void *epos, uint64_t nbytes,
                             const float *range_lo, const float *range_hi,
                             const uint8_t *range_type, int32_t nranges, const float *triangles,
                             uint64_t ntri, float R, float cell_width, pp_cloud **out);

/* kNN range of MKNN / 1NN jobs without CNTNB: per neighbour type, per-center
 * histograms cover only the distance within which factor x max(k, (k +
 * 6 sqrt(k) + 10) / 4) neighbours are expected at the cloud's mean density
 * over its occupied volume (default 4); centers with fewer than k neighbours
 * there are re-run over the full range, so results do not depend on the
 * factor.  0 = full-range per-center histograms always. */
pp_status pp_set_knn_range(pp_ctx *ctx, float factor);

/* Launch form of the query kernel (no counterpart in the reference, whose
 * OpenMP loop has no launches): 1 (default) = each launch group of a run of
 * several jobs (a pp_batch_run batch, a serving batch) -- the persistent
 * query launches of its chunks and their k-th neighbour reductions -- is
 * captured into one HIP graph and launched as a unit; 0 = the same launches
 * issued one by one.  Single-job runs always launch one by one (a graph's
 * instantiation costs more than it saves there).  Results are identical
 * either way. */
pp_status pp_set_graph_launch(pp_ctx *ctx, int32_t on);

/* ---- serving: a request loop over resident clouds (continuous batching) ----
 * Requests are queued by pp_serve_submit and admitted by their device-memory
 * footprint against hbm_budget_bytes; each of `slots` workers takes the
 * admitted requests waiting at that moment, in arrival order (up to
 * max_batch_jobs and batch_work estimated candidate evaluations; a larger
 * request runs alone), and runs them as one batch on its own HIP stream: a batch's
 * preparation overlaps on the device with the query of the batch before it,
 * whose query launches it then follows (queries keep arrival order).  Clouds must be resident (uploaded with a cell width).
 * The caller keeps each request's pp_job arrays and pp_hist_out buffers alive
 * until its ticket is reported done. */
typedef struct pp_serve_opts {
    int32_t slots;              /* concurrent batches (1..8) */
    int32_t max_batch_jobs;     /* requests per batch (1..65535) */
    uint64_t hbm_budget_bytes;  /* admission budget for per-request device memory; 0: half the free HBM */
    double batch_work;          /* estimated candidate evaluations per batch (a batch's
                                   requests all complete when its last one does, so this
                                   bounds the wait a batch adds to its first request);
                                   0: 5e9, about 3 ms of query kernel */
} pp_serve_opts;

pp_status pp_serve_create(pp_ctx *ctx, const pp_serve_opts *opts, pp_server **out);
/* Enqueues one request; *ticket identifies it.  PP_ERR_ADMISSION_REJECTED when
 * its footprint exceeds the whole budget (it could never be admitted). */
pp_status pp_serve_submit(pp_server *srv, const pp_cloud *cloud, const pp_job *job,
                          pp_hist_out *out, uint64_t *ticket);
/* Returns up to cap completed tickets (with their status) not returned before. */
pp_status pp_serve_poll(pp_server *srv, uint64_t *tickets, int32_t *statuses, int32_t cap,
                        int32_t *n);
/* Blocks until every submitted request has completed. */
pp_status pp_serve_drain(pp_server *srv);
pp_status pp_serve_destroy(pp_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* PP_SPATSTAT_H */
