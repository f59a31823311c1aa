/*
 * pp_horderdist.h -- C++ host mirror of the reference's spatial-statistics task
 * interface, and C entry points to it.
 *
 * The reference drives the RDF / kNN path through
 *   solver::characterize_spatstat                  src/PARAPROBE_SolverHdl.cpp:745-776
 *   horderdist::initialize_descrstat_tasks         src/PARAPROBE_SolverHdl.cpp:3019-3028
 *   horderdist::compute_generic_spatstat3          src/PARAPROBE_SolverHdl.cpp:3685-4296
 *   horderdist::report_apriori_descrstat2 / 3      src/PARAPROBE_SolverHdl.cpp:4299-4453
 * with tasks parsed by parse_tasks / TypeSpecDescrStat::define_iontask4
 * (src/PARAPROBE_PeriodicTable.cpp:498-690) and settings read from the XML keys
 * of src/PARAPROBE_Settings.cpp:347-393, 473-503.  pph::Horderdist keeps those
 * method names and that behaviour; its compute step calls libpp_spatstat.so
 * through include/pp_spatstat.h.
 */
#ifndef PP_HORDERDIST_H
#define PP_HORDERDIST_H

#include <stdint.h>

#include "pp_spatstat.h"

#ifdef __cplusplus
#include <map>
#include <random>
#include <string>
#include <vector>

namespace pph {

/* Element symbols (the reference's NuclidTable keys, iterated in map order)
 * and the ion-type registry IonTypes (name -> type id, "Unknown" -> 0). */
struct PeriodicTable {
    std::map<std::string, int> nuclids;
    std::map<std::string, unsigned> ion_types;
    PeriodicTable();
    void add_iontype(const std::string &name, unsigned id);
    unsigned max_type_id() const;
    /* Ranging (src/PARAPROBE_PeriodicTable.cpp:65-91, 119-176, 324-420):
     * mq2ion[type id] holds that type's closed [lo, hi] mass-to-charge ranges;
     * Unknown (0) has the dummy range [0, 0.001]. */
    struct MQRange {
        float lo, hi;
    };
    std::vector<std::vector<MQRange>> mq2ion;
    std::map<unsigned, std::string> ion_names;
    bool rangefile_loaded = false;
    /* read_rangefile2: single and molecular ion types from the [Ranges]
     * lines of an RRNG file, in order of first appearance. */
    bool read_rangefile2(const std::string &path);
    bool read_rangefile2_text(const std::string &text);
    void add_iontype_single_or_molecular(const std::vector<std::string> &tokens, MQRange r);
    /* mq2ionname: the first type (id ascending) with a range holding mq. */
    unsigned mq2ionname(float mq) const;
    /* The ranges in mq2ionname's order, for pp_pos_decode. */
    void flat_ranges(std::vector<float> &lo, std::vector<float> &hi,
                     std::vector<uint8_t> &type) const;
};

/* TypeSpecDescrStat (src/PARAPROBE_PeriodicTable.h:100-127). */
struct TypeSpecDescrStat {
    std::map<std::string, unsigned> trgcandidates;
    std::map<std::string, unsigned> envcandidates;
    bool define_iontask4(const std::string &command, const PeriodicTable &pse);
};

/* parse_tasks (src/PARAPROBE_PeriodicTable.cpp:653-690): "T1,T2-E1;..." */
void parse_tasks(const std::string &command, std::vector<TypeSpecDescrStat> &these,
                 const PeriodicTable &pse);

/* The MKNN code of compute_generic_spatstat3 (S.cpp:3739-3767): "1;10;100"
 * -> k-1 values sorted descending.  Returns false when a value is below 1
 * (the reference aborts the whole computation then). */
bool parse_mknn_code(const std::string &code, std::vector<int64_t> &kminus1_desc);

/* The spatial-statistics subset of Settings (src/PARAPROBE_Settings.cpp). */
struct SpatstatSettings {
    unsigned simid = 0;
    bool do_rdf = false, do_1nn = false, do_ripk = false, do_mknn = false, do_npcorr = false,
         do_cntnb = false, add_label_rnd = false;
    std::string tasks_code;
    float rmin = 0.f, rincr = 0.1f, rmax = 2.f;
    /* SurfaceCellsCarvingRadius (read only with AnalysisVolumeTessellation 1,
     * Settings.cpp:132, 410-419, 533-536); the surface distance uses
     * max(rmax, carving_radius) (S.cpp:387). */
    float carving_radius = 2.f;
    size_t knn_order = 0;
    std::string mknn_code;
    /* Reads <Parameters> children from an XML text (the keys the reference reads). */
    bool read_xml_text(const std::string &xml, std::string *why);
    /* Settings::checkUserInput for these keys (S.cpp:871-885). */
    bool check(std::string *why) const;
    uint32_t families() const;
};

/* RDF normalisation metadata the reference takes from its volume binner
 * (binner->metadata, S.cpp:4325-4336; size_t and apt_real there,
 * src/PARAPROBE_Datatypes.h:667-669). */
struct VolumeMeta {
    uint64_t nions_inside = 0;
    float volume_inside = 0.f;
};

class Horderdist {
public:
    Horderdist(pp_ctx *ctx, const SpatstatSettings &s, const PeriodicTable &pse);
    void initialize_descrstat_tasks();
    /* Runs the path on the GPU and writes the reference's CSV files into
     * outdir; returns the library status (no exception crosses). */
    pp_status compute_generic_spatstat3(const pp_cloud *cloud, const VolumeMeta &meta,
                                        const std::string &outdir, bool randomized = false);
    const std::vector<TypeSpecDescrStat> &tasks() const { return spatstat_tasks; }
    pp_hist_out last;                       /* statistics of the last run */
    std::vector<int64_t> rdf, nn1, ripk, mknn, cntnb;
    /* NPCORR counts [task][edge^3] (x fastest) of the last run, edge = npcorr_edge */
    std::vector<uint32_t> npcorr;
    int64_t npcorr_edge = 0;
    /* when set (a results file made by create_results_h5), each run also
     * writes its NPCORR cube there, as report_npc3d_histvalues_hdf5 does */
    std::string results_h5;

private:
    pp_ctx *ctx_;
    SpatstatSettings s_;
    const PeriodicTable &pse_;
    std::vector<TypeSpecDescrStat> spatstat_tasks;
};

/* HDF5 output through libpp_h5.so (include/pp_h5.h), found next to this
 * library at run time: whether it loads, and creating the results file with
 * the /DescrSpatStats groups (h5Hdl::create_paraprobe_results_file,
 * src/PARAPROBE_HDF5.cpp:369-408). */
bool hdf5_output_available();
bool create_results_h5(const std::string &fn, bool npcorr);

/* rndlabeler (S.cpp:2809-3000): label randomisation of the whole tip.  The
 * generator is a std::mt19937 seeded with Settings::RndDescrStatsPRNGSeed
 * ((unsigned) -1) and advanced by Settings::RndDescrStatsPRNGDiscard (700000)
 * at construction (Settings.cpp:166-167, S.cpp:2809-2816); reset() keeps its
 * state.  The global type array is the cloud's types in upload order and is
 * permuted with std::shuffle, so a cloud uploaded in the reference's order
 * (z-slabs by thread, KD-tree order inside each) gets the reference's
 * randomised labels when both use the same C++ standard library.  Applying the
 * labels builds a relabelled cloud on the device (pp_cloud_relabel); the
 * original cloud is left untouched. */
class RndLabeler {
public:
    static constexpr uint32_t kSeed = 4294967295u;   /* (unsigned int) -1 */
    static constexpr uint32_t kDiscard = 700000u;
    explicit RndLabeler(pp_ctx *ctx, uint32_t seed = kSeed, uint32_t discard = kDiscard);
    ~RndLabeler();
    RndLabeler(const RndLabeler &) = delete;
    RndLabeler &operator=(const RndLabeler &) = delete;
    /* S.cpp:2825-2835: clears the type arrays, keeps the generator. */
    void reset();
    /* S.cpp:2850-2880: the cloud's types in upload order. */
    pp_status learn_original_types_global(const pp_cloud *cloud);
    /* S.cpp:2882-2904: shuffles a copy once. */
    void shuffle_types_mt19937();
    /* S.cpp:2907-2942: the relabelled cloud (available through cloud()). */
    pp_status apply_shuffled_types_global(const pp_cloud *cloud);
    /* S.cpp:2951-2983: drops the relabelled cloud. */
    void reset_original_types_global();
    bool is_shuffled() const { return shuffled; }
    bool is_applied() const { return applied; }
    const pp_cloud *cloud() const { return rnd_; }
    std::vector<uint8_t> initial_iontypid, temporary_iontypid;

private:
    pp_ctx *ctx_;
    std::mt19937 urng;
    bool shuffled = false, applied = false;
    pp_cloud *rnd_ = nullptr;
};

/* CSV writers (S.cpp:4299-4453).  kind: 1 RDF, 2 1NN, 3 RIPK, 4 MKNN.
 * hist holds [lowest, bins[nbins], highest]. */
bool report_apriori_descrstat2(const std::string &path, int kind, const std::string &metric,
                               float rmin, float rincr, float rmax, const int64_t *hist,
                               const VolumeMeta &meta);
bool report_apriori_descrstat3(const std::string &path, const int64_t *cnts, int64_t len);
std::string report_filename(unsigned simid, const std::string &metric, const std::string &trg,
                            const std::string &env, bool randomized);

} // namespace pph

extern "C" {
#endif

/* C entry points (for callers without C++ and for the tests). */

/* Parses a task code against a registry of ion-type names/ids.  Writes up to
 * max_tasks masks (32 bytes each) and the concatenated trg / env names
 * (names_stride bytes each).  Returns the number of valid tasks. */
int32_t pph_parse_tasks(const char *code, int32_t ntypes, const char *const *type_names,
                        const uint32_t *type_ids, int32_t max_tasks, uint8_t *trg_masks,
                        uint8_t *env_masks, char *trg_names, char *env_names,
                        int32_t names_stride);
/* Parses an MKNN code; returns the number of values (k-1, descending) or -1. */
int32_t pph_parse_mknn_code(const char *code, int64_t *kminus1_desc, int32_t cap);
/* Reads the spatstat keys from XML text; out[0] = families, out[1] = knn order,
 * radii in radii[3]; copies the task and MKNN codes.  Returns 0 on success. */
int32_t pph_read_settings(const char *xml, uint32_t *families, float *radii, char *tasks_code,
                          int32_t tasks_cap, char *mknn_code, int32_t mknn_cap);
/* Writes one report file; returns 0 on success. */
int32_t pph_write_descrstat2(const char *path, int32_t kind, const char *metric, float rmin,
                             float rincr, float rmax, const int64_t *hist, uint64_t nions_inside,
                             float volume_inside);
int32_t pph_write_descrstat3(const char *path, const int64_t *cnts, int64_t len);
/* Reads an RRNG file (read_rangefile2).  Returns the number of ion types
 * (Unknown included) or -1; writes up to cap ranges in mq2ionname's order
 * (lo, hi, type) and *nranges, and the type names (names_stride bytes each,
 * type id order) when names is not NULL. */
int32_t pph_read_rangefile(const char *path, float *lo, float *hi, uint8_t *type, int32_t cap,
                           int32_t *nranges, char *names, int32_t names_stride);
/* End to end from files: settings XML text, a POS file (an EPOS file when its
 * name ends in "epos"), an RRNG file (NULL:
 * no ranging) and a triangle hull (ntri = 0: none) -> the POS decoded, ranged
 * and distanced on the GPU (pp_cloud_from_pos with R = max(SpatStatRadiusMax,
 * SurfaceCellsCarvingRadius)) -> pph_run_spatstat with the range file's ion
 * types.  Returns pp_status. */
int32_t pph_run_pos(pp_ctx *ctx, const char *xml, const char *pos_path, const char *rrng_path,
                    const float *triangles, uint64_t ntri, uint32_t simid,
                    uint64_t nions_inside, float volume_inside, const char *outdir);
/* SpatStatAdditionalLabelRandomization of the XML text: 1 when it reads 1
 * (Settings.cpp:496-501), 0 otherwise, -1 for malformed settings. */
int32_t pph_label_randomization(const char *xml);
/* std::shuffle of n types with a std::mt19937 seeded `seed` and advanced by
 * `discard` (the rndlabeler's shuffle); out may equal in. */
int32_t pph_shuffle_types(uint32_t seed, uint32_t discard, const uint8_t *in, uint64_t n,
                          uint8_t *out);
/* count outputs of a std::mt19937 seeded `seed` after `discard` draws. */
int32_t pph_mt19937(uint32_t seed, uint64_t discard, uint32_t *out, int32_t count);
/* End to end: settings XML text + type registry + a cloud already uploaded ->
 * GPU computation -> the reference's CSV files in outdir (solver::
 * characterize_spatstat, S.cpp:745-776: the original labels, then, when
 * SpatStatAdditionalLabelRandomization is 1, the randomised ones into
 * *.Rndmized.csv).  Returns pp_status. */
int32_t pph_run_spatstat(pp_ctx *ctx, const pp_cloud *cloud, const char *xml, uint32_t simid,
                         int32_t ntypes, const char *const *type_names, const uint32_t *type_ids,
                         uint64_t nions_inside, float volume_inside, const char *outdir);

#ifdef __cplusplus
}
#endif

#endif /* PP_HORDERDIST_H */
