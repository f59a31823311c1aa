// pp_horderdist.cpp -- C++ host mirror of the reference's spatial-statistics
// task interface (tasks, settings, compute driver, CSV writers) on top of the
// gfx950 C ABI.
#include "pp_horderdist.h"

#include <dlfcn.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

namespace pph {

namespace {

constexpr double kEpsilon = 1.0e-6;                  // src/PARAPROBE_Numerics.h:45
constexpr double kPi = 3.141592653589793238462643383279502884197169399375105820974;

// Element symbols of the reference's NuclidTable (hydrogen to uranium, plus
// plutonium); a std::map, so every scan below visits them in lexicographic
// key order, as the reference does.
const char *const kElements[] = {
    "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si",
    "P",  "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni",
    "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo",
    "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe", "Cs", "Ba",
    "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb",
    "Lu", "Hf", "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po",
    "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U",  "Pu"};

bool is_number(const std::string &s)
{
    if (s.empty())
        return false;
    for (char c : s)
        if (!std::isdigit((unsigned char)c))
            return false;
    return true;
}

// One comma-separated item of a task code resolved to an ion-type name, as in
// TypeSpecDescrStat::define_iontask4 (src/PARAPROBE_PeriodicTable.cpp:543-585):
// collect every element symbol the item contains (map order); if the item is
// exactly one of them it is a single-element ion, otherwise the molecular ion
// named by the concatenation of all of them.
std::string resolve_item(const std::string &item, const PeriodicTable &pse)
{
    std::vector<std::string> found;
    for (const auto &kv : pse.nuclids)
        if (item.find(kv.first) != std::string::npos)
            found.push_back(kv.first);
    for (const auto &k : found)
        if (item == k)
            return k;
    std::string cand;
    for (const auto &k : found)
        cand += k;
    return cand;
}

bool resolve_list(const std::string &list, const PeriodicTable &pse,
                  std::map<std::string, unsigned> &out)
{
    const size_t n = (size_t)std::count(list.begin(), list.end(), ',') + 1;
    std::istringstream in(list);
    std::string item;
    for (size_t i = 0; i < n; ++i) {
        std::getline(in, item, ',');  // an exhausted stream yields an empty item
        if (!in && item.size())
            item.clear();
        const std::string cand = resolve_item(item, pse);
        auto it = pse.ion_types.find(cand);
        if (it != pse.ion_types.end())
            out.insert(*it);
    }
    return !out.empty();
}

} // namespace

PeriodicTable::PeriodicTable()
{
    for (const char *e : kElements)
        nuclids.emplace(e, 0);
    ion_types["Unknown"] = 0;  // UNKNOWNTYPE (src/PARAPROBE_Numerics.h:78)
}

void PeriodicTable::add_iontype(const std::string &name, unsigned id) { ion_types[name] = id; }

unsigned PeriodicTable::max_type_id() const
{
    unsigned m = 0;
    for (const auto &kv : ion_types)
        m = std::max(m, kv.second);
    return m;
}

// ---------------------------------------------------------------------------
// ranging (src/PARAPROBE_PeriodicTable.cpp)

namespace {

// boost::split(tokens, line, boost::is_any_of(" \t")): every delimiter ends a
// token, so runs of delimiters give empty tokens
std::vector<std::string> split_blank_tab(const std::string &line)
{
    std::vector<std::string> out(1);
    for (char ch : line) {
        if (ch == ' ' || ch == '\t')
            out.emplace_back();
        else
            out.back().push_back(ch);
    }
    return out;
}

} // namespace

void PeriodicTable::add_iontype_single_or_molecular(const std::vector<std::string> &in, MQRange r)
{
    // element names found in the tokens between "Vol:..." and the last one,
    // concatenated in the table's key order; one-letter names need "X:"
    std::string cand;
    for (const auto &kv : nuclids) {
        const std::string &key = kv.first;
        const bool single = key.size() == 1;
        for (size_t t = 3; t + 1 < in.size(); ++t) {
            const bool hit = single ? in[t].find(key + ":") != std::string::npos
                                    : in[t].find(key) != std::string::npos;
            if (hit) {
                cand += key;
                break;
            }
        }
    }
    auto it = ion_types.find(cand);
    if (it == ion_types.end()) {
        const unsigned id = (unsigned)ion_types.size();
        ion_types[cand] = id;
        ion_names[id] = cand;
        if (mq2ion.size() <= id)
            mq2ion.resize(id + 1);
        mq2ion[id].push_back(r);
    } else {
        if (mq2ion.size() <= it->second)
            mq2ion.resize(it->second + 1);
        mq2ion[it->second].push_back(r);
    }
}

bool PeriodicTable::read_rangefile2_text(const std::string &text)
{
    ion_types.clear();
    ion_names.clear();
    mq2ion.assign(1, std::vector<MQRange>{{0.0f, (float)0.001}});  // Unknown's dummy range
    ion_types["Unknown"] = 0;
    ion_names[0] = "Unknown";
    std::istringstream f(text);
    std::string ln;
    while (f.good()) {
        std::getline(f, ln);
        if (ln.find("[Ranges]") != std::string::npos)
            break;
    }
    while (f.good()) {
        std::getline(f, ln);
        const std::vector<std::string> tk = split_blank_tab(ln);
        if (tk.size() >= 5 && tk[0].find("Range") != std::string::npos &&
            tk[2].find("Vol") != std::string::npos) {
            std::istringstream line(ln);
            std::string piece;
            std::getline(line, piece, '=');
            std::getline(line, piece, ' ');
            const float lo = std::stof(piece);
            std::getline(line, piece, ' ');
            const float hi = std::stof(piece);
            add_iontype_single_or_molecular(tk, MQRange{lo, hi});
        }
    }
    for (auto &v : mq2ion)
        std::stable_sort(v.begin(), v.end(),
                         [](const MQRange &a, const MQRange &b) { return a.hi < b.lo; });
    rangefile_loaded = true;
    return true;
}

bool PeriodicTable::read_rangefile2(const std::string &path)
{
    std::ifstream f(path);
    if (!f.is_open()) {
        rangefile_loaded = false;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return read_rangefile2_text(ss.str());
}

unsigned PeriodicTable::mq2ionname(float mq) const
{
    if (rangefile_loaded)
        for (size_t i = 0; i < mq2ion.size(); ++i)
            for (const MQRange &r : mq2ion[i])
                if (!(mq < r.lo) && !(mq > r.hi))
                    return (unsigned)i;
    return 0;
}

void PeriodicTable::flat_ranges(std::vector<float> &lo, std::vector<float> &hi,
                                std::vector<uint8_t> &type) const
{
    lo.clear();
    hi.clear();
    type.clear();
    if (!rangefile_loaded)
        return;
    for (size_t i = 0; i < mq2ion.size(); ++i)
        for (const MQRange &r : mq2ion[i]) {
            lo.push_back(r.lo);
            hi.push_back(r.hi);
            type.push_back((uint8_t)i);
        }
}

bool TypeSpecDescrStat::define_iontask4(const std::string &command, const PeriodicTable &pse)
{
    if (command.empty())
        return false;
    if (std::count(command.begin(), command.end(), '-') != 1)
        return false;
    std::istringstream line(command);
    std::string centers, nbors;
    std::getline(line, centers, '-');
    std::getline(line, nbors, '-');
    if (!resolve_list(centers, pse, trgcandidates))
        return false;
    if (!resolve_list(nbors, pse, envcandidates))
        return false;
    return true;
}

void parse_tasks(const std::string &command, std::vector<TypeSpecDescrStat> &these,
                 const PeriodicTable &pse)
{
    if (command.empty())
        return;
    const int n = (int)std::count(command.begin(), command.end(), ';') + 1;
    std::stringstream in;
    in << command;
    std::string piece;
    for (int t = 0; t < n; ++t) {
        std::getline(in, piece, ';');
        TypeSpecDescrStat task;
        if (task.define_iontask4(piece, pse))
            these.push_back(task);  // invalid tasks are dropped
    }
}

bool parse_mknn_code(const std::string &code, std::vector<int64_t> &out)
{
    out.clear();
    if (code.empty() || code.find('-') != std::string::npos)
        return true;  // no k values: no MKNN histograms
    const int nsemicola = (int)std::count(code.begin(), code.end(), ';');
    std::istringstream line(code);
    std::string piece;
    for (int i = 0; i <= nsemicola; ++i) {
        std::getline(line, piece, ';');
        if (!line && piece.size())
            piece.clear();
        if (is_number(piece)) {
            const unsigned long v = std::stoul(piece);
            if (v < 1)
                return false;  // "Instructing a kth order value smaller than 1"
            out.push_back((int64_t)v - 1);
        }
    }
    std::sort(out.begin(), out.end(), [](int64_t a, int64_t b) { return a > b; });
    return true;
}

// ---------------------------------------------------------------------------
// settings

namespace {

// Children <Key>text</Key> of the first element of an XML document (the
// reference reads tree.first_node() and its named children with RapidXML).
std::map<std::string, std::string> xml_children(const std::string &xml)
{
    std::string s;
    // drop comments and declarations
    for (size_t i = 0; i < xml.size();) {
        if (xml.compare(i, 4, "<!--") == 0) {
            size_t e = xml.find("-->", i + 4);
            i = (e == std::string::npos) ? xml.size() : e + 3;
        } else if (xml.compare(i, 2, "<?") == 0) {
            size_t e = xml.find("?>", i + 2);
            i = (e == std::string::npos) ? xml.size() : e + 2;
        } else {
            s += xml[i++];
        }
    }
    std::map<std::string, std::string> kv;
    size_t root = s.find('<');
    if (root == std::string::npos)
        return kv;
    size_t rend = s.find('>', root);
    if (rend == std::string::npos)
        return kv;
    size_t i = rend + 1;
    while (true) {
        size_t a = s.find('<', i);
        if (a == std::string::npos || s[a + 1] == '/')
            break;
        size_t b = s.find('>', a);
        if (b == std::string::npos)
            break;
        std::string name = s.substr(a + 1, b - a - 1);
        if (!name.empty() && name.back() == '/') {  // <Key/>
            kv.emplace(name.substr(0, name.size() - 1), "");
            i = b + 1;
            continue;
        }
        const std::string close = "</" + name + ">";
        size_t c = s.find(close, b + 1);
        if (c == std::string::npos)
            break;
        kv.emplace(name, s.substr(b + 1, c - b - 1));  // first occurrence wins
        i = c + close.size();
    }
    return kv;
}

} // namespace

bool SpatstatSettings::read_xml_text(const std::string &xml, std::string *why)
{
    auto kv = xml_children(xml);
    try {
        do_rdf = do_1nn = do_ripk = do_mknn = do_npcorr = do_cntnb = false;
        bool any = false;
        auto it = kv.find("AnalysisSpatDistrType");
        if (it != kv.end()) {
            const unsigned long mode = std::stoul(it->second);
            switch (mode) {
            case 1: do_rdf = any = true; break;
            case 2: do_1nn = any = true; break;
            case 3: do_ripk = any = true; break;
            case 4: do_mknn = any = true; break;
            case 5: do_npcorr = any = true; break;
            case 6: do_cntnb = any = true; break;
            default: {
                // several digits (src/PARAPROBE_Settings.cpp:380-391)
                const std::string m = it->second;
                if (m.find('0') != std::string::npos || m.size() > 4 || m.size() < 1)
                    break;
                if (m.find('1') != std::string::npos) do_rdf = any = true;
                if (m.find('2') != std::string::npos) do_1nn = any = true;
                if (m.find('3') != std::string::npos) do_ripk = any = true;
                if (m.find('4') != std::string::npos) do_mknn = any = true;
                if (m.find('5') != std::string::npos) do_npcorr = any = true;
                if (m.find('6') != std::string::npos) do_cntnb = any = true;
            }
            }
        }
        if (any) {
            tasks_code = "";
            if ((it = kv.find("DescrStatTasksCode")) != kv.end())
                tasks_code = it->second;
            if ((it = kv.find("SpatStatRadiusMin")) != kv.end())
                rmin = std::stof(it->second);
            if ((it = kv.find("SpatStatRadiusIncr")) != kv.end())
                rincr = std::stof(it->second);
            if ((it = kv.find("SpatStatRadiusMax")) != kv.end())
                rmax = std::stof(it->second);
            if ((it = kv.find("SpatStatKNNOrder")) != kv.end()) {
                const unsigned long v = std::stoul(it->second);
                if (v > 0)
                    knn_order = v - 1;
            } else {
                knn_order = 0;
            }
            if (do_mknn && (it = kv.find("SpatStatMKNNCode")) != kv.end())
                mknn_code = it->second;
            add_label_rnd = false;
            if ((it = kv.find("SpatStatAdditionalLabelRandomization")) != kv.end())
                add_label_rnd = std::stoul(it->second) == 1;
        }
        // the carving radius is read only for a Voronoi tessellation run
        if ((it = kv.find("AnalysisVolumeTessellation")) != kv.end() &&
            std::stoul(it->second) == 1 &&
            (it = kv.find("SurfaceCellsCarvingRadius")) != kv.end())
            carving_radius = std::stof(it->second);
    } catch (const std::exception &e) {
        if (why)
            *why = std::string("malformed setting: ") + e.what();
        return false;
    }
    return true;
}

bool SpatstatSettings::check(std::string *why) const
{
    const char *msg = nullptr;
    if (rmin < 0.f)
        msg = "SpatStatRadiusMin must be positive or zero!";
    else if (rincr < kEpsilon)
        msg = "SpatStatRadiusIncr must be positive!";
    else if (rmax < kEpsilon)
        msg = "SpatStatRadiusMax must be positive!";
    else if (rmin >= rmax)
        msg = "SpatStatRadiusMin is larger than SpatStatRadiusMax!";
    if (msg && why)
        *why = msg;
    return msg == nullptr;
}

uint32_t SpatstatSettings::families() const
{
    return (do_rdf ? PP_FAM_RDF : 0u) | (do_1nn ? PP_FAM_1NN : 0u) | (do_ripk ? PP_FAM_RIPK : 0u) |
           (do_mknn ? PP_FAM_MKNN : 0u) | (do_npcorr ? PP_FAM_NPCORR : 0u) |
           (do_cntnb ? PP_FAM_CNTNB : 0u);
}

// ---------------------------------------------------------------------------
// writers (S.cpp:4299-4453)

std::string report_filename(unsigned simid, const std::string &metric, const std::string &trg,
                            const std::string &env, bool randomized)
{
    return "PARAPROBE.SimID." + std::to_string(simid) + "." + metric + "." + trg + "." + env +
           (randomized ? ".Rndmized.csv" : ".Original.csv");
}

bool report_apriori_descrstat2(const std::string &path, int kind, const std::string &metric,
                               float rmin, float rincr, float rmax, const int64_t *hist,
                               const VolumeMeta &meta)
{
    // histogram geometry as histogram(start, width, end) holds it (doubles of
    // the float settings, src/PARAPROBE_Histogram.h:54-89)
    const double start = rmin, width = rincr, end = rmax;
    const size_t nb = (size_t)((end - start) / width);
    const unsigned nbins = (unsigned)nb;
    auto report = [&](unsigned b) -> double { return b < nbins ? (double)hist[1 + b] : 0.0; };
    const double lowest = (double)hist[0], highest = (double)hist[nbins + 1];
    std::ofstream o(path);
    if (!o.is_open())
        return false;
    o.precision(18);
    double bend = start + width;
    if (kind == 1) {  // RDF
        o << "NumberOfIonsInside" << ";" << meta.nions_inside << "\n";
        o << "VolumeInside" << ";" << meta.volume_inside << " (nm^3)\n";
        if (meta.volume_inside > kEpsilon) {
            double gd = (double)meta.nions_inside;
            gd /= (double)meta.volume_inside;
            o << "InsideAtomicDensityEst" << ";" << gd << "\n";
            const double norm = (1.f / gd) * (1.f / ((4.f / 3.f) * kPi));
            o << "NormalizationFactor" << ";" << norm << "\n";
            o << "r;AccumulatedCount;nRDF(r);ShellVolume;RDF(r)\n";
            o << "nm;1;1;nm^3;1\n";
            o << "r;AccumulatedCount;nRDF(r);ShellVolume;RDF(r)\n";
            o << start << ";" << lowest << ";;;\n";
            const double half_dr = 0.5 * width;
            for (unsigned b = 0; b < nbins; ++b, bend += width) {
                const double r = start + (0.5 + (double)b) * width;
                const double diff = (b > 0) ? (report(b) - report(b - 1)) : 0.f;
                const double hi = r + half_dr, lo = r - half_dr;
                const double sphvol = hi * hi * hi - lo * lo * lo;
                const double rdfval = (sphvol > kEpsilon) ? norm * diff / sphvol : 0.f;
                o << r << ";" << report(b) << ";" << diff << ";" << sphvol << ";" << rdfval << "\n";
            }
            o << "" << ";" << highest << ";;;\n";
        }
    }
    if (kind == 2 || kind == 4) {  // 1NN, MKNN
        o << (kind == 2 ? std::string("1NN") : metric) << "\n";
        o << "BinEnd(r);Counts;ECDFon[BinMinBinMax];CDFon[BinMinBinMax]\n";
        o << "nm;1;1;1\n";
        o << "BinEnd(r);Counts;ECDFon[BinMinBinMax];CDFon[BinMinBinMax]\n";
        o << start << ";" << lowest << ";;\n";
        double cum_sum = 0.f;
        for (unsigned b = 0; b < nbins; ++b)
            cum_sum += report(b);
        double sum = 0.f;
        for (unsigned b = 0; b < nbins; ++b, bend += width) {
            sum += report(b);
            const double csum = (cum_sum > kEpsilon) ? (sum / cum_sum) : 0.f;
            o << bend << ";" << report(b) << ";" << sum << ";" << csum << "\n";
        }
        o << "" << ";" << highest << ";;\n";
    }
    if (kind == 3) {  // RIPK
        o << "BinEnd(r);AccumulatedCounts\n";
        o << "nm;1\n";
        o << "BinEnd(r);AccumulatedCounts\n";
        o << start << ";" << lowest << "\n";
        double sum = 0;
        for (unsigned b = 0; b < nbins; ++b, bend += width) {
            sum += report(b);
            o << bend << ";" << sum << "\n";
        }
        o << "" << ";" << highest << "\n";
    }
    o.flush();
    return (bool)o;
}

bool report_apriori_descrstat3(const std::string &path, const int64_t *cnts, int64_t len)
{
    std::ofstream o(path);
    if (!o.is_open())
        return false;
    o.precision(18);
    o << "NeighborCountInRMax\n";
    o << "Value;Counts;ECDFon[BinMinBinMax];CDFon[BinMinBinMax]\n";
    o << "1;1;1;1\n";
    o << "Value;Counts;ECDFon[BinMinBinMax];CDFon[BinMinBinMax]\n";
    // discrhistogram keeps float counts (src/PARAPROBE_DiscreteHistogram.h:122)
    double cum_sum = 0.f;
    for (int64_t b = 0; b < len; ++b)
        cum_sum += (float)cnts[b];
    double sum = 0.f;
    for (int64_t b = 0; b < len; ++b) {
        const float v = (float)cnts[b];
        sum += v;
        const double csum = (cum_sum > kEpsilon) ? (sum / cum_sum) : 0.f;
        o << (size_t)b << ";" << v << ";" << sum << ";" << csum << "\n";
    }
    o.flush();
    return (bool)o;
}

// ---------------------------------------------------------------------------
// label randomisation (S.cpp:2809-3000)

RndLabeler::RndLabeler(pp_ctx *ctx, uint32_t seed, uint32_t discard) : ctx_(ctx)
{
    urng.seed(seed);
    urng.discard(discard);
}

RndLabeler::~RndLabeler() { reset_original_types_global(); }

void RndLabeler::reset()
{
    initial_iontypid.clear();
    temporary_iontypid.clear();
    shuffled = false;
    applied = false;
}

pp_status RndLabeler::learn_original_types_global(const pp_cloud *cloud)
{
    if (shuffled) {
        std::cerr << "WARNING::Refusing to randomize iontypes as they currently are randomized "
                     "already!"
                  << std::endl;
        return PP_OK;
    }
    uint64_t n = 0;
    if (!cloud || pp_cloud_count(cloud, &n) != PP_OK)
        return PP_ERR_INVALID_ARG;
    initial_iontypid.assign(n, 0);
    return pp_cloud_types(ctx_, cloud, initial_iontypid.data());
}

void RndLabeler::shuffle_types_mt19937()
{
    if (shuffled)
        return;
    temporary_iontypid = initial_iontypid;
    std::shuffle(temporary_iontypid.begin(), temporary_iontypid.end(), urng);
    shuffled = true;
}

pp_status RndLabeler::apply_shuffled_types_global(const pp_cloud *cloud)
{
    if (!shuffled || applied)
        return PP_OK;
    uint64_t n = 0;
    if (!cloud || pp_cloud_count(cloud, &n) != PP_OK)
        return PP_ERR_INVALID_ARG;
    if (temporary_iontypid.size() != n) {
        std::cerr << "WARNING::We have not consistently learned type of all ions, therefore "
                     "performing no randomization!"
                  << std::endl;
        return PP_ERR_INVALID_ARG;
    }
    pp_status st = pp_cloud_relabel(ctx_, cloud, temporary_iontypid.data(), &rnd_);
    if (st == PP_OK)
        applied = true;
    return st;
}

void RndLabeler::reset_original_types_global()
{
    if (rnd_)
        pp_cloud_free(ctx_, rnd_);
    rnd_ = nullptr;
    applied = false;
}

// ---------------------------------------------------------------------------
// HDF5 output: libpp_h5.so (next to this library), loaded when first needed;
// without it the NPCORR counts stay in memory only

namespace {

struct H5Writer {
    int (*create)(const char *, int32_t) = nullptr;
    int (*centers)(const char *, uint64_t, float) = nullptr;
    int (*hist)(const char *, const char *, const char *, int32_t, const uint32_t *,
                uint64_t) = nullptr;
    bool ok = false;
};

const H5Writer &h5_writer()
{
    static H5Writer w = [] {
        H5Writer r;
        Dl_info info;
        if (!dladdr((const void *)&h5_writer, &info) || !info.dli_fname)
            return r;
        std::string dir = info.dli_fname;
        const size_t slash = dir.rfind('/');
        dir = slash == std::string::npos ? "." : dir.substr(0, slash);
        void *h = dlopen((dir + "/libpp_h5.so").c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!h)
            return r;
        r.create = (int (*)(const char *, int32_t))dlsym(h, "pph5_create_results");
        r.centers = (int (*)(const char *, uint64_t, float))dlsym(h, "pph5_report_npc3d_bincenters");
        r.hist = (int (*)(const char *, const char *, const char *, int32_t, const uint32_t *,
                          uint64_t))dlsym(h, "pph5_report_npc3d_histvalues");
        r.ok = r.create && r.centers && r.hist;
        return r;
    }();
    return w;
}

} // namespace

bool hdf5_output_available() { return h5_writer().ok; }

bool create_results_h5(const std::string &fn, bool npcorr)
{
    const H5Writer &w = h5_writer();
    return w.ok && w.create(fn.c_str(), npcorr ? 1 : 0) == 0;
}

// ---------------------------------------------------------------------------
// driver

Horderdist::Horderdist(pp_ctx *ctx, const SpatstatSettings &s, const PeriodicTable &pse)
    : ctx_(ctx), s_(s), pse_(pse)
{
    std::memset(&last, 0, sizeof(last));
}

void Horderdist::initialize_descrstat_tasks()
{
    spatstat_tasks.clear();
    parse_tasks(s_.tasks_code, spatstat_tasks, pse_);
}

pp_status Horderdist::compute_generic_spatstat3(const pp_cloud *cloud, const VolumeMeta &meta,
                                                const std::string &outdir, bool randomized)
{
    uint32_t fam = s_.families();
    if (fam == 0) {
        std::cerr << "WARNING::No spatial statistics task to do" << std::endl;
        return PP_ERR_INVALID_ARG;
    }
    // avoid that in case of 2-point statistics the containers get too large:
    // the reference then skips the whole computation (S.cpp:3705-3712)
    if ((fam & PP_FAM_NPCORR) && pp_npcorr_geometry(s_.rmax, s_.rincr) < 0) {
        std::cerr << "WARNING::Computation of 2-point statistics would require too large "
                     "buffers, skipping!"
                  << std::endl;
        return PP_OK;
    }
    std::vector<int64_t> kdesc;
    if (fam & PP_FAM_MKNN) {
        if (!parse_mknn_code(s_.mknn_code, kdesc)) {
            std::cerr << "WARNING::Instructing a kth order value smaller than 1" << std::endl;
            return PP_ERR_INVALID_ARG;
        }
    }
    const int32_t T = (int32_t)spatstat_tasks.size();
    std::vector<uint8_t> trg(32 * (size_t)std::max(T, 1), 0), env(32 * (size_t)std::max(T, 1), 0);
    for (int32_t t = 0; t < T; ++t) {
        for (const auto &kv : spatstat_tasks[t].trgcandidates)
            if (kv.second < 256)
                trg[32 * t + (kv.second >> 3)] |= (uint8_t)(1u << (kv.second & 7));
        for (const auto &kv : spatstat_tasks[t].envcandidates)
            if (kv.second < 256)
                env[32 * t + (kv.second >> 3)] |= (uint8_t)(1u << (kv.second & 7));
    }
    int64_t nb = 0;
    pp_status st = pp_hist_geometry(s_.rmin, s_.rincr, s_.rmax, &nb);
    if (st != PP_OK)
        return st;
    const size_t C = (size_t)nb + 2;
    pp_job job;
    std::memset(&job, 0, sizeof(job));
    job.ntasks = T;
    job.trg_mask = trg.data();
    job.env_mask = env.data();
    job.rmin = s_.rmin;
    job.rincr = s_.rincr;
    job.rmax = s_.rmax;
    job.families = fam;
    job.nk = (int32_t)kdesc.size();
    job.kminus1_desc = kdesc.data();
    rdf.assign((fam & PP_FAM_RDF) ? T * C : 0, 0);
    nn1.assign((fam & PP_FAM_1NN) ? T * C : 0, 0);
    ripk.assign((fam & PP_FAM_RIPK) ? T * C : 0, 0);
    mknn.assign((fam & PP_FAM_MKNN) ? T * kdesc.size() * C : 0, 0);
    const int64_t cap = 1 << 20;
    cntnb.assign((fam & PP_FAM_CNTNB) ? T * (size_t)cap : 0, 0);
    std::memset(&last, 0, sizeof(last));
    last.rdf = rdf.empty() ? nullptr : rdf.data();
    last.nn1 = nn1.empty() ? nullptr : nn1.data();
    last.ripk = ripk.empty() ? nullptr : ripk.data();
    last.mknn = mknn.empty() ? nullptr : mknn.data();
    last.cntnb = cntnb.empty() ? nullptr : cntnb.data();
    last.cntnb_cap = (fam & PP_FAM_CNTNB) ? cap : 0;
    npcorr.clear();
    npcorr_edge = 0;
    if (fam & PP_FAM_NPCORR) {
        // n-point correlations: their own kernel; the counts stay in npcorr
        // (the reference writes them to HDF5, S.cpp:4254-4288)
        npcorr_edge = pp_npcorr_geometry(s_.rmax, s_.rincr);
        npcorr.assign((size_t)T * npcorr_edge * npcorr_edge * npcorr_edge, 0u);
        st = pp_npcorr_run(ctx_, cloud, T, trg.data(), env.data(), s_.rmax, s_.rincr,
                           (int64_t)s_.knn_order, npcorr.data(), nullptr);
        if (st != PP_OK) {
            std::cerr << "ERROR::n-point correlations on the GPU failed: "
                      << pp_last_error_string(ctx_) << std::endl;
            return st;
        }
        if (!results_h5.empty()) {
            // report_npc3d_bincenters_hdf5 + report_npc3d_histvalues_hdf5
            // (S.cpp:4270-4288): the bin centers once per file (the
            // randomised rerun finds them written), one column per task
            const H5Writer &w = h5_writer();
            if (w.ok) {
                (void)w.centers(results_h5.c_str(), (uint64_t)npcorr_edge, s_.rincr);
                const uint64_t nxyz = (uint64_t)npcorr_edge * npcorr_edge * npcorr_edge;
                for (int32_t t = 0; t < T; ++t) {
                    std::string tn, en;
                    for (const auto &kv : spatstat_tasks[t].trgcandidates)
                        tn += kv.first;
                    for (const auto &kv : spatstat_tasks[t].envcandidates)
                        en += kv.first;
                    if (w.hist(results_h5.c_str(), tn.c_str(), en.c_str(), randomized ? 1 : 0,
                               &npcorr[(size_t)t * nxyz], nxyz) != 0)
                        std::cerr << "WARNING::2-point statistics of " << tn << "_" << en
                                  << " not written to " << results_h5 << std::endl;
                }
            }
        }
        fam &= ~PP_FAM_NPCORR;
        job.families = fam;
        if (fam == 0)
            return PP_OK;
    }
    st = pp_spatstat_run(ctx_, cloud, &job, &last);
    if (st != PP_OK) {
        std::cerr << "ERROR::spatial statistics on the GPU failed: " << pp_last_error_string(ctx_)
                  << std::endl;
        return st;
    }
    // reports, in the reference's order (S.cpp:4188-4291)
    auto names = [&](int t, bool target) {
        std::string n;
        for (const auto &kv : target ? spatstat_tasks[t].trgcandidates
                                     : spatstat_tasks[t].envcandidates)
            n += kv.first;
        return n;
    };
    auto out = [&](const std::string &fn) { return outdir.empty() ? fn : outdir + "/" + fn; };
    bool ok = true;
    if (fam & PP_FAM_RDF)
        for (int t = 0; t < T; ++t)
            ok &= report_apriori_descrstat2(
                out(report_filename(s_.simid, "RDF", names(t, true), names(t, false), randomized)),
                1, "RDF", s_.rmin, s_.rincr, s_.rmax, &rdf[t * C], meta);
    if (fam & PP_FAM_1NN)
        for (int t = 0; t < T; ++t)
            ok &= report_apriori_descrstat2(
                out(report_filename(s_.simid, "1NN", names(t, true), names(t, false), randomized)),
                2, "1NN", s_.rmin, s_.rincr, s_.rmax, &nn1[t * C], meta);
    if (fam & PP_FAM_RIPK)
        for (int t = 0; t < T; ++t)
            ok &= report_apriori_descrstat2(
                out(report_filename(s_.simid, "RIPK", names(t, true), names(t, false), randomized)),
                3, "RIPK", s_.rmin, s_.rincr, s_.rmax, &ripk[t * C], meta);
    if (fam & PP_FAM_MKNN)
        for (int t = 0; t < T; ++t)
            for (size_t k = 0; k < kdesc.size(); ++k) {
                const std::string what = "MKNN" + std::to_string(kdesc[k] + 1);
                ok &= report_apriori_descrstat2(
                    out(report_filename(s_.simid, what, names(t, true), names(t, false),
                                        randomized)),
                    4, what, s_.rmin, s_.rincr, s_.rmax, &mknn[(t * kdesc.size() + k) * C], meta);
            }
    if (fam & PP_FAM_CNTNB)
        for (int t = 0; t < T; ++t) {
            int64_t len = 0;
            for (int64_t b = cap - 1; b >= 0; --b)
                if (cntnb[t * (size_t)cap + b]) {
                    len = b + 1;
                    break;
                }
            ok &= report_apriori_descrstat3(
                out(report_filename(s_.simid, "NBCNTinRMAX", names(t, true), names(t, false),
                                    randomized)),
                &cntnb[t * (size_t)cap], len);
        }
    if (!ok)
        std::cerr << "WARNING::some report files could not be written" << std::endl;
    return PP_OK;
}

} // namespace pph

// ---------------------------------------------------------------------------
// C entry points

namespace {

void copy_name(char *dst, int32_t cap, const std::string &s)
{
    if (!dst || cap <= 0)
        return;
    const size_t n = std::min<size_t>(s.size(), (size_t)cap - 1);
    std::memcpy(dst, s.data(), n);
    dst[n] = 0;
}

pph::PeriodicTable make_table(int32_t ntypes, const char *const *names, const uint32_t *ids)
{
    pph::PeriodicTable pse;
    for (int32_t i = 0; i < ntypes; ++i)
        pse.add_iontype(names[i], ids[i]);
    return pse;
}

} // namespace

extern "C" {

int32_t pph_parse_tasks(const char *code, int32_t ntypes, const char *const *type_names,
                        const uint32_t *type_ids, int32_t max_tasks, uint8_t *trg_masks,
                        uint8_t *env_masks, char *trg_names, char *env_names, int32_t names_stride)
{
    if (!code)
        return 0;
    pph::PeriodicTable pse = make_table(ntypes, type_names, type_ids);
    std::vector<pph::TypeSpecDescrStat> tasks;
    pph::parse_tasks(code, tasks, pse);
    const int32_t n = std::min<int32_t>((int32_t)tasks.size(), max_tasks);
    for (int32_t t = 0; t < n; ++t) {
        std::string tn, en;
        if (trg_masks)
            std::memset(trg_masks + 32 * t, 0, 32);
        if (env_masks)
            std::memset(env_masks + 32 * t, 0, 32);
        for (const auto &kv : tasks[t].trgcandidates) {
            tn += kv.first;
            if (trg_masks && kv.second < 256)
                trg_masks[32 * t + (kv.second >> 3)] |= (uint8_t)(1u << (kv.second & 7));
        }
        for (const auto &kv : tasks[t].envcandidates) {
            en += kv.first;
            if (env_masks && kv.second < 256)
                env_masks[32 * t + (kv.second >> 3)] |= (uint8_t)(1u << (kv.second & 7));
        }
        copy_name(trg_names ? trg_names + (size_t)t * names_stride : nullptr, names_stride, tn);
        copy_name(env_names ? env_names + (size_t)t * names_stride : nullptr, names_stride, en);
    }
    return (int32_t)tasks.size();
}

int32_t pph_parse_mknn_code(const char *code, int64_t *out, int32_t cap)
{
    std::vector<int64_t> v;
    if (!code || !pph::parse_mknn_code(code, v))
        return -1;
    for (int32_t i = 0; i < (int32_t)v.size() && i < cap; ++i)
        out[i] = v[i];
    return (int32_t)v.size();
}

int32_t pph_read_settings(const char *xml, uint32_t *families, float *radii, char *tasks_code,
                          int32_t tasks_cap, char *mknn_code, int32_t mknn_cap)
{
    if (!xml)
        return 1;
    pph::SpatstatSettings s;
    if (!s.read_xml_text(xml, nullptr))
        return 1;
    if (families)
        *families = s.families();
    if (radii) {
        radii[0] = s.rmin;
        radii[1] = s.rincr;
        radii[2] = s.rmax;
    }
    copy_name(tasks_code, tasks_cap, s.tasks_code);
    copy_name(mknn_code, mknn_cap, s.mknn_code);
    return s.check(nullptr) ? 0 : 2;
}

int32_t pph_write_descrstat2(const char *path, int32_t kind, const char *metric, float rmin,
                             float rincr, float rmax, const int64_t *hist, uint64_t nions_inside,
                             float volume_inside)
{
    if (!path || !hist)
        return 1;
    pph::VolumeMeta m;
    m.nions_inside = nions_inside;
    m.volume_inside = volume_inside;
    return pph::report_apriori_descrstat2(path, kind, metric ? metric : "", rmin, rincr, rmax,
                                          hist, m)
               ? 0
               : 1;
}

int32_t pph_write_descrstat3(const char *path, const int64_t *cnts, int64_t len)
{
    if (!path || (!cnts && len > 0))
        return 1;
    return pph::report_apriori_descrstat3(path, cnts, len) ? 0 : 1;
}

int32_t pph_run_spatstat(pp_ctx *ctx, const pp_cloud *cloud, const char *xml, uint32_t simid,
                         int32_t ntypes, const char *const *type_names, const uint32_t *type_ids,
                         uint64_t nions_inside, float volume_inside, const char *outdir)
{
    if (!ctx || !cloud || !xml)
        return PP_ERR_INVALID_ARG;
    pph::SpatstatSettings s;
    std::string why;
    if (!s.read_xml_text(xml, &why) || !s.check(&why)) {
        std::cerr << "ERROR::" << why << std::endl;
        return PP_ERR_INVALID_ARG;
    }
    s.simid = simid;
    pph::PeriodicTable pse = make_table(ntypes, type_names, type_ids);
    pph::Horderdist h(ctx, s, pse);
    h.initialize_descrstat_tasks();
    pph::VolumeMeta m;
    m.nions_inside = nions_inside;
    m.volume_inside = volume_inside;
    const std::string dir = outdir ? outdir : "";
    if (s.do_npcorr) {
        // the results file of solverHdl::generate_hdf5_resultsfile (S.cpp:
        // 7440-7453) with the /DescrSpatStats groups, for the NPCORR cubes
        const std::string fn = (dir.empty() ? std::string() : dir + "/") + "PARAPROBE.SimID." +
                               std::to_string(simid) + ".Results.h5";
        if (pph::create_results_h5(fn, true))
            h.results_h5 = fn;
        else
            std::cerr << "WARNING::no HDF5 output (libpp_h5.so missing or " << fn
                      << " not writable); 2-point statistics stay in memory" << std::endl;
    }
    pp_status st = h.compute_generic_spatstat3(cloud, m, dir);
    if (st != PP_OK || !s.add_label_rnd)
        return st;
    // randomise the labels over the whole tip, characterise again, restore
    pph::RndLabeler rnd(ctx);
    rnd.reset();
    if ((st = rnd.learn_original_types_global(cloud)) != PP_OK)
        return st;
    rnd.shuffle_types_mt19937();
    if ((st = rnd.apply_shuffled_types_global(cloud)) != PP_OK)
        return st;
    st = h.compute_generic_spatstat3(rnd.cloud(), m, dir, true);
    rnd.reset_original_types_global();
    return st;
}

int32_t pph_read_rangefile(const char *path, float *lo, float *hi, uint8_t *type, int32_t cap,
                           int32_t *nranges, char *names, int32_t names_stride)
{
    if (!path)
        return -1;
    pph::PeriodicTable pse;
    try {
        if (!pse.read_rangefile2(path))
            return -1;
    } catch (const std::exception &) {
        return -1;
    }
    std::vector<float> l, h;
    std::vector<uint8_t> t;
    pse.flat_ranges(l, h, t);
    const int32_t nr = (int32_t)l.size();
    if (nranges)
        *nranges = nr;
    for (int32_t r = 0; r < std::min(nr, cap); ++r) {
        if (lo)
            lo[r] = l[r];
        if (hi)
            hi[r] = h[r];
        if (type)
            type[r] = t[r];
    }
    const int32_t ntypes = (int32_t)pse.ion_names.size();
    if (names && names_stride > 0)
        for (const auto &kv : pse.ion_names) {
            char *dst = names + (size_t)kv.first * names_stride;
            std::strncpy(dst, kv.second.c_str(), names_stride - 1);
            dst[names_stride - 1] = 0;
        }
    return ntypes;
}

int32_t pph_run_pos(pp_ctx *ctx, const char *xml, const char *pos_path, const char *rrng_path,
                    const float *triangles, uint64_t ntri, uint32_t simid, uint64_t nions_inside,
                    float volume_inside, const char *outdir)
{
    if (!ctx || !xml || !pos_path)
        return PP_ERR_INVALID_ARG;
    pph::SpatstatSettings s;
    std::string why;
    if (!s.read_xml_text(xml, &why) || !s.check(&why)) {
        std::cerr << "ERROR::" << why << std::endl;
        return PP_ERR_INVALID_ARG;
    }
    // load_pos_sequentially (S.cpp:6797-6930): the whole file, 16 bytes a record
    std::ifstream f(pos_path, std::ios::binary);
    if (!f.is_open()) {
        std::cerr << "ERROR::File " << pos_path << " cannot be read!" << std::endl;
        return PP_ERR_INVALID_ARG;
    }
    std::vector<char> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    // a name ending in "epos" holds 44-byte EPOS records (load_epos_sequentially,
    // S.cpp:6933-7066), anything else 16-byte POS records
    const std::string pp(pos_path);
    const bool epos = pp.size() >= 4 && pp.compare(pp.size() - 4, 4, "epos") == 0;
    const size_t rec = epos ? 44 : 16;
    if (bytes.size() % rec != 0) {
        std::cerr << "ERROR::Size of file " << pos_path
                  << " is not an integer multiple of expected datatype" << std::endl;
        return PP_ERR_INVALID_ARG;
    }
    // identify_ions (S.cpp:7340-7438): a range file, or every ion Unknown
    pph::PeriodicTable pse;
    try {
        if (rrng_path && !pse.read_rangefile2(rrng_path))
            std::cerr << "WARNING::Unable to load RRNG file " << rrng_path << std::endl;
    } catch (const std::exception &e) {
        std::cerr << "ERROR::malformed range file: " << e.what() << std::endl;
        return PP_ERR_INVALID_ARG;
    }
    if (pse.ion_types.size() > 256) {
        std::cerr << "ERROR::more than 256 ion types" << std::endl;
        return PP_ERR_UNSUPPORTED;
    }
    std::vector<float> lo, hi;
    std::vector<uint8_t> rt;
    pse.flat_ranges(lo, hi, rt);
    const float R = std::max(s.rmax, s.carving_radius);
    pp_cloud *cl = nullptr;
    pp_status st = (epos ? pp_cloud_from_epos : pp_cloud_from_pos)(
        ctx, bytes.data(), bytes.size(), lo.data(), hi.data(), rt.data(), (int32_t)lo.size(),
        triangles, ntri, R, 0.f, &cl);
    if (st != PP_OK)
        return st;
    std::vector<const char *> names;
    std::vector<uint32_t> ids;
    for (const auto &kv : pse.ion_types) {
        names.push_back(kv.first.c_str());
        ids.push_back(kv.second);
    }
    st = (pp_status)pph_run_spatstat(ctx, cl, xml, simid, (int32_t)names.size(), names.data(),
                                     ids.data(), nions_inside, volume_inside, outdir);
    pp_cloud_free(ctx, cl);
    return st;
}

int32_t pph_label_randomization(const char *xml)
{
    if (!xml)
        return -1;
    pph::SpatstatSettings s;
    std::string why;
    if (!s.read_xml_text(xml, &why))
        return -1;
    return s.add_label_rnd ? 1 : 0;
}

int32_t pph_shuffle_types(uint32_t seed, uint32_t discard, const uint8_t *in, uint64_t n,
                          uint8_t *out)
{
    if (n > 0 && (!in || !out))
        return 1;
    std::mt19937 urng(seed);
    urng.discard(discard);
    std::vector<uint8_t> v(in, in + n);
    std::shuffle(v.begin(), v.end(), urng);
    std::copy(v.begin(), v.end(), out);
    return 0;
}

int32_t pph_mt19937(uint32_t seed, uint64_t discard, uint32_t *out, int32_t count)
{
    if (count < 0 || (count > 0 && !out))
        return 1;
    std::mt19937 urng(seed);
    urng.discard(discard);
    for (int32_t i = 0; i < count; ++i)
        out[i] = (uint32_t)urng();
    return 0;
}

} // extern "C"
