// gfpdal.cpp -- the PDAL-compatible stages over the gpupdal_amd C ABI.
#include "gfpdal.hpp"

#include "gpupdal_amd.h"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <mutex>
#include <sstream>

namespace gfpdal {

// --------------------------------------------------------------- naming --
const char* Dimension::name(Id id)
{
    switch (id) {
    case Id::X: return "X";
    case Id::Y: return "Y";
    case Id::Z: return "Z";
    case Id::Linearity: return "Linearity";
    case Id::Planarity: return "Planarity";
    case Id::Scattering: return "Scattering";
    case Id::Verticality: return "Verticality";
    case Id::Omnivariance: return "Omnivariance";
    case Id::Anisotropy: return "Anisotropy";
    case Id::Eigenentropy: return "Eigenentropy";
    case Id::EigenvalueSum: return "EigenvalueSum";
    case Id::SurfaceVariation: return "SurfaceVariation";
    case Id::DemantkeVerticality: return "DemantkeVerticality";
    case Id::Density: return "Density";
    case Id::OptimalKNN: return "OptimalKNN";
    case Id::OptimalRadius: return "OptimalRadius";
    case Id::NormalX: return "NormalX";
    case Id::NormalY: return "NormalY";
    case Id::NormalZ: return "NormalZ";
    case Id::Curvature: return "Curvature";
    case Id::Eigenvalue0: return "Eigenvalue0";
    case Id::Eigenvalue1: return "Eigenvalue1";
    case Id::Eigenvalue2: return "Eigenvalue2";
    case Id::Coplanar: return "Coplanar";
    case Id::Rank: return "Rank";
    case Id::NNDistance: return "NNDistance";
    case Id::Classification: return "Classification";
    }
    return "";
}

void Log::write(LogLevel l, const std::string& msg) const
{
    if (m_sink && static_cast<int>(l) <= static_cast<int>(m_level)) *m_sink << msg;
}

namespace {

std::string lower(std::string s)
{
    for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

std::string trim(const std::string& s)
{
    size_t b = s.find_first_not_of(" \t\r\n");
    if (b == std::string::npos) return "";
    size_t e = s.find_last_not_of(" \t\r\n");
    return s.substr(b, e - b + 1);
}

struct ArgError {
    std::string msg;
};

bool convert(const std::string& s, int& v)
{
    errno = 0;
    char* end = nullptr;
    long r = std::strtol(s.c_str(), &end, 10);
    if (errno || end == s.c_str() || *end || r < INT32_MIN || r > INT32_MAX) return false;
    v = static_cast<int>(r);
    return true;
}

bool convert(const std::string& s, size_t& v)
{
    std::string t = trim(s);
    if (t.empty() || t[0] == '-') return false;
    errno = 0;
    char* end = nullptr;
    unsigned long long r = std::strtoull(t.c_str(), &end, 10);
    if (errno || *end) return false;
    v = static_cast<size_t>(r);
    return true;
}

bool convert(const std::string& s, double& v)
{
    errno = 0;
    char* end = nullptr;
    double r = std::strtod(s.c_str(), &end);
    if (errno || end == s.c_str() || *end) return false;
    v = r;
    return true;
}

bool convert(const std::string& s, bool& v)
{
    std::string t = lower(trim(s));
    if (t == "true" || t == "1") v = true;
    else if (t == "false" || t == "0") v = false;
    else return false;
    return true;
}

bool convert(const std::string& s, std::string& v)
{
    v = s;
    return true;
}

// CovarianceFeaturesFilter's mode: case-insensitive raw | sqrt | normalized
// (filters/CovarianceFeaturesFilter.cpp:71-86).
bool convert(const std::string& s, CovarianceFeaturesFilter::Mode& m)
{
    std::string t = lower(trim(s));
    if (t == "raw") m = CovarianceFeaturesFilter::Mode::Raw;
    else if (t == "sqrt") m = CovarianceFeaturesFilter::Mode::SQRT;
    else if (t == "normalized") m = CovarianceFeaturesFilter::Mode::Normalized;
    else return false;
    return true;
}

bool convert(const std::string& s, uint8_t& v)
{
    int r = 0;
    if (!convert(trim(s), r) || r < 0 || r > 255) return false;
    v = static_cast<uint8_t>(r);
    return true;
}

// NNDistanceFilter's mode: case-insensitive kth | avg
// (filters/NNDistanceFilter.cpp:64-78).
bool convert(const std::string& s, NNDistanceFilter::Mode& m)
{
    std::string t = lower(trim(s));
    if (t == "kth") m = NNDistanceFilter::Mode::Kth;
    else if (t == "avg") m = NNDistanceFilter::Mode::Average;
    else return false;
    return true;
}

template <class T> class TArg : public Arg {
public:
    TArg(const std::string& name, T& var, T def) : Arg(name), m_var(var), m_def(def)
    {
        m_var = m_def;
    }
    void setValue(const std::string& s) override
    {
        if (m_set) throw ArgError{"Attempted to set value twice for argument '" + m_name + "'."};
        if (s.empty())
            throw ArgError{"Argument '" + m_name + "' needs a value and none was provided."};
        if (!convert(s, m_var))
            throw ArgError{"Invalid value '" + s + "' for argument '" + m_name + "'."};
        m_set = true;
    }
    void reset() override
    {
        m_var = m_def;
        m_set = false;
    }

private:
    T& m_var;
    T m_def;
};

// A list argument: every value adds its comma-separated items.
class ListArg : public Arg {
public:
    ListArg(const std::string& name, StringList& var, StringList def)
        : Arg(name), m_var(var), m_def(std::move(def))
    {
        m_var = m_def;
    }
    void setValue(const std::string& s) override
    {
        if (!m_set) m_var.clear();
        std::string item;
        std::istringstream in(s);
        while (std::getline(in, item, ','))
            if (!trim(item).empty()) m_var.push_back(item);
        m_set = true;
    }
    void reset() override
    {
        m_var = m_def;
        m_set = false;
    }

private:
    StringList& m_var;
    StringList m_def;
};

} // namespace

// --------------------------------------------------------------- options --
void Options::add(const std::string& name, const std::string& value)
{
    m_items.emplace_back(name, value);
}

void Options::add(const std::string& name, double value)
{
    std::ostringstream o;
    o.precision(17);
    o << value;
    add(name, o.str());
}

template <class T>
Arg& ProgramArgs::add(const std::string& name, const std::string&, T& var, T def)
{
    m_args.emplace_back(new TArg<T>(name, var, def));
    return *m_args.back();
}

Arg& ProgramArgs::add(const std::string& name, const std::string&, double& var)
{
    m_args.emplace_back(new TArg<double>(name, var, 0.0));
    return *m_args.back();
}

Arg& ProgramArgs::add(const std::string& name, const std::string&, StringList& var,
                      StringList def)
{
    m_args.emplace_back(new ListArg(name, var, std::move(def)));
    return *m_args.back();
}

template Arg& ProgramArgs::add<int>(const std::string&, const std::string&, int&, int);
template Arg& ProgramArgs::add<size_t>(const std::string&, const std::string&, size_t&, size_t);
template Arg& ProgramArgs::add<bool>(const std::string&, const std::string&, bool&, bool);
template Arg& ProgramArgs::add<std::string>(const std::string&, const std::string&,
                                            std::string&, std::string);
template Arg& ProgramArgs::add<uint8_t>(const std::string&, const std::string&, uint8_t&,
                                        uint8_t);
template Arg& ProgramArgs::add<NNDistanceFilter::Mode>(const std::string&, const std::string&,
                                                       NNDistanceFilter::Mode&,
                                                       NNDistanceFilter::Mode);
template Arg& ProgramArgs::add<CovarianceFeaturesFilter::Mode>(
    const std::string&, const std::string&, CovarianceFeaturesFilter::Mode&,
    CovarianceFeaturesFilter::Mode);

Arg* ProgramArgs::find(const std::string& name)
{
    for (auto& a : m_args)
        if (a->name() == name) return a.get();
    return nullptr;
}

void ProgramArgs::parse(const Options& opts)
{
    for (const auto& kv : opts.items()) {
        Arg* a = find(kv.first);
        if (!a) throw ArgError{"Unexpected argument '" + kv.first + "'."};
        a->setValue(kv.second);
    }
}

// ------------------------------------------------------------ point data --
void PointLayout::registerDim(Dimension::Id id)
{
    if (!hasDim(id)) m_dims.push_back(id);
}

void PointLayout::registerDims(const std::vector<Dimension::Id>& ids)
{
    for (auto id : ids) registerDim(id);
}

bool PointLayout::hasDim(Dimension::Id id) const
{
    return std::find(m_dims.begin(), m_dims.end(), id) != m_dims.end();
}

struct ResidentCloud {
    gf_cloud* cloud = nullptr;
    ~ResidentCloud()
    {
        if (cloud) gf_cloud_release(cloud);
    }
};

PointView::PointView(PointTable& table) : m_table(&table) {}

std::vector<double>& PointView::column(Dimension::Id dim)
{
    auto& c = m_cols[dim];
    if (c.size() < m_size) c.resize(m_size, 0.0);
    return c;
}

const std::vector<double>* PointView::findColumn(Dimension::Id dim) const
{
    auto it = m_cols.find(dim);
    return it == m_cols.end() ? nullptr : &it->second;
}

void PointView::setField(Dimension::Id dim, PointId idx, double val)
{
    if (idx >= m_size) {
        m_size = idx + 1;
        for (auto& kv : m_cols) kv.second.resize(m_size, 0.0);
    }
    column(dim)[idx] = val;
    if (dim == Dimension::Id::X || dim == Dimension::Id::Y || dim == Dimension::Id::Z)
        m_resident.reset();
}

void PointView::resize(point_count_t n)
{
    m_size = n;
    for (auto& kv : m_cols) kv.second.resize(m_size, 0.0);
}

template <class T> T PointView::getFieldAs(Dimension::Id dim, PointId idx) const
{
    const std::vector<double>* c = findColumn(dim);
    if (!c || idx >= c->size()) return T(0);
    return static_cast<T>((*c)[idx]);
}
template double PointView::getFieldAs<double>(Dimension::Id, PointId) const;
template float PointView::getFieldAs<float>(Dimension::Id, PointId) const;
template uint64_t PointView::getFieldAs<uint64_t>(Dimension::Id, PointId) const;

gf_ctx* device_context()
{
    static std::once_flag once;
    static gf_ctx* ctx = nullptr;
    static int rc = GF_OK;
    std::call_once(once, [] {
        const char* d = std::getenv("GFPDAL_DEVICE");
        rc = gf_ctx_create(d ? std::atoi(d) : 0, &ctx);
    });
    if (rc != GF_OK || !ctx)
        throw pdal_error("gpupdal_amd: no GPU context (gf_ctx_create returned " +
                         std::to_string(rc) + "); these stages have no CPU fallback");
    return ctx;
}

std::shared_ptr<ResidentCloud> PointView::resident()
{
    if (m_resident) return m_resident;
    gf_ctx* ctx = device_context();
    auto r = std::make_shared<ResidentCloud>();
    const double* x = column(Dimension::Id::X).data();
    const double* y = column(Dimension::Id::Y).data();
    const double* z = column(Dimension::Id::Z).data();
    int rc = gf_cloud_create(ctx, x, y, z, m_size, GF_WHERE_HOST, &r->cloud);
    if (rc != GF_OK) throw pdal_error(std::string("gpupdal_amd: ") + gf_last_error(ctx));
    m_resident = r;
    return r;
}

// ----------------------------------------------------------------- stages --
void Stage::setLog(std::ostream* sink, LogLevel level)
{
    m_log.setSink(sink);
    m_log.setLevel(level);
}

void Stage::throwError(const std::string& s) const
{
    throw pdal_error(getName() + ": " + s);
}

void Stage::prepare(PointTable& table)
{
    if (m_input) m_input->prepare(table);
    m_args.reset(new ProgramArgs);
    try {
        addArgs(*m_args);
        m_args->parse(m_options);
    } catch (const ArgError& e) {
        throw pdal_error(getName() + ": " + e.msg);
    }
    addDimensions(table.layout());
    prepared(table);
    m_prepared = true;
}

PointViewSet Stage::execute(PointTable& table)
{
    if (!m_prepared) prepare(table);
    return run(table);
}

PointViewSet Stage::run(PointTable& table)
{
    PointViewSet views = m_input ? m_input->run(table) : PointViewSet{};
    for (const auto& v : views) filter(*v);
    return views;
}

namespace {

// One stage on the GPU: columns start as the view's current values (rows a
// stage skips keep them), status is one byte per row.
struct StageResult {
    std::vector<std::vector<double>> cols;
    std::vector<uint8_t> status;
};

// row_k (optional): each row's own neighbor count instead of the stage's
// knn (covariancefeatures' optimized=true).
StageResult run_on_gpu(PointView& view, const gf_stage& st,
                       const std::vector<Dimension::Id>& outdims, const Stage& stage,
                       const std::vector<uint32_t>* row_k = nullptr)
{
    StageResult r;
    const uint64_t n = view.size();
    r.status.assign(n, 0);
    for (auto d : outdims) r.cols.push_back(view.column(d));
    std::vector<double*> ptrs;
    for (auto& c : r.cols) ptrs.push_back(c.data());
    auto res = view.resident();
    uint8_t* sp = r.status.data();
    gf_ctx* ctx = device_context();
    double* const* cp = ptrs.empty() ? nullptr : ptrs.data();
    int rc = row_k ? gf_run_rowk(ctx, res->cloud, &st, 1, row_k->data(), cp, &sp, GF_WHERE_HOST)
                   : gf_run(ctx, res->cloud, &st, 1, cp, &sp, GF_WHERE_HOST);
    if (rc != GF_OK) throw pdal_error(stage.getName() + ": " + gf_last_error(ctx));
    return r;
}

constexpr uint8_t kSkipOrFail = GF_ST_ZERO_COV | GF_ST_EIGEN_FAIL | GF_ST_FEATURE_INVALID |
                                GF_ST_FEW_NEIGHBORS;

// Neighbor counts within r of every point (for PDAL's "Found n neighbors but
// required min_k" messages), fetched only when some row was skipped for it.
std::vector<uint32_t> radius_counts_if_skipped(PointView& view, double r,
                                               const std::vector<uint8_t>& status,
                                               const Stage& stage)
{
    std::vector<uint32_t> counts;
    bool any = false;
    for (uint8_t s : status) any |= (s & GF_ST_FEW_NEIGHBORS) != 0;
    if (!any) return counts;
    counts.assign(view.size(), 0);
    gf_ctx* ctx = device_context();
    const int rc = gf_radius_counts(ctx, view.resident()->cloud, r, counts.data(), GF_WHERE_HOST);
    if (rc != GF_OK) throw pdal_error(stage.getName() + ": " + gf_last_error(ctx));
    return counts;
}

} // namespace

// --------------------------------------------------- covariancefeatures --
void CovarianceFeaturesFilter::addArgs(ProgramArgs& args)
{
    args.add("knn", "k-Nearest neighbors", m_knn, 10);
    args.add("threads", "Number of threads used to run this filter", m_threads, 1);
    args.add("feature_set", "Set of features to be computed", m_featureSetString,
             StringList{"dimensionality"});
    args.add("stride", "Compute features on strided neighbors", m_stride, size_t(1));
    m_radiusArg = &args.add("radius", "Radius for nearest neighbor search", m_radius);
    args.add("min_k", "Minimum number of neighbors in radius", m_minK, 3);
    args.add("mode", "Raw, normalized, or sqrt of eigenvalues", m_mode, Mode::SQRT);
    args.add("optimized", "Use OptimalKNN or OptimalRadius?", m_optimal, false);
}

void CovarianceFeaturesFilter::addDimensions(PointLayout* layout)
{
    using Id = Dimension::Id;
    m_extraDims.clear();
    for (auto& feat : m_featureSetString) {
        std::string f = trim(lower(feat));
        if (f == "dimensionality")
            m_extraDims.insert(m_extraDims.end(),
                               {Id::Linearity, Id::Planarity, Id::Scattering, Id::Verticality});
        else if (f == "all")
            m_extraDims.insert(m_extraDims.end(),
                               {Id::Linearity, Id::Planarity, Id::Scattering, Id::Verticality,
                                Id::Omnivariance, Id::Anisotropy, Id::Eigenentropy,
                                Id::EigenvalueSum, Id::SurfaceVariation,
                                Id::DemantkeVerticality, Id::Density});
        else if (f == "linearity") m_extraDims.push_back(Id::Linearity);
        else if (f == "planarity") m_extraDims.push_back(Id::Planarity);
        else if (f == "scattering") m_extraDims.push_back(Id::Scattering);
        else if (f == "verticality") m_extraDims.push_back(Id::Verticality);
        else if (f == "omnivariance") m_extraDims.push_back(Id::Omnivariance);
        else if (f == "anisotropy") m_extraDims.push_back(Id::Anisotropy);
        else if (f == "eigenentropy") m_extraDims.push_back(Id::Eigenentropy);
        else if (f == "eigenvaluesum") m_extraDims.push_back(Id::EigenvalueSum);
        else if (f == "surfacevariation") m_extraDims.push_back(Id::SurfaceVariation);
        else if (f == "demantkeverticality") m_extraDims.push_back(Id::DemantkeVerticality);
        else if (f == "density") m_extraDims.push_back(Id::Density);
    }
    layout->registerDims(m_extraDims);
}

void CovarianceFeaturesFilter::prepared(PointTable& table)
{
    using Id = Dimension::Id;
    PointLayout* layout = table.layout();
    if (std::count(m_extraDims.begin(), m_extraDims.end(), Id::Density)) {
        if (!(layout->hasDim(Id::OptimalKNN) && layout->hasDim(Id::OptimalRadius)))
            throwError("Density feature requires OptimalKNN and OptimalRadius "
                       "dimensions, which are missing in the input PointView.");
    }
    if (m_optimal) {
        if (!layout->hasDim(Id::OptimalKNN))
            throwError("Missing OptimalKNN dimension in input PointView.");
    }
}

void CovarianceFeaturesFilter::filter(PointView& view)
{
    using Id = Dimension::Id;
    const uint64_t n = view.size();
    if (!n) return;
    // GPU columns in feature-bit order (include/gpupdal_amd.h GF_F_*).
    static const Id kBitDims[10] = {Id::Linearity, Id::Planarity, Id::Scattering,
                                    Id::Verticality, Id::Omnivariance, Id::Anisotropy,
                                    Id::Eigenentropy, Id::EigenvalueSum, Id::SurfaceVariation,
                                    Id::DemantkeVerticality};
    uint32_t mask = 0;
    for (auto d : m_extraDims)
        for (int b = 0; b < 10; ++b)
            if (d == kBitDims[b]) mask |= 1u << b;
    std::vector<Id> outdims;
    for (int b = 0; b < 10; ++b)
        if (mask & (1u << b)) outdims.push_back(kBitDims[b]);
    gf_stage st{};
    st.stage = GF_STAGE_COVFEAT;
    st.knn = m_knn;
    st.stride = static_cast<int32_t>(m_stride);
    st.mode = m_mode == Mode::Raw ? GF_MODE_RAW
                                  : (m_mode == Mode::SQRT ? GF_MODE_SQRT : GF_MODE_NORMALIZED);
    st.features = mask;
    st.always_up = 1;
    // optimized=true takes each point's OptimalKNN nearest (stride 1) and
    // wins over radius (CovarianceFeaturesFilter.cpp:204-224).
    const bool radius = !m_optimal && m_radiusArg->set();
    if (radius) {
        st.radius = m_radius;
        st.min_k = m_minK;
    }
    std::vector<uint32_t> row_k;
    if (m_optimal) {
        row_k.resize(n);
        for (uint64_t i = 0; i < n; ++i) {
            const uint64_t k = view.getFieldAs<uint64_t>(Id::OptimalKNN, i);
            if (k > 1024) throwError("OptimalKNN above 1024 is outside the device kNN envelope");
            row_k[i] = static_cast<uint32_t>(k);
        }
    }
    StageResult r = run_on_gpu(view, st, outdims, *this, m_optimal ? &row_k : nullptr);
    const std::vector<uint32_t> counts =
        radius ? radius_counts_if_skipped(view, m_radius, r.status, *this) : std::vector<uint32_t>{};

    // Rows are applied in order; the first row PDAL would throw on ends the
    // stage after the rows before it were written.
    uint64_t end = n;
    const char* err = nullptr;
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t s = r.status[i];
        if (s & GF_ST_EIGEN_FAIL) {
            end = i;
            err = "Cannot perform eigen decomposition.";
            break;
        }
        if (s & GF_ST_FEATURE_INVALID) {
            end = i;
            err = "Eigenvalues are all 0. Can't compute local features.";
            break;
        }
    }
    const bool density = std::count(m_extraDims.begin(), m_extraDims.end(), Id::Density) > 0;
    for (uint64_t i = 0; i < end; ++i) {
        if (r.status[i] & GF_ST_FEW_NEIGHBORS) {
            log().write(LogLevel::Info, "Skipping point " + std::to_string(i) + ". Found " +
                                            std::to_string(counts[i]) +
                                            " neighbors but required " + std::to_string(m_minK) +
                                            ".\n");
            continue;
        }
        if (r.status[i] & kSkipOrFail) {
            log().write(LogLevel::Info,
                        "Skipping point " + std::to_string(i) +
                            ". Covariance matrix is all zeros. This suggests a large number "
                            "of redundant points. Consider using filters.sample with a "
                            "small radius to remove redundant points.\n");
            continue;
        }
        for (size_t c = 0; c < outdims.size(); ++c) view.setField(outdims[c], i, r.cols[c][i]);
        if (density) {
            const double kopt = view.getFieldAs<double>(Id::OptimalKNN, i);
            const double ropt = view.getFieldAs<double>(Id::OptimalRadius, i);
            const double pi = 3.14159265;
            view.setField(Id::Density, i,
                          (kopt + 1.0) / ((4.0 / 3.0) * pi * std::pow(ropt, 3)));
        }
    }
    if (err) throwError(err);
}

// ----------------------------------------------------------------- normal --
void NormalFilter::addArgs(ProgramArgs& args)
{
    m_knnArg = &args.add("knn", "k-Nearest Neighbors", m_knn, 8);
    m_radiusArg = &args.add("radius", "Radius to use for neighbor search", m_radius);
    m_viewpointArg = &args.add("viewpoint", "Viewpoint as WKT or GeoJSON", m_viewpoint,
                               std::string());
    args.add("always_up", "Normals always oriented with positive Z?", m_up, true);
    args.add("refine", "Refine normals using minimum spanning tree propagation?", m_refine,
             false);
}

void NormalFilter::addDimensions(PointLayout* layout)
{
    using Id = Dimension::Id;
    layout->registerDims({Id::NormalX, Id::NormalY, Id::NormalZ, Id::Curvature});
}

namespace {

// The viewpoint as WKT POINT (x y [z]) or GeoJSON {"coordinates": [x, y, z]}.
bool parse_viewpoint(const std::string& s, double v[3])
{
    std::string t = lower(s);
    size_t p = t.find("point");
    size_t q = t.find("coordinates");
    size_t from = p != std::string::npos ? p + 5 : (q != std::string::npos ? q + 11 : 0);
    if (p == std::string::npos && q == std::string::npos) return false;
    int got = 0;
    v[2] = 0.0;
    const char* c = s.c_str() + from;
    while (*c && got < 3) {
        if (std::isdigit(static_cast<unsigned char>(*c)) || *c == '-' || *c == '+' || *c == '.') {
            char* end = nullptr;
            double d = std::strtod(c, &end);
            if (end == c) {
                ++c;
                continue;
            }
            v[got++] = d;
            c = end;
        } else {
            ++c;
        }
    }
    return got >= 2;
}

} // namespace

void NormalFilter::prepared(PointTable&)
{
    if (m_up && m_viewpointArg->set()) {
        log().write(LogLevel::Warning, "Viewpoint provided. Ignoring always_up = TRUE.\n");
        m_up = false;
    }
    if (m_radiusArg->set() && m_knnArg->set()) throwError("Cannot set both knn and radius.");
    if (m_refine)
        throwError("refine (minimum spanning tree propagation) is not supported by the GPU "
                   "stage.");
    double vp[3];
    if (m_viewpointArg->set() && !parse_viewpoint(m_viewpoint, vp))
        throwError("Invalid viewpoint '" + m_viewpoint + "'.");
}

void NormalFilter::filter(PointView& view)
{
    using Id = Dimension::Id;
    const uint64_t n = view.size();
    if (!n) return;
    gf_stage st{};
    st.stage = GF_STAGE_NORMAL;
    st.knn = m_knn;
    st.stride = 1;
    st.always_up = m_up ? 1 : 0;
    if (m_radiusArg->set()) st.radius = m_radius;
    const std::vector<Id> outdims{Id::NormalX, Id::NormalY, Id::NormalZ, Id::Curvature};
    StageResult r = run_on_gpu(view, st, outdims, *this);
    double vp[3] = {0, 0, 0};
    const bool toward = m_viewpointArg->set() && parse_viewpoint(m_viewpoint, vp);
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t s = r.status[i];
        if (s & kSkipOrFail) {
            const char* msg =
                (s & GF_ST_FEW_NEIGHBORS)
                    ? "Not enough neighbors to compute normal."
                    : (s & GF_ST_ZERO_COV)
                          ? "Covariance matrix is all zeros. This suggests a large number of "
                            "redundant points."
                          : "Cannot perform eigen decomposition during normal calculation.";
            log().write(LogLevel::Info,
                        "Skipping point " + std::to_string(i) + ": " + msg + "\n");
            continue;
        }
        double nx = r.cols[0][i], ny = r.cols[1][i], nz = r.cols[2][i];
        if (toward) {
            // math::orientToViewpoint (pdal/private/MathUtils.hpp:326-329)
            const double dx = vp[0] - view.getFieldAs<double>(Id::X, i);
            const double dy = vp[1] - view.getFieldAs<double>(Id::Y, i);
            const double dz = vp[2] - view.getFieldAs<double>(Id::Z, i);
            if (dx * nx + dy * ny + dz * nz < 0) {
                nx = -nx;
                ny = -ny;
                nz = -nz;
            }
        }
        view.setField(Id::NormalX, i, nx);
        view.setField(Id::NormalY, i, ny);
        view.setField(Id::NormalZ, i, nz);
        view.setField(Id::Curvature, i, r.cols[3][i]);
    }
}

// ------------------------------------------------------------ eigenvalues --
void EigenvaluesFilter::addArgs(ProgramArgs& args)
{
    args.add("knn", "k-Nearest neighbors", m_knn, 8);
    args.add("normalize", "Normalize eigenvalues?", m_normalize, false);
    args.add("stride", "Compute features on strided neighbors", m_stride, size_t(1));
    m_radiusArg = &args.add("radius", "Radius for nearest neighbor search", m_radius);
    args.add("min_k", "Minimum number of neighbors in radius", m_minK, 3);
}

void EigenvaluesFilter::addDimensions(PointLayout* layout)
{
    using Id = Dimension::Id;
    layout->registerDim(Id::Eigenvalue0);
    layout->registerDim(Id::Eigenvalue1);
    layout->registerDim(Id::Eigenvalue2);
}

void EigenvaluesFilter::prepared(PointTable&)
{
    if (m_radiusArg->set()) {
        log().write(LogLevel::Warning, "Radius has been set. Ignoring knn and stride values.\n");
        if (m_radius <= 0.0) log().write(LogLevel::Error, "Radius must be greater than 0.\n");
        return;
    }
    log().write(LogLevel::Warning,
                "No radius specified. Proceeding with knn and stride, but ignoring min_k.\n");
}

void EigenvaluesFilter::filter(PointView& view)
{
    using Id = Dimension::Id;
    const uint64_t n = view.size();
    if (!n) return;
    gf_stage st{};
    st.stage = GF_STAGE_EIGENVALUES;
    st.knn = m_knn;
    st.stride = static_cast<int32_t>(m_stride);
    st.normalize = m_normalize ? 1 : 0;
    const bool radius = m_radiusArg->set();
    if (radius) {
        st.radius = m_radius;
        st.min_k = m_minK;
    }
    const std::vector<Id> outdims{Id::Eigenvalue0, Id::Eigenvalue1, Id::Eigenvalue2};
    StageResult r = run_on_gpu(view, st, outdims, *this);
    const std::vector<uint32_t> counts =
        radius ? radius_counts_if_skipped(view, m_radius, r.status, *this) : std::vector<uint32_t>{};
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t s = r.status[i];
        if (s & GF_ST_FEW_NEIGHBORS) {
            log().write(LogLevel::Info, "Skipping point " + std::to_string(i) + ". Found " +
                                            std::to_string(counts[i]) +
                                            " neighbors but required " + std::to_string(m_minK) +
                                            ".\n");
            continue;
        }
        if (s & GF_ST_EIGEN_FAIL) throwError("Cannot perform eigen decomposition.");
        if (s & kSkipOrFail) {
            log().write(LogLevel::Info,
                        "Skipping point " + std::to_string(i) +
                            ". Covariance matrix is all zeros. This suggests a large number "
                            "of redundant points. Consider using filters.sample with a "
                            "small radius to remove redundant points.\n");
            continue;
        }
        for (int c = 0; c < 3; ++c) view.setField(outdims[c], i, r.cols[c][i]);
    }
}

// ------------------------------------------------- optimalneighborhood --
void OptimalNeighborhoodFilter::addArgs(ProgramArgs& args)
{
    args.add("min_k", "Minimum k-Nearest Neighbors", m_kMin, uint64_t(10));
    args.add("max_k", "Maximum k-Nearest Neighbors", m_kMax, uint64_t(14));
}

void OptimalNeighborhoodFilter::addDimensions(PointLayout* layout)
{
    layout->registerDim(Dimension::Id::OptimalKNN);
    layout->registerDim(Dimension::Id::OptimalRadius);
}

void OptimalNeighborhoodFilter::filter(PointView& view)
{
    const uint64_t n = view.size();
    if (!n) return;
    std::vector<uint64_t> kopt(n);
    std::vector<double> ropt(n);
    auto res = view.resident();
    gf_ctx* ctx = device_context();
    const int rc = gf_optimal_neighborhood(ctx, res->cloud, static_cast<uint32_t>(m_kMin),
                                           static_cast<uint32_t>(m_kMax), kopt.data(),
                                           ropt.data(), GF_WHERE_HOST);
    if (rc != GF_OK) throwError(gf_last_error(ctx));
    for (uint64_t i = 0; i < n; ++i) {
        view.setField(Dimension::Id::OptimalKNN, i, static_cast<double>(kopt[i]));
        view.setField(Dimension::Id::OptimalRadius, i, ropt[i]);
    }
}

// --------------------------------------------------- approximatecoplanar --
void ApproximateCoplanarFilter::addArgs(ProgramArgs& args)
{
    args.add("knn", "k-Nearest Neighbors", m_knn, 8);
    args.add("thresh1", "Threshold 1", m_thresh1, 25.0);
    args.add("thresh2", "Threshold 2", m_thresh2, 6.0);
}

void ApproximateCoplanarFilter::addDimensions(PointLayout* layout)
{
    layout->registerDim(Dimension::Id::Coplanar);
}

void ApproximateCoplanarFilter::filter(PointView& view)
{
    const uint64_t n = view.size();
    if (!n) return;
    gf_stage st{};
    st.stage = GF_STAGE_COPLANAR;
    st.knn = m_knn;
    st.stride = 1;
    st.p1 = m_thresh1;
    st.p2 = m_thresh2;
    StageResult r = run_on_gpu(view, st, {Dimension::Id::Coplanar}, *this);
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t s = r.status[i];
        if (s & GF_ST_EIGEN_FAIL) throwError("Cannot perform eigen decomposition.");
        if (s & kSkipOrFail) {
            log().write(LogLevel::Info,
                        "Skipping point " + std::to_string(i) +
                            ". Covariance matrix is all zeros. This suggests a large number "
                            "of redundant points. Consider using filters.sample with a "
                            "small radius to remove redundant points.\n");
            continue;
        }
        view.setField(Dimension::Id::Coplanar, i, r.cols[0][i]);
    }
}

// ---------------------------------------------------------- estimaterank --
void EstimateRankFilter::addArgs(ProgramArgs& args)
{
    args.add("knn", "k-Nearest Neighbors", m_knn, 8);
    args.add("thresh", "Threshold", m_thresh, 0.01);
}

void EstimateRankFilter::addDimensions(PointLayout* layout)
{
    layout->registerDim(Dimension::Id::Rank);
}

void EstimateRankFilter::filter(PointView& view)
{
    const uint64_t n = view.size();
    if (!n) return;
    gf_stage st{};
    st.stage = GF_STAGE_RANK;
    st.knn = m_knn;
    st.stride = 1;
    st.p1 = m_thresh;
    StageResult r = run_on_gpu(view, st, {Dimension::Id::Rank}, *this);
    for (uint64_t i = 0; i < n; ++i) {
        // The device takes the singular values of the symmetric covariance
        // from its eigen solve; one that does not converge has no rank.
        if (r.status[i] & kSkipOrFail) throwError("Cannot perform eigen decomposition.");
        view.setField(Dimension::Id::Rank, i, r.cols[0][i]);
    }
}

// ------------------------------------------------------------ nndistance --
void NNDistanceFilter::addArgs(ProgramArgs& args)
{
    args.add("mode", "Distance computation mode (kth, avg)", m_mode, Mode::Kth);
    args.add("k", "k neighbors", m_k, size_t(10));
}

void NNDistanceFilter::addDimensions(PointLayout* layout)
{
    layout->registerDim(Dimension::Id::NNDistance);
}

void NNDistanceFilter::filter(PointView& view)
{
    const uint64_t n = view.size();
    if (!n) return;
    if (m_k > UINT32_MAX - 1) throwError("k is outside the device kNN envelope");
    std::vector<double>& out = view.column(Dimension::Id::NNDistance);
    gf_ctx* ctx = device_context();
    const int rc = gf_nndistance(ctx, view.resident()->cloud, static_cast<uint32_t>(m_k),
                                 m_mode == Mode::Kth ? GF_NND_KTH : GF_NND_AVG, out.data(),
                                 GF_WHERE_HOST);
    if (rc != GF_OK) throwError(gf_last_error(ctx));
}

// --------------------------------------------------------------- outlier --
void OutlierFilter::addArgs(ProgramArgs& args)
{
    args.add("method", "Method [default: statistical]", m_method, std::string("statistical"));
    args.add("min_k", "Minimum number of neighbors in radius", m_minK, 2);
    args.add("radius", "Radius", m_radius, 1.0);
    args.add("mean_k", "Mean number of neighbors", m_meanK, 8);
    args.add("multiplier", "Standard deviation threshold", m_multiplier, 2.0);
    args.add("class", "Class to use for noise points", m_class, uint8_t(7));
}

void OutlierFilter::addDimensions(PointLayout* layout)
{
    layout->registerDim(Dimension::Id::Classification);
}

PointViewSet OutlierFilter::run(PointTable& table)
{
    PointViewSet out;
    for (const auto& v : Stage::run(table))
        if (v->size()) out.insert(v);
    return out;
}

void OutlierFilter::filter(PointView& view)
{
    const uint64_t n = view.size();
    if (!n) return;
    const std::string m = lower(m_method);
    int method;
    if (m == "statistical") {
        method = GF_OUTLIER_STATISTICAL;
    } else if (m == "radius") {
        method = GF_OUTLIER_RADIUS;
    } else {
        log().write(LogLevel::Warning, "Requested method is unrecognized. Please choose from "
                                       "\"statistical\" or \"radius\".\n");
        return;
    }
    if (method == GF_OUTLIER_STATISTICAL && m_meanK < 0)
        throwError("mean_k must not be negative");
    std::vector<uint8_t> flags(n, 0);
    gf_ctx* ctx = device_context();
    const int rc = gf_outlier(ctx, view.resident()->cloud, method,
                              static_cast<uint32_t>(std::max(m_meanK, 0)), m_multiplier,
                              m_radius, m_minK, flags.data(), nullptr);
    if (rc != GF_OK) throwError(gf_last_error(ctx));
    uint64_t nout = 0;
    for (uint8_t f : flags) nout += f;
    if (nout == n) {
        log().write(LogLevel::Warning, "Requested filter would remove all points. Try a larger "
                                       "radius/smaller minimum neighbors.\n");
        return;
    }
    if (!nout) {
        log().write(LogLevel::Warning, "Filtered cloud has no outliers!\n");
        return;
    }
    log().write(LogLevel::Debug, "Labeled " + std::to_string(nout) + " outliers as noise!\n");
    for (uint64_t i = 0; i < n; ++i)
        if (flags[i]) view.setField(Dimension::Id::Classification, i, m_class);
}

// ------------------------------------------------------------------- LAS --
namespace {

// LAS data type (1..10 = u8, i8, u16, i16, u32, i32, u64, i64, f32, f64) and
// description of a dimension, from pdal/Dimension.json; descriptions are
// what the Extra Bytes descriptor holds (32 bytes).
struct DimInfo {
    Dimension::Id id;
    uint8_t las_type;
    const char* description;
};

const DimInfo kDimInfo[] = {
    {Dimension::Id::Linearity, 10, "Linearity of a point; larger val"},
    {Dimension::Id::Planarity, 10, "Planarity of a point; larger val"},
    {Dimension::Id::Scattering, 10, "Scattering of a point; larger va"},
    {Dimension::Id::Verticality, 10, "Verticality of a point; larger v"},
    {Dimension::Id::Omnivariance, 10, "Omnivariance of a point; cube ro"},
    {Dimension::Id::Anisotropy, 10, "Anisotropy of a point; larger va"},
    {Dimension::Id::Eigenentropy, 10, "Eigenentropy of a point; small v"},
    {Dimension::Id::EigenvalueSum, 10, "Sum of computed eigenvalues."},
    {Dimension::Id::SurfaceVariation, 10, "Surface variation of a point; la"},
    {Dimension::Id::DemantkeVerticality, 10, "Verticality of a point; larger v"},
    {Dimension::Id::Density, 10, "Estimate of point density"},
    {Dimension::Id::OptimalKNN, 7, "Optimal number of k nearest neig"},
    {Dimension::Id::OptimalRadius, 10, "Radius corresponding to optimal "},
    {Dimension::Id::NormalX, 10, "X component of a vector normal t"},
    {Dimension::Id::NormalY, 10, "Y component of a vector normal t"},
    {Dimension::Id::NormalZ, 10, "Z component of a vector normal t"},
    {Dimension::Id::Curvature, 10, "Curvature of surface at this poi"},
    {Dimension::Id::Eigenvalue0, 10, "Smallest eigenvalue obtained for"},
    {Dimension::Id::Eigenvalue1, 10, "Middle eigenvalue obtained form "},
    {Dimension::Id::Eigenvalue2, 10, "Largest eigenvalue obtained form"},
    {Dimension::Id::Coplanar, 1, "Indicator of whether or not a po"},
    {Dimension::Id::Rank, 1, "Estimated rank of neighborhood o"},
    {Dimension::Id::NNDistance, 10, "Distance metric related to a poi"},
};

const DimInfo* dim_info(Dimension::Id id)
{
    for (const DimInfo& d : kDimInfo)
        if (d.id == id) return &d;
    return nullptr;
}

// Dimension::type's names (pdal/DimUtil.hpp) as LAS data types.
int las_type_named(const std::string& t)
{
    static const std::pair<const char*, int> names[] = {
        {"uint8", 1},  {"uint8_t", 1},  {"int8", 2},    {"int8_t", 2},   {"uint16", 3},
        {"uint16_t", 3}, {"int16", 4},  {"int16_t", 4}, {"uint32", 5},   {"uint32_t", 5},
        {"int32", 6},  {"int32_t", 6},  {"uint64", 7},  {"uint64_t", 7}, {"int64", 8},
        {"int64_t", 8}, {"float", 9},   {"float32", 9}, {"double", 10},  {"float64", 10}};
    const std::string l = lower(trim(t));
    for (const auto& nt : names)
        if (l == nt.first) return nt.second;
    return 0;
}

bool dim_named(const std::string& name, Dimension::Id& out)
{
    const std::string l = lower(trim(name));
    for (const DimInfo& d : kDimInfo)
        if (lower(Dimension::name(d.id)) == l) {
            out = d.id;
            return true;
        }
    return false;
}

} // namespace

void LasReader::addArgs(ProgramArgs& args)
{
    args.add("filename", "Name of file to read", m_filename, std::string());
}

void LasReader::addDimensions(PointLayout* layout)
{
    layout->registerDim(Dimension::Id::Classification);
}

PointViewSet LasReader::run(PointTable& table)
{
    std::ifstream in(m_filename, std::ios::binary);
    if (!in) throwError("Unable to open stream for '" + m_filename + "'.");
    auto img = std::make_shared<std::vector<uint8_t>>(std::istreambuf_iterator<char>(in),
                                                      std::istreambuf_iterator<char>());
    gf_las_info info{};
    char msg[256] = {0};
    if (gf_las_parse(img->data(), img->size(), &info, msg, sizeof(msg)) != GF_OK)
        throwError(msg);
    gf_ctx* ctx = device_context();
    auto r = std::make_shared<ResidentCloud>();
    if (gf_cloud_create_las(ctx, img->data(), img->size(), &info, &r->cloud) != GF_OK)
        throwError(gf_last_error(ctx));
    PointViewPtr view(new PointView(table));
    const uint64_t n = info.point_count;
    view->resize(n);
    std::vector<double>& x = view->column(Dimension::Id::X);
    std::vector<double>& y = view->column(Dimension::Id::Y);
    std::vector<double>& z = view->column(Dimension::Id::Z);
    if (n && gf_debug_cloud_xyz(ctx, r->cloud, x.data(), y.data(), z.data()) != GF_OK)
        throwError(gf_last_error(ctx));
    // Classification: the low five bits of byte 15 (formats 0-5), byte 16
    // (formats 6-10) (io/private/las/Utils.cpp point field layouts).
    std::vector<double>& cls = view->column(Dimension::Id::Classification);
    const uint8_t* rec = img->data() + info.point_offset;
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t* p = rec + i * info.point_size;
        cls[i] = info.point_format < 6 ? (p[15] & 0x1F) : p[16];
    }
    view->adoptResident(r);
    view->setLasSource(img);
    return PointViewSet{view};
}

void LasWriter::addArgs(ProgramArgs& args)
{
    args.add("filename", "Output filename", m_filename, std::string());
    args.add("extra_dims", "List of dimension names to write in addition to those of the "
                           "point format or 'all' for all available dimensions",
             m_extraDims, StringList{});
}

PointViewSet LasWriter::run(PointTable& table)
{
    PointViewSet views = m_input ? Stage::run(table) : PointViewSet{};
    if (views.size() != 1) throwError("writes exactly one point view");
    PointView& view = **views.begin();
    const auto src = view.lasSource();
    if (!src) throwError("writes the points of a view read by readers.las");
    // extra_dims: "all" or names, each optionally "=type".
    std::vector<Dimension::Id> ids;
    std::vector<uint8_t> types;
    for (const std::string& item : m_extraDims) {
        const std::string t = trim(item);
        if (lower(t) == "all") {
            for (const DimInfo& d : kDimInfo)
                if (view.findColumn(d.id) && table.layout()->hasDim(d.id)) {
                    ids.push_back(d.id);
                    types.push_back(d.las_type);
                }
            continue;
        }
        const size_t eq = t.find('=');
        Dimension::Id id;
        if (!dim_named(t.substr(0, eq), id))
            throwError("Invalid extra dimension specified: '" + t + "'.  Need <dimension>=<type>.");
        int lt = dim_info(id)->las_type;
        if (eq != std::string::npos) {
            lt = las_type_named(t.substr(eq + 1));
            if (!lt) throwError("Invalid extra dimension type specified: '" + t + "'.");
        }
        ids.push_back(id);
        types.push_back(static_cast<uint8_t>(lt));
    }
    gf_las_info info{};
    char msg[256] = {0};
    if (gf_las_parse(src->data(), src->size(), &info, msg, sizeof(msg)) != GF_OK)
        throwError(msg);
    const uint64_t n = info.point_count;
    if (view.size() != n) throwError("the view no longer holds the file's points");
    std::vector<uint8_t> img(*src);
    // Classification back into the records.
    if (const std::vector<double>* cls = view.findColumn(Dimension::Id::Classification)) {
        uint8_t* rec = img.data() + info.point_offset;
        for (uint64_t i = 0; i < n; ++i) {
            uint8_t* p = rec + i * info.point_size;
            const uint8_t c = static_cast<uint8_t>((*cls)[i]);
            if (info.point_format < 6)
                p[15] = static_cast<uint8_t>((p[15] & 0xE0) | (c & 0x1F));
            else
                p[16] = c;
        }
    }
    std::vector<uint8_t> out;
    if (ids.empty()) {
        out.swap(img);
    } else {
        std::vector<std::string> names, descs;
        std::vector<const double*> cols;
        for (Dimension::Id id : ids) {
            names.push_back(Dimension::name(id));
            descs.push_back(dim_info(id)->description);
            cols.push_back(view.column(id).data());
        }
        std::vector<const char*> np, dp;
        for (size_t i = 0; i < names.size(); ++i) {
            np.push_back(names[i].c_str());
            dp.push_back(descs[i].c_str());
        }
        gf_ctx* ctx = device_context();
        uint64_t size = 0;
        const int k = static_cast<int>(ids.size());
        if (gf_las_write_extra(ctx, img.data(), img.size(), k, np.data(), dp.data(), types.data(),
                               cols.data(), GF_WHERE_HOST, nullptr, 0, &size) != GF_OK)
            throwError(gf_last_error(ctx));
        out.resize(size);
        if (gf_las_write_extra(ctx, img.data(), img.size(), k, np.data(), dp.data(), types.data(),
                               cols.data(), GF_WHERE_HOST, out.data(), out.size(), &size) !=
            GF_OK)
            throwError(gf_last_error(ctx));
    }
    std::ofstream f(m_filename, std::ios::binary | std::ios::trunc);
    if (!f) throwError("Couldn't open file '" + m_filename + "' for output.");
    f.write(reinterpret_cast<const char*>(out.data()), static_cast<std::streamsize>(out.size()));
    if (!f) throwError("Couldn't write file '" + m_filename + "'.");
    return views;
}

std::unique_ptr<Stage> createStage(const std::string& name)
{
    if (name == "readers.las") return std::unique_ptr<Stage>(new LasReader);
    if (name == "writers.las") return std::unique_ptr<Stage>(new LasWriter);
    if (name == "filters.approximatecoplanar")
        return std::unique_ptr<Stage>(new ApproximateCoplanarFilter);
    if (name == "filters.estimaterank") return std::unique_ptr<Stage>(new EstimateRankFilter);
    if (name == "filters.nndistance") return std::unique_ptr<Stage>(new NNDistanceFilter);
    if (name == "filters.outlier") return std::unique_ptr<Stage>(new OutlierFilter);
    if (name == "filters.optimalneighborhood")
        return std::unique_ptr<Stage>(new OptimalNeighborhoodFilter);
    if (name == "filters.covariancefeatures")
        return std::unique_ptr<Stage>(new CovarianceFeaturesFilter);
    if (name == "filters.normal") return std::unique_ptr<Stage>(new NormalFilter);
    if (name == "filters.eigenvalues") return std::unique_ptr<Stage>(new EigenvaluesFilter);
    if (name == "readers.buffer") return std::unique_ptr<Stage>(new BufferReader);
    return nullptr;
}

} // namespace gfpdal
