// gfpdal.hpp -- PDAL-compatible stage layer over the gpupdal_amd C ABI.
//
// The neighborhood stages with PDAL's names, options, defaults,
// dimension registration order, skip and error behaviour, running on one
// AMD GPU through include/gpupdal_amd.h:
//
//   filters.covariancefeatures  (filters/CovarianceFeaturesFilter.cpp:102-367)
//   filters.normal              (filters/NormalFilter.cpp:85-182)
//   filters.eigenvalues         (filters/EigenvaluesFilter.cpp:77-172)
//
// and the other consumers of KD3Index neighborhoods beside them:
// filters.optimalneighborhood, filters.approximatecoplanar,
// filters.estimaterank, filters.nndistance and filters.outlier; readers.las
// and writers.las (uncompressed LAS, extra_dims) on either side.
//
// libpdal itself is not available to link against, so the small part of its
// Stage API these stages use is provided here under the same names (Options,
// PointTable/PointLayout, PointView, BufferReader, Stage::prepare/execute,
// pdal_error).  A maintainer adding the GPU stages to a real PDAL build keeps
// the filter classes and swaps this API layer for libpdal's.
//
// Not supported on the GPU (raises pdal_error naming the option):
// filters.normal's refine.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <ostream>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

struct gf_ctx;
struct gf_cloud;

namespace gfpdal {

using PointId = uint64_t;
using point_count_t = uint64_t;
using StringList = std::vector<std::string>;

class pdal_error : public std::runtime_error {
public:
    explicit pdal_error(const std::string& s) : std::runtime_error(s) {}
};

namespace Dimension {
enum class Id {
    X, Y, Z,
    Linearity, Planarity, Scattering, Verticality, Omnivariance, Anisotropy,
    Eigenentropy, EigenvalueSum, SurfaceVariation, DemantkeVerticality, Density,
    OptimalKNN, OptimalRadius,
    NormalX, NormalY, NormalZ, Curvature,
    Eigenvalue0, Eigenvalue1, Eigenvalue2,
    Coplanar, Rank, NNDistance, Classification,
};
const char* name(Id id);
} // namespace Dimension

enum class LogLevel { Error = 0, Warning, Info, Debug };

// A stage's log: messages at or below the level go to the sink (none by
// default).
class Log {
public:
    void setLevel(LogLevel l) { m_level = l; }
    void setSink(std::ostream* s) { m_sink = s; }
    void write(LogLevel l, const std::string& msg) const;

private:
    LogLevel m_level = LogLevel::Error;
    std::ostream* m_sink = nullptr;
};

// ---------------------------------------------------------------- options --
class Options {
public:
    void add(const std::string& name, const std::string& value);
    void add(const std::string& name, const char* value) { add(name, std::string(value)); }
    void add(const std::string& name, bool value) { add(name, std::string(value ? "true" : "false")); }
    void add(const std::string& name, int value) { add(name, std::to_string(value)); }
    void add(const std::string& name, long value) { add(name, std::to_string(value)); }
    void add(const std::string& name, unsigned long value) { add(name, std::to_string(value)); }
    void add(const std::string& name, double value);
    const std::vector<std::pair<std::string, std::string>>& items() const { return m_items; }

private:
    std::vector<std::pair<std::string, std::string>> m_items;
};

// Typed command-line style arguments (the subset of PDAL's ProgramArgs the
// stages use): values arrive as strings and are converted with the same
// error texts.
class Arg {
public:
    virtual ~Arg() = default;
    bool set() const { return m_set; }
    const std::string& name() const { return m_name; }
    virtual void setValue(const std::string& s) = 0;
    virtual void reset() = 0;

protected:
    explicit Arg(std::string name) : m_name(std::move(name)) {}
    std::string m_name;
    bool m_set = false;
};

class ProgramArgs {
public:
    template <class T> Arg& add(const std::string& name, const std::string& desc, T& var, T def);
    Arg& add(const std::string& name, const std::string& desc, double& var);
    Arg& add(const std::string& name, const std::string& desc, StringList& var, StringList def);
    void parse(const Options& opts);

private:
    std::vector<std::unique_ptr<Arg>> m_args;
    Arg* find(const std::string& name);
};

// ------------------------------------------------------------ point data --
class PointLayout {
public:
    void registerDim(Dimension::Id id);
    void registerDims(const std::vector<Dimension::Id>& ids);
    bool hasDim(Dimension::Id id) const;
    const std::vector<Dimension::Id>& dims() const { return m_dims; }

private:
    std::vector<Dimension::Id> m_dims{Dimension::Id::X, Dimension::Id::Y, Dimension::Id::Z};
};

class PointTable {
public:
    PointLayout* layout() { return &m_layout; }

private:
    PointLayout m_layout;
};

struct ResidentCloud;

// Columnar point storage (FP64 columns).  The XYZ columns are uploaded to the
// GPU once and stay resident for every stage that runs on the view until X,
// Y or Z change.
class PointView {
public:
    explicit PointView(PointTable& table);
    point_count_t size() const { return m_size; }
    void setField(Dimension::Id dim, PointId idx, double val);
    template <class T> T getFieldAs(Dimension::Id dim, PointId idx) const;
    std::vector<double>& column(Dimension::Id dim);
    const std::vector<double>* findColumn(Dimension::Id dim) const;
    PointTable& table() const { return *m_table; }
    std::shared_ptr<ResidentCloud> resident();
    // n points, every column zero-filled to n.
    void resize(point_count_t n);
    // The cloud as already resident on the GPU (readers.las decodes there);
    // X, Y, Z of the view must hold the same coordinates.
    void adoptResident(std::shared_ptr<ResidentCloud> r) { m_resident = std::move(r); }
    // The LAS file image the view was read from (readers.las), which
    // writers.las writes the points of.
    std::shared_ptr<const std::vector<uint8_t>> lasSource() const { return m_las; }
    void setLasSource(std::shared_ptr<const std::vector<uint8_t>> img) { m_las = std::move(img); }

private:
    PointTable* m_table;
    point_count_t m_size = 0;
    std::map<Dimension::Id, std::vector<double>> m_cols;
    std::shared_ptr<ResidentCloud> m_resident;
    std::shared_ptr<const std::vector<uint8_t>> m_las;
};
using PointViewPtr = std::shared_ptr<PointView>;
using PointViewSet = std::set<PointViewPtr>;

// ----------------------------------------------------------------- stages --
class Stage {
public:
    virtual ~Stage() = default;
    virtual std::string getName() const = 0;
    void setOptions(const Options& o) { m_options = o; }
    void setInput(Stage& s) { m_input = &s; }
    void setLog(std::ostream* sink, LogLevel level);
    // Parses the options and registers the stage's dimensions.
    void prepare(PointTable& table);
    // Runs the stage over the input's views.
    PointViewSet execute(PointTable& table);

protected:
    virtual void addArgs(ProgramArgs&) {}
    virtual void addDimensions(PointLayout*) {}
    virtual void prepared(PointTable&) {}
    virtual PointViewSet run(PointTable& table);
    virtual void filter(PointView&) {}
    [[noreturn]] void throwError(const std::string& s) const;
    const Log& log() const { return m_log; }

    Stage* m_input = nullptr;

private:
    Options m_options;
    // The stage's arguments live as long as the stage: stages keep pointers
    // to them (to ask later whether an option was set).
    std::unique_ptr<ProgramArgs> m_args;
    Log m_log;
    bool m_prepared = false;
};

class Filter : public Stage {};

// Holds views for a pipeline's first stage (PDAL's BufferReader).
class BufferReader : public Stage {
public:
    std::string getName() const override { return "readers.buffer"; }
    void addView(PointViewPtr v) { m_views.insert(v); }

protected:
    PointViewSet run(PointTable&) override { return m_views; }

private:
    PointViewSet m_views;
};

class CovarianceFeaturesFilter : public Filter {
public:
    enum class Mode { Raw, SQRT, Normalized };
    std::string getName() const override { return "filters.covariancefeatures"; }

protected:
    void addArgs(ProgramArgs& args) override;
    void addDimensions(PointLayout* layout) override;
    void prepared(PointTable& table) override;
    void filter(PointView& view) override;

private:
    int m_knn = 10;
    int m_threads = 1;
    StringList m_featureSetString;
    size_t m_stride = 1;
    double m_radius = 0.0;
    Arg* m_radiusArg = nullptr;
    int m_minK = 3;
    Mode m_mode = Mode::SQRT;
    bool m_optimal = false;
    std::vector<Dimension::Id> m_extraDims;
};

class NormalFilter : public Filter {
public:
    std::string getName() const override { return "filters.normal"; }

protected:
    void addArgs(ProgramArgs& args) override;
    void addDimensions(PointLayout* layout) override;
    void prepared(PointTable& table) override;
    void filter(PointView& view) override;

private:
    int m_knn = 8;
    Arg* m_knnArg = nullptr;
    double m_radius = 0.0;
    Arg* m_radiusArg = nullptr;
    std::string m_viewpoint;
    Arg* m_viewpointArg = nullptr;
    bool m_up = true;
    bool m_refine = false;
};

class EigenvaluesFilter : public Filter {
public:
    std::string getName() const override { return "filters.eigenvalues"; }

protected:
    void addArgs(ProgramArgs& args) override;
    void addDimensions(PointLayout* layout) override;
    void prepared(PointTable& table) override;
    void filter(PointView& view) override;

private:
    int m_knn = 8;
    bool m_normalize = false;
    size_t m_stride = 1;
    double m_radius = 0.0;
    Arg* m_radiusArg = nullptr;
    int m_minK = 3;
};

// filters.optimalneighborhood (filters/OptimalNeighborhoodFilter.cpp:53-163):
// OptimalKNN and OptimalRadius per point, which covariancefeatures' Density
// and optimized mode read.
class OptimalNeighborhoodFilter : public Filter {
public:
    std::string getName() const override { return "filters.optimalneighborhood"; }

protected:
    void addArgs(ProgramArgs& args) override;
    void addDimensions(PointLayout* layout) override;
    void filter(PointView& view) override;

private:
    uint64_t m_kMin = 10;
    uint64_t m_kMax = 14;
};

// filters.approximatecoplanar (filters/ApproximateCoplanarFilter.cpp:66-121):
// Coplanar = 1 where the neighborhood's eigenvalues satisfy
// ev1 > thresh1 * ev0 and thresh2 * ev1 > ev2.
class ApproximateCoplanarFilter : public Filter {
public:
    std::string getName() const override { return "filters.approximatecoplanar"; }

protected:
    void addArgs(ProgramArgs& args) override;
    void addDimensions(PointLayout* layout) override;
    void filter(PointView& view) override;

private:
    int m_knn = 8;
    double m_thresh1 = 25.0;
    double m_thresh2 = 6.0;
};

// filters.estimaterank (filters/EstimateRankFilter.cpp:64-83): Rank, the
// numerical rank of each neighborhood's covariance.
class EstimateRankFilter : public Filter {
public:
    std::string getName() const override { return "filters.estimaterank"; }

protected:
    void addArgs(ProgramArgs& args) override;
    void addDimensions(PointLayout* layout) override;
    void filter(PointView& view) override;

private:
    int m_knn = 8;
    double m_thresh = 0.01;
};

// filters.nndistance (filters/NNDistanceFilter.cpp:64-133): NNDistance, the
// distance to the k-th neighbor or the mean distance to the k nearest.
class NNDistanceFilter : public Filter {
public:
    enum class Mode { Kth, Average };
    std::string getName() const override { return "filters.nndistance"; }

protected:
    void addArgs(ProgramArgs& args) override;
    void addDimensions(PointLayout* layout) override;
    void filter(PointView& view) override;

private:
    Mode m_mode = Mode::Kth;
    size_t m_k = 10;
};

// filters.outlier (filters/OutlierFilter.cpp:62-230): outliers by the
// statistical or radius method get Classification = class (default 7, low
// point); an empty view is dropped, as PDAL's run does.
class OutlierFilter : public Filter {
public:
    std::string getName() const override { return "filters.outlier"; }

protected:
    void addArgs(ProgramArgs& args) override;
    void addDimensions(PointLayout* layout) override;
    PointViewSet run(PointTable& table) override;
    void filter(PointView& view) override;

private:
    std::string m_method = "statistical";
    int m_minK = 2;
    double m_radius = 1.0;
    int m_meanK = 8;
    double m_multiplier = 2.0;
    uint8_t m_class = 7;
};

// readers.las (io/LasReader.cpp) for uncompressed LAS: the points are
// decoded on the GPU (X, Y, Z stay resident for the stages that follow and
// are copied to the view's columns) and Classification is read from the
// records; the view keeps the file image for writers.las.
class LasReader : public Stage {
public:
    std::string getName() const override { return "readers.las"; }

protected:
    void addArgs(ProgramArgs& args) override;
    void addDimensions(PointLayout* layout) override;
    PointViewSet run(PointTable& table) override;

private:
    std::string m_filename;
};

// writers.las (io/LasWriter.cpp) for the points of a readers.las view: the
// source file's header, VLRs and records, with the view's Classification
// written back into the records and the extra_dims appended as Extra Bytes
// (each as its dimension's type, or the type the option names).  The
// source's point format, version, scale and offset are kept.
class LasWriter : public Stage {
public:
    std::string getName() const override { return "writers.las"; }

protected:
    void addArgs(ProgramArgs& args) override;
    PointViewSet run(PointTable& table) override;

private:
    std::string m_filename;
    StringList m_extraDims;
};

// Creates a stage by its PDAL name (nullptr for other names), as
// StageFactory::createStage does for these stages.
std::unique_ptr<Stage> createStage(const std::string& name);

// The GPU the stages run on (default: device 0, or GFPDAL_DEVICE).
gf_ctx* device_context();

} // namespace gfpdal
