// stage_test.cpp -- the PDAL-compatible GPU stages (gpupdal_amd/pdal),
// tested the way the reference tests the CPU stages.
//
// Suites named Cpu* need no GPU (option parsing, dimension registration,
// preconditions); suites named Gpu* execute on GPU 0.  Run as
//   tests/cpp/stage_test Cpu      or      tests/cpp/stage_test Gpu
//
// The known answers are those of test/unit/filters/CovarianceFeaturesTest.cpp
// :46-188, test/unit/filters/NormalFilterTest.cpp:47-253 and
// test/unit/filters/NNDistanceTest.cpp:43-125.
#include "gfpdal.hpp"
#include "gpupdal_amd.h"
#include "mini_test.hpp"

#include <unistd.h>

#include <array>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>

using namespace gfpdal;
using Id = Dimension::Id;

namespace {

PointViewPtr make_view(PointTable& table, const std::vector<std::array<double, 3>>& pts)
{
    PointViewPtr view(new PointView(table));
    for (size_t i = 0; i < pts.size(); ++i) {
        view->setField(Id::X, i, pts[i][0]);
        view->setField(Id::Y, i, pts[i][1]);
        view->setField(Id::Z, i, pts[i][2]);
    }
    return view;
}

// FauxReader mode=grid over integer bounds: one point per integer coordinate
// with lo <= v < hi, x fastest (io/FauxReader.cpp), the first count points.
std::vector<std::array<double, 3>> faux_grid(int x0, int x1, int y0, int y1, int z0, int z1,
                                             size_t count)
{
    std::vector<std::array<double, 3>> p;
    for (int z = z0; z < std::max(z1, z0 + 1); ++z)
        for (int y = y0; y < std::max(y1, y0 + 1); ++y)
            for (int x = x0; x < std::max(x1, x0 + 1); ++x)
                if (p.size() < count) p.push_back({double(x), double(y), double(z)});
    return p;
}

// A LAS fixture of the test suite (tests/golden/las), found from this
// program's own path (tests/cpp/stage_test).
std::string fixture(const std::string& name)
{
    char buf[4096] = {0};
    const ssize_t len = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
    std::string exe(buf, len > 0 ? static_cast<size_t>(len) : 0);
    return exe.substr(0, exe.rfind('/')) + "/../golden/las/" + name;
}

std::vector<uint8_t> read_file(const std::string& path)
{
    std::ifstream in(path, std::ios::binary);
    return std::vector<uint8_t>(std::istreambuf_iterator<char>(in),
                                std::istreambuf_iterator<char>());
}

std::string temp_las(const char* tag)
{
    return std::string("/tmp/gfpdal_") + tag + "_" + std::to_string(getpid()) + ".las";
}

const char* kAllButDensity =
    "Dimensionality,Omnivariance,Anisotropy,Eigenentropy,EigenvalueSum,SurfaceVariation,"
    "DemantkeVerticality";

} // namespace

// ------------------------------------------------------------------ CPU --
TEST(CpuOptions, UnknownOption)
{
    PointTable table;
    CovarianceFeaturesFilter f;
    Options o;
    o.add("knnn", 3);
    f.setOptions(o);
    EXPECT_THROW_MSG(f.prepare(table), pdal_error,
                     "filters.covariancefeatures: Unexpected argument 'knnn'.");
}

TEST(CpuOptions, InvalidMode)
{
    PointTable table;
    CovarianceFeaturesFilter f;
    Options o;
    o.add("mode", "cube");
    f.setOptions(o);
    EXPECT_THROW_MSG(f.prepare(table), pdal_error,
                     "filters.covariancefeatures: Invalid value 'cube' for argument 'mode'.");
}

TEST(CpuOptions, ModeIsCaseInsensitive)
{
    for (const char* m : {"RAW", "Sqrt", "normalized"}) {
        PointTable table;
        CovarianceFeaturesFilter f;
        Options o;
        o.add("mode", m);
        f.setOptions(o);
        f.prepare(table);
    }
}

TEST(CpuOptions, InvalidKnn)
{
    PointTable table;
    EigenvaluesFilter f;
    Options o;
    o.add("knn", "eight");
    f.setOptions(o);
    EXPECT_THROW_MSG(f.prepare(table), pdal_error,
                     "filters.eigenvalues: Invalid value 'eight' for argument 'knn'.");
}

TEST(CpuOptions, StrideIsUnsigned)
{
    PointTable table;
    CovarianceFeaturesFilter f;
    Options o;
    o.add("stride", -2);
    f.setOptions(o);
    EXPECT_THROW_MSG(f.prepare(table), pdal_error,
                     "filters.covariancefeatures: Invalid value '-2' for argument 'stride'.");
}

TEST(CpuDims, DefaultFeatureSetIsDimensionality)
{
    PointTable table;
    CovarianceFeaturesFilter f;
    f.prepare(table);
    const auto& d = table.layout()->dims();
    ASSERT_EQ(d.size(), size_t(7));
    ASSERT_TRUE(d[3] == Id::Linearity && d[4] == Id::Planarity && d[5] == Id::Scattering &&
                d[6] == Id::Verticality);
}

TEST(CpuDims, FeatureSetOrderAndCase)
{
    PointTable table;
    CovarianceFeaturesFilter f;
    Options o;
    o.add("feature_set", " SurfaceVariation, linearity,bogus");
    f.setOptions(o);
    f.prepare(table);
    const auto& d = table.layout()->dims();
    ASSERT_EQ(d.size(), size_t(5));  // unknown names are ignored, as upstream
    ASSERT_TRUE(d[3] == Id::SurfaceVariation && d[4] == Id::Linearity);
}

TEST(CpuDims, NormalAndEigenvalueDims)
{
    PointTable table;
    NormalFilter n;
    n.prepare(table);
    EigenvaluesFilter e;
    e.prepare(table);
    const auto& d = table.layout()->dims();
    ASSERT_EQ(d.size(), size_t(10));
    ASSERT_TRUE(d[3] == Id::NormalX && d[6] == Id::Curvature && d[7] == Id::Eigenvalue0 &&
                d[9] == Id::Eigenvalue2);
}

TEST(CpuPrepared, DensityNeedsOptimalDims)
{
    PointTable table;
    CovarianceFeaturesFilter f;
    Options o;
    o.add("feature_set", "all");
    f.setOptions(o);
    EXPECT_THROW_MSG(f.prepare(table), pdal_error,
                     "filters.covariancefeatures: Density feature requires OptimalKNN and "
                     "OptimalRadius dimensions, which are missing in the input PointView.");
}

TEST(CpuPrepared, OptimizedNeedsOptimalKnn)
{
    PointTable table;
    CovarianceFeaturesFilter f;
    Options o;
    o.add("optimized", true);
    f.setOptions(o);
    EXPECT_THROW_MSG(f.prepare(table), pdal_error,
                     "filters.covariancefeatures: Missing OptimalKNN dimension in input "
                     "PointView.");
}

TEST(CpuPrepared, NormalKnnAndRadius)
{
    PointTable table;
    NormalFilter f;
    Options o;
    o.add("knn", 8);
    o.add("radius", 1.5);
    f.setOptions(o);
    EXPECT_THROW_MSG(f.prepare(table), pdal_error,
                     "filters.normal: Cannot set both knn and radius.");
}

TEST(CpuPrepared, EigenvaluesRadiusWarns)
{
    // EigenvaluesFilter.cpp:96-112: radius set -> a warning that knn and
    // stride are ignored, no error.
    PointTable table;
    EigenvaluesFilter f;
    Options o;
    o.add("radius", 2.0);
    f.setOptions(o);
    std::ostringstream log;
    f.setLog(&log, LogLevel::Warning);
    f.prepare(table);
    ASSERT_TRUE(log.str().find("Radius has been set. Ignoring knn and stride values.") !=
                std::string::npos);
}

TEST(CpuFactory, StageNames)
{
    ASSERT_EQ(createStage("filters.covariancefeatures")->getName(),
              std::string("filters.covariancefeatures"));
    ASSERT_EQ(createStage("filters.normal")->getName(), std::string("filters.normal"));
    ASSERT_EQ(createStage("filters.eigenvalues")->getName(), std::string("filters.eigenvalues"));
    ASSERT_EQ(createStage("filters.approximatecoplanar")->getName(),
              std::string("filters.approximatecoplanar"));
    ASSERT_EQ(createStage("filters.estimaterank")->getName(), std::string("filters.estimaterank"));
    ASSERT_EQ(createStage("filters.nndistance")->getName(), std::string("filters.nndistance"));
    ASSERT_EQ(createStage("filters.outlier")->getName(), std::string("filters.outlier"));
    ASSERT_EQ(createStage("readers.las")->getName(), std::string("readers.las"));
    ASSERT_EQ(createStage("writers.las")->getName(), std::string("writers.las"));
    ASSERT_TRUE(createStage("filters.smrf") == nullptr);
}

TEST(CpuOptions, NNDistanceMode)
{
    PointTable table;
    NNDistanceFilter f;
    Options o;
    o.add("mode", "median");
    f.setOptions(o);
    EXPECT_THROW_MSG(f.prepare(table), pdal_error,
                     "filters.nndistance: Invalid value 'median' for argument 'mode'.");
    PointTable t2;
    NNDistanceFilter f2;
    Options o2;
    o2.add("mode", "AVG");
    f2.setOptions(o2);
    f2.prepare(t2);
    ASSERT_TRUE(t2.layout()->hasDim(Id::NNDistance));
}

TEST(CpuDims, ConsumerDims)
{
    PointTable table;
    ApproximateCoplanarFilter c;
    EstimateRankFilter r;
    OutlierFilter o;
    c.prepare(table);
    r.prepare(table);
    o.prepare(table);
    ASSERT_TRUE(table.layout()->hasDim(Id::Coplanar));
    ASSERT_TRUE(table.layout()->hasDim(Id::Rank));
    ASSERT_TRUE(table.layout()->hasDim(Id::Classification));
    PointTable t2;
    OutlierFilter bad;
    Options ob;
    ob.add("class", 300);
    bad.setOptions(ob);
    EXPECT_THROW_MSG(bad.prepare(t2), pdal_error,
                     "filters.outlier: Invalid value '300' for argument 'class'.");
}

// ------------------------------------------------------------------ GPU --
TEST(GpuDimensionality, Linearity)
{
    PointTable table;
    BufferReader reader;
    CovarianceFeaturesFilter filter;
    Options ops;
    ops.add("knn", 3);
    ops.add("feature_set", kAllButDensity);
    filter.setInput(reader);
    filter.setOptions(ops);
    filter.prepare(table);
    PointViewPtr view = make_view(table, {{0, 0, 0}, {0, 0, 1}, {0, 0, 2}});
    reader.addView(view);
    PointViewSet set = filter.execute(table);
    PointViewPtr out = *set.begin();
    for (point_count_t i = 0; i < out->size(); ++i) {
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::Linearity, i), 1);
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::Verticality, i), 1);
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::Planarity, i), 0);
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::Scattering, i), 0);
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::Omnivariance, i), 0);
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::Anisotropy, i), 1);
        ASSERT_TRUE(std::isnan(out->getFieldAs<float>(Id::Eigenentropy, i)));
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::EigenvalueSum, i), 1);
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::SurfaceVariation, i), 0);
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::DemantkeVerticality, i), 1);
    }
}

TEST(GpuDimensionality, Planarity)
{
    PointTable table;
    BufferReader reader;
    CovarianceFeaturesFilter filter;
    Options ops;
    ops.add("feature_set", kAllButDensity);
    filter.setInput(reader);
    filter.setOptions(ops);
    filter.prepare(table);
    reader.addView(make_view(table, {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}}));
    PointViewPtr out = *filter.execute(table).begin();
    for (point_count_t i = 0; i < out->size(); ++i) {
        ASSERT_LE(out->getFieldAs<float>(Id::Linearity, i), 0.5);
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::Verticality, i), 0);
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::Planarity, i), 1);
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::Scattering, i), 0);
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::Omnivariance, i), 0);
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::Anisotropy, i), 1);
        ASSERT_TRUE(std::isnan(out->getFieldAs<float>(Id::Eigenentropy, i)));
        ASSERT_NEAR(out->getFieldAs<float>(Id::EigenvalueSum, i), 0.667, 0.001);
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::SurfaceVariation, i), 0);
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::DemantkeVerticality, i), 0);
    }
}

TEST(GpuDimensionality, Scattering)
{
    PointTable table;
    BufferReader reader;
    CovarianceFeaturesFilter filter;
    Options ops;
    ops.add("feature_set", kAllButDensity);
    filter.setInput(reader);
    filter.setOptions(ops);
    filter.prepare(table);
    reader.addView(make_view(table, {{0, 0, 0}, {1, 0, 0}, {1, 1, 1}, {0, 1, 2}}));
    PointViewPtr out = *filter.execute(table).begin();
    for (point_count_t i = 0; i < out->size(); ++i) {
        ASSERT_LE(out->getFieldAs<float>(Id::Linearity, i), 0.5);
        ASSERT_GE(out->getFieldAs<float>(Id::Verticality, i), 0.5);
        ASSERT_LE(out->getFieldAs<float>(Id::Planarity, i), 1);
        ASSERT_GE(out->getFieldAs<float>(Id::Scattering, i), 0.1);
        ASSERT_NEAR(out->getFieldAs<float>(Id::Omnivariance, i), 0.458, 0.001);
        ASSERT_NEAR(out->getFieldAs<float>(Id::Anisotropy, i), 0.864, 0.001);
        ASSERT_NEAR(out->getFieldAs<float>(Id::Eigenentropy, i), 0.488, 0.001);
        ASSERT_NEAR(out->getFieldAs<float>(Id::EigenvalueSum, i), 1.583, 0.001);
        ASSERT_NEAR(out->getFieldAs<float>(Id::SurfaceVariation, i), 0.095, 0.001);
        ASSERT_NEAR(out->getFieldAs<float>(Id::DemantkeVerticality, i), 0.492, 0.001);
    }
}

namespace {

void check_plane(const std::vector<std::array<double, 3>>& pts, double nx, double ny, double nz)
{
    PointTable table;
    BufferReader reader;
    NormalFilter filter;
    Options ops;
    ops.add("knn", 3);
    filter.setInput(reader);
    filter.setOptions(ops);
    filter.prepare(table);
    reader.addView(make_view(table, pts));
    PointViewPtr out = *filter.execute(table).begin();
    for (point_count_t i = 0; i < out->size(); ++i) {
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::NormalX, i), nx);
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::NormalY, i), ny);
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::NormalZ, i), nz);
        ASSERT_FLOAT_EQ(out->getFieldAs<float>(Id::Curvature, i), 0.0);
    }
}

} // namespace

TEST(GpuNormal, XYPlane) { check_plane(faux_grid(0, 2, 0, 2, 0, 0, 4), 0, 0, 1); }
TEST(GpuNormal, XZPlane) { check_plane(faux_grid(0, 2, 0, 0, 0, 2, 4), 0, 1, 0); }
TEST(GpuNormal, YZPlane) { check_plane(faux_grid(0, 0, 0, 2, 0, 2, 4), 1, 0, 0); }

TEST(GpuNormal, RampPlaneAlwaysUp)
{
    // Points on z = x: normal (-sqrt(2)/2, 0, sqrt(2)/2), NormalFilterTest.cpp:158-205.
    std::vector<std::array<double, 3>> p;
    for (int y = 0; y < 3; ++y)
        for (int x = 0; x < 3; ++x) p.push_back({double(x), double(y), double(x)});
    PointTable table;
    BufferReader reader;
    NormalFilter filter;
    filter.setInput(reader);
    filter.prepare(table);
    reader.addView(make_view(table, p));
    PointViewPtr out = *filter.execute(table).begin();
    const double h = std::sqrt(2.0) / 2;
    for (point_count_t i = 0; i < out->size(); ++i) {
        ASSERT_NEAR(out->getFieldAs<double>(Id::NormalX, i), -h, 1e-9);
        ASSERT_NEAR(out->getFieldAs<double>(Id::NormalY, i), 0.0, 1e-9);
        ASSERT_NEAR(out->getFieldAs<double>(Id::NormalZ, i), h, 1e-9);
    }
}

TEST(GpuNormal, ViewpointOrientation)
{
    // A horizontal plane seen from below: normals point down.
    PointTable table;
    BufferReader reader;
    NormalFilter filter;
    Options ops;
    ops.add("viewpoint", "POINT (1 1 -10)");
    filter.setInput(reader);
    filter.setOptions(ops);
    filter.prepare(table);
    reader.addView(make_view(table, faux_grid(0, 3, 0, 3, 0, 0, 9)));
    PointViewPtr out = *filter.execute(table).begin();
    for (point_count_t i = 0; i < out->size(); ++i)
        ASSERT_NEAR(out->getFieldAs<double>(Id::NormalZ, i), -1.0, 1e-12);
}

TEST(GpuEigenvalues, PlaneAndNormalize)
{
    PointTable table;
    BufferReader reader;
    EigenvaluesFilter filter;
    Options ops;
    ops.add("knn", 8);
    ops.add("normalize", true);
    filter.setInput(reader);
    filter.setOptions(ops);
    filter.prepare(table);
    reader.addView(make_view(table, faux_grid(0, 3, 0, 3, 0, 0, 9)));
    PointViewPtr out = *filter.execute(table).begin();
    for (point_count_t i = 0; i < out->size(); ++i) {
        ASSERT_NEAR(out->getFieldAs<double>(Id::Eigenvalue0, i), 0.0, 1e-15);
        ASSERT_NEAR(out->getFieldAs<double>(Id::Eigenvalue1, i) +
                        out->getFieldAs<double>(Id::Eigenvalue2, i),
                    1.0, 1e-12);
    }
}

TEST(GpuChain, NormalThenFeaturesOnOneView)
{
    // Two stages over one view: the second reuses the resident cloud.
    std::vector<std::array<double, 3>> p;
    for (int y = 0; y < 20; ++y)
        for (int x = 0; x < 20; ++x) p.push_back({x * 0.5, y * 0.5, 0.1 * x + 0.01 * (y % 3)});
    PointTable table;
    BufferReader reader;
    NormalFilter normal;
    CovarianceFeaturesFilter feats;
    Options a, b;
    a.add("knn", 8);
    b.add("knn", 8);
    normal.setOptions(a);
    feats.setOptions(b);
    normal.setInput(reader);
    feats.setInput(normal);
    feats.prepare(table);
    PointViewPtr view = make_view(table, p);
    reader.addView(view);
    PointViewPtr out = *feats.execute(table).begin();
    for (point_count_t i = 0; i < out->size(); ++i) {
        ASSERT_GE(out->getFieldAs<double>(Id::NormalZ, i), 0.0);
        ASSERT_GE(out->getFieldAs<double>(Id::Planarity, i), 0.0);
        ASSERT_LE(out->getFieldAs<double>(Id::Planarity, i), 1.0);
    }
}

TEST(GpuSkip, ZeroCovarianceKeepsValuesAndLogs)
{
    PointTable table;
    BufferReader reader;
    EigenvaluesFilter filter;
    Options ops;
    ops.add("knn", 3);
    filter.setInput(reader);
    filter.setOptions(ops);
    std::ostringstream log;
    filter.setLog(&log, LogLevel::Info);
    filter.prepare(table);
    PointViewPtr view = make_view(table, {{5, 5, 5}, {5, 5, 5}, {5, 5, 5}, {5, 5, 5}});
    for (point_count_t i = 0; i < 4; ++i) view->setField(Id::Eigenvalue0, i, -1.0);
    reader.addView(view);
    PointViewPtr out = *filter.execute(table).begin();
    for (point_count_t i = 0; i < 4; ++i)
        ASSERT_EQ(out->getFieldAs<double>(Id::Eigenvalue0, i), -1.0);
    ASSERT_TRUE(log.str().find("Skipping point 3. Covariance matrix is all zeros.") !=
                std::string::npos);
}

TEST(GpuSkip, StridePaddingSkipsRowZero)
{
    // Two points, knn=1, stride 3: KD3Index pads each strided list with id 0
    // (KDImpl.hpp:351-381), so row 0 sees only itself (zero covariance, a
    // skip) and row 1 sees points 1 and 0 (a line).
    PointTable table;
    BufferReader reader;
    CovarianceFeaturesFilter filter;
    Options ops;
    ops.add("knn", 1);
    ops.add("stride", 3);
    filter.setInput(reader);
    filter.setOptions(ops);
    filter.prepare(table);
    reader.addView(make_view(table, {{0, 0, 0}, {1, 0, 0}}));
    PointViewPtr out = *filter.execute(table).begin();
    ASSERT_EQ(out->getFieldAs<double>(Id::Linearity, 0), 0.0);  // skipped, prior value
    ASSERT_FLOAT_EQ(out->getFieldAs<double>(Id::Linearity, 1), 1.0);
}

TEST(GpuRadius, LineSkipsEndsBelowMinK)
{
    // Five points on a line, spacing 1, radius 1.5: the ends have two
    // neighbors (themselves included), the others three.  min_k = 3 skips
    // the ends with PDAL's message; the middle rows are a line.
    PointTable table;
    BufferReader reader;
    CovarianceFeaturesFilter filter;
    Options ops;
    ops.add("radius", 1.5);
    filter.setInput(reader);
    filter.setOptions(ops);
    std::ostringstream log;
    filter.setLog(&log, LogLevel::Info);
    filter.prepare(table);
    reader.addView(make_view(table, {{0, 0, 0}, {1, 0, 0}, {2, 0, 0}, {3, 0, 0}, {4, 0, 0}}));
    PointViewPtr out = *filter.execute(table).begin();
    ASSERT_EQ(out->getFieldAs<double>(Id::Linearity, 0), 0.0);  // skipped, prior value
    ASSERT_EQ(out->getFieldAs<double>(Id::Linearity, 4), 0.0);
    for (point_count_t i = 1; i < 4; ++i) {
        ASSERT_FLOAT_EQ(out->getFieldAs<double>(Id::Linearity, i), 1.0);
        ASSERT_FLOAT_EQ(out->getFieldAs<double>(Id::Planarity, i), 0.0);
    }
    ASSERT_TRUE(log.str().find("Skipping point 0. Found 2 neighbors but required 3.") !=
                std::string::npos);
    ASSERT_TRUE(log.str().find("Skipping point 4. Found 2 neighbors but required 3.") !=
                std::string::npos);
}

TEST(GpuRadius, NormalOfAPlane)
{
    // A 4 x 4 lattice in z = 0 with radius 1.1: every row has at least three
    // neighbors and the normal is +z.
    PointTable table;
    BufferReader reader;
    NormalFilter filter;
    Options ops;
    ops.add("radius", 1.1);
    filter.setInput(reader);
    filter.setOptions(ops);
    filter.prepare(table);
    reader.addView(make_view(table, faux_grid(0, 4, 0, 4, 0, 0, 16)));
    PointViewPtr out = *filter.execute(table).begin();
    for (point_count_t i = 0; i < 16; ++i) {
        ASSERT_NEAR(out->getFieldAs<double>(Id::NormalX, i), 0.0, 1e-12);
        ASSERT_NEAR(out->getFieldAs<double>(Id::NormalY, i), 0.0, 1e-12);
        ASSERT_NEAR(out->getFieldAs<double>(Id::NormalZ, i), 1.0, 1e-12);
    }
}

TEST(GpuChain, OptimalNeighborhoodThenDensity)
{
    // filters.optimalneighborhood registers OptimalKNN/OptimalRadius, so
    // covariancefeatures can compute Density from them
    // (CovarianceFeaturesFilter.cpp:159-167, 357-362).
    std::vector<std::array<double, 3>> p;
    for (int y = 0; y < 15; ++y)
        for (int x = 0; x < 15; ++x)
            p.push_back({x * 0.5 + 0.03 * (y % 4), y * 0.5, 0.2 * ((x * 7 + y * 3) % 5)});
    PointTable table;
    BufferReader reader;
    OptimalNeighborhoodFilter opt;
    CovarianceFeaturesFilter feats;
    Options b;
    b.add("knn", 8);
    b.add("feature_set", "Density");
    feats.setOptions(b);
    opt.setInput(reader);
    feats.setInput(opt);
    feats.prepare(table);
    PointViewPtr view = make_view(table, p);
    reader.addView(view);
    PointViewPtr out = *feats.execute(table).begin();
    int chosen = 0;
    for (point_count_t i = 0; i < out->size(); ++i) {
        const double k = out->getFieldAs<double>(Id::OptimalKNN, i);
        const double r = out->getFieldAs<double>(Id::OptimalRadius, i);
        if (k == 0) continue;  // every entropy NaN: PDAL keeps 0
        ++chosen;
        ASSERT_GE(k, 10.0);
        ASSERT_LE(k, 14.0);
        ASSERT_TRUE(r > 0.0);
        ASSERT_NEAR(out->getFieldAs<double>(Id::Density, i),
                    (k + 1) / (4.0 / 3.0 * 3.14159265 * r * r * r), 1e-9);
    }
    ASSERT_TRUE(chosen > 100);
}

TEST(GpuChain, OptimalNeighborhoodThenOptimized)
{
    // optimized=true: each row's neighborhood is its OptimalKNN nearest
    // (CovarianceFeaturesFilter.cpp:204-210).  Checked against the plain
    // stage run with knn = OptimalKNN - 1 for the rows of each k (the same
    // neighbor set; the sums may differ in the last bits).
    std::vector<std::array<double, 3>> p;
    for (int y = 0; y < 20; ++y)
        for (int x = 0; x < 20; ++x)
            p.push_back({x * 0.5 + 0.03 * (y % 4), y * 0.5 + 0.01 * (x % 3),
                         0.2 * ((x * 7 + y * 3) % 5)});
    PointTable table;
    BufferReader reader;
    OptimalNeighborhoodFilter opt;
    CovarianceFeaturesFilter feats;
    Options b;
    b.add("optimized", true);
    b.add("feature_set", "Linearity,Planarity,Omnivariance");
    feats.setOptions(b);
    opt.setInput(reader);
    feats.setInput(opt);
    feats.prepare(table);
    reader.addView(make_view(table, p));
    PointViewPtr out = *feats.execute(table).begin();
    int checked = 0;
    for (int k = 10; k <= 14; ++k) {
        PointTable t2;
        BufferReader r2;
        CovarianceFeaturesFilter plain;
        Options o2;
        o2.add("knn", k - 1);
        o2.add("feature_set", "Linearity,Planarity,Omnivariance");
        plain.setOptions(o2);
        plain.setInput(r2);
        plain.prepare(t2);
        r2.addView(make_view(t2, p));
        PointViewPtr ref = *plain.execute(t2).begin();
        for (PointId i = 0; i < out->size(); ++i) {
            if (out->getFieldAs<double>(Id::OptimalKNN, i) != k) continue;
            for (Id d : {Id::Linearity, Id::Planarity, Id::Omnivariance}) {
                const double a = out->getFieldAs<double>(d, i), e = ref->getFieldAs<double>(d, i);
                ASSERT_NEAR(a, e, 1e-9 * std::max(1.0, std::fabs(e)));
            }
            ++checked;
        }
    }
    ASSERT_TRUE(checked > 200);
}

TEST(GpuNNDistance, ReferenceGrid)
{
    // test/unit/filters/NNDistanceTest.cpp:43-125: a 10 x 10 x 10 lattice.
    auto run = [](const char* mode, int k) {
        PointTable table;
        BufferReader reader;
        NNDistanceFilter filter;
        Options o;
        o.add("mode", mode);
        o.add("k", k);
        filter.setInput(reader);
        filter.setOptions(o);
        filter.prepare(table);
        reader.addView(make_view(table, faux_grid(0, 10, 0, 10, 0, 10, 1000)));
        return *filter.execute(table).begin();
    };
    PointViewPtr v = run("kth", 3);
    for (PointId i = 0; i < v->size(); ++i)
        ASSERT_EQ(v->getFieldAs<double>(Id::NNDistance, i), 1.0);
    v = run("kth", 4);
    int corners = 0;
    for (PointId i = 0; i < v->size(); ++i) {
        const double x = v->getFieldAs<double>(Id::X, i), y = v->getFieldAs<double>(Id::Y, i),
                     z = v->getFieldAs<double>(Id::Z, i);
        if ((x == 9 || x == 0) && (y == 9 || y == 0) && (z == 9 || z == 0)) {
            ASSERT_EQ(v->getFieldAs<double>(Id::NNDistance, i), std::sqrt(2));
            ++corners;
        }
    }
    ASSERT_EQ(corners, 8);
    v = run("avg", 6);
    for (PointId i = 0; i < v->size(); ++i) {
        const double d = v->getFieldAs<double>(Id::NNDistance, i);
        const double x = v->getFieldAs<double>(Id::X, i), y = v->getFieldAs<double>(Id::Y, i),
                     z = v->getFieldAs<double>(Id::Z, i);
        const int e = (x == 9 || x == 0) + (y == 9 || y == 0) + (z == 9 || z == 0);
        if (e == 3) ASSERT_EQ(d, (1 + std::sqrt(2)) / 2.0);
        else if (e == 2) ASSERT_EQ(d, (2 + std::sqrt(2)) / 3.0);
        else if (e == 1) ASSERT_EQ(d, (5 + std::sqrt(2)) / 6.0);
        else ASSERT_EQ(d, 1.0);
    }
}

TEST(GpuCoplanarRank, PlaneLineBox)
{
    auto run = [](Stage& f, const std::vector<std::array<double, 3>>& pts) {
        PointTable table;
        BufferReader reader;
        f.setInput(reader);
        f.prepare(table);
        reader.addView(make_view(table, pts));
        return *f.execute(table).begin();
    };
    std::vector<std::array<double, 3>> line;
    for (int i = 0; i < 20; ++i) line.push_back({0.5 * i, 0.25 * i, 0.0});
    const auto plane = faux_grid(0, 10, 0, 10, 0, 0, 100);
    const auto box = faux_grid(0, 6, 0, 6, 0, 6, 216);
    {
        // Interior lattice rows: the point, its 4 axis neighbors and 3 of
        // its 4 diagonal ones (ties), eigenvalues 0 and about 0.55, 0.84.
        ApproximateCoplanarFilter f;
        PointViewPtr v = run(f, plane);
        int interior = 0;
        for (PointId i = 0; i < v->size(); ++i) {
            const double x = v->getFieldAs<double>(Id::X, i), y = v->getFieldAs<double>(Id::Y, i);
            if (x < 2 || x > 7 || y < 2 || y > 7) continue;
            ASSERT_EQ(v->getFieldAs<double>(Id::Coplanar, i), 1.0);
            ++interior;
        }
        ASSERT_EQ(interior, 36);
    }
    {
        ApproximateCoplanarFilter f;
        PointViewPtr v = run(f, line);
        for (PointId i = 0; i < v->size(); ++i) ASSERT_EQ(v->getFieldAs<double>(Id::Coplanar, i), 0.0);
    }
    const std::pair<const std::vector<std::array<double, 3>>*, double> cases[] = {
        {&line, 1.0}, {&plane, 2.0}, {&box, 3.0}};
    for (const auto& c : cases) {
        EstimateRankFilter f;
        Options o;
        o.add("knn", 12);
        f.setOptions(o);
        PointViewPtr v = run(f, *c.first);
        for (PointId i = 0; i < v->size(); ++i) ASSERT_EQ(v->getFieldAs<double>(Id::Rank, i), c.second);
    }
    {
        EstimateRankFilter f;  // duplicates: a zero covariance has rank 0
        PointViewPtr v = run(f, {{1, 2, 3}, {1, 2, 3}, {1, 2, 3}, {1, 2, 3}});
        for (PointId i = 0; i < v->size(); ++i) ASSERT_EQ(v->getFieldAs<double>(Id::Rank, i), 0.0);
    }
}

TEST(GpuOutlier, FarPointsAreNoise)
{
    // A 10 x 10 lattice plus two far points: both methods label exactly the
    // far points (class 18 here), and leave the others' classification.
    auto pts = faux_grid(0, 10, 0, 10, 0, 0, 100);
    pts.push_back({100, 100, 0});
    pts.push_back({-80, 50, 3});
    for (const char* method : {"statistical", "radius"}) {
        PointTable table;
        BufferReader reader;
        OutlierFilter f;
        Options o;
        o.add("method", method);
        o.add("class", 18);
        o.add("radius", 2.0);
        f.setInput(reader);
        f.setOptions(o);
        f.prepare(table);
        PointViewPtr view = make_view(table, pts);
        for (PointId i = 0; i < view->size(); ++i) view->setField(Id::Classification, i, 2.0);
        reader.addView(view);
        PointViewPtr v = *f.execute(table).begin();
        for (PointId i = 0; i < v->size(); ++i)
            ASSERT_EQ(v->getFieldAs<double>(Id::Classification, i), i >= 100 ? 18.0 : 2.0);
    }
}

TEST(GpuOutlier, WarningsAndEmptyView)
{
    PointTable table;
    BufferReader reader;
    OutlierFilter f;
    Options o;
    o.add("method", "radius");
    o.add("radius", 0.5);  // no lattice point has a neighbor within 0.5
    f.setInput(reader);
    f.setOptions(o);
    std::ostringstream log;
    f.setLog(&log, LogLevel::Warning);
    f.prepare(table);
    reader.addView(make_view(table, faux_grid(0, 5, 0, 5, 0, 0, 25)));
    PointViewSet out = f.execute(table);
    ASSERT_EQ(out.size(), size_t(1));
    PointViewPtr v = *out.begin();
    for (PointId i = 0; i < v->size(); ++i)
        ASSERT_EQ(v->getFieldAs<double>(Id::Classification, i), 0.0);
    ASSERT_TRUE(log.str().find("Requested filter would remove all points.") != std::string::npos);

    PointTable t2;
    BufferReader r2;
    OutlierFilter f2;
    Options o2;
    o2.add("method", "median");
    f2.setInput(r2);
    f2.setOptions(o2);
    std::ostringstream log2;
    f2.setLog(&log2, LogLevel::Warning);
    f2.prepare(t2);
    r2.addView(make_view(t2, {{0, 0, 0}, {1, 0, 0}}));
    r2.addView(PointViewPtr(new PointView(t2)));  // empty: dropped
    ASSERT_EQ(f2.execute(t2).size(), size_t(1));
    ASSERT_TRUE(log2.str().find("Requested method is unrecognized.") != std::string::npos);
}

TEST(GpuLas, OutlierChainReferenceCount)
{
    // readers.las -> filters.outlier -> writers.las, read back: the points
    // not classified 7 are the reference's 96 (test/unit/OldPCLBlockTest.cpp
    // :43-80, statistical, mean_k 2, multiplier 1.5).
    const std::string out = temp_las("outlier");
    {
        PointTable table;
        LasReader reader;
        OutlierFilter outlier;
        LasWriter writer;
        Options ro, fo, wo;
        ro.add("filename", fixture("autzen-point-format-3.las"));
        fo.add("method", "statistical");
        fo.add("multiplier", 1.5);
        fo.add("mean_k", 2);
        wo.add("filename", out);
        reader.setOptions(ro);
        outlier.setOptions(fo);
        writer.setOptions(wo);
        outlier.setInput(reader);
        writer.setInput(outlier);
        writer.prepare(table);
        writer.execute(table);
    }
    PointTable t2;
    LasReader back;
    Options bo;
    bo.add("filename", out);
    back.setOptions(bo);
    back.prepare(t2);
    PointViewPtr v = *back.execute(t2).begin();
    ASSERT_EQ(v->size(), point_count_t(106));
    int kept = 0;
    for (PointId i = 0; i < v->size(); ++i)
        kept += v->getFieldAs<double>(Id::Classification, i) != 7.0;
    ASSERT_EQ(kept, 96);
    std::remove(out.c_str());
}

TEST(GpuLas, FeaturesAsExtraBytes)
{
    // readers.las -> covariancefeatures -> writers.las extra_dims: the
    // records grow by a double and a float, holding the view's values.
    const std::string src = fixture("1.2-with-color.las");
    const std::string out = temp_las("features");
    PointTable table;
    LasReader reader;
    CovarianceFeaturesFilter feats;
    LasWriter writer;
    Options ro, wo;
    ro.add("filename", src);
    wo.add("filename", out);
    wo.add("extra_dims", "Linearity, Planarity=float");
    reader.setOptions(ro);
    writer.setOptions(wo);
    feats.setInput(reader);
    writer.setInput(feats);
    writer.prepare(table);
    PointViewPtr v = *writer.execute(table).begin();
    const std::vector<uint8_t> a = read_file(src), b = read_file(out);
    gf_las_info ha{}, hb{};
    ASSERT_EQ(gf_las_parse(a.data(), a.size(), &ha, nullptr, 0), GF_OK);
    ASSERT_EQ(gf_las_parse(b.data(), b.size(), &hb, nullptr, 0), GF_OK);
    ASSERT_EQ(hb.point_size, ha.point_size + 12);
    ASSERT_EQ(hb.point_count, ha.point_count);
    for (PointId i = 0; i < v->size(); ++i) {
        const uint8_t* ra = a.data() + ha.point_offset + i * ha.point_size;
        const uint8_t* rb = b.data() + hb.point_offset + i * hb.point_size;
        ASSERT_TRUE(std::memcmp(ra, rb, ha.point_size) == 0);
        double lin;
        float pla;
        std::memcpy(&lin, rb + ha.point_size, 8);
        std::memcpy(&pla, rb + ha.point_size + 8, 4);
        const double want = v->getFieldAs<double>(Id::Linearity, i);
        ASSERT_TRUE(std::memcmp(&lin, &want, 8) == 0);
        const float wantf = static_cast<float>(v->getFieldAs<double>(Id::Planarity, i));
        ASSERT_TRUE(std::memcmp(&pla, &wantf, 4) == 0);
    }
    std::remove(out.c_str());
}

TEST(CpuLas, WriterOptions)
{
    PointTable table;
    LasWriter w;
    Options o;
    o.add("dataformat_id", 3);
    w.setOptions(o);
    EXPECT_THROW_MSG(w.prepare(table), pdal_error,
                     "writers.las: Unexpected argument 'dataformat_id'.");
}

int main(int argc, char** argv) { return mini::run_all(argc, argv); }
