// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A thin C ABI over the reference's own vendored libraries, compiled from
// their sources where they lie (vendor/nanoflann/nanoflann.hpp, the
// reference's nanoflann 1.3.1 with its concurrent-build patch, and
// vendor/eigen, Eigen 3.4.0).  The output lands in oracle/_ref/ and is used
// only to pin the C restatement in oracle/*.c; nothing in the product loads it.
//
// The dataset adaptor below is this harness's own code.  It gives nanoflann
// what PDAL's KD3ImplT gives it (pdal/private/KDImpl.hpp:209-436): three FP64
// axes, the data bounds as the root box, leaf size 100, L2_Simple_Adaptor with
// double distances, dynamic dimension, size_t indices, a KNNResultSet over
// 64-bit ids with capacity stride*min(n,k) and default SearchParams.  The
// covariance entry point evaluates PDAL's covariance expression
// (pdal/private/MathUtils.cpp:139-192) with the reference's Eigen so the
// restatement's summation order can be compared bit for bit.
#include <nanoflann/nanoflann.hpp>
#include <Eigen/Dense>

#include <algorithm>
#include <cstdint>
#include <vector>

namespace {

struct Cloud
{
    const double* p[3];
    std::size_t n;

    std::size_t kdtree_get_point_count() const { return n; }
    double kdtree_get_pt(std::size_t idx, int dim) const { return p[dim][idx]; }

    template <class BBOX>
    bool kdtree_get_bbox(BBOX& bb) const
    {
        for (int d = 0; d < 3; ++d)
        {
            double lo = p[d][0], hi = p[d][0];
            for (std::size_t i = 1; i < n; ++i)
            {
                lo = (std::min)(lo, p[d][i]);
                hi = (std::max)(hi, p[d][i]);
            }
            bb[d].low = lo;
            bb[d].high = hi;
        }
        return true;
    }
};

using Tree = nanoflann::KDTreeSingleIndexAdaptor<
    nanoflann::L2_Simple_Adaptor<double, Cloud, double>, Cloud, -1,
    std::size_t>;

struct Handle
{
    Cloud cloud;
    Tree tree;
    explicit Handle(const Cloud& c)
        : cloud(c), tree(3, cloud, nanoflann::KDTreeSingleIndexAdaptorParams(100))
    {}
};

void preorder(const Tree::Node* node, std::vector<const Tree::Node*>& out)
{
    out.push_back(node);
    if (node->child1 || node->child2)
    {
        preorder(node->child1, out);
        preorder(node->child2, out);
    }
}

} // namespace

extern "C" {

void* ref_kd3_build(const double* x, const double* y, const double* z,
                    uint64_t n, int threads)
{
    Cloud c{{x, y, z}, static_cast<std::size_t>(n)};
    Handle* h = new Handle(c);
    h->tree.set_build_threads(threads);
    h->tree.buildIndex();
    return h;
}

void ref_kd3_free(void* h) { delete static_cast<Handle*>(h); }

uint64_t ref_kd3_neighbors(void* hv, double qx, double qy, double qz,
                           uint64_t k, uint64_t stride, uint64_t* ids)
{
    Handle* h = static_cast<Handle*>(hv);
    k = (std::min)(static_cast<uint64_t>(h->cloud.n), k);
    uint64_t k2 = stride * k;
    std::vector<uint64_t> out(k2);
    std::vector<double> d2(k2);
    nanoflann::KNNResultSet<double, uint64_t, uint64_t> rs(k2);
    rs.init(out.data(), d2.data());
    double q[3] = {qx, qy, qz};
    h->tree.findNeighbors(rs, q, nanoflann::SearchParams());
    if (stride > 1)
        for (uint64_t i = 1; i < k; ++i)
            out[i] = out[i * stride];
    for (uint64_t i = 0; i < k; ++i)
        ids[i] = out[i];
    return k;
}

// KD3Index::radius (pdal/private/KDImpl.hpp:398-415): radiusSearch with
// r*r and sorted results; returns the count, fills up to cap ids and squared
// distances.
uint64_t ref_kd3_radius(void* hv, double qx, double qy, double qz, double r, uint64_t cap,
                        uint64_t* ids, double* d2)
{
    auto* h = static_cast<Handle*>(hv);
    std::vector<std::pair<std::size_t, double>> matches;
    nanoflann::SearchParams params;
    params.sorted = true;
    const double q[3] = {qx, qy, qz};
    const std::size_t count = h->tree.radiusSearch(q, r * r, matches, params);
    for (std::size_t i = 0; i < count && i < cap; ++i) {
        ids[i] = matches[i].first;
        d2[i] = matches[i].second;
    }
    return count;
}

// math::computeRank (pdal/private/MathUtils.cpp:197-208) on a given 3x3:
// Eigen's JacobiSVD with the threshold narrowed to float.
int ref_jacobi_rank(const double* cov, double threshold)
{
    Eigen::Matrix3d B;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) B(i, j) = cov[3 * i + j];
    Eigen::JacobiSVD<Eigen::Matrix3d> svd(B);
    svd.setThreshold((float)threshold);
    return static_cast<int>(svd.rank());
}

uint64_t ref_kd3_node_count(void* hv)
{
    Handle* h = static_cast<Handle*>(hv);
    std::vector<const Tree::Node*> nodes;
    if (h->tree.root_node)
        preorder(h->tree.root_node, nodes);
    return nodes.size();
}

void ref_kd3_export(void* hv, uint64_t* vind, uint64_t* left, uint64_t* right,
                    int32_t* divfeat, double* divlow, double* divhigh,
                    int32_t* is_leaf)
{
    Handle* h = static_cast<Handle*>(hv);
    for (std::size_t i = 0; i < h->cloud.n; ++i)
        vind[i] = h->tree.vind[i];
    std::vector<const Tree::Node*> nodes;
    if (h->tree.root_node)
        preorder(h->tree.root_node, nodes);
    for (std::size_t i = 0; i < nodes.size(); ++i)
    {
        const Tree::Node* nd = nodes[i];
        bool leaf = !nd->child1 && !nd->child2;
        is_leaf[i] = leaf ? 1 : 0;
        if (leaf)
        {
            left[i] = nd->node_type.lr.left;
            right[i] = nd->node_type.lr.right;
            divfeat[i] = -1;
            divlow[i] = divhigh[i] = 0.0;
        }
        else
        {
            left[i] = right[i] = 0;
            divfeat[i] = nd->node_type.sub.divfeat;
            divlow[i] = nd->node_type.sub.divlow;
            divhigh[i] = nd->node_type.sub.divhigh;
        }
    }
}

int ref_eigen_sym3(const double* cov, double* evals, double* evecs)
{
    Eigen::Matrix3d m;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            m(r, c) = cov[3 * r + c];
    Eigen::SelfAdjointEigenSolver<Eigen::Matrix3d> s(m);
    const auto& v = s.eigenvalues();
    const auto& e = s.eigenvectors();
    for (int i = 0; i < 3; ++i)
        evals[i] = v[i];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            evecs[3 * c + r] = e(r, c);
    return s.info() == Eigen::Success ? 0 : 1;
}

void ref_pdal_covariance(const double* x, const double* y, const double* z,
                         const uint64_t* ids, uint64_t n, double* cov)
{
    double mean[3] = {0.0, 0.0, 0.0};
    const double* axes[3] = {x, y, z};
    for (uint64_t j = 0; j < n; ++j)
        for (int a = 0; a < 3; ++a)
        {
            double delta = axes[a][ids[j]] - mean[a];
            mean[a] = mean[a] + delta / static_cast<double>(j + 1);
        }
    Eigen::MatrixXd A(3, static_cast<Eigen::Index>(n));
    for (uint64_t j = 0; j < n; ++j)
        for (int a = 0; a < 3; ++a)
            A(a, static_cast<Eigen::Index>(j)) =
                static_cast<float>(axes[a][ids[j]] - mean[a]);
    Eigen::Matrix3d B = A * A.transpose() / (n - 1);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            cov[3 * r + c] = B(r, c);
}

} // extern "C"
