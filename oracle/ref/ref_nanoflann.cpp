// Competitor baseline (test and measurement infrastructure): the reference's vendored nanoflann
// KD-tree (inc/structures/nanoflann.hpp, NANOFLANN_VERSION 0x171), built over a cloud and searched the
// way the reference's benchmark drives it: KDTreeSingleIndexAdaptor<L2_Simple_Adaptor<double, cloud>,
// cloud, 3, size_t> with leaf size maxPointsLeaf (nanoflann_wrappers.hpp:42-47, neighbor_benchmarks.hpp:658),
// radiusSearch(pt, r * r, matches) per centre (:340-358) and knnSearch(pt, k, idx, d2) (:361-380), OpenMP
// parallel over centres. Compiled from the reference's own header by oracle/Makefile into _ref/.
#include <omp.h>

#include <cstdint>
#include <vector>

#include "nanoflann.hpp"

namespace {

struct FlatCloud {  // x, y, z triples (the reference adapts its Point containers the same way)
    const double *xyz;
    size_t n;
    size_t kdtree_get_point_count() const { return n; }
    double kdtree_get_pt(size_t i, size_t dim) const { return xyz[3 * i + dim]; }
    template <class BBOX>
    bool kdtree_get_bbox(BBOX &) const {
        return false;
    }
};

using Tree = nanoflann::KDTreeSingleIndexAdaptor<nanoflann::L2_Simple_Adaptor<double, FlatCloud>, FlatCloud, 3, size_t>;

struct Handle {
    FlatCloud cloud;
    Tree tree;
    Handle(const double *xyz, size_t n, size_t leaf) : cloud{xyz, n}, tree(3, cloud, {leaf}) {}
};

}  // namespace

extern "C" {

void *refnf_build(const double *xyz, size_t n, uint32_t leaf) { return new Handle(xyz, n, leaf); }

void refnf_free(void *h) { delete static_cast<Handle *>(h); }

void refnf_radius_counts(void *h, const double *centers, size_t nq, double radius, uint64_t *counts) {
    const Tree &t = static_cast<Handle *>(h)->tree;
#pragma omp parallel for schedule(dynamic, 256)
    for (size_t q = 0; q < nq; q++) {
        std::vector<nanoflann::ResultItem<size_t, double>> matches;
        counts[q] = t.radiusSearch(centers + 3 * q, radius * radius, matches);
    }
}

void refnf_knn(void *h, const double *centers, size_t nq, uint32_t k, int64_t *idx, double *d2) {
    const Tree &t = static_cast<Handle *>(h)->tree;
#pragma omp parallel for schedule(dynamic, 256)
    for (size_t q = 0; q < nq; q++) {
        std::vector<size_t> ix(k);
        std::vector<double> ds(k);
        const size_t m = t.knnSearch(centers + 3 * q, k, ix.data(), ds.data());
        for (uint32_t j = 0; j < k; j++) {
            idx[q * k + j] = j < m ? (int64_t)ix[j] : -1;
            d2[q * k + j] = j < m ? ds[j] : -1.0;
        }
    }
}

}  // extern "C"
