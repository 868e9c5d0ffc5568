// Drop-in LinearOctree (reference inc/structures/linear_octree.hpp:614-1111): the same constructor
// and search signatures, backed by the gfx950 build and searches through the C ABI. Each octree
// owns one device context (its cloud, keys and tree stay resident on the GPU).
//
// Per-query calls made concurrently from several threads (the reference harness' OpenMP loop) are
// coalesced into batched launches (octree_amd_batcher.hpp); the batch calls at the end
// (neighborsCount, neighborsCountAll) are the throughput API. Results are identical to
// the reference's: indices into the sorted cloud, lists in the reference's emission order.
#pragma once

#include <concepts>
#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <utility>
#include <vector>

#include "../encoding/point_encoder.hpp"
#include "../geometry/box.hpp"
#include "../kernels/kernel_factory.hpp"
#include "../main_options.hpp"
#include "../octree_amd_batcher.hpp"
#include "../octree_amd_points.hpp"
#include "../octree_amd_runtime.hpp"
#include "neighbor_set.hpp"

struct BuildLog {
    size_t cloudSize = 0;
    EncoderType encoding = EncoderType::NO_ENCODING;
    size_t maxLeafPoints = 0;
    double linearOctreeLeafTime = 0.0, linearOctreeInternalTime = 0.0, totalTime = 0.0;
    size_t octreeLeaves = 0, octreeTotalNodes = 0;
};

template <class Container>
class LinearOctree {
    using key_t = PointEncoding::key_t;
    static constexpr int kPrecision = OCT_PREC_FP32_RECHECK;  // identical sets to FP64, faster

public:
    explicit LinearOctree(Container &points, std::vector<key_t> &codes, Box box, PointEncoding::PointEncoder &enc,
                          std::shared_ptr<BuildLog> log = nullptr)
        : points_(points), enc_(enc), codes_(codes), box_(box), ctx_(new octamd::Context()),
          batcher_(new octamd::SearchBatcher(ctx_->get(), ctx_->mutex())) {
        if (codes.size() != points.size()) throw std::invalid_argument("octree_amd: codes and points differ in size");
        const double ctr[3] = {box.center().getX(), box.center().getY(), box.center().getZ()};
        const double rad[3] = {box.radii().getX(), box.radii().getY(), box.radii().getZ()};
        octamd::check(ctx(), octamd::set_sorted(ctx(), points, reinterpret_cast<const uint64_t *>(codes.data()),
                                                (int)enc.getEncoder(), ctr, rad),
                      "oct_set_sorted");
        octamd::check(ctx(), oct_build(ctx(), (uint32_t)mainOptions.maxPointsLeaf, &info_), "oct_build");
        if (log) {
            log->cloudSize = points.size();
            log->encoding = enc.getEncoder();
            log->maxLeafPoints = mainOptions.maxPointsLeaf;
            log->linearOctreeLeafTime = info_.build_ms * 1e-3;
            log->totalTime = info_.build_ms * 1e-3;
            log->octreeLeaves = info_.n_leaf;
            log->octreeTotalNodes = info_.n_total;
        }
    }

    // ---- searches with a kernel object (linear_octree.hpp:695-816, 980-1035): the kernel's shape,
    // centre and radii go to the device, which evaluates the same predicates as the object
    template <class Kernel>
        requires std::derived_from<Kernel, KernelAbstract>
    NeighborSet<Container> neighborsStruct(const Kernel &k) const {
        return as_set(list_k(k, OCT_ALGO_STRUCT));
    }
    template <class Kernel>
        requires std::derived_from<Kernel, KernelAbstract>
    std::vector<size_t> neighborsPrune(const Kernel &k) const {
        return as_indices(list_k(k, OCT_ALGO_PRUNE));
    }
    template <class Kernel, class Function>
        requires std::derived_from<Kernel, KernelAbstract>
    std::vector<size_t> neighbors(const Kernel &k, Function &&condition) const {
        return filter(as_indices(list_k(k, OCT_ALGO_NEIGHBORS)), condition);
    }
    template <class Kernel>
        requires std::derived_from<Kernel, KernelAbstract>
    NeighborSet<Container> neighborsApprox(const Kernel &k, uint32_t precisionLevel, bool upperBound) const {
        const KernelAbstract &ka = k;
        const bool box = is_box(ka.type());
        return as_set(approx((int)ka.type(), ka.center(), ka.radii().getX(), box ? &ka.radii() : nullptr, precisionLevel,
                             upperBound));
    }

    // ---- radius searches (linear_octree.hpp:1038-1087)
    template <Kernel_t K = Kernel_t::square>
    NeighborSet<Container> neighborsStruct(const Point &p, double radius) const {
        return as_set(list((int)K, OCT_ALGO_STRUCT, p, radius, nullptr));
    }
    template <Kernel_t K = Kernel_t::cube>
    NeighborSet<Container> neighborsStruct(const Point &p, const Vector &radii) const {
        return as_set(list((int)K, OCT_ALGO_STRUCT, p, 0.0, &radii));
    }
    template <Kernel_t K = Kernel_t::square>
    std::vector<size_t> neighborsPrune(const Point &p, double radius) const {
        return as_indices(list((int)K, OCT_ALGO_PRUNE, p, radius, nullptr));
    }
    template <Kernel_t K = Kernel_t::cube>
    std::vector<size_t> neighborsPrune(const Point &p, const Vector &radii) const {
        return as_indices(list((int)K, OCT_ALGO_PRUNE, p, 0.0, &radii));
    }
    template <Kernel_t K = Kernel_t::square>
    std::vector<size_t> neighbors(const Point &p, double radius) const {
        return as_indices(list((int)K, OCT_ALGO_NEIGHBORS, p, radius, nullptr));
    }
    template <Kernel_t K = Kernel_t::cube>
    std::vector<size_t> neighbors(const Point &p, const Vector &radii) const {
        return as_indices(list((int)K, OCT_ALGO_NEIGHBORS, p, 0.0, &radii));
    }
    // isInside && condition(point) (linear_octree.hpp:1076-1087): the condition is the caller's
    // host function, applied to the device's kernel hits in their order
    template <Kernel_t K = Kernel_t::square, class Function>
    std::vector<size_t> neighbors(const Point &p, double radius, Function &&condition) const {
        return filter(neighbors<K>(p, radius), condition);
    }
    template <Kernel_t K = Kernel_t::cube, class Function>
    std::vector<size_t> neighbors(const Point &p, const Vector &radii, Function &&condition) const {
        return filter(neighbors<K>(p, radii), condition);
    }

    // ---- approximate search (linear_octree.hpp:945-1035, 1089-1099): neighborsStruct with a depth
    // cutoff; nodes deeper than the precision level are added whole (upperBound) or dropped. As in
    // the reference, the tolerance argument of these overloads is passed on as the uint32_t precision
    // level itself (the double converts to the level parameter of the kernel overload).
    template <Kernel_t K = Kernel_t::square>
    NeighborSet<Container> neighborsApprox(const Point &p, double radius, double tolerancePercentage, bool upperBound) const {
        return as_set(approx((int)K, p, radius, nullptr, (uint32_t)tolerancePercentage, upperBound));
    }
    template <Kernel_t K = Kernel_t::cube>
    NeighborSet<Container> neighborsApprox(const Point &p, const Vector &radii, double tolerancePercentage,
                                           bool upperBound) const {
        return as_set(approx((int)K, p, 0.0, &radii, (uint32_t)tolerancePercentage, upperBound));
    }
    // getPrecisionLevel (linear_octree.hpp:945-973): the first level whose node radii are below the
    // tolerance; the z condition is the reference's `radii.getZ() && toleranceVector.getZ()`
    uint32_t getPrecisionLevel(const Vector &toleranceVector) const {
        fetch_tree();
        uint32_t precisionLevel = info_.max_depth_seen;
        for (uint32_t i = 0; i < info_.max_depth_seen; i++) {
            const double *r = &radii_[3 * i];
            if (r[0] < toleranceVector.getX() && r[1] < toleranceVector.getY() && r[2] && toleranceVector.getZ()) {
                precisionLevel = i;
                break;
            }
        }
        return precisionLevel;
    }
    uint32_t getPrecisionLevel(double tolerance) const { return getPrecisionLevel(Vector{tolerance, tolerance, tolerance}); }
    uint32_t getPrecisionLevel(const Vector &kernelRadii, double tolerancePercentage) const {
        return getPrecisionLevel(Vector{kernelRadii.getX() * tolerancePercentage / 100.0,
                                        kernelRadii.getY() * tolerancePercentage / 100.0,
                                        kernelRadii.getZ() * tolerancePercentage / 100.0});
    }

    // ---- kNN (linear_octree.hpp:899-942): min(n, k) nearest, ascending squared distance
    size_t knn(const Point &p, const size_t k, std::vector<size_t> &indexes, std::vector<double> &dists) const {
        const size_t m = std::min(points_.size(), k);
        if (m == 0) return 0;
        octamd::SearchRequest r;
        r.kind = octamd::SearchRequest::KNN;
        r.precision = OCT_PREC_FP64;
        r.k = (uint32_t)k;
        r.c[0] = p.getX(), r.c[1] = p.getY(), r.c[2] = p.getZ();
        batcher_->run(r);
        if (indexes.size() < m) indexes.resize(m);
        if (dists.size() < m) dists.resize(m);
        for (size_t i = 0; i < m; i++) indexes[i] = (size_t)r.knn_idx[i], dists[i] = r.knn_d2[i];
        return m;
    }

    // ---- batch API: the throughput path
    // per-centre neighbour counts for a batch of centres
    template <Kernel_t K = Kernel_t::sphere>
    std::vector<uint32_t> neighborsCount(const std::vector<Point> &centers, double radius) const {
        std::vector<double> c(3 * centers.size());
        for (size_t i = 0; i < centers.size(); i++)
            c[3 * i] = centers[i].getX(), c[3 * i + 1] = centers[i].getY(), c[3 * i + 2] = centers[i].getZ();
        std::vector<uint32_t> counts(centers.size());
        std::lock_guard<std::mutex> g(ctx_->mutex());
        octamd::check(ctx(), oct_radius(ctx(), (int)K, OCT_ALGO_STRUCT, kPrecision, c.data(), nullptr, nullptr, radius,
                                        centers.size(), 0, counts.data(), nullptr),
                      "oct_radius");
        return counts;
    }
    // every point of the (sorted) cloud as a centre, in SFC order (the harness' -s all mode)
    template <Kernel_t K = Kernel_t::sphere>
    std::vector<uint32_t> neighborsCountAll(double radius) const {
        std::vector<uint32_t> counts(points_.size());
        std::lock_guard<std::mutex> g(ctx_->mutex());
        octamd::check(ctx(), oct_radius(ctx(), (int)K, OCT_ALGO_STRUCT, kPrecision, nullptr, nullptr, nullptr, radius,
                                        points_.size(), 0, counts.data(), nullptr),
                      "oct_radius");
        return counts;
    }

    // ---- structure queries (linear_octree.hpp:626-651)
    size_t computeMemoryFootprint() const {
        return sizeof(LinearOctree) + (size_t)info_.n_total * (2 * sizeof(size_t) + sizeof(uint32_t) + sizeof(Point)) +
               sizeof(uint32_t) + 22 * sizeof(Vector);
    }
    double getDensity() const {
        return (double)points_.size() / (box_.radii().getX() * box_.radii().getY() * box_.radii().getZ() * 8.0f);
    }
    const oct_tree_info &treeInfo() const { return info_; }

    // singleTraversal (linear_octree.hpp:653-686) with the caller's host callbacks, over a copy of the
    // device tree: continuationCriterion(node, depth) -> bool, endpointAction(leaf)
    template <class C, class A>
    void singleTraversal(C &&continuationCriterion, A &&endpointAction) const {
        fetch_tree();
        if (!continuationCriterion(0u, 0u)) return;
        if (offsets_[0] == 0) {
            endpointAction(0u);
            return;
        }
        std::vector<std::pair<uint32_t, uint32_t>> stack{{0u, 0u}};
        uint32_t node = 0, depth = 0;
        do {
            for (uint32_t o = 0; o < 8; o++) {
                const uint32_t child = offsets_[node] + o;
                if (continuationCriterion(child, depth + 1)) {
                    if (offsets_[child] == 0)
                        endpointAction(child);
                    else
                        stack.emplace_back(child, depth + 1);
                }
            }
            std::tie(node, depth) = stack.back();
            stack.pop_back();
        } while (node != 0);
    }
    // node arrays of the reference (centers, internalRanges, offsets), copied from the device on first use
    const std::vector<Point> &centers() const { return fetch_tree(), centers_; }
    const std::vector<std::pair<size_t, size_t>> &internalRanges() const { return fetch_tree(), ranges_; }
    const std::vector<uint32_t> &offsets() const { return fetch_tree(), offsets_; }

private:
    Container &points_;
    PointEncoding::PointEncoder &enc_;
    std::vector<key_t> &codes_;
    Box box_;
    std::unique_ptr<octamd::Context> ctx_;
    std::unique_ptr<octamd::SearchBatcher> batcher_;  // per-query calls from many threads share launches
    oct_tree_info info_{};
    mutable std::vector<Point> centers_;
    mutable std::vector<std::pair<size_t, size_t>> ranges_;
    mutable std::vector<uint32_t> offsets_;
    mutable std::vector<double> radii_;  // precomputedRadii, 22 x 3
    std::unique_ptr<std::once_flag> tree_once_ = std::make_unique<std::once_flag>();

    oct_ctx *ctx() const { return ctx_->get(); }
    static bool is_box(Kernel_t t) { return t == Kernel_t::cube || t == Kernel_t::square; }

    struct Lists {
        int algo;
        std::vector<uint32_t> entries;
    };
    // one query's lists, through the batcher: concurrent callers share one launch
    Lists list(int kernel, int algo, const Point &p, double radius, const Vector *radii) const {
        octamd::SearchRequest r;
        r.kind = radii ? octamd::SearchRequest::BOX : octamd::SearchRequest::RADIUS;
        r.kernel = kernel, r.algo = algo, r.precision = kPrecision;
        r.c[0] = p.getX(), r.c[1] = p.getY(), r.c[2] = p.getZ();
        r.radius = radius;
        if (radii) r.r3[0] = radii->getX(), r.r3[1] = radii->getY(), r.r3[2] = radii->getZ();
        batcher_->run(r);
        return Lists{algo, std::move(r.entries)};
    }
    // a kernel object's search: the ball shapes by their radius, the box shapes by their radii
    Lists list_k(const KernelAbstract &k, int algo) const {
        if (is_box(k.type())) return list((int)k.type(), algo, k.center(), 0.0, &k.radii());
        return list((int)k.type(), algo, k.center(), k.radii().getX(), nullptr);
    }
    Lists approx(int kernel, const Point &p, double radius, const Vector *radii, uint32_t level, bool upper) const {
        const double c[3] = {p.getX(), p.getY(), p.getZ()};
        std::lock_guard<std::mutex> g(ctx_->mutex());
        double r3[3] = {0, 0, 0};
        if (radii) r3[0] = radii->getX(), r3[1] = radii->getY(), r3[2] = radii->getZ();
        octamd::check(ctx(), oct_radius_approx(ctx(), kernel, kPrecision, c, nullptr, nullptr, radius, radii ? r3 : nullptr, 1,
                                               level, upper ? 1 : 0, 1, nullptr, nullptr),
                      "oct_radius_approx");
        uint64_t offs[2];
        octamd::check(ctx(), oct_radius_fetch(ctx(), offs, nullptr, 0), "oct_radius_fetch");
        Lists l{OCT_ALGO_STRUCT, std::vector<uint32_t>((size_t)offs[1] * 2)};
        if (!l.entries.empty())
            octamd::check(ctx(), oct_radius_fetch(ctx(), offs, l.entries.data(), l.entries.size()), "oct_radius_fetch");
        return l;
    }
    NeighborSet<Container> as_set(const Lists &l) const {
        NeighborSet<Container> s(&points_);
        for (size_t i = 0; i + 1 < l.entries.size(); i += 2) s.addRange(l.entries[i], l.entries[i + 1]);
        return s;
    }
    static std::vector<size_t> as_indices(const Lists &l) { return std::vector<size_t>(l.entries.begin(), l.entries.end()); }
    template <class Function>
    std::vector<size_t> filter(std::vector<size_t> v, Function &&cond) const {
        std::vector<size_t> out;
        out.reserve(v.size());
        for (size_t i : v)
            if (cond(points_[i])) out.push_back(i);
        return out;
    }
    // Copies the device tree once; concurrent first callers wait until the copy is complete (a copy
    // that throws leaves the flag unset, so the next call tries again).
    void fetch_tree() const {
        std::call_once(*tree_once_, [this] {
            const size_t nt = info_.n_total;
            std::vector<uint64_t> rg(2 * nt);
            std::vector<double> ce(3 * nt);
            std::vector<uint32_t> off(nt);
            std::vector<double> rad(66);
            {
                std::lock_guard<std::mutex> g(ctx_->mutex());
                octamd::check(ctx(), oct_get_tree(ctx(), nullptr, off.data(), rg.data(), ce.data(), rad.data()), "oct_get_tree");
            }
            centers_.resize(nt);
            ranges_.resize(nt);
            for (size_t i = 0; i < nt; i++) {
                centers_[i] = Point(ce[3 * i], ce[3 * i + 1], ce[3 * i + 2]);
                ranges_[i] = {(size_t)rg[2 * i], (size_t)rg[2 * i + 1]};
            }
            radii_ = std::move(rad);
            offsets_ = std::move(off);
        });
    }
};
