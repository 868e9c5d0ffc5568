// Drives the drop-in facade the way the reference's own example and tests do (examples/example.cpp,
// tests/structures/test_octree.cpp) and prints every result as text for comparison with the oracle.
// Usage: facade_check <encoder 1|5> <max_leaf> <points.bin | points.las> <centers.bin> <radius>
//   points.bin / centers.bin: raw little-endian FP64 x, y, z triples; a .las file is read with
//   readPoints (points and their attributes), and the attributes are printed after sortPoints.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <optional>
#include <string>
#include <thread>
#include <vector>

#include "encoding/hilbert_encoder_3d.hpp"
#include "encoding/morton_encoder_3d.hpp"
#include "encoding/point_encoder_factory.hpp"
#include "kernels/kernel_factory.hpp"
#include "readers/handlers.hpp"
#include "structures/linear_octree.hpp"

static std::vector<double> read_f64(const char *path) {
    std::ifstream f(path, std::ios::binary);
    std::vector<char> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<double> v(b.size() / 8);
    std::memcpy(v.data(), b.data(), v.size() * 8);
    return v;
}

template <Kernel_t K>
static void run_kernel(const LinearOctree<PointsAoS> &oct, const std::vector<Point> &centers, double r, const char *name) {
    for (size_t q = 0; q < centers.size(); q++) {
        auto s = oct.neighborsStruct<K>(centers[q], r);
        std::printf("struct %s %zu %zu", name, q, s.size());
        for (auto &rg : s.ranges) std::printf(" %zu:%zu", rg.first, rg.second);
        std::printf("\n");
        auto p = oct.neighborsPrune<K>(centers[q], r);
        std::printf("prune %s %zu %zu", name, q, p.size());
        for (size_t i : p) std::printf(" %zu", i);
        std::printf("\n");
        auto n = oct.neighbors<K>(centers[q], r);
        std::printf("neighbors %s %zu %zu", name, q, n.size());
        for (size_t i : n) std::printf(" %zu", i);
        std::printf("\n");
        size_t it = 0;
        for (auto e : s) it += (e.second.getX() == e.second.getX());
        if (it != s.size()) std::printf("iterator mismatch %zu %zu\n", it, s.size());
    }
}

// Per-query calls from 8 threads at once (the reference harness' OpenMP loop over queries): the
// batched launches must give every caller its own results, the same as sequential calls.
static void check_concurrent(const LinearOctree<PointsAoS> &oct, const std::vector<Point> &centers, double r) {
    const size_t reps = 40, nq = centers.size() * reps;
    std::vector<std::vector<std::pair<size_t, size_t>>> st(nq);
    std::vector<std::vector<size_t>> pr(nq), kn(nq);
    auto one = [&](size_t i, std::vector<std::pair<size_t, size_t>> &a, std::vector<size_t> &b, std::vector<size_t> &c) {
        const Point &p = centers[i % centers.size()];
        const double ri = r * (1.0 + 0.25 * (double)(i % 3));  // radii differ within a batch
        a = oct.neighborsStruct<Kernel_t::sphere>(p, ri).ranges;
        b = oct.neighborsPrune<Kernel_t::cube>(p, Vector(ri, 0.5 * ri, ri));
        std::vector<double> d2(5 + i % 4);
        c.assign(d2.size(), 0);
        oct.knn(p, d2.size(), c, d2);
    };
    std::vector<std::thread> th;
    for (int t = 0; t < 8; t++)
        th.emplace_back([&, t] {
            for (size_t i = (size_t)t; i < nq; i += 8) one(i, st[i], pr[i], kn[i]);
        });
    for (auto &x : th) x.join();
    size_t bad = 0;
    for (size_t i = 0; i < nq; i += 7) {
        std::vector<std::pair<size_t, size_t>> a;
        std::vector<size_t> b, c;
        one(i, a, b, c);
        bad += (a != st[i]) + (b != pr[i]) + (c != kn[i]);
    }
    std::printf("concurrent %zu %zu\n", (nq + 6) / 7, bad);
}

static bool same_set(const NeighborSet<PointsAoS> &a, const NeighborSet<PointsAoS> &b) {
    return a.ranges == b.ranges && a.size() == b.size();
}

// The kernel-object overloads (linear_octree.hpp:695, 746, 796, 980) against the (point, radius)
// overloads, for kernels made by kernelFactory, kernelFactory_ptr (through KernelAbstract&) and the
// box kernels with per-axis radii; mismatches are printed, the checks counted.
template <Kernel_t K>
static void check_kernel_objects(const LinearOctree<PointsAoS> &oct, const PointsAoS &points, const std::vector<Point> &centers,
                                 double r, const char *name) {
    size_t checks = 0;
    auto even = [](const Point &p) { return p.id() % 2 == 0; };
    for (size_t q = 0; q < centers.size(); q++) {
        const auto k = kernelFactory<K>(centers[q], r);
        const auto kp = kernelFactory_ptr<K>(centers[q], r);
        const KernelAbstract &ka = *kp;
        const auto st = oct.neighborsStruct<K>(centers[q], r);
        if (!same_set(oct.neighborsStruct(k), st) || !same_set(oct.neighborsStruct(ka), st))
            std::printf("kobj_mismatch struct %s %zu\n", name, q);
        const auto pr = oct.neighborsPrune<K>(centers[q], r);
        if (oct.neighborsPrune(k) != pr || oct.neighborsPrune(ka) != pr) std::printf("kobj_mismatch prune %s %zu\n", name, q);
        std::vector<size_t> nb = oct.neighbors<K>(centers[q], r), nbe;
        for (size_t i : nb)
            if (even(points[i])) nbe.push_back(i);
        if (oct.neighbors(k, even) != nbe || oct.neighbors<K>(centers[q], r, even) != nbe)
            std::printf("kobj_mismatch neighbors %s %zu\n", name, q);
        // every listed point is inside by the host kernel object, and the counts agree
        size_t inside = 0;
        for (size_t i : nb) inside += k.isInside(points[i]);
        if (inside != nb.size() || st.size() != nb.size()) std::printf("kobj_mismatch isInside %s %zu\n", name, q);
        for (uint32_t lv = 0; lv < 3; lv++)
            if (!same_set(oct.neighborsApprox(k, lv, lv == 1), oct.neighborsApprox<K>(centers[q], r, (double)lv, lv == 1)))
                std::printf("kobj_mismatch approx %s %zu\n", name, q);
        checks += 8;
        if constexpr (K == Kernel_t::cube || K == Kernel_t::square) {
            const Vector rv(r, 0.6 * r, 1.4 * r);
            const auto kb = kernelFactory<K>(centers[q], rv);
            if (!same_set(oct.neighborsStruct(kb), oct.neighborsStruct<K>(centers[q], rv)) ||
                oct.neighborsPrune(kb) != oct.neighborsPrune<K>(centers[q], rv))
                std::printf("kobj_mismatch box %s %zu\n", name, q);
            checks += 2;
        }
    }
    std::printf("kobj_ok %s %zu\n", name, checks);
}

// The same cloud as PointsSoA: sortPoints, LinearOctree<PointsSoA> and its searches must equal the
// AoS run's (same keys, same order, same tree, same lists).
static void check_soa(const std::vector<double> &xyz, const PointsAoS &sorted_aos, const std::vector<PointEncoding::key_t> &codes,
                      const LinearOctree<PointsAoS> &oct, const std::vector<Point> &centers, double r,
                      PointEncoding::PointEncoder &enc) {
    PointsSoA soa;
    for (size_t i = 0; i < xyz.size() / 3; i++) soa.push_back(Point(i, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
    std::optional<std::vector<PointMetadata>> none = std::nullopt;
    auto [scodes, sbox] = enc.sortPoints(soa, none);
    size_t same = 0;
    for (size_t i = 0; i < soa.size(); i++) same += (soa[i] == sorted_aos[i] && soa[i].id() == sorted_aos[i].id());
    LinearOctree<PointsSoA> soct(soa, scodes, sbox, enc);
    size_t lists = 0;
    for (size_t q = 0; q < centers.size(); q++) {
        const auto a = oct.neighborsStruct<Kernel_t::sphere>(centers[q], r);
        const auto b = soct.neighborsStruct<Kernel_t::sphere>(centers[q], r);
        size_t it = 0;
        for (auto e : b) it += (e.second == soa[e.first]);
        lists += (a.ranges == b.ranges && it == b.size() && oct.neighborsPrune(kernelFactory<Kernel_t::cube>(centers[q], r)) ==
                                                               soct.neighborsPrune(kernelFactory<Kernel_t::cube>(centers[q], r)));
    }
    std::printf("soa %zu %zu %d %d %zu\n", soa.size(), same, (int)(scodes == codes), (int)(soct.treeInfo().n_total == oct.treeInfo().n_total),
                lists);
}

int main(int argc, char **argv) {
    if (argc < 6) return 2;
    const int enc_id = std::atoi(argv[1]);
    mainOptions.maxPointsLeaf = (size_t)std::atoi(argv[2]);
    const std::vector<double> cq = read_f64(argv[4]);
    const double r = std::atof(argv[5]);
    PointsAoS points;
    std::optional<std::vector<PointMetadata>> metadata = std::nullopt;
    const std::string src = argv[3];
    if (src.size() > 4 && src.compare(src.size() - 4, 4, ".las") == 0) {
        auto read = readPoints<PointsAoS>(src);
        points = std::move(read.first);
        metadata = std::move(read.second);
    } else {
        const std::vector<double> xyz = read_f64(argv[3]);
        for (size_t i = 0; i < xyz.size() / 3; i++) points.push_back(Point(i, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
    }
    std::vector<Point> centers;
    for (size_t i = 0; i < cq.size() / 3; i++) centers.push_back(Point(cq[3 * i], cq[3 * i + 1], cq[3 * i + 2]));

    PointEncoding::PointEncoder &enc = PointEncoding::getEncoder((EncoderType)enc_id);
    auto [codes, box] = enc.sortPoints(points, metadata);
    if (metadata)
        for (size_t i = 0; i < metadata->size(); i++) {
            const PointMetadata &m = (*metadata)[i];
            std::printf("meta %zu %.17g %u %u %u %u %u %d %u %u %u %u %u\n", i, m.intensity, m.returnNumber, m.numberOfReturns,
                        m.scanDirection, m.edgeOfFlightLine, m.classification, (int)m.scanAngleRank, m.userData,
                        m.pointSourceId, m.r, m.g, m.b);
        }
    std::printf("box %.17g %.17g %.17g %.17g %.17g %.17g\n", box.minX(), box.minY(), box.minZ(), box.maxX(), box.maxY(), box.maxZ());
    for (size_t i = 0; i < points.size(); i++) std::printf("pt %zu %zu %llu\n", i, points[i].id(), (unsigned long long)codes[i]);

    // encodePoints / encodeFromPoint / getAnchorCoords (point_encoder.hpp:37-63) on the sorted cloud
    // give back its sorted keys
    {
        const auto again = enc.encodePoints(points, box);
        size_t anchors = 0;
        for (size_t i = 0; i < points.size(); i += points.size() / 7 + 1) {
            PointEncoding::coords_t x, y, z, ax, ay, az;
            enc.decode(codes[i], x, y, z);
            enc.getAnchorCoords(points[i], box, ax, ay, az);
            anchors += (ax == x && ay == y && az == z && enc.encodeFromPoint(points[i], box) == codes[i]);
        }
        std::printf("encodepoints %d %zu\n", (int)(again == codes), anchors);
    }
    LinearOctree<PointsAoS> oct(points, codes, box, enc);
    std::printf("tree %u %u %u\n", oct.treeInfo().n_leaf, oct.treeInfo().n_internal, oct.treeInfo().n_total);
    run_kernel<Kernel_t::sphere>(oct, centers, r, "sphere");
    run_kernel<Kernel_t::circle>(oct, centers, r, "circle");
    run_kernel<Kernel_t::cube>(oct, centers, r, "cube");
    run_kernel<Kernel_t::square>(oct, centers, r, "square");
    // neighborsApprox: the tolerance argument is the precision level itself (reference overload)
    for (size_t q = 0; q < centers.size(); q++)
        for (int lv = 0; lv < 4; lv++)
            for (int up = 0; up < 2; up++) {
                auto s = oct.neighborsApprox<Kernel_t::sphere>(centers[q], r, (double)lv, up != 0);
                std::printf("approx %d:%d %zu %zu", lv, up, q, s.size());
                for (auto &rg : s.ranges) std::printf(" %zu:%zu", rg.first, rg.second);
                std::printf("\n");
            }
    std::printf("precision_level %u %u\n", oct.getPrecisionLevel(r), oct.getPrecisionLevel(Vector(r, r, r), 50.0));
    for (size_t q = 0; q < centers.size(); q++) {
        std::vector<size_t> idx(7);
        std::vector<double> d2(7);
        size_t m = oct.knn(centers[q], 7, idx, d2);
        std::printf("knn %zu %zu", q, m);
        for (size_t i = 0; i < m; i++) std::printf(" %zu:%.17g", idx[i], d2[i]);
        std::printf("\n");
    }
    auto counts = oct.neighborsCountAll<Kernel_t::sphere>(r);
    unsigned long long tot = 0;
    for (auto c : counts) tot += c;
    std::printf("countall %llu\n", tot);
    size_t leaves = 0;
    oct.singleTraversal([](uint32_t, uint32_t) { return true; }, [&](uint32_t) { leaves++; });
    std::printf("traversal_leaves %zu\n", leaves);
    // getCenter (point_encoder.hpp:517-539) of each node from its first point's key equals the tree's centre
    {
        double hl[3];
        const double eps = enc.eps();
        hl[0] = 0.5f * eps * (box.maxX() - box.minX());
        hl[1] = 0.5f * eps * (box.maxY() - box.minY());
        hl[2] = 0.5f * eps * (box.maxZ() - box.minZ());
        std::vector<Vector> radii(enc.maxDepth() + 1);
        for (uint32_t l = 0; l <= enc.maxDepth(); l++) {
            const PointEncoding::coords_t side = 1u << (enc.maxDepth() - l);
            radii[l] = Vector(side * hl[0], side * hl[1], side * hl[2]);
        }
        size_t nodes = 0, exact = 0;
        oct.singleTraversal(
            [&](uint32_t node, uint32_t depth) {
                const auto rg = oct.internalRanges()[node];
                if (rg.first < rg.second) {
                    const Point c = enc.getCenter(codes[rg.first], depth, box, hl, radii);
                    const Point &t = oct.centers()[node];
                    nodes++;
                    exact += (c.getX() == t.getX() && c.getY() == t.getY() && c.getZ() == t.getZ());
                }
                return true;
            },
            [](uint32_t) {});
        std::printf("getcenter %zu %zu\n", nodes, exact);
    }
    check_kernel_objects<Kernel_t::sphere>(oct, points, centers, r, "sphere");
    check_kernel_objects<Kernel_t::circle>(oct, points, centers, r, "circle");
    check_kernel_objects<Kernel_t::cube>(oct, points, centers, r, "cube");
    check_kernel_objects<Kernel_t::square>(oct, points, centers, r, "square");
    if (!metadata) check_soa(read_f64(argv[3]), points, codes, oct, centers, r, enc);
    check_concurrent(oct, centers, r);
    PointEncoding::key_t k = enc.encode(1000, 2000, 3000);
    PointEncoding::coords_t x, y, z;
    enc.decode(k, x, y, z);
    std::printf("codec %llu %u %u %u\n", (unsigned long long)k, (unsigned)x, (unsigned)y, (unsigned)z);
    return 0;
}
