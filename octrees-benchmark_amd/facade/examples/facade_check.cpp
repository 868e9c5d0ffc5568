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
#include <vector>

#include "encoding/hilbert_encoder_3d.hpp"
#include "encoding/morton_encoder_3d.hpp"
#include "encoding/point_encoder_factory.hpp"
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
    PointEncoding::key_t k = enc.encode(1000, 2000, 3000);
    PointEncoding::coords_t x, y, z;
    enc.decode(k, x, y, z);
    std::printf("codec %llu %u %u %u\n", (unsigned long long)k, (unsigned)x, (unsigned)y, (unsigned)z);
    return 0;
}
