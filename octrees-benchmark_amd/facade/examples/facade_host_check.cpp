// Host-side parts of the drop-in facade, exercised without a GPU: the kernel objects (kernelFactory,
// kernelFactory_ptr, isInside / boxOverlap / boxIntersect), the point containers (PointsAoS,
// PointsSoA, NeighborSet over both), the 2-D encoders and NoEncoding. The kernel objects drive the
// reference's traversals (linear_octree.hpp:653-816) over a tree read from files, so their decisions
// can be compared with the oracle's.
// Usage: facade_host_check <dir> <radius>
//   dir holds raw little-endian arrays: pts.f64 (sorted cloud, n x 3), q.f64 (centres, m x 3),
//   off.u32 (first child per node, 0 = leaf), rng.u64 (node point ranges, 2 per node),
//   ctr.f64 (node centres, 3 per node), rad.f64 (level radii, 22 x 3).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <utility>
#include <vector>

#include "encoding/point_encoder_factory.hpp"
#include "geometry/point_containers.hpp"
#include "kernels/kernel_factory.hpp"
#include "structures/neighbor_set.hpp"

template <class T>
static std::vector<T> read_raw(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    std::vector<char> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<T> v(b.size() / sizeof(T));
    std::memcpy(v.data(), b.data(), v.size() * sizeof(T));
    return v;
}

struct Tree {
    std::vector<uint32_t> off;
    std::vector<uint64_t> rng;
    std::vector<Point> ctr;
    std::vector<Vector> rad;
};

// singleTraversal (linear_octree.hpp:653-686) with the reference's callbacks
template <class Crit, class Leaf>
static void traverse(const Tree &t, Crit &&crit, Leaf &&leaf) {
    if (!crit(0u, 0u)) return;
    if (t.off[0] == 0) {
        leaf(0u);
        return;
    }
    std::vector<std::pair<uint32_t, uint32_t>> stack;
    uint32_t node = 0, depth = 0;
    for (;;) {
        for (uint32_t o = 0; o < 8; o++) {
            const uint32_t c = t.off[node] + o;
            if (crit(c, depth + 1)) {
                if (t.off[c] == 0) leaf(c);
                else stack.emplace_back(c, depth + 1);
            }
        }
        if (stack.empty()) break;
        std::tie(node, depth) = stack.back();
        stack.pop_back();
    }
}

// neighborsPrune (linear_octree.hpp:746-784) and neighbors (:796-816) with a kernel object
template <class K>
static std::vector<size_t> prune(const Tree &t, const PointsAoS &pts, const K &k) {
    std::vector<size_t> out;
    traverse(
        t,
        [&](uint32_t n, uint32_t d) {
            switch (k.boxIntersect(t.ctr[n], t.rad[d])) {
                case KernelAbstract::IntersectionJudgement::INSIDE:
                    for (size_t i = t.rng[2 * n]; i < t.rng[2 * n + 1]; i++) out.push_back(i);
                    return false;
                case KernelAbstract::IntersectionJudgement::OVERLAP: return true;
                default: return false;
            }
        },
        [&](uint32_t n) {
            for (size_t i = t.rng[2 * n]; i < t.rng[2 * n + 1]; i++)
                if (k.isInside(pts[i])) out.push_back(i);
        });
    return out;
}
template <class K>
static std::vector<size_t> overlap_only(const Tree &t, const PointsAoS &pts, const K &k) {
    std::vector<size_t> out;
    traverse(
        t, [&](uint32_t n, uint32_t d) { return k.boxOverlap(t.ctr[n], t.rad[d]); },
        [&](uint32_t n) {
            for (size_t i = t.rng[2 * n]; i < t.rng[2 * n + 1]; i++)
                if (k.isInside(pts[i])) out.push_back(i);
        });
    return out;
}

static void print_list(const char *tag, const char *name, size_t q, const std::vector<size_t> &v) {
    std::printf("%s %s %zu %zu", tag, name, q, v.size());
    for (size_t i : v) std::printf(" %zu", i);
    std::printf("\n");
}

template <Kernel_t K>
static void run_kernel(const Tree &t, const PointsAoS &pts, const std::vector<Point> &qs, double r, const char *name) {
    for (size_t q = 0; q < qs.size(); q++) {
        const auto k = kernelFactory<K>(qs[q], r);
        print_list("prune", name, q, prune(t, pts, k));
        print_list("neighbors", name, q, overlap_only(t, pts, k));
        std::vector<size_t> brute;
        for (size_t i = 0; i < pts.size(); i++)
            if (k.isInside(pts[i])) brute.push_back(i);
        print_list("brute", name, q, brute);
        // the same kernel through kernelFactory_ptr (virtual dispatch) and the scalar boxIntersect overload
        const auto kp = kernelFactory_ptr<K>(qs[q], r);
        if (kp->type() != K) std::printf("factory_ptr type mismatch\n");
        for (size_t n = 0; n < t.off.size(); n += 97) {
            const Vector &rv = t.rad[0];
            if (kp->boxIntersect(t.ctr[n], rv.getX()) != k.boxIntersect(t.ctr[n], Vector(rv.getX(), rv.getX(), rv.getX())))
                std::printf("scalar boxIntersect mismatch %zu\n", n);
        }
    }
}

template <Kernel_t K>
static void run_box(const Tree &t, const PointsAoS &pts, const std::vector<Point> &qs, const Vector &r, const char *name) {
    for (size_t q = 0; q < qs.size(); q++) print_list("boxprune", name, q, prune(t, pts, kernelFactory<K>(qs[q], r)));
}

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    const std::string dir = argv[1];
    const double r = std::atof(argv[2]);
    const auto xyz = read_raw<double>(dir + "/pts.f64");
    const auto qv = read_raw<double>(dir + "/q.f64");
    Tree t;
    t.off = read_raw<uint32_t>(dir + "/off.u32");
    t.rng = read_raw<uint64_t>(dir + "/rng.u64");
    const auto ctr = read_raw<double>(dir + "/ctr.f64");
    const auto rad = read_raw<double>(dir + "/rad.f64");
    for (size_t i = 0; i < ctr.size() / 3; i++) t.ctr.emplace_back(ctr[3 * i], ctr[3 * i + 1], ctr[3 * i + 2]);
    for (size_t i = 0; i < rad.size() / 3; i++) t.rad.emplace_back(rad[3 * i], rad[3 * i + 1], rad[3 * i + 2]);

    // containers: the same cloud as AoS (push_back) and SoA (resize + set), read back through the
    // PointContainer operations
    PointsAoS aos;
    PointsSoA soa;
    soa.resize(xyz.size() / 3);
    for (size_t i = 0; i < xyz.size() / 3; i++) {
        const Point p(i, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        aos.push_back(p);
        soa.set(i, p);
    }
    size_t same = 0, i = 0;
    for (auto it = soa.begin(); it != soa.end(); ++it, ++i) {
        const Point p = *it;
        same += (p == aos[i] && p.id() == aos[i].id() && soa.dataX()[i] == p.getX() && soa.dataIds()[i] == i);
    }
    std::printf("containers %zu %zu %zu\n", aos.size(), soa.size(), same);

    std::vector<Point> qs;
    for (size_t q = 0; q < qv.size() / 3; q++) qs.emplace_back(qv[3 * q], qv[3 * q + 1], qv[3 * q + 2]);
    run_kernel<Kernel_t::sphere>(t, aos, qs, r, "sphere");
    run_kernel<Kernel_t::circle>(t, aos, qs, r, "circle");
    run_kernel<Kernel_t::cube>(t, aos, qs, r, "cube");
    run_kernel<Kernel_t::square>(t, aos, qs, r, "square");
    const Vector rv(r, 0.6 * r, 1.4 * r);
    run_box<Kernel_t::cube>(t, aos, qs, rv, "cube");
    run_box<Kernel_t::square>(t, aos, qs, rv, "square");

    // NeighborSet over both containers: the same ranges yield the same (index, point) sequence
    NeighborSet<PointsAoS> na(&aos);
    NeighborSet<PointsSoA> ns(&soa);
    for (size_t a = 0; a + 7 < aos.size(); a += aos.size() / 5 + 1) na.addRange(a, a + 7), ns.addRange(a, a + 7);
    na.addRange(3, 3), ns.addRange(3, 3);
    size_t visited = 0, agree = 0;
    auto ia = na.begin();
    for (auto e : ns) {
        const auto ea = *ia;
        visited++;
        agree += (e.first == ea.first && e.second == ea.second);
        ++ia;
    }
    std::printf("neighborset %zu %zu %zu %d\n", na.size(), visited, agree, (int)(ia == na.end()));

    // 2-D encoders and NoEncoding
    using namespace PointEncoding;
    const uint64_t pairs[][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}, {5, 9}, {2147483647, 0}, {0, 2147483647},
                                 {2147483647, 2147483647}, {123456789, 987654321}, {1u << 30, 3}};
    for (auto &p : pairs) {
        PointEncoder &mz = getEncoder(MORTON_ENCODER_2D_Z), &hz = getEncoder(HILBERT_ENCODER_2D_Z);
        coords_t x, y, z;
        const PointEncoding::key_t km = mz.encode(p[0], p[1], 77), kh = hz.encode(p[0], p[1], 77);
        mz.decode(km, x, y, z);
        const bool mok = x == p[0] && y == p[1] && z == 0;
        hz.decode(kh, x, y, z);
        const bool hok = x == p[0] && y == p[1] && z == 0;
        std::printf("enc2d %llu %llu %llu %llu %d %d\n", (unsigned long long)p[0], (unsigned long long)p[1],
                    (unsigned long long)km, (unsigned long long)kh, (int)mok, (int)hok);
    }
    // axis selection: axis X keeps (y, z), axis Y keeps (x, z)
    {
        coords_t x, y, z;
        const PointEncoding::key_t kx = getEncoder(MORTON_ENCODER_2D_X).encode(11, 5, 9), ky = getEncoder(HILBERT_ENCODER_2D_Y).encode(5, 11, 9);
        getEncoder(MORTON_ENCODER_2D_X).decode(kx, x, y, z);
        std::printf("axis %llu %u %u %u", (unsigned long long)kx, (unsigned)x, (unsigned)y, (unsigned)z);
        getEncoder(HILBERT_ENCODER_2D_Y).decode(ky, x, y, z);
        std::printf(" %llu %u %u %u\n", (unsigned long long)ky, (unsigned)x, (unsigned)y, (unsigned)z);
    }
    // the first 4^4 Hilbert 2-D keys of a 16 x 16 corner, decoded: consecutive cells are unit steps
    {
        const HilbertEncoder2D &h = static_cast<const HilbertEncoder2D &>(getEncoder(HILBERT_ENCODER_2D_Z));
        int steps = 0, inside = 0;
        coords_t px = 0, py = 0;
        for (PointEncoding::key_t d = 0; d < 256; d++) {
            coords_t a, b;
            h.decode2D(d, a, b);
            inside += (a < 16 && b < 16 && h.encode2D(a, b) == d);
            if (d) steps += ((a > px ? a - px : px - a) + (b > py ? b - py : py - b)) == 1;
            px = a, py = b;
        }
        std::printf("hilbert2d_walk %d %d\n", steps, inside);
    }
    // keys of caller pairs (pairs.u64, n x 2) under MortonEncoder2D / HilbertEncoder2D, axis Z (x, y)
    {
        const auto pr = read_raw<uint64_t>(dir + "/pairs.u64");
        const PointEncoder &mz = getEncoder(MORTON_ENCODER_2D_Z), &hz = getEncoder(HILBERT_ENCODER_2D_Z);
        for (size_t j = 0; j + 1 < pr.size(); j += 2)
            std::printf("pair %llu %llu\n", (unsigned long long)mz.encode(pr[j], pr[j + 1], 0),
                        (unsigned long long)hz.encode(pr[j], pr[j + 1], 0));
    }
    for (int e = 0; e <= 9; e++) {
        const PointEncoder &enc = getEncoder((EncoderType)e);
        // (the 3-D encoders' codec runs on the device: not called here)
        coords_t x = 0, y = 0, z = 0;
        if (!enc.is3D()) enc.decode(enc.encode(3, 4, 5), x, y, z);
        std::printf("encoder %d %d %s %s %u %u %llu %d %.17g %u %u %u\n", e, (int)enc.getEncoder(),
                    enc.getEncoderName().c_str(), enc.getShortEncoderName().c_str(), enc.maxDepth(), enc.unusedBits(),
                    (unsigned long long)enc.upperBound(), (int)enc.is3D(), enc.eps(), (unsigned)x, (unsigned)y, (unsigned)z);
    }
    return 0;
}
