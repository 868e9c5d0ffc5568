// Deterministic synthetic point-cloud generators for the benchmark workloads.
//
// Every point is produced from its own counter-based random stream (splitmix64 seeded with
// (seed, point index)), using only IEEE basic operations and an in-file sine polynomial, so the
// same cloud comes out bit-for-bit on any x86-64 host and for any OpenMP thread count. Coordinates
// are LAS-style: an integer millimetre grid, then X * 0.001 + offset, which is how the reference
// reads LAS points (inc/readers/las_file_reader_parallel.hpp:88-92).
//
// Built with g++ -O2 -fopenmp -ffp-contract=off (no FMA contraction, so results do not depend on
// the host CPU's FMA support).
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    Rng(uint64_t seed, uint64_t stream) : s(seed * 0xD1B54A32D192ED03ull ^ (stream + 0x632BE59BD9B4E019ull)) {
        splitmix64(s);
    }
    // U[0,1) with 53 random bits
    double uni() { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }
    // Irwin-Hall(12) - 6: approximately N(0, 1)
    double normal() {
        double a = 0.0;
        for (int i = 0; i < 12; i++) a += uni();
        return a - 6.0;
    }
};

const double kPi = 3.14159265358979323846;

// sin(x): reduce to [-pi, pi] with one rounded step, then a degree-15 odd Taylor polynomial
// (|error| < 1e-6 on the reduced range; only determinism matters here).
double det_sin(double x) {
    double k = std::floor(x * (1.0 / (2.0 * kPi)) + 0.5);
    double r = x - k * (2.0 * kPi);
    double r2 = r * r;
    double p = -1.0 / 1307674368000.0;               // -1/15!
    p = p * r2 + 1.0 / 6227020800.0;                 // 1/13!
    p = p * r2 - 1.0 / 39916800.0;                   // -1/11!
    p = p * r2 + 1.0 / 362880.0;                     // 1/9!
    p = p * r2 - 1.0 / 5040.0;                       // -1/7!
    p = p * r2 + 1.0 / 120.0;                        // 1/5!
    p = p * r2 - 1.0 / 6.0;                          // -1/3!
    return r + r * r2 * p;
}

inline double las_mm(double v, double offset) {
    double X = std::floor(v * 1000.0 + 0.5);
    return X * 0.001 + offset;
}

struct Terrain {
    double a[8], k[8], l[8], phi[8];
    explicit Terrain(uint64_t seed) {
        Rng r(seed, 0xA11CE5ull);
        double w[8], sw = 0.0;
        for (int i = 0; i < 8; i++) sw += (w[i] = 0.2 + 0.8 * r.uni());
        for (int i = 0; i < 8; i++) {
            a[i] = 14.0 * w[i] / sw;  // sum |a_i| = 14 m <= 15 m
            double lambda = 40.0 + 560.0 * r.uni();
            double th = 2.0 * kPi * r.uni();
            k[i] = 2.0 * kPi * det_sin(th + 0.5 * kPi) / lambda;
            l[i] = 2.0 * kPi * det_sin(th) / lambda;
            phi[i] = 2.0 * kPi * r.uni();
        }
    }
    double z(double x, double y) const {
        double s = 0.0;
        for (int i = 0; i < 8; i++) s += a[i] * det_sin(k[i] * x + l[i] * y + phi[i]);
        return s;
    }
};

}  // namespace

extern "C" {

// Uniform box [0,sx) x [0,sy) x [0,sz) shifted by (ox,oy,oz), LAS millimetre grid. xyz: n x 3.
void oct_synth_uniform(uint64_t seed, size_t n, double sx, double sy, double sz, double ox, double oy, double oz,
                       double *xyz) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        Rng r(seed, i);
        double x = sx * r.uni(), y = sy * r.uni(), z = sz * r.uni();
        xyz[3 * i + 0] = las_mm(x, ox);
        xyz[3 * i + 1] = las_mm(y, oy);
        xyz[3 * i + 2] = las_mm(z, oz);
    }
}

// Synthetic airborne-LiDAR tile over [0,extent)^2 (offset ox, oy):
//  - ground: z = sum of 8 random sinusoids (sum |a_i| = 14 m) + N(0, 0.05 m);
//  - n_buildings axis-aligned boxes (footprint U[10,40] m, height U[5,30] m) standing on the
//    terrain, points on roofs and walls in proportion to area (frac_building of the points);
//  - Gaussian vegetation blobs (sigma = 2 m) 3-15 m above ground holding frac_veg of the points.
void oct_synth_lidar_tile(uint64_t seed, size_t n, double extent, double ox, double oy, int n_buildings,
                          int n_blobs, double frac_building, double frac_veg, double *xyz) {
    Terrain terr(seed);
    struct Bld {
        double cx, cy, w, d, h, base, roof, wall, cum;
    };
    std::vector<Bld> b((size_t)(n_buildings > 0 ? n_buildings : 0));
    {
        Rng r(seed, 0xB0B0ull);
        double cum = 0.0;
        for (auto &q : b) {
            q.w = 10.0 + 30.0 * r.uni();
            q.d = 10.0 + 30.0 * r.uni();
            q.h = 5.0 + 25.0 * r.uni();
            q.cx = 0.5 * q.w + (extent - q.w) * r.uni();
            q.cy = 0.5 * q.d + (extent - q.d) * r.uni();
            q.base = terr.z(q.cx, q.cy);
            q.roof = q.w * q.d;
            q.wall = 2.0 * (q.w + q.d) * q.h;
            cum += q.roof + q.wall;
            q.cum = cum;
        }
    }
    struct Blob {
        double x, y, z;
    };
    std::vector<Blob> bl((size_t)(n_blobs > 0 ? n_blobs : 0));
    {
        Rng r(seed, 0xB10Bull);
        for (auto &q : bl) {
            q.x = extent * r.uni();
            q.y = extent * r.uni();
            q.z = terr.z(q.x, q.y) + 3.0 + 12.0 * r.uni();
        }
    }
    const double totalArea = b.empty() ? 0.0 : b.back().cum;
    const double lim = extent - 0.001;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        Rng r(seed, 0x100000000ull + i);
        double u = r.uni();
        double x, y, z;
        if (u < frac_veg && !bl.empty()) {
            const Blob &q = bl[(size_t)(r.uni() * (double)bl.size())];
            x = q.x + 2.0 * r.normal();
            y = q.y + 2.0 * r.normal();
            z = q.z + 2.0 * r.normal();
        } else if (u < frac_veg + frac_building && !b.empty()) {
            double t = r.uni() * totalArea;
            size_t lo = 0, hi = b.size() - 1;
            while (lo < hi) {
                size_t mid = (lo + hi) / 2;
                if (b[mid].cum <= t)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            const Bld &q = b[lo];
            double s = r.uni() * (q.roof + q.wall);
            if (s < q.roof) {
                x = q.cx + (r.uni() - 0.5) * q.w;
                y = q.cy + (r.uni() - 0.5) * q.d;
                z = q.base + q.h;
            } else {
                double per = r.uni() * 2.0 * (q.w + q.d);
                double x0 = q.cx - 0.5 * q.w, y0 = q.cy - 0.5 * q.d;
                if (per < q.w) {
                    x = x0 + per, y = y0;
                } else if (per < q.w + q.d) {
                    x = x0 + q.w, y = y0 + (per - q.w);
                } else if (per < 2.0 * q.w + q.d) {
                    x = x0 + (per - q.w - q.d), y = y0 + q.d;
                } else {
                    x = x0, y = y0 + (per - 2.0 * q.w - q.d);
                }
                z = q.base + q.h * r.uni();
            }
        } else {
            x = extent * r.uni();
            y = extent * r.uni();
            z = terr.z(x, y) + 0.05 * r.normal();
        }
        x = x < 0.0 ? 0.0 : (x > lim ? lim : x);
        y = y < 0.0 ? 0.0 : (y > lim ? lim : y);
        xyz[3 * i + 0] = las_mm(x, ox);
        xyz[3 * i + 1] = las_mm(y, oy);
        xyz[3 * i + 2] = las_mm(z, 0.0);
    }
}

// Synthetic mobile-laser-scan sweep along a vehicle trajectory of about `length` metres: a smooth
// random curve (Catmull-Rom spline through control points 50 m apart, heading a random walk) with a
// gently rolling road elevation. Returns, in proportion:
//  - 0.1 %: outliers, uniform in the corridor's bounding box (+-100 m, z in [-20, 60] m);
//  - 5 %: pole cylinders (radius 0.15 m, height 8 m) every 30 m on both sides, 6.5 m from the path;
//  - 30 %: facade planes on both sides, 9-12 m from the path (offset and height 6-20 m per 20 m block);
//  - the rest: ground returns at lateral range d drawn with density proportional to 1/d^2 on
//    [1, 80] m (inverse CDF d = 1 / (1 - u (1 - 1/80))), height = road elevation + N(0, 0.02 m).
// Scan rings are not traced; the mix reproduces their range falloff and surface types.
void oct_synth_mls_sweep(uint64_t seed, size_t n, double length, double ox, double oy, double *xyz) {
    const int M = (int)(length / 50.0) + 3;  // control points (with one extra at each end)
    std::vector<double> px((size_t)M), py((size_t)M);
    {
        Rng r(seed, 0x5EEDull);
        double h = 2.0 * kPi * r.uni(), x = 0.0, y = 0.0;
        for (int i = 0; i < M; i++) {
            px[(size_t)i] = x, py[(size_t)i] = y;
            h += 0.25 * r.normal();
            x += 50.0 * det_sin(h + 0.5 * kPi);
            y += 50.0 * det_sin(h);
        }
    }
    const int S = M - 3;  // spline segments between control points 1 .. M-2
    auto curve = [&](double t, double &x, double &y, double &tx, double &ty) {
        int j = (int)t;
        if (j > S - 1) j = S - 1;
        const double u = t - (double)j, u2 = u * u, u3 = u2 * u;
        const double *X = &px[(size_t)j], *Y = &py[(size_t)j];
        x = 0.5 * ((2.0 * X[1]) + (-X[0] + X[2]) * u + (2.0 * X[0] - 5.0 * X[1] + 4.0 * X[2] - X[3]) * u2 +
                   (-X[0] + 3.0 * X[1] - 3.0 * X[2] + X[3]) * u3);
        y = 0.5 * ((2.0 * Y[1]) + (-Y[0] + Y[2]) * u + (2.0 * Y[0] - 5.0 * Y[1] + 4.0 * Y[2] - Y[3]) * u2 +
                   (-Y[0] + 3.0 * Y[1] - 3.0 * Y[2] + Y[3]) * u3);
        tx = 0.5 * ((-X[0] + X[2]) + 2.0 * (2.0 * X[0] - 5.0 * X[1] + 4.0 * X[2] - X[3]) * u +
                    3.0 * (-X[0] + 3.0 * X[1] - 3.0 * X[2] + X[3]) * u2);
        ty = 0.5 * ((-Y[0] + Y[2]) + 2.0 * (2.0 * Y[0] - 5.0 * Y[1] + 4.0 * Y[2] - Y[3]) * u +
                    3.0 * (-Y[0] + 3.0 * Y[1] - 3.0 * Y[2] + Y[3]) * u2);
        const double nrm = std::sqrt(tx * tx + ty * ty);
        tx /= nrm, ty /= nrm;
    };
    auto road_z = [&](double t) { return 4.0 * det_sin(0.21 * t) + 1.5 * det_sin(0.73 * t + 1.0); };
    double mnx = px[0], mxx = px[0], mny = py[0], mxy = py[0];
    for (int i = 0; i < M; i++) {
        mnx = std::fmin(mnx, px[(size_t)i]), mxx = std::fmax(mxx, px[(size_t)i]);
        mny = std::fmin(mny, py[(size_t)i]), mxy = std::fmax(mxy, py[(size_t)i]);
    }
    const double span = (double)S;           // spline parameter range [0, S)
    const int n_poles = (int)(length / 30.0);  // per side
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        Rng r(seed, 0x300000000ull + i);
        const double u = r.uni();
        double x, y, z;
        if (u < 0.001) {
            x = mnx - 100.0 + (mxx - mnx + 200.0) * r.uni();
            y = mny - 100.0 + (mxy - mny + 200.0) * r.uni();
            z = -20.0 + 80.0 * r.uni();
        } else {
            const double side = r.uni() < 0.5 ? -1.0 : 1.0;
            double t, cx, cy, tx, ty, lat;
            if (u < 0.051) {  // pole
                const int k = (int)(r.uni() * (double)n_poles);
                t = ((double)k + 0.5) * span / (double)n_poles;
                curve(t, cx, cy, tx, ty);
                const double a = 2.0 * kPi * r.uni();
                lat = side * 6.5;
                x = cx - ty * lat + 0.15 * det_sin(a + 0.5 * kPi);
                y = cy + tx * lat + 0.15 * det_sin(a);
                z = road_z(t) + 8.0 * r.uni();
            } else if (u < 0.351) {  // facade
                t = span * r.uni();
                curve(t, cx, cy, tx, ty);
                Rng b(seed, 0xFACADEull + (uint64_t)(t * 2.5) * 2 + (side > 0.0 ? 1 : 0));  // 20 m blocks
                const double off = 9.0 + 3.0 * b.uni(), h = 6.0 + 14.0 * b.uni();
                lat = side * off;
                x = cx - ty * lat;
                y = cy + tx * lat;
                z = road_z(t) + h * r.uni();
            } else {  // ground, range density ~ 1/d^2 on [1, 80] m
                t = span * r.uni();
                curve(t, cx, cy, tx, ty);
                const double d = 1.0 / (1.0 - r.uni() * (1.0 - 1.0 / 80.0));
                lat = side * d;
                x = cx - ty * lat;
                y = cy + tx * lat;
                z = road_z(t) + 0.02 * r.normal();
            }
        }
        xyz[3 * i + 0] = las_mm(x, ox);
        xyz[3 * i + 1] = las_mm(y, oy);
        xyz[3 * i + 2] = las_mm(z, 40.0);
    }
}

// nq indices drawn uniformly with replacement from [0, n) (the reference's SearchSet random mode,
// inc/benchmarking/search_set.hpp:37-41, with a portable generator instead of std::mt19937).
void oct_synth_indices(uint64_t seed, size_t nq, size_t n, uint32_t *out) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < nq; i++) {
        Rng r(seed, 0x200000000ull + i);
        size_t v = (size_t)(r.uni() * (double)n);
        out[i] = (uint32_t)(v < n ? v : n - 1);
    }
}

}  // extern "C"
