// Hand-off of the facade's point containers to the C ABI: uploads in the container's own layout
// (PointsAoS / std::vector<Point> as 32-byte records, PointsSoA as three coordinate arrays, any other
// PointContainer through a packed copy) and the device reorder of sortPoints applied back to the
// container (point_encoder.hpp:321-323, 359-375: the points move with their keys).
#pragma once

#include <cstdint>
#include <type_traits>
#include <vector>

#include "geometry/point.hpp"
#include "geometry/point_containers.hpp"
#include "octree_amd_runtime.hpp"

namespace octamd {

template <class C>
inline constexpr bool is_aos_v = std::is_same_v<C, PointsAoS> || std::is_same_v<C, std::vector<Point>>;

// x, y, z of any container, packed (the layout oct_set_points reads at stride 24)
template <class Container>
std::vector<double> packed_xyz(const Container &points) {
    std::vector<double> xyz(3 * points.size());
    size_t i = 0;
    for (auto it = points.begin(); it != points.end(); ++it, ++i) {
        const Point p = *it;
        xyz[3 * i] = p.getX(), xyz[3 * i + 1] = p.getY(), xyz[3 * i + 2] = p.getZ();
    }
    return xyz;
}

template <class Container>
int set_points(oct_ctx *c, const Container &points) {
    if constexpr (is_aos_v<Container>) {
        return oct_set_points(c, xyz_ptr(points), points.size(), sizeof(Point));
    } else if constexpr (std::is_same_v<Container, PointsSoA>) {
        return oct_set_points_soa(c, points.dataX(), points.dataY(), points.dataZ(), points.size());
    } else {
        const std::vector<double> xyz = packed_xyz(points);
        return oct_set_points(c, xyz.data(), points.size(), 24);
    }
}

template <class Container>
int set_sorted(oct_ctx *c, const Container &points, const uint64_t *codes, int encoder, const double ctr[3],
               const double rad[3]) {
    if constexpr (is_aos_v<Container>) {
        return oct_set_sorted(c, xyz_ptr(points), points.size(), sizeof(Point), codes, encoder, ctr, rad);
    } else if constexpr (std::is_same_v<Container, PointsSoA>) {
        return oct_set_sorted_soa(c, points.dataX(), points.dataY(), points.dataZ(), points.size(), codes, encoder, ctr, rad);
    } else {
        const std::vector<double> xyz = packed_xyz(points);
        return oct_set_sorted(c, xyz.data(), points.size(), 24, codes, encoder, ctr, rad);
    }
}

// points[i] <- points[perm[i]] with perm the context's last sort: trivially copyable record arrays
// are gathered on the device, other containers through their set() on the host permutation.
template <class Container>
void reorder(oct_ctx *c, Container &points, const std::vector<uint32_t> &perm) {
    const size_t n = points.size();
    if (n == 0) return;
    if constexpr (is_aos_v<Container>) {
        Container sorted(n);
        check(c, oct_gather_records(c, points.data(), sorted.data(), sizeof(Point), 0), "oct_gather_records");
        points.swap(sorted);
    } else if constexpr (std::is_same_v<Container, PointsSoA>) {
        PointsSoA sorted(n);
        check(c, oct_gather_records(c, points.dataX(), sorted.dataX(), sizeof(double), 0), "oct_gather_records");
        check(c, oct_gather_records(c, points.dataY(), sorted.dataY(), sizeof(double), 0), "oct_gather_records");
        check(c, oct_gather_records(c, points.dataZ(), sorted.dataZ(), sizeof(double), 0), "oct_gather_records");
        check(c, oct_gather_records(c, points.dataIds(), sorted.dataIds(), sizeof(size_t), 0), "oct_gather_records");
        points.swap(sorted);
    } else {
        Container sorted(points);
        for (size_t i = 0; i < n; i++) sorted.set(i, points[perm[i]]);
        points = std::move(sorted);
    }
}

}  // namespace octamd
