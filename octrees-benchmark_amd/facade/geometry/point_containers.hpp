// Drop-in point containers: the array-of-structs container (reference point_containers.hpp).
#pragma once

#include <type_traits>
#include <vector>

#include "point.hpp"

using PointsAoS = std::vector<Point>;

// First coordinate of an AoS cloud: Point is standard-layout with x first, so the C ABI reads
// x, y, z at a 32-byte stride straight from the caller's vector.
static_assert(std::is_standard_layout<Point>::value, "Point must be standard-layout");
inline const double *xyz_ptr(const PointsAoS &pts) {
    return pts.empty() ? nullptr : reinterpret_cast<const double *>(pts.data());
}
