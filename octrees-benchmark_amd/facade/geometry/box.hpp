// Drop-in Box and mbb (reference inc/geometry/box.hpp:17-21, 94-137).
#pragma once

#include <vector>

#include "../octree_amd_points.hpp"
#include "../octree_amd_runtime.hpp"
#include "point.hpp"
#include "point_containers.hpp"

class Box {
    Point center_;
    Vector radii_;
    Point min_, max_;

public:
    Box() = default;
    Box(const Point &center, const Vector &radii) : center_(center), radii_(radii) {
        min_ = center_ - radii_;
        max_ = center_ + radii_;
    }
    const Point &center() const { return center_; }
    const Vector &radii() const { return radii_; }
    double minX() const { return min_.getX(); }
    double minY() const { return min_.getY(); }
    double minZ() const { return min_.getZ(); }
    double maxX() const { return max_.getX(); }
    double maxY() const { return max_.getY(); }
    double maxZ() const { return max_.getZ(); }
    bool isInside(const Point &p) const {
        return p.getX() > minX() && p.getX() < maxX() && p.getY() > minY() && p.getY() < maxY() && p.getZ() > minZ() &&
               p.getZ() < maxZ();
    }
};

inline bool insideBox2D(const Point &p, const Vector &mn, const Vector &mx) {
    return p.getX() > mn.getX() && p.getY() > mn.getY() && p.getX() < mx.getX() && p.getY() < mx.getY();
}
inline bool insideBox3D(const Point &p, const Vector &mn, const Vector &mx) {
    return insideBox2D(p, mn, mx) && p.getZ() > mn.getZ() && p.getZ() < mx.getZ();
}

// mbb: bounding box of the cloud on the device; returns the centre, radii in maxRadius.
template <class Container>
inline Point mbb(const Container &points, Vector &maxRadius) {
    auto &c = octamd::shared();
    std::lock_guard<std::mutex> g(c.mutex());
    octamd::check(c.get(), octamd::set_points(c.get(), points), "oct_set_points");
    double ctr[3], rad[3];
    octamd::check(c.get(), oct_bbox(c.get(), ctr, rad), "oct_bbox");
    maxRadius = Vector(rad[0], rad[1], rad[2]);
    return Point(ctr[0], ctr[1], ctr[2]);
}
