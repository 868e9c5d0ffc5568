// Drop-in Point / Vector (reference inc/geometry/point.hpp:15-60): x, y, z and a size_t id, 32
// bytes, the layout the C ABI reads directly (stride 32).
#pragma once

#include <cmath>
#include <cstddef>

class Point {
protected:
    double x_{}, y_{}, z_{};
    size_t id_{};

public:
    Point() = default;
    explicit Point(size_t id) : id_(id) {}
    Point(double x, double y) : x_(x), y_(y) {}
    Point(double x, double y, double z) : x_(x), y_(y), z_(z) {}
    Point(size_t id, double x, double y, double z) : x_(x), y_(y), z_(z), id_(id) {}

    double getX() const { return x_; }
    double getY() const { return y_; }
    double getZ() const { return z_; }
    size_t id() const { return id_; }
    void setX(double v) { x_ = v; }
    void setY(double v) { y_ = v; }
    void setZ(double v) { z_ = v; }
    void setId(size_t v) { id_ = v; }

    Point operator+(const Point &o) const { return Point(x_ + o.x_, y_ + o.y_, z_ + o.z_); }
    Point operator-(const Point &o) const { return Point(x_ - o.x_, y_ - o.y_, z_ - o.z_); }
    Point operator*(double f) const { return Point(x_ * f, y_ * f, z_ * f); }
    Point operator/(double f) const { return Point(x_ / f, y_ / f, z_ / f); }
    bool operator==(const Point &o) const { return x_ == o.x_ && y_ == o.y_ && z_ == o.z_; }

    double distance3D(const Point &p) const {
        const double dx = x_ - p.x_, dy = y_ - p.y_, dz = z_ - p.z_;
        return std::sqrt(dx * dx + dy * dy + dz * dz);
    }
};

using Vector = Point;

static_assert(sizeof(Point) == 32, "the C ABI reads points with a 32-byte stride");

// Squared distance as the reference's knn reports it (linear_octree.hpp:844-851).
inline double squaredDistance(const Point &a, const Point &b) {
    const double dx = a.getX() - b.getX(), dy = a.getY() - b.getY(), dz = a.getZ() - b.getZ();
    return dx * dx + dy * dy + dz * dz;
}
