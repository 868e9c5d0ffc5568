// Drop-in point containers (reference inc/geometry/point_containers.hpp:13-163): the PointContainer
// concept, the array-of-structs PointsAoS and the structure-of-arrays PointsSoA, with the reference's
// member functions. The device path reads an AoS cloud straight from its 32-byte records and an SoA
// cloud from its three coordinate arrays (oct_set_points / oct_set_points_soa); sortPoints reorders
// either layout with the permutation the device computed.
#pragma once

#include <concepts>
#include <cstddef>
#include <cstdlib>
#include <iterator>
#include <new>
#include <type_traits>
#include <utility>
#include <vector>

#include "point.hpp"

template <typename T>
concept PointContainer = std::copyable<T> && std::movable<T> && requires(T c, size_t i, const Point &p) {
    { c.size() } -> std::convertible_to<size_t>;
    { c.resize(i) };
    { c.clear() };
    { c.set(i, p) };
    { c[i] } -> std::convertible_to<Point>;
    { c.begin() } -> std::input_iterator;
    { c.end() } -> std::input_iterator;
    { c.push_back(p) };
};

// Array of structs: the points themselves, 32 bytes each (x, y, z, id).
class PointsAoS {
    std::vector<Point> pts_;

public:
    using value_type = Point;
    using Iterator = std::vector<Point>::const_iterator;

    PointsAoS(size_t n = 0) : pts_(n) {}
    PointsAoS(std::vector<Point> v) : pts_(std::move(v)) {}

    Point &operator[](size_t i) { return pts_[i]; }
    const Point &operator[](size_t i) const { return pts_[i]; }
    void set(size_t i, const Point &p) { pts_[i] = p; }
    void resize(size_t n) { pts_.resize(n); }
    size_t size() const { return pts_.size(); }
    bool empty() const { return pts_.empty(); }
    void clear() { pts_.clear(); }
    void push_back(const Point &p) { pts_.push_back(p); }
    void reserve(size_t n) { pts_.reserve(n); }

    Iterator begin() const { return pts_.begin(); }
    Iterator end() const { return pts_.end(); }
    Iterator cbegin() const { return pts_.cbegin(); }
    Iterator cend() const { return pts_.cend(); }

    // contiguous records, for the device transfers
    Point *data() { return pts_.data(); }
    const Point *data() const { return pts_.data(); }
    void swap(PointsAoS &o) noexcept { pts_.swap(o.pts_); }
    std::vector<Point> &vector() { return pts_; }
    const std::vector<Point> &vector() const { return pts_; }
};

namespace octamd {
// 64-byte aligned storage for the SoA arrays (the reference uses Eigen's aligned allocator).
template <class T>
struct AlignedAlloc {
    using value_type = T;
    AlignedAlloc() = default;
    template <class U>
    AlignedAlloc(const AlignedAlloc<U> &) {}
    T *allocate(size_t n) {
        void *p = ::operator new(n * sizeof(T), std::align_val_t(64));
        return static_cast<T *>(p);
    }
    void deallocate(T *p, size_t) { ::operator delete(p, std::align_val_t(64)); }
    template <class U>
    bool operator==(const AlignedAlloc<U> &) const { return true; }
    template <class U>
    bool operator!=(const AlignedAlloc<U> &) const { return false; }
};
}  // namespace octamd

// Structure of arrays: x, y, z and id in four arrays; element access yields Point copies.
class PointsSoA {
    std::vector<double, octamd::AlignedAlloc<double>> xs_, ys_, zs_;
    std::vector<size_t, octamd::AlignedAlloc<size_t>> ids_;

public:
    using value_type = Point;

    PointsSoA(size_t n = 0) : xs_(n), ys_(n), zs_(n), ids_(n) {}

    Point operator[](size_t i) const { return Point{ids_[i], xs_[i], ys_[i], zs_[i]}; }
    void set(size_t i, const Point &p) { xs_[i] = p.getX(), ys_[i] = p.getY(), zs_[i] = p.getZ(), ids_[i] = p.id(); }
    size_t size() const { return xs_.size(); }
    bool empty() const { return xs_.empty(); }
    void resize(size_t n) { xs_.resize(n), ys_.resize(n), zs_.resize(n), ids_.resize(n); }
    void clear() { xs_.clear(), ys_.clear(), zs_.clear(), ids_.clear(); }
    void push_back(const Point &p) {
        xs_.push_back(p.getX()), ys_.push_back(p.getY()), zs_.push_back(p.getZ()), ids_.push_back(p.id());
    }

    double *dataX() { return xs_.data(); }
    double *dataY() { return ys_.data(); }
    double *dataZ() { return zs_.data(); }
    size_t *dataIds() { return ids_.data(); }
    const double *dataX() const { return xs_.data(); }
    const double *dataY() const { return ys_.data(); }
    const double *dataZ() const { return zs_.data(); }
    const size_t *dataIds() const { return ids_.data(); }
    void swap(PointsSoA &o) noexcept { xs_.swap(o.xs_), ys_.swap(o.ys_), zs_.swap(o.zs_), ids_.swap(o.ids_); }

    // forward iterator yielding Point copies
    class Iterator {
        const PointsSoA *s_ = nullptr;
        size_t i_ = 0;

    public:
        using value_type = Point;
        using reference = Point;
        using pointer = void;
        using difference_type = std::ptrdiff_t;
        using iterator_category = std::forward_iterator_tag;

        Iterator() = default;
        Iterator(const PointsSoA *s, size_t i) : s_(s), i_(i) {}
        Point operator*() const { return (*s_)[i_]; }
        Iterator &operator++() {
            ++i_;
            return *this;
        }
        Iterator operator++(int) {
            Iterator t = *this;
            ++i_;
            return t;
        }
        bool operator==(const Iterator &o) const { return i_ == o.i_ && s_ == o.s_; }
        bool operator!=(const Iterator &o) const { return !(*this == o); }
    };
    Iterator begin() const { return Iterator(this, 0); }
    Iterator end() const { return Iterator(this, size()); }
    Iterator cbegin() const { return begin(); }
    Iterator cend() const { return end(); }
};

static_assert(PointContainer<PointsAoS>, "PointsAoS must satisfy PointContainer");
static_assert(PointContainer<PointsSoA>, "PointsSoA must satisfy PointContainer");

// First coordinate of an AoS cloud: Point is standard-layout with x first, so the C ABI reads x, y, z
// at a 32-byte stride straight from the caller's records.
static_assert(std::is_standard_layout<Point>::value, "Point must be standard-layout");
inline const double *xyz_ptr(const PointsAoS &pts) {
    return pts.empty() ? nullptr : reinterpret_cast<const double *>(pts.data());
}
inline const double *xyz_ptr(const std::vector<Point> &pts) {
    return pts.empty() ? nullptr : reinterpret_cast<const double *>(pts.data());
}
