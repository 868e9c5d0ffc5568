// Drop-in search kernels (reference inc/kernels/kernel_abstract.hpp:13-66): a centre, per-axis radii
// and the kernel's bounding box c -/+ r (:41-45). The concrete kernels evaluate the reference's FP64
// predicates on the host, for callers that use the objects directly; LinearOctree's searches hand the
// kernel's type, centre and radii to the device, which evaluates the same predicates.
#pragma once

#include <cstdint>

#include "../geometry/box.hpp"
#include "../geometry/point.hpp"
#include "../main_options.hpp"

class KernelAbstract {
    Vector radii_{};
    Point center_{};
    Point boxMin_{}, boxMax_{};

public:
    enum class IntersectionJudgement { OUTSIDE = 0, OVERLAP = 1, INSIDE = 2 };

    KernelAbstract(const Point &center, const double radius) : radii_(radius, radius, radius), center_(center) { makeBox(); }
    KernelAbstract(const Point &center, const Vector &radii) : radii_(radii), center_(center) { makeBox(); }
    virtual ~KernelAbstract() = default;

    void makeBox() {
        boxMin_ = center_ - radii_;
        boxMax_ = center_ + radii_;
    }

    const Vector &radii() const { return radii_; }
    const Point &center() const { return center_; }
    const Point &boxMin() const { return boxMin_; }
    const Point &boxMax() const { return boxMax_; }

    // the kernel's shape, as the device's kernel id (Kernel_t values)
    virtual Kernel_t type() const = 0;

    virtual bool isInside(const Point &p) const = 0;
    virtual bool boxOverlap(const Point &center, double radius) const = 0;
    virtual bool boxOverlap(const Point &center, const Vector &radii) const = 0;
    virtual IntersectionJudgement boxIntersect(const Point &center, double radius) const = 0;
    virtual IntersectionJudgement boxIntersect(const Point &center, const Vector &radii) const = 0;
};

namespace octamd {
// n in [lo, hi] (the reference's onInterval, util.hpp:134-137)
inline bool within(double n, double lo, double hi) { return n >= lo && n <= hi; }
}  // namespace octamd
