// Drop-in 2-D kernels (reference inc/kernels/kernel_2d.hpp:10-218): KernelCircle (|p - c|^2 <= r^2
// in x, y) and KernelSquare (c - r <= p <= c + r in x, y); z is ignored.
#pragma once

#include <algorithm>
#include <cmath>

#include "kernel_abstract.hpp"

class Kernel2D : public KernelAbstract {
public:
    Kernel2D(const Point &center, const double radius) : KernelAbstract(center, radius) {}
    Kernel2D(const Point &center, const Vector &radii) : KernelAbstract(center, radii) {}

    // kernel_2d.hpp:16-41
    bool boxOverlap(const Point &c, const double r) const override { return boxOverlap(c, Vector(r, r, r)); }
    bool boxOverlap(const Point &c, const Vector &r) const override {
        const Point &lo = boxMin(), &hi = boxMax();
        if (c.getX() + r.getX() < lo.getX() || c.getY() + r.getY() < lo.getY()) return false;
        if (c.getX() - r.getX() > hi.getX() || c.getY() - r.getY() > hi.getY()) return false;
        return true;
    }
};

class KernelCircle : public Kernel2D {
    double radius_, radiusSq_;

public:
    KernelCircle(const Point &center, const double radius) : Kernel2D(center, radius), radius_(radius), radiusSq_(radius * radius) {}

    double radius() const { return radius_; }
    Kernel_t type() const override { return Kernel_t::circle; }

    bool isInside(const Point &p) const override {
        const double dx = p.getX() - center().getX(), dy = p.getY() - center().getY();
        return dx * dx + dy * dy <= radiusSq_;
    }
    // the sphere's decisions restricted to x, y (kernel_2d.hpp:69-144)
    IntersectionJudgement boxIntersect(const Point &oc, const double orad) const override { return judge(oc, orad, orad); }
    IntersectionJudgement boxIntersect(const Point &oc, const Vector &orad) const override {
        return judge(oc, orad.getX(), orad.getY());
    }

private:
    IntersectionJudgement judge(const Point &oc, double hx, double hy) const {
        double ax = std::abs(center().getX() - oc.getX());
        double ay = std::abs(center().getY() - oc.getY());
        if (ax > radius_ + hx || ay > radius_ + hy) return IntersectionJudgement::OUTSIDE;
        const double fx = ax + hx, fy = ay + hy;
        if (fx * fx + fy * fy <= radiusSq_) return IntersectionJudgement::INSIDE;
        if ((ax < hx) + (ay < hy) > 1) return IntersectionJudgement::OVERLAP;
        ax = std::max(ax - hx, 0.0), ay = std::max(ay - hy, 0.0);
        return ax * ax + ay * ay <= radiusSq_ ? IntersectionJudgement::OVERLAP : IntersectionJudgement::OUTSIDE;
    }
};

class KernelSquare : public Kernel2D {
public:
    KernelSquare(const Point &center, const double radius) : Kernel2D(center, radius) {}
    KernelSquare(const Point &center, const Vector &radii) : Kernel2D(center, radii) {}

    Kernel_t type() const override { return Kernel_t::square; }

    bool isInside(const Point &p) const override {
        return octamd::within(p.getX(), boxMin().getX(), boxMax().getX()) &&
               octamd::within(p.getY(), boxMin().getY(), boxMax().getY());
    }
    // kernel_2d.hpp:160-217
    IntersectionJudgement boxIntersect(const Point &c, const double r) const override { return boxIntersect(c, Vector(r, r, r)); }
    IntersectionJudgement boxIntersect(const Point &c, const Vector &r) const override {
        const double hiX = c.getX() + r.getX(), loX = c.getX() - r.getX();
        const double hiY = c.getY() + r.getY(), loY = c.getY() - r.getY();
        const Point &kmin = boxMin(), &kmax = boxMax();
        if (hiX < kmin.getX() || hiY < kmin.getY() || loX > kmax.getX() || loY > kmax.getY())
            return IntersectionJudgement::OUTSIDE;
        if (hiX <= kmax.getX() && hiY <= kmax.getY() && loX >= kmin.getX() && loY >= kmin.getY())
            return IntersectionJudgement::INSIDE;
        return IntersectionJudgement::OVERLAP;
    }
};
