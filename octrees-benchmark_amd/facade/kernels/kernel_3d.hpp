// Drop-in 3-D kernels (reference inc/kernels/kernel_3d.hpp:10-235): KernelSphere (|p - c|^2 < r^2)
// and KernelCube (c - r <= p <= c + r per axis), with the reference's boxOverlap / boxIntersect
// decisions and FP64 operation order.
#pragma once

#include <algorithm>
#include <cmath>

#include "kernel_abstract.hpp"

class Kernel3D : public KernelAbstract {
public:
    Kernel3D(const Point &center, const double radius) : KernelAbstract(center, radius) {}
    Kernel3D(const Point &center, const Vector &radii) : KernelAbstract(center, radii) {}

    // false iff the box c -/+ r lies beyond the kernel box along some axis (kernel_3d.hpp:16-48)
    bool boxOverlap(const Point &c, const double r) const override { return boxOverlap(c, Vector(r, r, r)); }
    bool boxOverlap(const Point &c, const Vector &r) const override {
        const Point &lo = boxMin(), &hi = boxMax();
        if (c.getX() + r.getX() < lo.getX() || c.getY() + r.getY() < lo.getY() || c.getZ() + r.getZ() < lo.getZ()) return false;
        if (c.getX() - r.getX() > hi.getX() || c.getY() - r.getY() > hi.getY() || c.getZ() - r.getZ() > hi.getZ()) return false;
        return true;
    }
};

class KernelSphere : public Kernel3D {
    double radius_, radiusSq_;

public:
    KernelSphere(const Point &center, const double radius) : Kernel3D(center, radius), radius_(radius), radiusSq_(radius * radius) {}

    double radius() const { return radius_; }
    Kernel_t type() const override { return Kernel_t::sphere; }

    bool isInside(const Point &p) const override {
        const double dx = p.getX() - center().getX(), dy = p.getY() - center().getY(), dz = p.getZ() - center().getZ();
        return dx * dx + dy * dy + dz * dz < radiusSq_;
    }
    // node box (centre, half-side per axis) against the ball (kernel_3d.hpp:76-156): outside when
    // some axis gap exceeds r + half-side; inside when the farthest corner is within r; otherwise
    // overlapping when at least two axes straddle, or when the nearest point is within r
    IntersectionJudgement boxIntersect(const Point &oc, const double orad) const override {
        return judge(oc, orad, orad, orad);
    }
    IntersectionJudgement boxIntersect(const Point &oc, const Vector &orad) const override {
        return judge(oc, orad.getX(), orad.getY(), orad.getZ());
    }

private:
    IntersectionJudgement judge(const Point &oc, double hx, double hy, double hz) const {
        double ax = std::abs(center().getX() - oc.getX());
        double ay = std::abs(center().getY() - oc.getY());
        double az = std::abs(center().getZ() - oc.getZ());
        if (ax > radius_ + hx || ay > radius_ + hy || az > radius_ + hz) return IntersectionJudgement::OUTSIDE;
        const double fx = ax + hx, fy = ay + hy, fz = az + hz;
        if (fx * fx + fy * fy + fz * fz <= radiusSq_) return IntersectionJudgement::INSIDE;
        if ((ax < hx) + (ay < hy) + (az < hz) > 1) return IntersectionJudgement::OVERLAP;
        ax = std::max(ax - hx, 0.0), ay = std::max(ay - hy, 0.0), az = std::max(az - hz, 0.0);
        return ax * ax + ay * ay + az * az <= radiusSq_ ? IntersectionJudgement::OVERLAP : IntersectionJudgement::OUTSIDE;
    }
};

class KernelCube : public Kernel3D {
public:
    KernelCube(const Point &center, const double radius) : Kernel3D(center, radius) {}
    KernelCube(const Point &center, const Vector &radii) : Kernel3D(center, radii) {}

    Kernel_t type() const override { return Kernel_t::cube; }

    bool isInside(const Point &p) const override {
        return octamd::within(p.getX(), boxMin().getX(), boxMax().getX()) &&
               octamd::within(p.getY(), boxMin().getY(), boxMax().getY()) &&
               octamd::within(p.getZ(), boxMin().getZ(), boxMax().getZ());
    }
    // node box against the kernel box (kernel_3d.hpp:182-234): outside when disjoint along an axis,
    // inside when contained on every axis, else overlapping
    IntersectionJudgement boxIntersect(const Point &c, const double r) const override { return boxIntersect(c, Vector(r, r, r)); }
    IntersectionJudgement boxIntersect(const Point &c, const Vector &r) const override {
        const double hiX = c.getX() + r.getX(), loX = c.getX() - r.getX();
        const double hiY = c.getY() + r.getY(), loY = c.getY() - r.getY();
        const double hiZ = c.getZ() + r.getZ(), loZ = c.getZ() - r.getZ();
        const Point &kmin = boxMin(), &kmax = boxMax();
        if (hiX < kmin.getX() || hiY < kmin.getY() || hiZ < kmin.getZ() || loX > kmax.getX() || loY > kmax.getY() ||
            loZ > kmax.getZ())
            return IntersectionJudgement::OUTSIDE;
        if (hiX <= kmax.getX() && hiY <= kmax.getY() && hiZ <= kmax.getZ() && loX >= kmin.getX() && loY >= kmin.getY() &&
            loZ >= kmin.getZ())
            return IntersectionJudgement::INSIDE;
        return IntersectionJudgement::OVERLAP;
    }
};
