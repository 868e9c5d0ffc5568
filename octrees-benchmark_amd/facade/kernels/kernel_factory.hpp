// Drop-in kernelFactory / kernelFactory_ptr (reference inc/kernels/kernel_factory.hpp:14-41): a kernel
// object of the Kernel_t shape around a centre; Vector radii only for the box shapes.
#pragma once

#include <memory>

#include "../geometry/point.hpp"
#include "../main_options.hpp"
#include "kernel_2d.hpp"
#include "kernel_3d.hpp"

template <Kernel_t kernel>
inline std::unique_ptr<KernelAbstract> kernelFactory_ptr(const Point &center, const double radius) {
    if constexpr (kernel == Kernel_t::circle) return std::make_unique<KernelCircle>(center, radius);
    else if constexpr (kernel == Kernel_t::square) return std::make_unique<KernelSquare>(center, radius);
    else if constexpr (kernel == Kernel_t::sphere) return std::make_unique<KernelSphere>(center, radius);
    else return std::make_unique<KernelCube>(center, radius);
}

template <Kernel_t kernel>
inline auto kernelFactory(const Point &center, const double radius) {
    if constexpr (kernel == Kernel_t::circle) return KernelCircle(center, radius);
    else if constexpr (kernel == Kernel_t::square) return KernelSquare(center, radius);
    else if constexpr (kernel == Kernel_t::sphere) return KernelSphere(center, radius);
    else return KernelCube(center, radius);
}

template <Kernel_t kernel>
inline auto kernelFactory(const Point &center, const Vector &radii) {
    static_assert(kernel == Kernel_t::square || kernel == Kernel_t::cube, "Vector radii need a box kernel (square or cube)");
    if constexpr (kernel == Kernel_t::square) return KernelSquare(center, radii);
    else return KernelCube(center, radii);
}
