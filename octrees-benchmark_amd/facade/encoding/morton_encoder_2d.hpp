// Drop-in MortonEncoder2D(axis) (reference inc/encoding/morton_encoder_2d.hpp:7-94): 2-D Morton keys of
// the two coordinates other than `axis` (31 bits each, the first in the even bits). An API stub as in
// the reference: its keys are computed on the host, and it builds no LinearOctree (the device path
// takes the 3-D encoders only).
#pragma once

#include <cstdint>
#include <string>

#include "point_encoder.hpp"

namespace PointEncoding {

namespace detail {
// bit j of v -> bit 2j
inline uint64_t spread2(uint64_t v) {
    v &= 0xffffffffull;
    v = (v | (v << 16)) & 0x0000ffff0000ffffull;
    v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
    v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
    v = (v | (v << 2)) & 0x3333333333333333ull;
    v = (v | (v << 1)) & 0x5555555555555555ull;
    return v;
}
// bit 2j of v -> bit j
inline uint64_t compact2(uint64_t v) {
    v &= 0x5555555555555555ull;
    v = (v | (v >> 1)) & 0x3333333333333333ull;
    v = (v | (v >> 2)) & 0x0f0f0f0f0f0f0f0full;
    v = (v | (v >> 4)) & 0x00ff00ff00ff00ffull;
    v = (v | (v >> 8)) & 0x0000ffff0000ffffull;
    v = (v | (v >> 16)) & 0x00000000ffffffffull;
    return v;
}
// the (first, second) coordinates a 2-D encoder keeps for `axis` (0: y, z; 1: x, z; else x, y)
inline void plane(uint8_t axis, coords_t x, coords_t y, coords_t z, coords_t &a, coords_t &b) {
    if (axis == 0) a = y, b = z;
    else if (axis == 1) a = x, b = z;
    else a = x, b = y;
}
inline void unplane(uint8_t axis, coords_t a, coords_t b, coords_t &x, coords_t &y, coords_t &z) {
    x = y = z = 0;
    if (axis == 0) y = a, z = b;
    else if (axis == 1) x = a, z = b;
    else x = a, y = b;
}
inline EncoderType encoder_2d(bool hilbert, uint8_t axis) {
    const int a = axis > 2 ? 2 : axis;
    return (EncoderType)((hilbert ? HILBERT_ENCODER_2D_X : MORTON_ENCODER_2D_X) + a);
}
}  // namespace detail

class MortonEncoder2D : public PointEncoder {
public:
    static constexpr unsigned MAX_DEPTH = 31;
    // The reference's expression: under C++20 the int 1 << 31 is -2^31, so this eps is negative, as there.
    static constexpr double EPS = 1.0f / (1 << MAX_DEPTH);
    static constexpr key_t UPPER_BOUND = 0x4000000000000000ull;
    static constexpr uint32_t UNUSED_BITS = 2;

    const uint8_t axis = 0;
    explicit MortonEncoder2D(uint8_t axis) : axis(axis) {}

    key_t encode(coords_t x, coords_t y, coords_t z) const override {
        coords_t a, b;
        detail::plane(axis, x, y, z, a, b);
        return (key_t)(detail::spread2(a) | (detail::spread2(b) << 1));
    }
    void decode(key_t code, coords_t &x, coords_t &y, coords_t &z) const override {
        detail::unplane(axis, (coords_t)detail::compact2(code), (coords_t)detail::compact2(code >> 1), x, y, z);
    }
    void encodeVectorized(const uint32_t *, const uint32_t *, const uint32_t *, std::vector<key_t> &, size_t) const override {}

    uint32_t maxDepth() const override { return MAX_DEPTH; }
    double eps() const override { return EPS; }
    key_t upperBound() const override { return UPPER_BOUND; }
    uint32_t unusedBits() const override { return UNUSED_BITS; }
    bool is3D() const override { return false; }
    EncoderType getEncoder() const override { return detail::encoder_2d(false, axis); }
    std::string getEncoderName() const override { return "MortonEncoder2D Axis " + std::string(1, (char)('X' + axis)); }
    std::string getShortEncoderName() const override { return "mort_2d_axis_" + std::string(1, (char)('X' + axis)); }
};

}  // namespace PointEncoding
