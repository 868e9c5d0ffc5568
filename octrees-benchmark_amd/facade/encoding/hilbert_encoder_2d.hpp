// Drop-in HilbertEncoder2D(axis) (reference inc/encoding/hilbert_encoder_2d.hpp:7-135): the 2-D Hilbert
// curve over the two coordinates other than `axis`, 31 levels, in the reference's orientation (the
// curve starts at (0, 0) and its first quadrant step goes to (0, 1), digit (3 rx) ^ ry per level). An
// API stub as in the reference: keys are computed on the host, and it builds no LinearOctree.
#pragma once

#include <cstdint>
#include <string>
#include <utility>

#include "morton_encoder_2d.hpp"
#include "point_encoder.hpp"

namespace PointEncoding {

class HilbertEncoder2D : public PointEncoder {
public:
    static constexpr uint32_t MAX_DEPTH = 31;
    static constexpr double EPS = 1.0 / (1ull << MAX_DEPTH);
    static constexpr key_t UPPER_BOUND = 0x8000000000000000ull;
    static constexpr uint32_t UNUSED_BITS = 2;

    const uint8_t axis;
    explicit HilbertEncoder2D(uint8_t axis) : axis(axis) {}

    // Walks the quadrants from the top level down: each level contributes its digit, then the
    // remaining coordinates are mapped into the child's frame (transpose, and reflect for digit 3).
    key_t encode2D(coords_t a, coords_t b) const {
        const uint64_t side = 1ull << MAX_DEPTH;
        uint64_t x = a, y = b;
        key_t key = 0;
        for (uint64_t half = side >> 1; half; half >>= 1) {
            const uint64_t qa = (x & half) ? 1 : 0, qb = (y & half) ? 1 : 0;
            key += (key_t)(half * half) * ((3 * qa) ^ qb);
            if (!qb) {
                if (qa) x = side - 1 - x, y = side - 1 - y;
                std::swap(x, y);
            }
        }
        return key;
    }
    // Inverse of encode2D: digits from the lowest level up, each undoing one level's frame change.
    void decode2D(key_t key, coords_t &a, coords_t &b) const {
        uint64_t x = 0, y = 0, t = key;
        for (uint64_t s = 1; s < (1ull << MAX_DEPTH); s <<= 1, t >>= 2) {
            const uint64_t qa = 1 & (t >> 1), qb = 1 & (t ^ qa);
            if (!qb) {
                if (qa) x = s - 1 - x, y = s - 1 - y;
                std::swap(x, y);
            }
            x += s * qa, y += s * qb;
        }
        a = (coords_t)x, b = (coords_t)y;
    }

    key_t encode(coords_t x, coords_t y, coords_t z) const override {
        coords_t a, b;
        detail::plane(axis, x, y, z, a, b);
        return encode2D(a, b);
    }
    void decode(key_t code, coords_t &x, coords_t &y, coords_t &z) const override {
        coords_t a, b;
        decode2D(code, a, b);
        detail::unplane(axis, a, b, x, y, z);
    }
    void encodeVectorized(const uint32_t *, const uint32_t *, const uint32_t *, std::vector<key_t> &, size_t) const override {}

    uint32_t maxDepth() const override { return MAX_DEPTH; }
    double eps() const override { return EPS; }
    key_t upperBound() const override { return UPPER_BOUND; }
    uint32_t unusedBits() const override { return UNUSED_BITS; }
    bool is3D() const override { return false; }
    EncoderType getEncoder() const override { return detail::encoder_2d(true, axis); }
    std::string getEncoderName() const override { return "HilbertEncoder2D Axis " + std::string(1, (char)('X' + axis)); }
    std::string getShortEncoderName() const override { return "hilb_2d_axis_" + std::string(1, (char)('X' + axis)); }
};

}  // namespace PointEncoding
