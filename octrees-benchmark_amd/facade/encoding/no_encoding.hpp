// Drop-in NoEncoding (reference inc/encoding/no_encoding.hpp:7-40): every key 0, every decode (0, 0, 0).
// An API stub as in the reference; it builds no LinearOctree.
#pragma once

#include <string>

#include "point_encoder.hpp"

namespace PointEncoding {

class NoEncoding : public PointEncoder {
public:
    static constexpr unsigned MAX_DEPTH = 21;
    static constexpr double EPS = 1.0f / (1 << MAX_DEPTH);
    static constexpr key_t UPPER_BOUND = 0x8000000000000000ull;
    static constexpr uint32_t UNUSED_BITS = 1;

    key_t encode(coords_t, coords_t, coords_t) const override { return 0; }
    void decode(key_t, coords_t &x, coords_t &y, coords_t &z) const override { x = y = z = 0; }
    void encodeVectorized(const uint32_t *, const uint32_t *, const uint32_t *, std::vector<key_t> &, size_t) const override {}
    key_t encodeFromPoint(const Point &, const Box &) const override { return 0; }

    uint32_t maxDepth() const override { return MAX_DEPTH; }
    double eps() const override { return EPS; }
    key_t upperBound() const override { return UPPER_BOUND; }
    uint32_t unusedBits() const override { return UNUSED_BITS; }
    bool is3D() const override { return false; }
    EncoderType getEncoder() const override { return EncoderType::NO_ENCODING; }
    std::string getEncoderName() const override { return "NoEncoding"; }
    std::string getShortEncoderName() const override { return "none"; }
};

}  // namespace PointEncoding
