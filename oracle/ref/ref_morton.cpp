// Thin C entry points over the reference's vendored libmorton (inc/encoding/libmorton/morton.h),
// compiled from the reference sources with the reference's Release flags. Used only as a checker
// for the oracle's Morton restatement (morton_encoder_3d.hpp:35-46 -> morton_BMI.h:40-58).
#include <cstdint>
#include "morton.h"

extern "C" uint64_t ref_morton3d_encode(uint64_t x, uint64_t y, uint64_t z) {
    return libmorton::morton3D_64_encode((uint_fast32_t)x, (uint_fast32_t)y, (uint_fast32_t)z);
}

extern "C" void ref_morton3d_decode(uint64_t key, uint64_t out[3]) {
    uint_fast32_t x, y, z;
    libmorton::morton3D_64_decode((uint_fast64_t)key, x, y, z);
    out[0] = x, out[1] = y, out[2] = z;
}

// 2-D keys as MortonEncoder2D uses them (morton_encoder_2d.hpp:27-40 -> libmorton morton2D_64_encode)
extern "C" uint64_t ref_morton2d_encode(uint64_t x, uint64_t y) {
    return libmorton::morton2D_64_encode((uint_fast32_t)x, (uint_fast32_t)y);
}
