// Drop-in getEncoder(EncoderType) (reference point_encoder_factory.hpp:12-44). Only the 3-D
// encoders build a LinearOctree; the 2-D encoders and NoEncoding are not provided on the GPU path.
#pragma once

#include <stdexcept>

#include "hilbert_encoder_3d.hpp"
#include "morton_encoder_3d.hpp"

namespace PointEncoding {

inline PointEncoder &getEncoder(EncoderType t) {
    static MortonEncoder3D morton;
    static HilbertEncoder3D hilbert;
    if (t == EncoderType::MORTON_ENCODER_3D) return morton;
    if (t == EncoderType::HILBERT_ENCODER_3D) return hilbert;
    throw std::invalid_argument("octree_amd: only the 3-D Morton and Hilbert encoders are available");
}

}  // namespace PointEncoding
