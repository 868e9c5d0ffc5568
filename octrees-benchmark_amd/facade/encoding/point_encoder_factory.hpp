// Drop-in getEncoder(EncoderType) (reference inc/encoding/point_encoder_factory.hpp:12-44): one static
// encoder per type; NoEncoding for NO_ENCODING and unknown values. Only the 3-D encoders sort and
// build on the device.
#pragma once

#include "hilbert_encoder_2d.hpp"
#include "hilbert_encoder_3d.hpp"
#include "morton_encoder_2d.hpp"
#include "morton_encoder_3d.hpp"
#include "no_encoding.hpp"

namespace PointEncoding {

inline PointEncoder &getEncoder(EncoderType t) {
    static MortonEncoder3D morton3D;
    static HilbertEncoder3D hilbert3D;
    static MortonEncoder2D morton2DX(0), morton2DY(1), morton2DZ(2);
    static HilbertEncoder2D hilbert2DX(0), hilbert2DY(1), hilbert2DZ(2);
    static NoEncoding none;
    switch (t) {
        case EncoderType::MORTON_ENCODER_3D: return morton3D;
        case EncoderType::HILBERT_ENCODER_3D: return hilbert3D;
        case EncoderType::MORTON_ENCODER_2D_X: return morton2DX;
        case EncoderType::MORTON_ENCODER_2D_Y: return morton2DY;
        case EncoderType::MORTON_ENCODER_2D_Z: return morton2DZ;
        case EncoderType::HILBERT_ENCODER_2D_X: return hilbert2DX;
        case EncoderType::HILBERT_ENCODER_2D_Y: return hilbert2DY;
        case EncoderType::HILBERT_ENCODER_2D_Z: return hilbert2DZ;
        default: return none;
    }
}

}  // namespace PointEncoding
