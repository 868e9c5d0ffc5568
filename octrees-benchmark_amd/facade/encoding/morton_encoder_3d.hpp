// Drop-in MortonEncoder3D (reference inc/encoding/morton_encoder_3d.hpp).
#pragma once

#include "point_encoder.hpp"

namespace PointEncoding {

class MortonEncoder3D : public PointEncoder {
public:
    EncoderType getEncoder() const override { return EncoderType::MORTON_ENCODER_3D; }
    std::string getEncoderName() const override { return "MortonEncoder3D"; }
    std::string getShortEncoderName() const override { return "mort"; }
};

}  // namespace PointEncoding
