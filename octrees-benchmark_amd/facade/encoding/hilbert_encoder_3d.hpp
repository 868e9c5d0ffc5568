// Drop-in HilbertEncoder3D (reference inc/encoding/hilbert_encoder_3d.hpp).
#pragma once

#include "point_encoder.hpp"

namespace PointEncoding {

class HilbertEncoder3D : public PointEncoder {
public:
    EncoderType getEncoder() const override { return EncoderType::HILBERT_ENCODER_3D; }
    std::string getEncoderName() const override { return "HilbertEncoder3D"; }
    std::string getShortEncoderName() const override { return "hilb"; }
};

}  // namespace PointEncoding
