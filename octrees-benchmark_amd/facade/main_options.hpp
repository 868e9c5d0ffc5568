// Drop-in for the reference's main_options.hpp (enums and the build parameter the octree reads).
// Enum values are the reference's (inc/main_options.hpp:29-35, 73-83, 97-103), so configuration
// carries over unchanged; the C ABI uses the same numbers.
#pragma once

#include <cstddef>

enum SearchAlgo {
    NEIGHBORS_PTR, NEIGHBORS, NEIGHBORS_PRUNE, NEIGHBORS_STRUCT, NEIGHBORS_APPROX, NEIGHBORS_UNIBN,
    NEIGHBORS_PCLKD, NEIGHBORS_PCLOCT, NEIGHBORS_NANOFLANN, NEIGHBORS_PICOTREE, KNN_V2, KNN_NANOFLANN,
    KNN_PCLKD, KNN_PCLOCT, KNN_PICOTREE
};

enum EncoderType {
    NO_ENCODING, MORTON_ENCODER_3D, MORTON_ENCODER_2D_X, MORTON_ENCODER_2D_Y, MORTON_ENCODER_2D_Z,
    HILBERT_ENCODER_3D, HILBERT_ENCODER_2D_X, HILBERT_ENCODER_2D_Y, HILBERT_ENCODER_2D_Z
};

enum class Kernel_t { circle, sphere, square, cube };

// The fields of the reference's options struct that the octree path reads (main_options.hpp:210).
struct main_options {
    size_t maxPointsLeaf = 128;
    bool searchAll = false;
};

inline main_options mainOptions;
