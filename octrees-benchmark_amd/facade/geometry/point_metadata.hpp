// Per-point LAS attributes carried alongside the cloud (reference point_metadata.hpp:9-105). The
// facade's sortPoints co-permutes a vector of them with the points.
#pragma once

#include <cstdint>

struct PointMetadata {
    double intensity{};
    uint8_t returnNumber{}, numberOfReturns{}, scanDirection{}, edgeOfFlightLine{};
    uint8_t classification{};
    int8_t scanAngleRank{};
    uint8_t userData{};
    uint16_t pointSourceId{};
    uint32_t r{}, g{}, b{};
};
