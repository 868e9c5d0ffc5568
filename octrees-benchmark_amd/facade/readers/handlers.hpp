// Drop-in readPoints (reference readers/handlers.hpp:57-72): reads a point cloud file with its
// per-point attributes. LAS files go through the device-decoding LasFileReaderParallel; any other
// extension prints "Uncompatible file format" and exits with -1, as the reference does for formats it
// has no reader for.
#pragma once

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <optional>
#include <string>
#include <utility>
#include <vector>

#include "las_file_reader_parallel.hpp"

template <class Container>
std::pair<Container, std::optional<std::vector<PointMetadata>>> readPoints(const fs::path &fileName) {
    std::string ext = fileName.extension().string();
    for (auto &c : ext) c = (char)std::tolower((unsigned char)c);
    if (ext != ".las") {
        std::printf("Uncompatible file format\n");
        std::exit(-1);
    }
    LasFileReaderParallel<Container> reader(fileName);
    auto [points, metadata] = reader.readMeta();
    return std::make_pair(std::move(points), std::optional<std::vector<PointMetadata>>(std::move(metadata)));
}
