// Drop-in LAS reader (reference readers/las_file_reader_parallel.hpp:23-145): the same class, constructor
// and read() / readMeta(). The file's header is parsed and its point records are decoded on the device
// by the C ABI (oct_las_load, oct_las_fetch) instead of LASlib; points are Point(i, X * scale + offset,
// ...) in file order, attributes PointMetadata. Uncompressed LAS only: a LAZ file raises
// std::runtime_error (LASzip is not available).
#pragma once

#include <filesystem>
#include <utility>
#include <vector>

#include "../geometry/point.hpp"
#include "../geometry/point_containers.hpp"
#include "../geometry/point_metadata.hpp"
#include "../octree_amd_runtime.hpp"

namespace fs = std::filesystem;

template <class Container>
class LasFileReaderParallel {
public:
    explicit LasFileReaderParallel(const fs::path &path) : path_(path) {}

    Container read() { return load(false).first; }
    std::pair<Container, std::vector<PointMetadata>> readMeta() { return load(true); }

private:
    fs::path path_;

    std::pair<Container, std::vector<PointMetadata>> load(bool with_meta) {
        octamd::Context ctx;  // the decoded cloud lives only for this call
        oct_las_header h;
        octamd::check(ctx.get(), oct_las_load(ctx.get(), path_.c_str(), &h), "oct_las_load");
        const size_t n = (size_t)h.n_points;
        std::vector<double> xyz(3 * n);
        std::vector<oct_point_meta> m(with_meta ? n : 0);
        octamd::check(ctx.get(), oct_las_fetch(ctx.get(), n ? xyz.data() : nullptr, (with_meta && n) ? m.data() : nullptr),
                      "oct_las_fetch");
        Container points(n);
        for (size_t i = 0; i < n; i++) points.set(i, Point(i, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
        std::vector<PointMetadata> meta;
        if (with_meta) {
            meta.resize(n);
            for (size_t i = 0; i < n; i++) {
                PointMetadata &d = meta[i];
                d.intensity = m[i].intensity;
                d.returnNumber = m[i].return_number, d.numberOfReturns = m[i].number_of_returns;
                d.scanDirection = m[i].scan_direction, d.edgeOfFlightLine = m[i].edge_of_flight_line;
                d.classification = m[i].classification;
                d.scanAngleRank = m[i].scan_angle_rank;
                d.userData = m[i].user_data;
                d.pointSourceId = m[i].point_source_id;
                d.r = m[i].r, d.g = m[i].g, d.b = m[i].b;
            }
        }
        return {std::move(points), std::move(meta)};
    }
};
