// Drop-in PointEncoder (reference inc/encoding/point_encoder.hpp:15-52, 285-290, 484-503, 542-645).
// Encoding and the SFC reorder run on the GPU through the C ABI; the constexpr key helpers below
// are the reference's integer key algebra, kept for callers that use them.
#pragma once

#include <algorithm>
#include <cstdint>
#include <memory>
#include <optional>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../geometry/box.hpp"
#include "../geometry/point.hpp"
#include "../geometry/point_containers.hpp"
#include "../geometry/point_metadata.hpp"
#include "../main_options.hpp"
#include "../octree_amd_points.hpp"
#include "../octree_amd_runtime.hpp"

struct EncodingLog {
    double boundingBoxTime = 0.0, encodingTime = 0.0, sortingTime = 0.0;
};

namespace PointEncoding {

using coords_t = uint_fast32_t;
using key_t = uint_fast64_t;
static_assert(sizeof(key_t) == sizeof(uint64_t) && std::is_unsigned<key_t>::value, "64-bit keys");

class PointEncoder {
public:
    virtual ~PointEncoder() = default;

    virtual EncoderType getEncoder() const = 0;
    virtual std::string getEncoderName() const = 0;
    virtual std::string getShortEncoderName() const = 0;
    virtual bool is3D() const { return true; }
    virtual uint32_t maxDepth() const { return 21; }
    virtual double eps() const { return 1.0f / (1 << 21); }
    virtual key_t upperBound() const { return 0x8000000000000000ull; }
    virtual uint32_t unusedBits() const { return 1; }

    // encode / decode one cell coordinate triple (batched calls: encodeCoords / decodeKeys)
    virtual key_t encode(coords_t x, coords_t y, coords_t z) const {
        uint32_t c[3] = {(uint32_t)x, (uint32_t)y, (uint32_t)z};
        uint64_t k = 0;
        run_codec(c, &k, 1, false);
        return (key_t)k;
    }
    virtual void decode(key_t code, coords_t &x, coords_t &y, coords_t &z) const {
        uint32_t c[3];
        uint64_t k = code;
        run_codec(c, &k, 1, true);
        x = c[0], y = c[1], z = c[2];
    }
    // encodeVectorized (point_encoder.hpp:24-35): keys[i .. i + 8) for the coordinates x/y/z[0 .. 8)
    virtual void encodeVectorized(const uint32_t *x, const uint32_t *y, const uint32_t *z, std::vector<key_t> &keys,
                                  size_t i) const {
        uint32_t c[24];
        uint64_t k[8];
        for (int j = 0; j < 8; j++) c[3 * j] = x[j], c[3 * j + 1] = y[j], c[3 * j + 2] = z[j];
        run_codec(c, k, 8, false);
        for (int j = 0; j < 8; j++) keys[i + j] = (key_t)k[j];
    }

    // getAnchorCoords (point_encoder.hpp:37-49): the point's cell in the 2^maxDepth grid over bbox. The
    // 3-D encoders quantise on the device (the key of the point, decoded back: encode is a bijection
    // on the grid). The 2-D encoders' grid has 2^31 cells per axis; the reference forms that scale
    // as the int `1 << 31`, which overflows, so here it is the double 2^31.
    void getAnchorCoords(const Point &p, const Box &bbox, coords_t &x, coords_t &y, coords_t &z) const {
        if (is3D()) {
            decode(encodeFromPoint(p, bbox), x, y, z);
            return;
        }
        const double tx = ((p.getX() - bbox.center().getX()) + bbox.radii().getX()) / (2 * bbox.radii().getX());
        const double ty = ((p.getY() - bbox.center().getY()) + bbox.radii().getY()) / (2 * bbox.radii().getY());
        const double tz = ((p.getZ() - bbox.center().getZ()) + bbox.radii().getZ()) / (2 * bbox.radii().getZ());
        const double scale = (double)(1ull << maxDepth());
        const coords_t top = (coords_t(1) << maxDepth()) - 1;
        x = std::min((coords_t)(tx * scale), top);
        y = std::min((coords_t)(ty * scale), top);
        z = std::min((coords_t)(tz * scale), top);
    }
    // encodeFromPoint (point_encoder.hpp:52): one point's key under bbox
    virtual key_t encodeFromPoint(const Point &p, const Box &bbox) const {
        if (!is3D()) {
            coords_t x, y, z;
            getAnchorCoords(p, bbox, x, y, z);
            return encode(x, y, z);
        }
        const Point one[1] = {p};
        return encode_cloud(reinterpret_cast<const double *>(one), 1, sizeof(Point), bbox)[0];
    }
    // encodePoints (point_encoder.hpp:54-63): keys of every point under bbox, in the container's order
    template <class Container>
    std::vector<key_t> encodePoints(const Container &points, const Box &bbox) const {
        const std::vector<double> xyz = octamd::packed_xyz(points);
        return encode_cloud(xyz.data(), points.size(), 24, bbox);
    }

    // getCenter (point_encoder.hpp:517-539): centre of the level-`level` cell holding `code`, in the
    // reference's FP64 operation order (low corner of the cell, then + the level's radii)
    Point getCenter(key_t code, uint32_t level, const Box &bbox, const double *halfLengths,
                    const std::vector<Vector> &precomputedRadii) const {
        coords_t x, y, z;
        decode(code, x, y, z);
        const coords_t mask = ((coords_t(1) << maxDepth()) - 1) ^ ((coords_t(1) << (maxDepth() - level)) - 1);
        x &= mask, y &= mask, z &= mask;
        return Point(bbox.minX() + x * halfLengths[0] * 2, bbox.minY() + y * halfLengths[1] * 2,
                     bbox.minZ() + z * halfLengths[2] * 2) +
               precomputedRadii[level];
    }

    void encodeCoords(const uint32_t *xyz, size_t n, uint64_t *keys) const {
        run_codec(const_cast<uint32_t *>(xyz), keys, n, false);
    }
    void decodeKeys(const uint64_t *keys, size_t n, uint32_t *xyz) const {
        run_codec(xyz, const_cast<uint64_t *>(keys), n, true);
    }

    // sortPoints(points, meta, bbox): encode with the given box and reorder points (and metadata) in place
    template <class Container>
    std::vector<key_t> sortPoints(Container &points, std::optional<std::vector<PointMetadata>> &meta_opt, const Box &bbox,
                                  std::shared_ptr<EncodingLog> log = nullptr) const {
        auto &c = octamd::shared();
        std::lock_guard<std::mutex> g(c.mutex());
        const size_t n = points.size();
        octamd::check(c.get(), octamd::set_points(c.get(), points), "oct_set_points");
        const double ctr[3] = {bbox.center().getX(), bbox.center().getY(), bbox.center().getZ()};
        const double rad[3] = {bbox.radii().getX(), bbox.radii().getY(), bbox.radii().getZ()};
        octamd::check(c.get(), oct_encode(c.get(), (int)getEncoder(), ctr, rad, nullptr), "oct_encode");
        std::vector<uint32_t> perm(n);
        std::vector<key_t> keys(n);
        octamd::check(c.get(), oct_sort(c.get(), perm.data(), reinterpret_cast<uint64_t *>(keys.data())), "oct_sort");
        apply(c.get(), points, meta_opt, perm);
        if (log) {
            double ms[8];
            oct_stage_times(c.get(), ms);
            log->encodingTime = ms[1] * 1e-3, log->sortingTime = (ms[2] + ms[3]) * 1e-3;
        }
        return keys;
    }

    // sortPoints(points, meta): bounding box, encode and reorder; returns the keys and the Box
    template <class Container>
    std::pair<std::vector<key_t>, Box> sortPoints(Container &points, std::optional<std::vector<PointMetadata>> &meta_opt,
                                                  std::shared_ptr<EncodingLog> log = nullptr) const {
        auto &c = octamd::shared();
        std::lock_guard<std::mutex> g(c.mutex());
        const size_t n = points.size();
        octamd::check(c.get(), octamd::set_points(c.get(), points), "oct_set_points");
        double ctr[3], rad[3];
        std::vector<uint32_t> perm(n);
        std::vector<key_t> keys(n);
        octamd::check(c.get(),
                      oct_sort_points(c.get(), (int)getEncoder(), ctr, rad, perm.data(), reinterpret_cast<uint64_t *>(keys.data())),
                      "oct_sort_points");
        apply(c.get(), points, meta_opt, perm);
        if (log) {
            double ms[8];
            oct_stage_times(c.get(), ms);
            log->boundingBoxTime = ms[0] * 1e-3, log->encodingTime = ms[1] * 1e-3, log->sortingTime = (ms[2] + ms[3]) * 1e-3;
        }
        return {std::move(keys), Box(Point(ctr[0], ctr[1], ctr[2]), Vector(rad[0], rad[1], rad[2]))};
    }

    // ---- key algebra (point_encoder.hpp:542-645)
    uint32_t countLeadingZeros(key_t x) const { return x == 0 ? 64u : (uint32_t)__builtin_clzll((unsigned long long)x); }
    key_t nodeRange(uint32_t level) const { return key_t(1) << (3u * (maxDepth() - level)); }
    uint32_t getSiblingId(key_t code, uint32_t level) const { return (uint32_t)((code >> (3u * (maxDepth() - level))) & 7u); }
    uint32_t octalDigit(key_t code, uint32_t position) const { return (uint32_t)((code >> (3u * (maxDepth() - position))) & 7u); }
    uint32_t decodePrefixLength(key_t code) const { return 63u - countLeadingZeros(code); }
    key_t encodePlaceholderBit(key_t code, int level) const {
        return (code >> (3u * (maxDepth() - (uint32_t)level))) | (key_t(1) << (3u * (uint32_t)level));
    }
    key_t decodePlaceholderBit(key_t code) const {
        const uint32_t len = decodePrefixLength(code);
        return (code ^ (key_t(1) << len)) << (3u * maxDepth() - len);
    }
    uint32_t getLevel(key_t range) const {
        if (range == 0) return 0;
        return (countLeadingZeros(range - 1) - unusedBits()) / 3u;
    }

private:
    std::vector<key_t> encode_cloud(const double *xyz, size_t n, size_t stride, const Box &bbox) const {
        std::vector<key_t> keys(n);
        if (n == 0) return keys;
        auto &c = octamd::shared();
        std::lock_guard<std::mutex> g(c.mutex());
        octamd::check(c.get(), oct_set_points(c.get(), xyz, n, stride), "oct_set_points");
        const double ctr[3] = {bbox.center().getX(), bbox.center().getY(), bbox.center().getZ()};
        const double rad[3] = {bbox.radii().getX(), bbox.radii().getY(), bbox.radii().getZ()};
        octamd::check(c.get(), oct_encode(c.get(), (int)getEncoder(), ctr, rad, reinterpret_cast<uint64_t *>(keys.data())),
                      "oct_encode");
        return keys;
    }
    void run_codec(uint32_t *xyz, uint64_t *keys, size_t n, bool decode) const {
        auto &c = octamd::shared();
        std::lock_guard<std::mutex> g(c.mutex());
        if (decode)
            octamd::check(c.get(), oct_decode_keys(c.get(), (int)getEncoder(), keys, n, xyz), "oct_decode_keys");
        else
            octamd::check(c.get(), oct_encode_coords(c.get(), (int)getEncoder(), xyz, n, keys), "oct_encode_coords");
    }

    // the points and (when present) their metadata reordered as the keys, by the permutation the
    // context holds (records gathered on the device)
    template <class Container>
    static void apply(oct_ctx *ctx, Container &points, std::optional<std::vector<PointMetadata>> &meta_opt,
                      const std::vector<uint32_t> &perm) {
        octamd::reorder(ctx, points, perm);
        if (meta_opt) {
            std::vector<PointMetadata> m(meta_opt->size());
            if (m.size() == perm.size()) {
                if (!m.empty())
                    octamd::check(ctx, oct_gather_records(ctx, meta_opt->data(), m.data(), sizeof(PointMetadata), 0),
                                  "oct_gather_records");
            } else {  // a metadata vector of another length: the reference's loop over the points' order
                for (size_t i = 0; i < perm.size() && i < m.size(); i++) m[i] = (*meta_opt)[perm[i]];
            }
            meta_opt->swap(m);
        }
    }
};

}  // namespace PointEncoding
