// LAS ingest: the reference's reader (las_file_reader_parallel.hpp:23-145, through LASlib) turns every
// point record into Point(i, X * xScale + xOffset, Y * yScale + yOffset, Z * zScale + zOffset)
// (:88-92, FP64, int32 X / Y / Z) and, with metadata, into PointMetadata(intensity, return number,
// number of returns, scan direction, edge of flight line, classification, scan angle, user data, point
// source id, R, G, B) (:124-137).
//
// Here the host parses the public header block (LAS 1.0-1.4 layouts) and streams the raw point records
// to the device through a pinned staging buffer; k_las_decode then turns 256 records per block into
// FP64 coordinates (the context's cloud, input order) and per-point attributes. Records are staged in
// LDS with aligned dword loads (a block's records start at a multiple of 256 * record length bytes)
// and each thread reads its record's fields from LDS bytes, so the global reads are coalesced
// whatever the record length; records longer than 128 bytes (many extra bytes) are read in place.
//
// Attribute rules, point formats 0-5: the record's own fields; classification is the 5-bit class
// (LASlib keeps the synthetic / key-point / withheld bits apart). Formats 6-10 (LAS 1.4) go through
// LASlib's legacy view: return number and number of returns folded to 3 bits (more than 7 returns:
// number 7, return 6 or 7), classification kept below 32 else 0, scan angle 0.006 degrees per unit
// converted to float and truncated to a char, as the reference's static_cast<char>(get_scan_angle()).
// Compressed (LAZ) files need LASzip and are rejected.
#include <cstdio>
#include <cstring>
#include <string>

#include "device.hpp"

namespace oct {

constexpr uint32_t kLasBlock = 256;
constexpr uint32_t kLasStageMax = 128;  // record bytes staged in LDS

struct LasDecodeArgs {
    const uint8_t *rec;
    uint32_t n, rl;
    int format;
    double scale[3], offset[3];
    double *xyz;
    oct_point_meta *meta;
};

__device__ inline uint32_t le16(const uint8_t *p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8); }
__device__ inline uint32_t le32(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

__global__ __launch_bounds__(kLasBlock) void k_las_decode(const LasDecodeArgs a) {
    __shared__ uint32_t s_tile[kLasBlock * kLasStageMax / 4];
    const uint32_t t = threadIdx.x;
    const uint32_t first = blockIdx.x * kLasBlock;
    const uint32_t cnt = (a.n - first) < kLasBlock ? (a.n - first) : kLasBlock;
    const uint8_t *r;
    if (a.rl <= kLasStageMax) {
        const uint32_t *src = reinterpret_cast<const uint32_t *>(a.rec + (size_t)first * a.rl);
        const uint32_t words = (cnt * a.rl + 3) / 4;  // the device copy is padded past the last record
        for (uint32_t w = t; w < words; w += kLasBlock) s_tile[w] = src[w];
        __syncthreads();
        r = reinterpret_cast<const uint8_t *>(s_tile) + t * a.rl;
    } else {
        r = a.rec + (size_t)(first + (t < cnt ? t : 0)) * a.rl;
    }
    if (t >= cnt) return;
    const size_t i = (size_t)first + t;
    const int32_t X = (int32_t)le32(r), Y = (int32_t)le32(r + 4), Z = (int32_t)le32(r + 8);
    a.xyz[3 * i + 0] = (double)X * a.scale[0] + a.offset[0];
    a.xyz[3 * i + 1] = (double)Y * a.scale[1] + a.offset[1];
    a.xyz[3 * i + 2] = (double)Z * a.scale[2] + a.offset[2];
    if (!a.meta) return;
    oct_point_meta m;
    m.intensity = (double)le16(r + 12);
    m.pad0 = 0, m.pad1 = 0;
    uint32_t rgb_at = 0;
    if (a.format <= 5) {
        const uint32_t f = r[14];
        m.return_number = (uint8_t)(f & 7u), m.number_of_returns = (uint8_t)((f >> 3) & 7u);
        m.scan_direction = (uint8_t)((f >> 6) & 1u), m.edge_of_flight_line = (uint8_t)(f >> 7);
        m.classification = (uint8_t)(r[15] & 31u);
        m.scan_angle_rank = (int8_t)r[16];
        m.user_data = r[17];
        m.point_source_id = (uint16_t)le16(r + 18);
        rgb_at = a.format == 2 ? 20u : (a.format == 3 || a.format == 5) ? 28u : 0u;
    } else {
        const uint32_t ern = r[14] & 15u, enr = r[14] >> 4;
        if (enr > 7) {
            m.return_number = (uint8_t)(ern > 6 ? (ern >= enr ? 7u : 6u) : ern);
            m.number_of_returns = 7;
        } else {
            m.return_number = (uint8_t)(ern & 7u), m.number_of_returns = (uint8_t)enr;
        }
        m.scan_direction = (uint8_t)((r[15] >> 6) & 1u), m.edge_of_flight_line = (uint8_t)(r[15] >> 7);
        m.classification = r[16] < 32 ? r[16] : 0;
        m.user_data = r[17];
        const int16_t sa = (int16_t)le16(r + 18);
        m.scan_angle_rank = (int8_t)(int)(0.006f * (float)sa);
        m.point_source_id = (uint16_t)le16(r + 20);
        rgb_at = (a.format == 7 || a.format == 8 || a.format == 10) ? 30u : 0u;
    }
    m.r = rgb_at ? le16(r + rgb_at) : 0u;
    m.g = rgb_at ? le16(r + rgb_at + 2) : 0u;
    m.b = rgb_at ? le16(r + rgb_at + 4) : 0u;
    a.meta[i] = m;
}

}  // namespace oct

using namespace oct;

namespace {

// smallest record of each point data format (LAS 1.4 R15, tables 7-17)
const uint16_t kLasMinRecord[11] = {20, 28, 26, 34, 57, 63, 30, 36, 38, 59, 67};

template <class T>
T get_le(const unsigned char *p) {
    T v;
    std::memcpy(&v, p, sizeof(T));  // the library runs on little-endian hosts (x86-64), as LAS is
    return v;
}

int las_fail(oct_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int parse_header(FILE *f, oct_las_header *h, std::string &err) {
    unsigned char b[375];
    std::memset(b, 0, sizeof(b));
    const size_t got = std::fread(b, 1, sizeof(b), f);
    if (got < 227 || std::memcmp(b, "LASF", 4) != 0) {
        err = "not a LAS file (signature)";
        return OCT_ERR_ARG;
    }
    std::memset(h, 0, sizeof(*h));
    h->version_major = b[24], h->version_minor = b[25];
    h->header_size = get_le<uint16_t>(b + 94);
    h->offset_to_points = get_le<uint32_t>(b + 96);
    h->n_vlr = get_le<uint32_t>(b + 100);
    const uint8_t fmt = b[104];
    h->record_length = get_le<uint16_t>(b + 105);
    uint64_t n = get_le<uint32_t>(b + 107);
    for (int a = 0; a < 3; a++) {
        h->scale[a] = get_le<double>(b + 131 + 8 * a);
        h->offset[a] = get_le<double>(b + 155 + 8 * a);
        h->max[a] = get_le<double>(b + 179 + 16 * a);
        h->min[a] = get_le<double>(b + 187 + 16 * a);
    }
    if (h->version_major != 1 || h->version_minor > 4) {
        err = "LAS version " + std::to_string(h->version_major) + "." + std::to_string(h->version_minor) + " not supported";
        return OCT_ERR_UNSUPPORTED;
    }
    if (fmt & 0xC0) {
        err = "compressed (LAZ) point records need LASzip";
        return OCT_ERR_UNSUPPORTED;
    }
    if (fmt > 10) {
        err = "point data record format " + std::to_string(fmt) + " not defined";
        return OCT_ERR_UNSUPPORTED;
    }
    h->point_format = fmt;
    if (h->version_minor >= 4 && n == 0) {
        if (got < 255 || h->header_size < 255) {
            err = "LAS 1.4 header too short";
            return OCT_ERR_ARG;
        }
        n = get_le<uint64_t>(b + 247);
    }
    h->n_points = n;
    if (h->record_length < kLasMinRecord[fmt]) {
        err = "point record length below the size of its format";
        return OCT_ERR_ARG;
    }
    if (h->offset_to_points < h->header_size) {
        err = "point data offset inside the header";
        return OCT_ERR_ARG;
    }
    return OCT_OK;
}

}  // namespace

extern "C" {

int oct_las_read_header(const char *path, oct_las_header *h) {
    if (!path || !h) return OCT_ERR_ARG;
    FILE *f = std::fopen(path, "rb");
    if (!f) return OCT_ERR_ARG;
    std::string err;
    const int r = parse_header(f, h, err);
    std::fclose(f);
    return r;
}

int oct_las_load(oct_ctx *ctx, const char *path, oct_las_header *h_out) {
    if (!ctx || !path) return las_fail(ctx, OCT_ERR_ARG, "null argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return las_fail(ctx, OCT_ERR_ARG, std::string("cannot open ") + path);
    struct Closer {
        FILE *f;
        ~Closer() { std::fclose(f); }
    } closer{f};
    oct_las_header h;
    std::string err;
    int r = parse_header(f, &h, err);
    if (r) return las_fail(ctx, r, err);
    if (h.n_points >= (1ull << 30)) return las_fail(ctx, OCT_ERR_ARG, "at most 2^30 - 1 points");
    const size_t n = (size_t)h.n_points, rl = h.record_length, bytes = n * rl;
    if (std::fseek(f, (long)h.offset_to_points, SEEK_SET) != 0) return las_fail(ctx, OCT_ERR_ARG, "seek to the point records");
    hipError_t e;
    if ((e = hipSetDevice(ctx->device)) != hipSuccess) return las_fail(ctx, OCT_ERR_HIP, hipGetErrorString(e));
    if ((e = ctx->las_rec.reserve(bytes + 8)) != hipSuccess) return las_fail(ctx, OCT_ERR_NOMEM, "alloc point records");
    if ((e = ctx->xyz.reserve(3 * n + 3)) != hipSuccess) return las_fail(ctx, OCT_ERR_NOMEM, "alloc points");
    if ((e = ctx->las_meta.reserve(n + 1)) != hipSuccess) return las_fail(ctx, OCT_ERR_NOMEM, "alloc point attributes");
    // file -> pinned staging (two halves, alternating) -> device
    const size_t half = 32u << 20;
    unsigned char *stage = nullptr;
    if ((e = hipHostMalloc((void **)&stage, 2 * half, hipHostMallocDefault)) != hipSuccess)
        return las_fail(ctx, OCT_ERR_NOMEM, "alloc pinned staging");
    hipEvent_t done[2] = {nullptr, nullptr};
    (void)hipEventCreate(&done[0]), (void)hipEventCreate(&done[1]);
    size_t pos = 0;
    int k = 0;
    r = OCT_OK;
    while (pos < bytes && r == OCT_OK) {
        const size_t len = bytes - pos < half ? bytes - pos : half;
        (void)hipEventSynchronize(done[k]);  // the copy that last used this half has finished
        if (std::fread(stage + k * half, 1, len, f) != len) {
            r = las_fail(ctx, OCT_ERR_ARG, "file shorter than its point records");
            break;
        }
        if ((e = hipMemcpyAsync(ctx->las_rec.p + pos, stage + k * half, len, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess ||
            (e = hipEventRecord(done[k], ctx->stream)) != hipSuccess) {
            r = las_fail(ctx, OCT_ERR_HIP, hipGetErrorString(e));
            break;
        }
        pos += len;
        k ^= 1;
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipEventDestroy(done[0]), (void)hipEventDestroy(done[1]);
    (void)hipHostFree(stage);
    if (r) return r;
    if (n) {
        LasDecodeArgs a;
        a.rec = ctx->las_rec.p;
        a.n = (uint32_t)n, a.rl = (uint32_t)rl, a.format = h.point_format;
        for (int c = 0; c < 3; c++) a.scale[c] = h.scale[c], a.offset[c] = h.offset[c];
        a.xyz = ctx->xyz.p, a.meta = ctx->las_meta.p;
        k_las_decode<<<(unsigned)((n + kLasBlock - 1) / kLasBlock), kLasBlock, 0, ctx->stream>>>(a);
        if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
            return las_fail(ctx, OCT_ERR_HIP, hipGetErrorString(e));
    }
    ctx->n = n;
    ctx->las_n = n;
    ctx->las_loaded = true;
    ctx->has_points = true;
    ctx->sort_lsd_learned = false;
    ctx->has_keys = ctx->has_sorted = ctx->has_tree = ctx->has_lists = false;
    if (h_out) *h_out = h;
    return OCT_OK;
}

int oct_las_fetch(oct_ctx *ctx, double *xyz, oct_point_meta *meta) {
    if (!ctx) return OCT_ERR_ARG;
    if (!ctx->las_loaded) return las_fail(ctx, OCT_ERR_STATE, "the context's cloud was not loaded from a LAS file");
    hipError_t e;
    if ((e = hipSetDevice(ctx->device)) != hipSuccess) return las_fail(ctx, OCT_ERR_HIP, hipGetErrorString(e));
    const size_t n = ctx->las_n;
    if (n && xyz && (e = hipMemcpyAsync(xyz, ctx->xyz.p, n * 24, hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess)
        return las_fail(ctx, OCT_ERR_HIP, hipGetErrorString(e));
    if (n && meta &&
        (e = hipMemcpyAsync(meta, ctx->las_meta.p, n * sizeof(oct_point_meta), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess)
        return las_fail(ctx, OCT_ERR_HIP, hipGetErrorString(e));
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return las_fail(ctx, OCT_ERR_HIP, hipGetErrorString(e));
    return OCT_OK;
}

}  // extern "C"
