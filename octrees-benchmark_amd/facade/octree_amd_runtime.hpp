// Device runtime shared by the drop-in facade headers: RAII ownership of C-ABI contexts and the
// error convention (status codes from the C ABI become std::runtime_error here; the reference
// itself has no error path on the search side).
#pragma once

#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>

#include "octree_amd.h"

namespace octamd {

inline int default_device() {
    const char *e = std::getenv("OCTREE_AMD_DEVICE");
    return e ? std::atoi(e) : 0;
}

inline void check(const oct_ctx *c, int status, const char *what) {
    if (status != OCT_OK)
        throw std::runtime_error(std::string("octree_amd: ") + what + ": " + (c ? oct_last_error(c) : "no context"));
}

// One C-ABI context: one HIP stream and the device buffers of one cloud / tree.
class Context {
public:
    explicit Context(int device = default_device()) { check(nullptr, oct_create(&ctx_, device), "oct_create"); }
    ~Context() {
        if (ctx_) oct_destroy(ctx_);
    }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    oct_ctx *get() const { return ctx_; }
    std::mutex &mutex() const { return mu_; }

private:
    oct_ctx *ctx_ = nullptr;
    mutable std::mutex mu_;
};

// Process-wide context for the calls that own no tree: encoders' key algebra, sortPoints, mbb.
inline Context &shared() {
    static Context c;
    return c;
}

}  // namespace octamd
