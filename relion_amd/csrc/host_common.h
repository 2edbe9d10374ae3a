// Host-side HIP helpers shared by the E-step engine and the pose proposal: status macros, a grow-only device
// buffer and the projector/backprojector objects of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "rln_common.h"

#define HIP_TRY(expr)                                                                                     \
    do {                                                                                                  \
        hipError_t _e = (expr);                                                                           \
        if (_e != hipSuccess) {                                                                           \
            ::rln::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
            return RLN_ERR_HIP;                                                                           \
        }                                                                                                 \
    } while (0)

#define RLN_TRY(expr)             \
    do {                          \
        int _rc = (expr);         \
        if (_rc != RLN_OK) return _rc; \
    } while (0)

namespace rln {

// Grow-only device buffer.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    int ensure(size_t need) {
        if (need <= bytes) return RLN_OK;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        size_t want = std::max<size_t>(need, 256);
        if (hipMalloc(&ptr, want) != hipSuccess) {
            ptr = nullptr;
            set_error("device allocation of %zu bytes failed", want);
            return RLN_ERR_MEMORY;
        }
        bytes = want;
        return RLN_OK;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <class T>
    T *as() const {
        return reinterpret_cast<T *>(ptr);
    }
};

// Grow-only page-locked host buffer (hipHostMalloc): copies from it run asynchronously to the host.
struct PinnedBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf() { release(); }
    int ensure(size_t need) {
        if (need <= bytes) return RLN_OK;
        release();
        const size_t want = std::max<size_t>(need + need / 4, 4096);
        if (hipHostMalloc(&ptr, want, hipHostMallocDefault) != hipSuccess) {
            ptr = nullptr;
            set_error("page-locked host allocation of %zu bytes failed", want);
            return RLN_ERR_MEMORY;
        }
        bytes = want;
        return RLN_OK;
    }
    void release() {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

struct MdlDims {
    int xdim = 0, ydim = 0, zdim = 0, inity = 0, initz = 0, maxr = 0;
    double pad = 0.;
    bool operator==(const MdlDims &o) const {
        return xdim == o.xdim && ydim == o.ydim && zdim == o.zdim && inity == o.inity && initz == o.initz &&
               maxr == o.maxr && pad == o.pad;
    }
    size_t voxels() const { return (size_t)xdim * ydim * zdim; }
};

inline int check_dims(const MdlDims &d) {
    if (d.xdim <= 1 || d.ydim <= 1 || d.zdim <= 1 || d.maxr <= 0 || !(d.pad >= 1.))
        RLN_FAIL(RLN_ERR_ARG, "invalid model dimensions %d x %d x %d (maxr %d, pad %g)", d.xdim, d.ydim, d.zdim,
                 d.maxr, d.pad);
    return RLN_OK;
}

}  // namespace rln

struct rln_projector {
    int device = 0;
    rln::MdlDims dims;
    rln::DevBuf data;  // corner-packed volume: 4 float4 per trilinear base voxel, in bricks
    unsigned long long version = 0;
    size_t bases() const { return rlnk::packed_bases(dims.xdim, dims.ydim, dims.zdim); }
};

struct rln_backprojector {
    int device = 0;
    rln::MdlDims dims;
    rln::DevBuf slots;  // corner-slot accumulators: kSlot floats per trilinear base voxel, in bricks
    size_t bases() const { return rlnk::packed_bases(dims.xdim, dims.ydim, dims.zdim); }
};

namespace rln {
// Packs a plain [z][y][x] float2 volume on the device (dims of p) into p's corner layout, on stream st.
int pack_device_into_projector(rln_projector *p, const float2 *dev_plain, hipStream_t st);
// Sums b's corner slots into plain [z][y][x] data (re, im) and weight volumes (dims of b) on the device, on stream st.
int fold_backprojector_device(rln_backprojector *b, float *re, float *im, float *w, hipStream_t st);
}  // namespace rln


