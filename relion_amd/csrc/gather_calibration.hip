// Reference kernels of known traffic for reading HBM counters: random aligned gathers of 64-byte blocks (the
// corner blocks the fine pass reads), random aligned gathers of 128-byte blocks, and a streaming copy, over a
// 4 GiB buffer. Printed: the bytes each kernel moves and its useful bandwidth. Run under
// `rocprofv3 --pmc FETCH_SIZE` (or WRITE_SIZE) to see what the counter reads for each request size.
//
// Usage: gather_calibration [samples_log2 = 24]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#define CK(x)                                                                      \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));           \
            return 1;                                                              \
        }                                                                          \
    } while (0)

// one 64-byte block per thread: block index idx[i] (< 2^26 for a 4 GiB buffer)
__global__ void k_gather64(const float4 *__restrict__ buf, const unsigned *__restrict__ idx, long long n,
                           float *__restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 *p = buf + (size_t)idx[i] * 4;
    const float4 a = p[0], b = p[1], c = p[2], d = p[3];
    out[i] = a.x + b.y + c.z + d.w;
}

// one 128-byte block per thread: block index idx[i] (< 2^25)
__global__ void k_gather128(const float4 *__restrict__ buf, const unsigned *__restrict__ idx, long long n,
                            float *__restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 *p = buf + (size_t)idx[i] * 8;
    const float4 a = p[0], b = p[1], c = p[2], d = p[3], e = p[4], f = p[5], g = p[6], h = p[7];
    out[i] = a.x + b.y + c.z + d.w + e.x + f.y + g.z + h.w;
}

// streaming copy of n float4
__global__ void k_stream_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int main(int argc, char **argv) {
    const int lg = argc > 1 ? std::atoi(argv[1]) : 24;
    if (lg < 10 || lg > 26) {
        std::fprintf(stderr, "samples_log2 must be in [10, 26]\n");
        return 2;
    }
    const size_t bytes = (size_t)1 << 32;
    const long long n = 1LL << lg;
    float4 *buf = nullptr, *dst = nullptr;
    unsigned *d_i64 = nullptr, *d_i128 = nullptr;
    float *out = nullptr;
    CK(hipMalloc(&buf, bytes));
    CK(hipMemset(buf, 0, bytes));
    std::vector<unsigned> i64(n), i128(n);
    std::mt19937_64 rng(7);
    for (long long i = 0; i < n; ++i) {
        i64[i] = (unsigned)(rng() % (bytes / 64));    // block of 64 B: 4 float4 from 4 * idx, below 2^28 float4
        i128[i] = (unsigned)(rng() % (bytes / 128));  // block of 128 B: 8 float4 from 8 * idx
    }
    CK(hipMalloc(&d_i64, n * sizeof(unsigned)));
    CK(hipMalloc(&d_i128, n * sizeof(unsigned)));
    CK(hipMalloc(&out, n * sizeof(float)));
    const size_t copy_bytes = (size_t)1 << 30;
    CK(hipMalloc(&dst, copy_bytes));
    CK(hipMemcpy(d_i64, i64.data(), n * sizeof(unsigned), hipMemcpyHostToDevice));
    CK(hipMemcpy(d_i128, i128.data(), n * sizeof(unsigned), hipMemcpyHostToDevice));
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    const unsigned grid = (unsigned)((n + 255) / 256);
    for (int rep = 0; rep < 3; ++rep) {
        float t64 = 0.f, t128 = 0.f, tc = 0.f;
        CK(hipEventRecord(a));
        k_gather64<<<grid, 256>>>(buf, d_i64, n, out);
        CK(hipEventRecord(b));
        CK(hipEventSynchronize(b));
        CK(hipEventElapsedTime(&t64, a, b));
        CK(hipEventRecord(a));
        k_gather128<<<grid, 256>>>(buf, d_i128, n, out);
        CK(hipEventRecord(b));
        CK(hipEventSynchronize(b));
        CK(hipEventElapsedTime(&t128, a, b));
        CK(hipEventRecord(a));
        k_stream_copy<<<4096, 256>>>(buf, dst, (long long)(copy_bytes / 16));
        CK(hipEventRecord(b));
        CK(hipEventSynchronize(b));
        CK(hipEventElapsedTime(&tc, a, b));
        std::printf("rep %d: gather64 %lld x 64 B = %.3f GB read in %.3f ms (%.0f GB/s); gather128 %lld x 128 B = %.3f GB "
                    "in %.3f ms (%.0f GB/s); copy %.3f GB read + written in %.3f ms (%.0f GB/s each way)\n",
                    rep, n, n * 64e-9, t64, n * 64.0 / t64 / 1e6, n, n * 128e-9, t128, n * 128.0 / t128 / 1e6,
                    copy_bytes * 1e-9, tc, copy_bytes / tc / 1e6);
    }
    CK(hipGetLastError());
    CK(hipFree(buf));
    CK(hipFree(dst));
    CK(hipFree(d_i64));
    CK(hipFree(d_i128));
    CK(hipFree(out));
    return 0;
}
