// Shared device geometry of the E-step and the pose-proposal kernels for CDNA4 (gfx950): image and volume
// geometry, projection frames, the corner-packed trilinear gather and complex products.
//
// Image geometry: an s x (s/2+1) half-complex image, pixel p = i*(s/2+1) + x, frequency
// (x, y) with y = i for i <= s/2 else i - s (FOR_ALL_ELEMENTS_IN_FFTW_TRANSFORM, src/fftw.h:99-102).
// Volume geometry: projector layout of src/projector.cpp:59-89, interleaved float2 [z][y][x].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rlnk {

constexpr int kBlock = 256;

// Geometry of one image size, precomputed on the host in double (exact integer tests).
struct ImgGeom {
    int s, xs, P;
    int proj_r2_max;  // pixels projected: x^2+y^2 <= proj_r2_max (Projector::project circle and r_max test)
    int bp_r2_max;    // pixels backprojected: x^2+y^2 <= bp_r2_max (backproject2Dto3D ellipse/r_max test)
};

struct VolGeom {
    int xdim, ydim, zdim, starty, startz;
};

// Per-orientation projection frame: 3-D coordinates of pixel (x, y) are x*e0 + y*e1 (padding included),
// e0 = pad * row 0 of A, e1 = pad * row 1 of A (Ainv = A^T in Projector::project, src/projector.cpp:638-640).
struct Frame {
    float e0x, e0y, e0z, e1x, e1y, e1z;
};

__host__ __device__ __forceinline__ void pixel_xy(int p, int s, int xs, int &x, int &y) {
    const int i = p / xs;
    x = p - i * xs;
    y = (i < xs) ? i : i - s;
}

// Corner-packed reference: for every trilinear base voxel (x0, y0, z0) with x0 < xdim-1, y0 < ydim-1,
// z0 < zdim-1, the 8 corner values are stored contiguously as 4 float4 (64 bytes):
//   [V(z0,y0,x0), V(z0,y0,x0+1)], [V(z0,y0+1,x0), V(z0,y0+1,x0+1)],
//   [V(z0+1,y0,x0), V(z0+1,y0,x0+1)], [V(z0+1,y0+1,x0), V(z0+1,y0+1,x0+1)]
// so that one trilinear sample is one aligned 64-byte read instead of four scattered cache lines.
// Base voxels are grouped in bricks of 4 x 4 x 4 (4 KB of the reference, 8 KB of backprojection slots), bricks
// and the voxels inside them in x-fastest order: the samples of a projection that neighbour each other in any
// direction of the volume then fall into few bricks, which keeps their random 64-byte reads on few DRAM pages.
constexpr int kBrickLog = 2;

__host__ __device__ __forceinline__ int brick_count(int dim) {
    return (dim - 1 + (1 << kBrickLog) - 1) >> kBrickLog;  // bricks covering the dim - 1 base positions
}

__host__ __device__ __forceinline__ size_t packed_index(const VolGeom &g, int x0, int y0, int z0) {
    constexpr int m = (1 << kBrickLog) - 1;
    const size_t brick = ((size_t)(z0 >> kBrickLog) * (size_t)brick_count(g.ydim) + (size_t)(y0 >> kBrickLog)) *
                             (size_t)brick_count(g.xdim) +
                         (size_t)(x0 >> kBrickLog);
    return (brick << (3 * kBrickLog)) + (size_t)((((z0 & m) << kBrickLog) + (y0 & m)) << kBrickLog) + (size_t)(x0 & m);
}

// Base-voxel slots to allocate (whole bricks).
__host__ __device__ __forceinline__ size_t packed_bases(int xdim, int ydim, int zdim) {
    return ((size_t)brick_count(xdim) * (size_t)brick_count(ydim) * (size_t)brick_count(zdim)) << (3 * kBrickLog);
}

// Trilinear gather of the padded reference at (xp, yp, zp), Hermitian flip for xp < 0
// (Projector::project TRILINEAR branch, src/projector.cpp:676-744). Out-of-box samples give 0.
__device__ __forceinline__ float2 sample_volume(const float4 *__restrict__ vol, const VolGeom &g, float xp, float yp,
                                                float zp) {
    const bool neg = xp < 0.f;
    if (neg) {
        xp = -xp;
        yp = -yp;
        zp = -zp;
    }
    const float fx0 = floorf(xp), fy0 = floorf(yp), fz0 = floorf(zp);
    const float fx = xp - fx0, fy = yp - fy0, fz = zp - fz0;
    const int x0 = (int)fx0;
    const int y0 = (int)fy0 - g.starty;
    const int z0 = (int)fz0 - g.startz;
    if (x0 < 0 || x0 + 1 >= g.xdim || y0 < 0 || y0 + 1 >= g.ydim || z0 < 0 || z0 + 1 >= g.zdim)
        return make_float2(0.f, 0.f);
    const float4 *c = vol + 4 * packed_index(g, x0, y0, z0);
    const float4 q0 = c[0], q1 = c[1], q2 = c[2], q3 = c[3];
    float2 r;
    {
        const float dx00 = q0.x + (q0.z - q0.x) * fx, dx01 = q2.x + (q2.z - q2.x) * fx;
        const float dx10 = q1.x + (q1.z - q1.x) * fx, dx11 = q3.x + (q3.z - q3.x) * fx;
        const float dxy0 = dx00 + (dx10 - dx00) * fy, dxy1 = dx01 + (dx11 - dx01) * fy;
        r.x = dxy0 + (dxy1 - dxy0) * fz;
    }
    {
        const float dx00 = q0.y + (q0.w - q0.y) * fx, dx01 = q2.y + (q2.w - q2.y) * fx;
        const float dx10 = q1.y + (q1.w - q1.y) * fx, dx11 = q3.y + (q3.w - q3.y) * fx;
        const float dxy0 = dx00 + (dx10 - dx00) * fy, dxy1 = dx01 + (dx11 - dx01) * fy;
        r.y = dxy0 + (dxy1 - dxy0) * fz;
    }
    if (neg) r.y = -r.y;
    return r;
}

__device__ __forceinline__ float2 project_pixel(const float4 *__restrict__ vol, const VolGeom &g, const Frame &f,
                                                int x, int y) {
    const float fxp = (float)x, fyp = (float)y;
    return sample_volume(vol, g, f.e0x * fxp + f.e1x * fyp, f.e0y * fxp + f.e1y * fyp, f.e0z * fxp + f.e1z * fyp);
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

}  // namespace rlnk
