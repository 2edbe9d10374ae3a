// Device kernels of the learned pose proposal for CDNA4 (gfx950). Included by proposal.hip only.
//
// The network maps a W x W real image (the particle's central Fourier window |k| < W/2, inverse transformed
// and normalised) to logits over direction x psi bins. Six 3x3 convolutions (pad 1, stride 1/2 alternating,
// C channels, ReLU) and one fully connected layer; activations are NHWC BF16, weights BF16 copies of FP32
// master weights, all products on v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
//
// Convolution as a product: col[m][k] (m = (image, oy, ox), k = tap * Cin + c with tap = 3*ky + kx, zero
// padded to Kp) times W[Cout][Kp]^T gives the NHWC output. Backward: dW = dZ^T col (split over m with FP32
// atomics); the data gradient is an implicit GEMM over the taps fused with the previous layer's ReLU mask
// (k_conv_bwd_data), or, for channel counts it does not take, dcol = dZ W and a gather col2im.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.h"

namespace rlnp {

using rlnk::Frame;
using rlnk::VolGeom;
constexpr int kB = 256;

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;  // BF16 bit pattern in memory

__device__ __forceinline__ bf16 to_bf16(float x) { return __builtin_bit_cast(bf16, (__bf16)x); }
__device__ __forceinline__ float from_bf16(bf16 h) { return __uint_as_float((unsigned int)h << 16); }

// counter-based normal deviates (splitmix64 finaliser + Box-Muller)
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float2 normal2(uint64_t key) {
    const uint64_t h = mix64(key);
    const float u1 = ((float)(h >> 40) + 0.5f) * (1.f / 16777216.f);
    const float u2 = ((float)((h >> 16) & 0xffffffULL) + 0.5f) * (1.f / 16777216.f);
    const float r = sqrtf(-2.f * logf(u1));
    float s, c;
    sincosf(6.283185307179586f * u2, &s, &c);
    return make_float2(r * c, r * s);
}

// --------------------------------------------------------------------------------------------------
// Training samples: the W x (W/2+1) Fourier window of a projection of the reference at a known pose, with the
// acquisition's CTF, a shift and Gaussian noise of the model's spectrum (per-component variance
// sigma2[ires], ires = round(|k|) in the original box), phase flipped (multiplied by the sign of the CTF).
// --------------------------------------------------------------------------------------------------
struct GenSample {
    Frame fr;                 // projection frame of the pose (padding included)
    float dx, dy;             // shift in pixels of the original box
    float defU, defV, defAng; // CTF (Angstrom, degrees)
};

struct CtfConst {
    float K1, K2, K3;  // pi*lambda, pi/2*Cs*lambda^3, atan(Q0/sqrt(1-Q0^2)) (src/ctf.h:184-240)
    float inv_box_A;   // 1 / (ori_size * pixel_size)
    float ori_size;
};

__device__ __forceinline__ float ctf_at(const CtfConst &k, float defU, float defV, float defAng, float x, float y) {
    const float fx = x * k.inv_box_A, fy = y * k.inv_box_A;
    float s, c;
    sincosf(defAng * 0.017453292519943295f, &s, &c);
    const float axx = -defU * c * c - defV * s * s;
    const float axy = -defU * c * s + defV * s * c;
    const float ayy = -defU * s * s - defV * c * c;
    const float u2 = fx * fx + fy * fy;
    const float gamma = k.K1 * (axx * fx * fx + 2.f * axy * fx * fy + ayy * fy * fy) + k.K2 * u2 * u2 - k.K3;
    return -sinf(gamma);
}

// grid (ceil(W*(W/2+1)/256), n); out [n][W][W/2+1]
__global__ void __launch_bounds__(kB) k_gen_window(const float4 *__restrict__ vol, VolGeom vg, int proj_r2_max,
                                                   const GenSample *__restrict__ smp, CtfConst ck,
                                                   const float *__restrict__ sigma2, int nshell, int W,
                                                   unsigned long long seed, long long first,
                                                   float2 *__restrict__ out) {
    const int xw = W / 2 + 1;
    const int idx = blockIdx.x * kB + threadIdx.x;
    const int b = blockIdx.y;
    if (idx >= W * xw) return;
    const int r = idx / xw, kx = idx - r * xw;
    const int ky = (r < W / 2) ? r : r - W;
    const GenSample g = smp[b];
    const int r2 = kx * kx + ky * ky;
    float2 v = make_float2(0.f, 0.f);
    if (r2 <= proj_r2_max) v = rlnk::project_pixel(vol, vg, g.fr, kx, ky);
    const float ctf = ctf_at(ck, g.defU, g.defV, g.defAng, (float)kx, (float)ky);
    float sn, cs;
    sincosf(6.283185307179586f * ((float)kx * g.dx + (float)ky * g.dy) / ck.ori_size, &sn, &cs);
    float2 f = rlnk::cmul(v, make_float2(cs, sn));
    f.x *= ctf;
    f.y *= ctf;
    const int ires = (int)rintf(sqrtf((float)r2));
    const float sg = (ires < nshell) ? sqrtf(sigma2[ires]) : 0.f;
    const float2 nz = normal2((seed * 0x100000001b3ULL) ^ ((uint64_t)(first + b) << 20) ^ (uint64_t)idx);
    // phase flipped, as the particle windows (k_window_image with the particle's CTF)
    const float sgn = ctf < 0.f ? -1.f : 1.f;
    out[(size_t)b * W * xw + idx] = make_float2(sgn * (f.x + sg * nz.x), sgn * (f.y + sg * nz.y));
}

// --------------------------------------------------------------------------------------------------
// Network input: img(u, v) = sum_{ky=-W/2}^{W/2-1} sum_{kx=0}^{W/2} h(kx) Re[F(kx,ky) e^{2 pi i (kx u + ky v)/W}],
// h(0) = h(W/2) = 1, else 2, from a half-complex source of s_src rows of xs_src (row of ky: ky >= 0 ? ky :
// ky + s_src), phase flipped by the sign of ctf (same layout, optional), normalised to zero mean and unit
// variance. One block per image; out bf16 [n][W][W] (v-major), optional FP32 copy.
// --------------------------------------------------------------------------------------------------
constexpr int kMaxW = 64;

__global__ void __launch_bounds__(kB) k_window_image(const float2 *__restrict__ src, const float *__restrict__ ctf,
                                                     long long src_stride, int s_src, int xs_src, int W,
                                                     bf16 *__restrict__ out, float *__restrict__ out_f32) {
    __shared__ float2 Fw[kMaxW][kMaxW / 2 + 1];
    __shared__ float2 G[kMaxW / 2 + 1][kMaxW];
    __shared__ float2 tw[kMaxW];
    __shared__ float img[kMaxW * kMaxW];
    __shared__ double red[2][kB / 64];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int xw = W / 2 + 1;
    const float2 *F = src + (size_t)b * src_stride;
    for (int j = tid; j < W; j += kB) {
        float s, c;
        sincosf(6.283185307179586f * (float)j / (float)W, &s, &c);
        tw[j] = make_float2(c, s);
    }
    for (int idx = tid; idx < W * xw; idx += kB) {
        const int r = idx / xw, kx = idx - r * xw;
        const int ky = (r < W / 2) ? r : r - W;
        const int row = (ky >= 0) ? ky : ky + s_src;
        float2 f = F[(size_t)row * xs_src + kx];
        if (ctf && ctf[(size_t)b * src_stride + (size_t)row * xs_src + kx] < 0.f) f = make_float2(-f.x, -f.y);
        Fw[r][kx] = f;
    }
    __syncthreads();
    // G(kx, v) = sum_ky F(kx, ky) e^{2 pi i ky v / W}
    for (int idx = tid; idx < xw * W; idx += kB) {
        const int kx = idx / W, v = idx - kx * W;
        float2 acc = make_float2(0.f, 0.f);
        // twiddle index (ky * v) mod W advances by v per row, also across the wrap to negative ky (-W v = 0 mod W)
        int ti = 0;
        for (int r = 0; r < W; ++r) {
            const float2 t = tw[ti];
            const float2 f = Fw[r][kx];
            acc.x += f.x * t.x - f.y * t.y;
            acc.y += f.x * t.y + f.y * t.x;
            ti += v;
            if (ti >= W) ti -= W;
        }
        G[kx][v] = acc;
    }
    __syncthreads();
    double s1 = 0., s2 = 0.;
    for (int idx = tid; idx < W * W; idx += kB) {
        const int v = idx / W, u = idx - v * W;
        float acc = 0.f;
        int ti = 0;  // (kx * u) mod W
        for (int kx = 0; kx < xw; ++kx) {
            const float h = (kx == 0 || kx == W / 2) ? 1.f : 2.f;
            const float2 t = tw[ti];
            const float2 g = G[kx][v];
            acc += h * (g.x * t.x - g.y * t.y);
            ti += u;
            if (ti >= W) ti -= W;
        }
        img[idx] = acc;
        s1 += acc;
        s2 += (double)acc * acc;
    }
    for (int o = 32; o > 0; o >>= 1) {
        s1 += __shfl_xor(s1, o);
        s2 += __shfl_xor(s2, o);
    }
    if ((tid & 63) == 0) {
        red[0][tid >> 6] = s1;
        red[1][tid >> 6] = s2;
    }
    __syncthreads();
    double t1 = 0., t2 = 0.;
    for (int w = 0; w < kB / 64; ++w) {
        t1 += red[0][w];
        t2 += red[1][w];
    }
    const double mean = t1 / (double)(W * W);
    const double var = t2 / (double)(W * W) - mean * mean;
    const float inv = var > 0. ? (float)(1. / sqrt(var)) : 0.f;
    const float mf = (float)mean;
    for (int idx = tid; idx < W * W; idx += kB) {
        const float y = (img[idx] - mf) * inv;
        out[(size_t)b * W * W + idx] = to_bf16(y);
        if (out_f32) out_f32[(size_t)b * W * W + idx] = y;
    }
}

// --------------------------------------------------------------------------------------------------
// im2col for a 3x3 convolution, pad 1: X [n][H][H][C] bf16 -> col [n*Ho*Ho][Kp], k = tap*C + c (k >= 9C: 0).
// C % 8 == 0: one thread per (m, tap, 8 channels); C == 1: one thread per (m, k).
// --------------------------------------------------------------------------------------------------
// Tiles of 64 pixels x 64 channels transposed through LDS: k_im2col_t64 writes the im2col columns of a 64-channel
// layer k-major (colT[k][m], k = tap * 64 + c, m the output pixel) and k_transpose64 a [M][64] gradient as [64][M],
// so that the weight gradient dW = dZ^T col reads both operands with the pixel index contiguous. Block: 64
// consecutive pixels (and one tap, blockIdx.y, for the columns); M % 64 == 0 is not required.
__device__ __forceinline__ void store_tile_t64(const bf16 (*t)[72], long long m0, long long M, bf16 *__restrict__ out,
                                               long long row0) {
    // thread q: channel q >> 3, pixels 8 (q & 7) .. + 7 of the tile -> out[(row0 + c) * M + m0 + ...]
    for (int q = threadIdx.x; q < 512; q += kB) {
        const int c = q >> 3, p8 = (q & 7) * 8;
        bf16 e[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) e[j] = t[p8 + j][c];
        bf16 *dst = out + (size_t)(row0 + c) * M + m0 + p8;
        if (m0 + p8 + 8 <= M) {
            *reinterpret_cast<uint4 *>(dst) =
                make_uint4((unsigned)e[0] | ((unsigned)e[1] << 16), (unsigned)e[2] | ((unsigned)e[3] << 16),
                           (unsigned)e[4] | ((unsigned)e[5] << 16), (unsigned)e[6] | ((unsigned)e[7] << 16));
        } else {
            for (int j = 0; j < 8 && m0 + p8 + j < M; ++j) dst[j] = e[j];
        }
    }
}

__global__ void __launch_bounds__(kB) k_im2col_t64(const bf16 *__restrict__ X, int n, int H, int stride, int Ho,
                                                   bf16 *__restrict__ colT) {
    __shared__ __attribute__((aligned(16))) bf16 t[64][72];
    const long long M = (long long)n * Ho * Ho;
    const long long m0 = (long long)blockIdx.x * 64;
    const int tap = blockIdx.y, ky = tap / 3, kx = tap - 3 * (tap / 3);
    for (int q = threadIdx.x; q < 512; q += kB) {
        const int pm = q >> 3, c8 = q & 7;
        const long long m = m0 + pm;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (m < M) {
            const int img = (int)(m / (Ho * Ho)), pix = (int)(m - (long long)img * Ho * Ho);
            const int oy = pix / Ho, ox = pix - oy * Ho;
            const int iy = oy * stride + ky - 1, ix = ox * stride + kx - 1;
            if (iy >= 0 && iy < H && ix >= 0 && ix < H)
                v = *reinterpret_cast<const uint4 *>(X + (((size_t)img * H + iy) * H + ix) * 64 + 8 * c8);
        }
        *reinterpret_cast<uint4 *>(&t[pm][8 * c8]) = v;
    }
    __syncthreads();
    store_tile_t64(t, m0, M, colT, (long long)tap * 64);
}

__global__ void __launch_bounds__(kB) k_transpose64(const bf16 *__restrict__ A, long long M, bf16 *__restrict__ AT) {
    __shared__ __attribute__((aligned(16))) bf16 t[64][72];
    const long long m0 = (long long)blockIdx.x * 64;
    for (int q = threadIdx.x; q < 512; q += kB) {
        const int pm = q >> 3, c8 = q & 7;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (m0 + pm < M) v = *reinterpret_cast<const uint4 *>(A + (size_t)(m0 + pm) * 64 + 8 * c8);
        *reinterpret_cast<uint4 *>(&t[pm][8 * c8]) = v;
    }
    __syncthreads();
    store_tile_t64(t, m0, M, AT, 0);
}

__global__ void __launch_bounds__(kB) k_im2col(const bf16 *__restrict__ X, int n, int H, int C, int stride, int Ho,
                                               int Kp, bf16 *__restrict__ col) {
    const long long idx = (long long)blockIdx.x * kB + threadIdx.x;
    const long long M = (long long)n * Ho * Ho;
    if (C % 8 == 0) {
        const int c8 = C / 8;
        if (idx >= M * 9 * c8) return;
        const long long m = idx / (9 * c8);
        const int rem = (int)(idx - m * 9 * c8), tap = rem / c8, cg = rem - tap * c8;
        const int img = (int)(m / (Ho * Ho)), pix = (int)(m - (long long)img * Ho * Ho);
        const int oy = pix / Ho, ox = pix - oy * Ho;
        const int iy = oy * stride + tap / 3 - 1, ix = ox * stride + tap % 3 - 1;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (iy >= 0 && iy < H && ix >= 0 && ix < H)
            v = *reinterpret_cast<const uint4 *>(X + (((size_t)img * H + iy) * H + ix) * C + cg * 8);
        *reinterpret_cast<uint4 *>(col + (size_t)m * Kp + tap * C + cg * 8) = v;
    } else {
        if (idx >= M * Kp) return;
        const long long m = idx / Kp;
        const int k = (int)(idx - m * Kp);
        bf16 v = 0;
        if (k < 9 * C) {
            const int tap = k / C, c = k - tap * C;
            const int img = (int)(m / (Ho * Ho)), pix = (int)(m - (long long)img * Ho * Ho);
            const int oy = pix / Ho, ox = pix - oy * Ho;
            const int iy = oy * stride + tap / 3 - 1, ix = ox * stride + tap % 3 - 1;
            if (iy >= 0 && iy < H && ix >= 0 && ix < H) v = X[(((size_t)img * H + iy) * H + ix) * C + c];
        }
        col[idx] = v;
    }
}

// dX[img][y][x][c] = sum over taps of dcol[(img, oy, ox)][tap*C + c] with oy*stride + ky - 1 = y (and x), times
// the ReLU mask of the layer below (Y > 0), as bf16. One thread per element of dX.
__global__ void __launch_bounds__(kB) k_col2im_mask(const float *__restrict__ dcol, int n, int H, int C, int stride,
                                                    int Ho, int Kp, const bf16 *__restrict__ Y,
                                                    bf16 *__restrict__ dZ) {
    const long long idx = (long long)blockIdx.x * kB + threadIdx.x;
    const long long total = (long long)n * H * H * C;
    if (idx >= total) return;
    const int c = (int)(idx % C);
    const long long pix = idx / C;
    const int x = (int)(pix % H), y = (int)((pix / H) % H), img = (int)(pix / ((long long)H * H));
    float acc = 0.f;
    for (int ky = 0; ky < 3; ++ky) {
        const int ty = y + 1 - ky;
        if (ty < 0 || ty % stride) continue;
        const int oy = ty / stride;
        if (oy >= Ho) continue;
        for (int kx = 0; kx < 3; ++kx) {
            const int tx = x + 1 - kx;
            if (tx < 0 || tx % stride) continue;
            const int ox = tx / stride;
            if (ox >= Ho) continue;
            acc += dcol[(((size_t)img * Ho + oy) * Ho + ox) * Kp + (3 * ky + kx) * C + c];
        }
    }
    dZ[idx] = to_bf16(from_bf16(Y[idx]) > 0.f ? acc : 0.f);
}

// dZ = bf16(d * (Y > 0)) elementwise
__global__ void __launch_bounds__(kB) k_mask_relu(const float *__restrict__ d, const bf16 *__restrict__ Y, long long n,
                                                  bf16 *__restrict__ dZ) {
    const long long i = (long long)blockIdx.x * kB + threadIdx.x;
    if (i < n) dZ[i] = to_bf16(from_bf16(Y[i]) > 0.f ? d[i] : 0.f);
}

// out[j] += sum_i Z[i][j] over rows [blockIdx.y*rows, +rows): bias gradients. grid (ceil(N/256), ceil(M/rows)).
__global__ void __launch_bounds__(kB) k_col_sum(const bf16 *__restrict__ Z, int M, int N, int rows,
                                                float *__restrict__ out) {
    const int j = blockIdx.x * kB + threadIdx.x;
    if (j >= N) return;
    const int r0 = blockIdx.y * rows, r1 = min(M, r0 + rows);
    float s = 0.f;
    for (int i = r0; i < r1; ++i) s += from_bf16(Z[(size_t)i * N + j]);
    atomicAdd(&out[j], s);
}

// out[j] += sum_i Z[i][j] over rows [blockIdx.y*rows, +rows), for N % 8 == 0 and N <= 8 * kB: N/8 threads per row
// read 16 bytes each (a row per N/8 lanes, kB/(N/8) rows per pass), FP32 partials per thread, then a block sum per
// column through LDS and one atomic per column and block. grid (1, ceil(M/rows)).
__global__ void __launch_bounds__(kB) k_col_sum8(const bf16 *__restrict__ Z, int M, int N, int rows,
                                                 float *__restrict__ out) {
    __shared__ float part[8 * kB];
    const int tpr = N >> 3, rpp = kB / tpr;  // threads per row, rows per pass
    const int tid = threadIdx.x, rr = tid / tpr, cg = tid - rr * tpr;
    float s[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) s[q] = 0.f;
    const int r0 = blockIdx.y * rows, r1 = min(M, r0 + rows);
    if (rr < rpp) {
        for (int i = r0 + rr; i < r1; i += rpp) {
            const uint4 v = *reinterpret_cast<const uint4 *>(Z + (size_t)i * N + 8 * cg);
            const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                s[2 * q] += __uint_as_float(w[q] << 16);
                s[2 * q + 1] += __uint_as_float(w[q] & 0xffff0000u);
            }
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) part[rr * N + 8 * cg + q] = s[q];
    }
    __syncthreads();
    for (int j = tid; j < N; j += kB) {
        float t = 0.f;
        for (int q = 0; q < rpp; ++q) t += part[q * N + j];
        atomicAdd(&out[j], t);
    }
}

// Weights of a layer rearranged for the data gradient: Wd[c][tap*Cout + co] = W[co][tap*Cin + c].
__global__ void __launch_bounds__(kB) k_wt_bwd(const bf16 *__restrict__ W, int Cin, int Cout, int Kp,
                                               bf16 *__restrict__ Wd) {
    const int idx = blockIdx.x * kB + threadIdx.x;
    const int Kd = 9 * Cout;
    if (idx >= Cin * Kd) return;
    const int c = idx / Kd, k = idx - c * Kd, tap = k / Cout, co = k - tap * Cout;
    Wd[idx] = W[(size_t)co * Kp + tap * Cin + c];
}

// Data gradient of a 3x3 convolution (pad 1), fused with the ReLU mask of the layer below:
// dX[img][y][x][c] = (Y > 0) * sum over taps (ky, kx) and co of dZ[img][oy][ox][co] W[co][tap*Cin + c], with
// oy = (y + 1 - ky) / stride where that divides and lies inside (likewise ox), as bf16. An implicit GEMM: rows
// (img, y, x), k = tap*Cout + co, columns c; each lane's A fragment (8 channels of one tap) is read straight from
// the NHWC dZ, the B fragments from Wd (k_wt_bwd). Block = 4 waves x 16 rows, NT column tiles (Cin = 16 NT),
// Cout % 8 == 0. For stride 2 (H % 16 == 0, Cout % 32 == 0) the rows run by parity class of (y, x): row
// ((img * 4 + 2 py + px) * H/2 + y') * H/2 + x' is pixel (2 y' + py, 2 x' + px), so that a block's 64 rows share
// one class, and only the taps that reach an output for that class are visited (ky = 1 for even y, 0 and 2 for odd
// y; likewise kx): 1, 2 or 4 of the 9 taps.
template <int NT>
__global__ void __launch_bounds__(kB) k_conv_bwd_data(const bf16 *__restrict__ dZ, int n, int H, int Cout,
                                                      int stride, int Ho, const bf16 *__restrict__ Wd,
                                                      const bf16 *__restrict__ Y, bf16 *__restrict__ dX) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, gq = lane >> 4;
    const long long HH = (long long)H * H;
    const long long M = (long long)n * HH;
    const bool par = stride == 2 && H % 16 == 0 && Cout % 32 == 0;
    // pixel (img, y, x) of row mm
    auto decode = [&](long long mm, int &im, int &yy, int &xx) {
        im = (int)(mm / HH);
        const int pix = (int)(mm - (long long)im * HH);
        if (par) {
            const int h = H >> 1, q = h * h;
            const int c = pix / q, r = pix - c * q, ry = r / h;
            yy = 2 * ry + (c >> 1);
            xx = 2 * (r - ry * h) + (c & 1);
        } else {
            yy = pix / H;
            xx = pix - yy * H;
        }
    };
    const long long m = ((long long)blockIdx.x * 4 + wave) * 16 + r16;  // this lane's A row: input pixel
    int img = 0, y = 0, x = 0;
    const bool mok = m < M;
    if (mok) decode(m, img, y, x);
    const int Kd = 9 * Cout;
    floatx4_t acc[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[j] = floatx4_t{0.f, 0.f, 0.f, 0.f};
    auto kstep = [&](int k) {  // this lane's 8 k values from k
        uint4 av = make_uint4(0u, 0u, 0u, 0u);
        if (mok && k < Kd) {
            const int tap = k / Cout, co = k - tap * Cout;
            const int ty = y + 1 - tap / 3, tx = x + 1 - tap % 3;
            if (ty >= 0 && tx >= 0 && ty % stride == 0 && tx % stride == 0) {
                const int oy = ty / stride, ox = tx / stride;
                if (oy < Ho && ox < Ho)
                    av = *reinterpret_cast<const uint4 *>(dZ + (((size_t)img * Ho + oy) * Ho + ox) * Cout + co);
            }
        }
        const bf16x8_t a = __builtin_bit_cast(bf16x8_t, av);
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const int col = 16 * j + r16;  // B column: input channel
            uint4 bv = make_uint4(0u, 0u, 0u, 0u);
            if (k < Kd) bv = *reinterpret_cast<const uint4 *>(Wd + (size_t)col * Kd + k);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(bf16x8_t, bv), acc[j], 0, 0, 0);
        }
    };
    if (par) {
        // the block's class (H^2 / 4 is a multiple of 64, so its rows share one)
        const int cls = (int)((((long long)blockIdx.x * 64) % HH) / (HH / 4));
        const int py = cls >> 1, px = cls & 1;
        for (int ky = 1 - py; ky < 3; ky += 2)
            for (int kx = 1 - px; kx < 3; kx += 2)
                for (int co0 = 0; co0 < Cout; co0 += 32) kstep((3 * ky + kx) * Cout + co0 + 8 * gq);
    } else {
        for (int k0 = 0; k0 < Kd; k0 += 32) kstep(k0 + 8 * gq);
    }
    const int Cin = 16 * NT;
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int col = 16 * j + r16;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const long long mm = ((long long)blockIdx.x * 4 + wave) * 16 + 4 * gq + r;
            if (mm < M) {
                int im, yy, xx;
                decode(mm, im, yy, xx);
                const size_t idx = (((size_t)im * H + yy) * H + xx) * Cin + col;
                dX[idx] = to_bf16(from_bf16(Y[idx]) > 0.f ? acc[j][r] : 0.f);
            }
        }
    }
}

// --------------------------------------------------------------------------------------------------
// GEMM on MFMA: C(m, n) = sum_k A(m, k) B(k, n), A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn], bf16.
// AK: A is k-contiguous (sak == 1, K % 8 == 0) else m-contiguous (sam == 1, M % 8 == 0); BK likewise with n.
// Tile 64 x 64 per block of 4 waves (each 32 x 32: 2 x 2 MFMA tiles), k in LDS stages of 64, the next stage's
// operands loaded into registers while the MFMAs consume the current one; blockIdx.z splits k into chunks of
// kchunk (a multiple of 32). Epilogues: kEpiF32 stores (split k: chunk z's partial sums at Cf + z * zstride, bias
// from chunk 0 only), kEpiAtomicF32 adds atomically (split k), kEpiAddF32 adds without atomics (one k chunk: every
// element has one writer), kEpiBiasReluBf16 stores bias + ReLU as bf16.
// --------------------------------------------------------------------------------------------------
enum { kEpiF32 = 0, kEpiAtomicF32 = 1, kEpiBiasReluBf16 = 2, kEpiAddF32 = 3 };

struct GemmArgs {
    const bf16 *A, *B;
    long long sam, sak, sbk, sbn;
    int M, N, K, kchunk;
    const float *bias;   // [N] or null
    float *Cf;           // kEpiF32 / kEpiAtomicF32 / kEpiAddF32
    bf16 *Cb;            // kEpiBiasReluBf16
    long long ldc;
    long long zstride;   // kEpiF32: offset of chunk z's partial sums (0 when k is not split)
};

template <bool AK, bool BK, int EPI>
__global__ void __launch_bounds__(kB) k_gemm(GemmArgs g) {
    constexpr int kKs = 64;          // k per LDS stage
    constexpr int kPitch = kKs + 8;  // bf16 per LDS row: 144-byte rows keep 16-byte alignment
    __shared__ __attribute__((aligned(16))) bf16 As[64 * kPitch];
    __shared__ __attribute__((aligned(16))) bf16 Bs[64 * kPitch];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int r16 = lane & 15, gq = lane >> 4;
    const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
    const int kb = blockIdx.z * g.kchunk, ke = min(g.K, kb + g.kchunk);
    floatx4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4_t{0.f, 0.f, 0.f, 0.f};
    // this thread's 16-byte pieces of a stage: k-contiguous operands as row tid / 4, k 8 (tid % 4) + 32 h;
    // m- or n-contiguous ones as k row tid / 8 + 32 h, columns 8 (tid % 8) .. + 7. In LDS, the 8-element chunk c of
    // row r sits at chunk c ^ ((r / 16) % 4): the transposing stores of rows 8 apart then fall into distinct banks,
    // and the 16 rows of an MFMA fragment read share one swizzle.
    auto sw = [](int row, int k) { return row * kPitch + ((((k >> 3) ^ (row >> 4)) & 7) << 3) + (k & 7); };
    uint4 ra[2], rb[2];
    auto gload = [&](int k0) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (AK) {
                const int m = tid >> 2, kk = (tid & 3) * 8 + 32 * h;
                if (m0 + m < g.M && k0 + kk < ke)
                    v = *reinterpret_cast<const uint4 *>(g.A + (size_t)(m0 + m) * g.sam + (k0 + kk));
            } else {
                const int k = (tid >> 3) + 32 * h, mm = (tid & 7) * 8;
                if (k0 + k < ke && m0 + mm < g.M)
                    v = *reinterpret_cast<const uint4 *>(g.A + (size_t)(k0 + k) * g.sak + (m0 + mm));
            }
            ra[h] = v;
            v = make_uint4(0u, 0u, 0u, 0u);
            if (BK) {
                const int n = tid >> 2, kk = (tid & 3) * 8 + 32 * h;
                if (n0 + n < g.N && k0 + kk < ke)
                    v = *reinterpret_cast<const uint4 *>(g.B + (size_t)(n0 + n) * g.sbn + (k0 + kk));
            } else {
                const int k = (tid >> 3) + 32 * h, nn = (tid & 7) * 8;
                if (k0 + k < ke && n0 + nn < g.N)
                    v = *reinterpret_cast<const uint4 *>(g.B + (size_t)(k0 + k) * g.sbk + (n0 + nn));
            }
            rb[h] = v;
        }
    };
    auto lstore = [&]() {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (AK) {
                *reinterpret_cast<uint4 *>(&As[sw(tid >> 2, (tid & 3) * 8 + 32 * h)]) = ra[h];
            } else {
                const int k = (tid >> 3) + 32 * h, mm = (tid & 7) * 8;
                const bf16 *e = reinterpret_cast<const bf16 *>(&ra[h]);
#pragma unroll
                for (int j = 0; j < 8; ++j) As[sw(mm + j, k)] = e[j];
            }
            if (BK) {
                *reinterpret_cast<uint4 *>(&Bs[sw(tid >> 2, (tid & 3) * 8 + 32 * h)]) = rb[h];
            } else {
                const int k = (tid >> 3) + 32 * h, nn = (tid & 7) * 8;
                const bf16 *e = reinterpret_cast<const bf16 *>(&rb[h]);
#pragma unroll
                for (int j = 0; j < 8; ++j) Bs[sw(nn + j, k)] = e[j];
            }
        }
    };
    if (kb < ke) gload(kb);
    for (int k0 = kb; k0 < ke; k0 += kKs) {
        lstore();
        __syncthreads();
        if (k0 + kKs < ke) gload(k0 + kKs);
#pragma unroll
        for (int ks = 0; ks < kKs; ks += 32) {
            bf16x8_t a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *reinterpret_cast<const bf16x8_t *>(&As[sw(32 * wm + 16 * i + r16, ks + 8 * gq)]);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                b[j] = *reinterpret_cast<const bf16x8_t *>(&Bs[sw(32 * wn + 16 * j + r16, ks + 8 * gq)]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    float *Cz = g.Cf;
    if (EPI == kEpiF32) Cz += (size_t)blockIdx.z * g.zstride;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int n = n0 + 32 * wn + 16 * j + r16;
            if (n >= g.N) continue;
            const float bias = (g.bias && blockIdx.z == 0) ? g.bias[n] : 0.f;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + 32 * wm + 16 * i + 4 * gq + r;
                if (m >= g.M) continue;
                const float v = acc[i][j][r] + bias;
                if (EPI == kEpiF32) Cz[(size_t)m * g.ldc + n] = v;
                else if (EPI == kEpiAtomicF32) atomicAdd(&g.Cf[(size_t)m * g.ldc + n], v);
                else if (EPI == kEpiAddF32) g.Cf[(size_t)m * g.ldc + n] += v;
                else g.Cb[(size_t)m * g.ldc + n] = to_bf16(fmaxf(v, 0.f));
            }
        }
}

// C[m][n] = sum over z < nz of P[z][m][n] (partial sums of a k-split kEpiF32 GEMM, z stride M * ldc), in z order.
__global__ void __launch_bounds__(kB) k_sum_parts(const float *__restrict__ P, int nz, int M, int N, long long ldc,
                                                  float *__restrict__ C) {
    const long long i = (long long)blockIdx.x * kB + threadIdx.x;
    if (i >= (long long)M * N) return;
    const int m = (int)(i / N), n = (int)(i - (long long)m * N);
    const size_t o = (size_t)m * ldc + n, zs = (size_t)M * ldc;
    float s = P[o];
    for (int z = 1; z < nz; ++z) s += P[o + z * zs];
    C[o] = s;
}

// --------------------------------------------------------------------------------------------------
// Forward 3x3 convolution as an implicit GEMM (inference): the A operand of each MFMA (8 consecutive k =
// 8 channels of one tap of one output pixel) is read straight from the NHWC input, so no im2col buffer is
// written or read. Block = 4 waves x 16 output pixels x all Cout (Cout / 16 MFMA column tiles, <= 4).
// Cin % 8 == 0, or Cin == 1 (then k = tap, 9 taps padded to Kp = 16 in the weights; with col, the lanes also store
// each pixel's Kp taps there: the im2col columns of the layer's weight gradient).
// --------------------------------------------------------------------------------------------------
template <int NT>  // column tiles of 16 output channels
__global__ void __launch_bounds__(kB) k_conv_fwd(const bf16 *__restrict__ X, int n, int H, int Cin, int stride, int Ho,
                                                 int Kp, const bf16 *__restrict__ Wt, const float *__restrict__ bias,
                                                 bf16 *__restrict__ Y, bf16 *__restrict__ col = nullptr) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, gq = lane >> 4;
    const long long M = (long long)n * Ho * Ho;
    const long long m = ((long long)blockIdx.x * 4 + wave) * 16 + r16;  // this lane's output pixel (A row)
    int img = 0, oy = 0, ox = 0;
    const bool mok = m < M;
    if (mok) {
        img = (int)(m / (Ho * Ho));
        const int pix = (int)(m - (long long)img * Ho * Ho);
        oy = pix / Ho;
        ox = pix - oy * Ho;
    }
    floatx4_t acc[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[j] = floatx4_t{0.f, 0.f, 0.f, 0.f};
    const int Kloop = (Cin == 1) ? 32 : Kp;  // Cin == 1: one K step of 32, k >= 9 are zeros
    for (int k0 = 0; k0 < Kloop; k0 += 32) {
        const int k = k0 + 8 * gq;
        uint4 av = make_uint4(0u, 0u, 0u, 0u);
        if (Cin == 1) {
            bf16 e[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int tap = k + q;
                bf16 v = 0;
                if (mok && tap < 9) {
                    const int iy = oy * stride + tap / 3 - 1, ix = ox * stride + tap % 3 - 1;
                    if (iy >= 0 && iy < H && ix >= 0 && ix < H) v = X[((size_t)img * H + iy) * H + ix];
                }
                e[q] = v;
            }
            av = make_uint4((unsigned)e[0] | ((unsigned)e[1] << 16), (unsigned)e[2] | ((unsigned)e[3] << 16),
                            (unsigned)e[4] | ((unsigned)e[5] << 16), (unsigned)e[6] | ((unsigned)e[7] << 16));
            if (col && mok && k < Kp) *reinterpret_cast<uint4 *>(col + (size_t)m * Kp + k) = av;  // im2col row
        } else if (mok) {
            const int tap = k / Cin, c = k - tap * Cin;
            const int iy = oy * stride + tap / 3 - 1, ix = ox * stride + tap % 3 - 1;
            if (iy >= 0 && iy < H && ix >= 0 && ix < H)
                av = *reinterpret_cast<const uint4 *>(X + (((size_t)img * H + iy) * H + ix) * Cin + c);
        }
        const bf16x8_t a = __builtin_bit_cast(bf16x8_t, av);
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const int col = 16 * j + r16;  // B column: output channel
            uint4 bv = make_uint4(0u, 0u, 0u, 0u);
            if (k < Kp) bv = *reinterpret_cast<const uint4 *>(Wt + (size_t)col * Kp + k);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(bf16x8_t, bv), acc[j], 0, 0, 0);
        }
    }
    // D rows 4*gq + r of the wave's 16 pixels, column r16 of each tile: bias and ReLU into LDS, then the wave's 16
    // consecutive NHWC pixels (16 x Cout contiguous bf16) leave in 16-byte stores
    constexpr int Cout = 16 * NT;
    __shared__ __align__(16) bf16 s_out[4][16][Cout + 8];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int col = 16 * j + r16;
        const float bj = bias[col];
#pragma unroll
        for (int r = 0; r < 4; ++r) s_out[wave][4 * gq + r][col] = to_bf16(fmaxf(acc[j][r] + bj, 0.f));
    }
    __syncthreads();
    const long long mb = ((long long)blockIdx.x * 4 + wave) * 16;
    for (int u = lane; u < 2 * Cout; u += 64) {
        const int px = u / (Cout / 8), c8 = u - px * (Cout / 8);
        if (mb + px < M)
            *reinterpret_cast<uint4 *>(Y + (size_t)(mb + px) * Cout + 8 * c8) =
                *reinterpret_cast<const uint4 *>(&s_out[wave][px][8 * c8]);
    }
}

// --------------------------------------------------------------------------------------------------
// Forward 3x3 convolution as an implicit GEMM with the weights resident in LDS (inference, Cin % 8 == 0).
// The block loads the whole weight matrix [Cout][Kp] into LDS once (rows padded by 8 elements against bank
// conflicts) and then walks output-pixel tiles of 256 (8 waves x 2 row tiles of 16) with a grid stride. Per
// K step of 32 a wave reads 2 A fragments (8 channels of one tap per lane) from the NHWC input and NT B
// fragments from LDS and issues 2 x NT MFMAs; the A fragments of the next step are loaded before the MFMAs of
// the current one. Eight waves share one copy of the weights, so that the two blocks that fit a CU's LDS keep
// 4 waves per SIMD in flight against the latency of the A loads.
// --------------------------------------------------------------------------------------------------
constexpr int kConvMT = 2;                      // row tiles (16 output pixels each) per wave
constexpr int kConvWaves = 8;                   // waves per block
constexpr int kConvThreads = 64 * kConvWaves;
constexpr int kConvTile = 16 * kConvMT * kConvWaves;  // output pixels per block tile
constexpr int kConvLdsPad = 8;                  // bf16 elements of padding per weight row in LDS

__device__ __forceinline__ uint4 conv_a_frag(const bf16 *__restrict__ X, int H, int Cin, int stride, int Ho,
                                             long long M, long long m, int k) {
    uint4 av = make_uint4(0u, 0u, 0u, 0u);
    if (m < M) {
        const int img = (int)(m / (Ho * Ho));
        const int pix = (int)(m - (long long)img * Ho * Ho);
        const int oy = pix / Ho, ox = pix - (pix / Ho) * Ho;
        const int tap = k / Cin, c = k - tap * Cin;
        const int iy = oy * stride + tap / 3 - 1, ix = ox * stride + tap % 3 - 1;
        if (iy >= 0 && iy < H && ix >= 0 && ix < H)
            av = *reinterpret_cast<const uint4 *>(X + (((size_t)img * H + iy) * H + ix) * Cin + c);
    }
    return av;
}

template <int NT>  // column tiles of 16 output channels
__global__ void __launch_bounds__(kConvThreads) k_conv_fwd_lds(const bf16 *__restrict__ X, int n, int H, int Cin, int stride,
                                                     int Ho, int Kp, const bf16 *__restrict__ Wt,
                                                     const float *__restrict__ bias, bf16 *__restrict__ Y) {
    extern __shared__ uint4 s_w4[];  // [16 * NT][Kp + kConvLdsPad] bf16
    bf16 *s_w = reinterpret_cast<bf16 *>(s_w4);
    const int ldw = Kp + kConvLdsPad;
    const int Cout = 16 * NT;
    // weights into LDS, 16 bytes at a time (Kp % 8 == 0)
    for (int q = threadIdx.x; q < Cout * (Kp / 8); q += kConvThreads) {
        const int row = q / (Kp / 8), c8 = q - row * (Kp / 8);
        *reinterpret_cast<uint4 *>(s_w + (size_t)row * ldw + 8 * c8) =
            *reinterpret_cast<const uint4 *>(Wt + (size_t)row * Kp + 8 * c8);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, gq = lane >> 4;
    const long long M = (long long)n * Ho * Ho;
    const long long n_tiles = (M + kConvTile - 1) / kConvTile;
    for (long long tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const long long mbase = (tile * kConvWaves + wave) * (16 * kConvMT);  // this wave's first output pixel
        floatx4_t acc[kConvMT][NT];
#pragma unroll
        for (int i = 0; i < kConvMT; ++i)
#pragma unroll
            for (int j = 0; j < NT; ++j) acc[i][j] = floatx4_t{0.f, 0.f, 0.f, 0.f};
        uint4 a_cur[kConvMT];
#pragma unroll
        for (int i = 0; i < kConvMT; ++i) a_cur[i] = conv_a_frag(X, H, Cin, stride, Ho, M, mbase + 16 * i + r16, 8 * gq);
        for (int k0 = 0; k0 < Kp; k0 += 32) {
            uint4 a_nxt[kConvMT];
            const bool more = k0 + 32 < Kp;
#pragma unroll
            for (int i = 0; i < kConvMT; ++i)
                a_nxt[i] = more ? conv_a_frag(X, H, Cin, stride, Ho, M, mbase + 16 * i + r16, k0 + 32 + 8 * gq)
                                : make_uint4(0u, 0u, 0u, 0u);
            bf16x8_t b[NT];
#pragma unroll
            for (int j = 0; j < NT; ++j)
                b[j] = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const uint4 *>(
                                                        s_w + (size_t)(16 * j + r16) * ldw + k0 + 8 * gq));
#pragma unroll
            for (int i = 0; i < kConvMT; ++i) {
                const bf16x8_t a = __builtin_bit_cast(bf16x8_t, a_cur[i]);
#pragma unroll
                for (int j = 0; j < NT; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[i][j], 0, 0, 0);
            }
#pragma unroll
            for (int i = 0; i < kConvMT; ++i) a_cur[i] = a_nxt[i];
        }
        // D rows 4*gq + r of each 16-pixel tile, column r16 of each channel tile: bias, ReLU, bf16 NHWC
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const int col = 16 * j + r16;
            const float bj = bias[col];
#pragma unroll
            for (int i = 0; i < kConvMT; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const long long mm = mbase + 16 * i + 4 * gq + r;
                    if (mm < M) Y[(size_t)mm * Cout + col] = to_bf16(fmaxf(acc[i][j][r] + bj, 0.f));
                }
        }
    }
}

// --------------------------------------------------------------------------------------------------
// Forward 3x3 convolution (inference) of the 64-channel layers (Cin = Cout = 64, K = 576) with both operands in
// LDS. The block keeps the whole weight matrix resident and walks tiles of output pixels with a grid stride; per
// tile it stages the input rows the tile reads (a patch: 64 channels per pixel, padded to 72 against bank
// conflicts; rows outside the image as zeros), so each input element comes from HBM or L2 once per tile instead of
// once per tap. A tile is 256 / S consecutive output pixels (S the stride): TH whole output rows of one image, or
// whole images when an image has fewer pixels. While the MFMAs of a tile run, the next tile's patch is loaded into
// registers; it goes to LDS after the barrier that ends the tile. Wave w owns two row tiles of 16 pixels and 4 / S
// column tiles of 16 channels; per K step of 32 (half the channels of one tap) it reads its A fragments from the
// patch and its B fragments from the weights and issues 8 MFMAs. The fragments and the k order are those of
// k_conv_fwd_lds, so both kernels round alike.
// --------------------------------------------------------------------------------------------------
constexpr int kPatchWaves = 8;
constexpr int kPatchThreads = 64 * kPatchWaves;
constexpr int kPatchPix = 72;   // bf16 per staged input pixel
constexpr int kPatchRegs = 9;   // 16-byte staging registers per thread: a patch holds at most 9 * 512 units
constexpr int kPatchK = 576;    // 9 taps x 64 channels

struct PatchGeom {
    int H, Ho;
    int TP;       // output pixels per tile
    int n_img_t;  // images per tile (1 when a tile is part of an image)
    int PR;       // staged input rows per image: (TH - 1) * S + 3
    int units;    // 16-byte units of a patch: n_img_t * PR * H * 8
};

// EPI 0: bias and ReLU (a forward layer). EPI 1 (stride 1): the data gradient of a stride-1 layer, X its output
// gradient dZ, Wt its weights transposed with the taps flipped (k_wt_layers), the result masked by the ReLU of the
// layer's input Ymask (dX = (Ymask > 0) * sum), no bias.
template <int S, int EPI = 0>
__global__ void __launch_bounds__(kPatchThreads) k_conv_fwd_patch(const bf16 *__restrict__ X, int n, PatchGeom g,
                                                                  const bf16 *__restrict__ Wt,
                                                                  const float *__restrict__ bias,
                                                                  bf16 *__restrict__ Y,
                                                                  const bf16 *__restrict__ Ymask = nullptr) {
    extern __shared__ uint4 s_dyn[];
    constexpr int kLdw = kPatchK + kConvLdsPad;
    bf16 *s_w = reinterpret_cast<bf16 *>(s_dyn);  // [64][kLdw]
    bf16 *s_p = s_w + 64 * kLdw;                  // [n_img_t][PR][H][kPatchPix]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, gq = lane >> 4;
    for (int q = tid; q < 64 * (kPatchK / 8); q += kPatchThreads) {
        const int row = q / (kPatchK / 8), c8 = q - row * (kPatchK / 8);
        *reinterpret_cast<uint4 *>(s_w + row * kLdw + 8 * c8) =
            *reinterpret_cast<const uint4 *>(Wt + (size_t)row * kPatchK + 8 * c8);
    }
    const int HoHo = g.Ho * g.Ho;
    const int M = n * HoHo;  // below 2^31: the host caps the batch
    const int n_tiles = (M + g.TP - 1) / g.TP;
    // first image and first output row of tile t (tiles start at row boundaries, or at image boundaries)
    auto tile_origin = [&](int t, int &img0, int &oy0) {
        const int m0 = t * g.TP;
        img0 = (int)(m0 / HoHo);
        oy0 = (int)((m0 - img0 * HoHo) / g.Ho);
    };
    uint4 pre[kPatchRegs];
    // unit u of the patch: 8 channels c8 of staged pixel (image it, row, col), col fastest; staged row r of an image
    // is input row oy0 * S - 1 + r
    auto load_patch = [&](int t) {
        int img0, oy0;
        tile_origin(t, img0, oy0);
#pragma unroll
        for (int r = 0; r < kPatchRegs; ++r) {
            const int u = tid + r * kPatchThreads;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (u < g.units) {
                const int c8 = u & 7, pix = u >> 3;
                const int rr = pix / g.H, col = pix - rr * g.H;
                const int it = rr / g.PR, row = rr - it * g.PR;
                const int img = img0 + it, iy = oy0 * S - 1 + row;
                if (img < n && iy >= 0 && iy < g.H)
                    v = *reinterpret_cast<const uint4 *>(X + (((size_t)img * g.H + iy) * g.H + col) * 64 + 8 * c8);
            }
            pre[r] = v;
        }
    };
    constexpr int NTW = 4 / S;                       // column tiles per wave
    const int rp = (S == 1) ? wave : (wave >> 1);    // the wave's row tiles: 2 rp, 2 rp + 1
    const int ct0 = (S == 1) ? 0 : NTW * (wave & 1); // its first column tile
    if ((int)blockIdx.x < n_tiles) load_patch(blockIdx.x);
    for (int t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        __syncthreads();  // weights stored (first tile); the previous tile's patch reads done
#pragma unroll
        for (int r = 0; r < kPatchRegs; ++r) {
            const int u = tid + r * kPatchThreads;
            if (u < g.units) *reinterpret_cast<uint4 *>(s_p + (size_t)(u >> 3) * kPatchPix + 8 * (u & 7)) = pre[r];
        }
        __syncthreads();
        if (t + gridDim.x < n_tiles) load_patch(t + gridDim.x);
        int img0, oy0;
        tile_origin(t, img0, oy0);
        const int m0 = t * g.TP;
        // this lane's two A rows: staged pixel (image, row, column) of tap (0, 0), or none past the last pixel
        int pb[2], px[2];
        bool pok[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int m = m0 + (2 * rp + i) * 16 + r16;
            pok[i] = m < M;
            const int img = (int)(m / HoHo);
            const int pl = (int)(m - img * HoHo);
            const int oy = pl / g.Ho, ox = pl - oy * g.Ho;
            pb[i] = ((img - img0) * g.PR + (oy - oy0) * S) * g.H;
            px[i] = ox * S - 1;
        }
        floatx4_t acc[2][NTW];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < NTW; ++j) acc[i][j] = floatx4_t{0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < kPatchK; k0 += 32) {
            const int tap = k0 >> 6, c = (k0 & 63) + 8 * gq;
            const int ky = tap / 3, kx = tap - 3 * ky;
            bf16x8_t a[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int ix = px[i] + kx;
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if (pok[i] && ix >= 0 && ix < g.H)
                    v = *reinterpret_cast<const uint4 *>(s_p + (size_t)(pb[i] + ky * g.H + ix) * kPatchPix + c);
                a[i] = __builtin_bit_cast(bf16x8_t, v);
            }
            bf16x8_t b[NTW];
#pragma unroll
            for (int j = 0; j < NTW; ++j)
                b[j] = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const uint4 *>(
                                                        s_w + (16 * (ct0 + j) + r16) * kLdw + k0 + 8 * gq));
            // weights as the A operand: D tiles are channels x pixels, so that a lane holds 4 consecutive channels
            // of one pixel
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < NTW; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[j], a[i], acc[i][j], 0, 0, 0);
        }
        // D rows 4 gq + r of each channel tile, column r16 of each 16-pixel tile: bias and ReLU (or the ReLU mask of
        // Ymask), bf16 NHWC, 8 bytes per lane
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int mm = m0 + (2 * rp + i) * 16 + r16;
            if (mm >= M) continue;
#pragma unroll
            for (int j = 0; j < NTW; ++j) {
                const int c0 = 16 * (ct0 + j) + 4 * gq;
                const size_t idx = (size_t)mm * 64 + c0;
                uint2 o;
                bf16 *oe = reinterpret_cast<bf16 *>(&o);
                if (EPI == 0) {
                    const float4 bj = *reinterpret_cast<const float4 *>(bias + c0);
                    oe[0] = to_bf16(fmaxf(acc[i][j][0] + bj.x, 0.f));
                    oe[1] = to_bf16(fmaxf(acc[i][j][1] + bj.y, 0.f));
                    oe[2] = to_bf16(fmaxf(acc[i][j][2] + bj.z, 0.f));
                    oe[3] = to_bf16(fmaxf(acc[i][j][3] + bj.w, 0.f));
                } else {
                    const uint2 mk = *reinterpret_cast<const uint2 *>(Ymask + idx);
                    const bf16 *me = reinterpret_cast<const bf16 *>(&mk);
#pragma unroll
                    for (int r = 0; r < 4; ++r) oe[r] = to_bf16(from_bf16(me[r]) > 0.f ? acc[i][j][r] : 0.f);
                }
                *reinterpret_cast<uint2 *>(Y + idx) = o;
            }
        }
    }
}

// The weights of up to 8 64-channel layers rearranged for their data gradients in one launch (blockIdx.y = layer
// entry; entry e's output at wt + off_out[e]): mode 0 for k_conv_bwd_s2_patch, transposed,
// Wd[c][t*64 + co] = W[co][t*64 + c] (as k_wt_bwd); mode 1 for the stride-1 layers' data gradient as a stride-1
// convolution of dZ, transposed with the taps flipped, Wf[c][t*64 + co] = W[co][(8 - t)*64 + c] (output pixel y takes
// dZ at y + 1 - ky = y + ky' - 1, ky' = 2 - ky).
struct WtPlan {
    long long w_off[8], off_out[8];
    int mode[8];
};
__global__ void __launch_bounds__(kB) k_wt_layers(const bf16 *__restrict__ W, WtPlan plan, bf16 *__restrict__ wt) {
    const int idx = blockIdx.x * kB + threadIdx.x;
    if (idx >= 64 * kPatchK) return;
    const int e = blockIdx.y;
    const int c = idx / kPatchK, k = idx - c * kPatchK, t = k >> 6, co = k & 63;
    const int src_tap = plan.mode[e] == 0 ? t : 8 - t;
    wt[plan.off_out[e] + idx] = W[plan.w_off[e] + (size_t)co * kPatchK + src_tap * 64 + c];
}

// Data gradient of a stride-2 64-channel layer (input H x H, output Ho x Ho, H = 2 Ho), masked by the ReLU of its
// input Ymask. Input pixel (2a + py, 2b + px) takes dZ at (a + dy, b + dx) through the taps ky with
// 2 oy + ky - 1 = 2a + py, i.e. ky = 1 (dy = 0) for py = 0 and ky = 0 (dy = 1), 2 (dy = 0) for py = 1; likewise kx.
// So every position (a, b) of the dZ grid yields the four pixels of its 2 x 2 block, one per parity class (py, px),
// from the dZ pixels at (a..a+1, b..b+1), through 1, 2, 2 and 4 taps: 9 taps of 64 channels per position. Tiles of
// 256 positions start at row or image boundaries (PatchGeom over the dZ grid: H the input size, Ho the dZ size, PR
// = the tile's dZ rows plus the next one, zero past the image); the weights Wd[c][tap*64 + co] = W[co][tap*64 + c]
// (k_wt_bwd) stay in LDS. Each wave: two 16-position row tiles x 64 channels, class by class, stored through LDS as
// 16-byte rows. With colsum_part, block b writes the column sums of its part of the masked dX to colsum_part[64 b ..]
// (k_colsum_parts adds them up: the bias gradient of the layer whose output gradient dX is).
__global__ void __launch_bounds__(kPatchThreads) k_conv_bwd_s2_patch(const bf16 *__restrict__ dZ, int n, PatchGeom g,
                                                                     const bf16 *__restrict__ Wd,
                                                                     const bf16 *__restrict__ Ymask,
                                                                     bf16 *__restrict__ dX,
                                                                     float *__restrict__ colsum_part) {
    extern __shared__ uint4 s_dyn[];
    constexpr int kLdw = kPatchK + kConvLdsPad;
    bf16 *s_w = reinterpret_cast<bf16 *>(s_dyn);  // [64][kLdw]
    bf16 *s_p = s_w + 64 * kLdw;                  // [n_img_t][PR][Ho][kPatchPix]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, gq = lane >> 4;
    for (int q = tid; q < 64 * (kPatchK / 8); q += kPatchThreads) {
        const int row = q / (kPatchK / 8), c8 = q - row * (kPatchK / 8);
        *reinterpret_cast<uint4 *>(s_w + row * kLdw + 8 * c8) =
            *reinterpret_cast<const uint4 *>(Wd + (size_t)row * kPatchK + 8 * c8);
    }
    const int Ho = g.Ho, H = g.H, HoHo = Ho * Ho;
    const int M = n * HoHo;  // below 2^31: the host caps the batch
    const int n_tiles = (M + g.TP - 1) / g.TP;
    auto tile_origin = [&](int t, int &img0, int &a0) {
        const int m0 = t * g.TP;
        img0 = (int)(m0 / HoHo);
        a0 = (int)((m0 - img0 * HoHo) / Ho);
    };
    uint4 pre[kPatchRegs];
    // unit u of the patch: 8 channels c8 of staged pixel (image it, row, col), col fastest; staged row r of an image
    // is dZ row a0 + r
    auto load_patch = [&](int t) {
        int img0, a0;
        tile_origin(t, img0, a0);
#pragma unroll
        for (int r = 0; r < kPatchRegs; ++r) {
            const int u = tid + r * kPatchThreads;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (u < g.units) {
                const int c8 = u & 7, pix = u >> 3;
                const int rr = pix / Ho, col = pix - rr * Ho;
                const int it = rr / g.PR, row = rr - it * g.PR;
                const int img = img0 + it, iy = a0 + row;
                if (img < n && iy < Ho)
                    v = *reinterpret_cast<const uint4 *>(dZ + (((size_t)img * Ho + iy) * Ho + col) * 64 + 8 * c8);
            }
            pre[r] = v;
        }
    };
    // this wave's output staging: 32 positions x 64 channels, kPatchPix apart
    bf16 *s_o = s_p + (size_t)g.n_img_t * g.PR * Ho * kPatchPix + (size_t)wave * 32 * kPatchPix;
    float csum[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};  // channels 8 (lane % 8) .. + 7 of dX
    if ((int)blockIdx.x < n_tiles) load_patch(blockIdx.x);
    for (int t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        __syncthreads();  // weights stored (first tile); the previous tile's patch reads done
#pragma unroll
        for (int r = 0; r < kPatchRegs; ++r) {
            const int u = tid + r * kPatchThreads;
            if (u < g.units) *reinterpret_cast<uint4 *>(s_p + (size_t)(u >> 3) * kPatchPix + 8 * (u & 7)) = pre[r];
        }
        __syncthreads();
        if (t + gridDim.x < n_tiles) load_patch(t + gridDim.x);
        int img0, a0;
        tile_origin(t, img0, a0);
        const int m0 = t * g.TP;
        // this lane's two A rows: the staged pixel of position (a, b) and its column b, or none past the last position
        int pb[2], pcol[2];
        bool pok[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int m = m0 + (2 * wave + i) * 16 + r16;
            pok[i] = m < M;
            const int img = (int)(m / HoHo);
            const int pl = (int)(m - img * HoHo);
            const int a = pl / Ho;
            pcol[i] = pl - a * Ho;
            pb[i] = ((img - img0) * g.PR + (a - a0)) * Ho;
        }
        const int c8 = lane & 7;
#pragma unroll
        for (int cls = 0; cls < 4; ++cls) {
            const int py = cls >> 1, px = cls & 1;
            // this lane's epilogue pixels (8 channels of positions m0 + 32 wave + lane / 8 + 8 q): their ReLU mask
            // is loaded before the class's products
            size_t oidx[4];
            uint4 mk[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int mm = m0 + 32 * wave + (lane >> 3) + 8 * q;
                oidx[q] = ~(size_t)0;
                mk[q] = make_uint4(0u, 0u, 0u, 0u);
                if (mm < M) {
                    const int img = (int)(mm / HoHo);
                    const int pl = (int)(mm - img * HoHo);
                    const int a = pl / Ho, b = pl - a * Ho;
                    oidx[q] = (((size_t)img * H + 2 * a + py) * H + 2 * b + px) * 64 + 8 * c8;
                    mk[q] = *reinterpret_cast<const uint4 *>(Ymask + oidx[q]);
                }
            }
            floatx4_t acc[2][4];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = floatx4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ky = 1 - py; ky < 3; ky += 2)
#pragma unroll
                for (int kx = 1 - px; kx < 3; kx += 2) {
                    const int dy = (py + 1 - ky) >> 1, dx = (px + 1 - kx) >> 1;
                    const int tap = 3 * ky + kx;
#pragma unroll
                    for (int c0 = 0; c0 < 64; c0 += 32) {
                        const int c = c0 + 8 * gq;
                        bf16x8_t a[2];
#pragma unroll
                        for (int i = 0; i < 2; ++i) {
                            uint4 v = make_uint4(0u, 0u, 0u, 0u);
                            if (pok[i] && pcol[i] + dx < Ho)
                                v = *reinterpret_cast<const uint4 *>(
                                    s_p + (size_t)(pb[i] + dy * Ho + pcol[i] + dx) * kPatchPix + c);
                            a[i] = __builtin_bit_cast(bf16x8_t, v);
                        }
                        bf16x8_t b[4];
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            b[j] = __builtin_bit_cast(
                                bf16x8_t, *reinterpret_cast<const uint4 *>(s_w + (16 * j + r16) * kLdw + tap * 64 + c));
#pragma unroll
                        for (int i = 0; i < 2; ++i)
#pragma unroll
                            for (int j = 0; j < 4; ++j)
                                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
                    }
                }
            // D rows 4 gq + r of each 16-position tile, column r16 of each channel tile, as bf16 into the wave's own
            // staging rows 16 i + 4 gq + r (position m0 + 32 wave + row); then each lane moves 8 channels of 4
            // positions: the class's pixel of the position, masked, 16-byte NHWC stores. The rows are the wave's
            // alone: wave-scope ordering of its LDS accesses suffices.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        s_o[(16 * i + 4 * gq + r) * kPatchPix + 16 * j + r16] = to_bf16(acc[i][j][r]);
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (oidx[q] == ~(size_t)0) continue;
                const int row = (lane >> 3) + 8 * q;
                const uint4 v = *reinterpret_cast<const uint4 *>(s_o + row * kPatchPix + 8 * c8);
                const bf16 *ve = reinterpret_cast<const bf16 *>(&v);
                const bf16 *me = reinterpret_cast<const bf16 *>(&mk[q]);
                uint4 o;
                bf16 *oe = reinterpret_cast<bf16 *>(&o);
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    oe[e] = from_bf16(me[e]) > 0.f ? ve[e] : (bf16)0;
                    csum[e] += from_bf16(oe[e]);
                }
                *reinterpret_cast<uint4 *>(dX + oidx[q]) = o;
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (colsum_part) {  // the block's column sums of dX: lanes of one channel group, then the 8 waves in LDS
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            csum[e] += __shfl_xor(csum[e], 8);
            csum[e] += __shfl_xor(csum[e], 16);
            csum[e] += __shfl_xor(csum[e], 32);
        }
        float *s_red = reinterpret_cast<float *>(s_w);  // [8][64]: the weights are no longer read
        __syncthreads();
        if (lane < 8)
#pragma unroll
            for (int e = 0; e < 8; ++e) s_red[wave * 64 + 8 * lane + e] = csum[e];
        __syncthreads();
        if (tid < 64) {
            float t = 0.f;
            for (int w = 0; w < kPatchWaves; ++w) t += s_red[w * 64 + tid];
            colsum_part[(size_t)blockIdx.x * 64 + tid] = t;
        }
    }
}

// out[c] += sum over blocks b < nb of part[64 b + c]; one block of 1024 threads: thread (g, c) sums the blocks
// b = g (mod 16), then the 16 groups are added in order.
__global__ void __launch_bounds__(1024) k_colsum_parts(const float *__restrict__ part, int nb,
                                                       float *__restrict__ out) {
    __shared__ float s[16][64];
    const int c = threadIdx.x & 63, grp = threadIdx.x >> 6;
    float t = 0.f;
    for (int b = grp; b < nb; b += 16) t += part[(size_t)b * 64 + c];
    s[grp][c] = t;
    __syncthreads();
    if (grp == 0) {
        float u = 0.f;
        for (int k = 0; k < 16; ++k) u += s[k][c];
        out[c] += u;
    }
}

// Weight and bias gradients of a 64-channel layer (3 x 3, pad 1, stride S) without im2col columns:
// gW[co][tap*64 + ci] = sum_m dZ[m][co] X[pixel(m, tap)][ci], gb[co] = sum_m dZ[m][co], m over the n Ho x Ho output
// positions. Tiles of the forward geometry (PatchGeom: 256 / S positions from a row or image boundary, PR input rows
// per image). The tile's input rows are staged in LDS transposed, once per column tap kx: copy kx holds
// X[row][S ox + kx - 1][ci] at [kx][ci][(image, row)][ox] (zero outside the image), so the 8 consecutive positions of
// a B fragment (one output row, Ho % 8 == 0) are 16 contiguous bytes; rows of a plane are CP = n_img_t PR Ho + 8
// bf16 apart. A fragments: the tile's slice of dZt [64][M] (dZ transposed by k_transpose64), staged in LDS with the
// input rows (both loaded into registers during the previous tile's products). Wave w holds the sums of
// channels 32 (w % 2) .. +31 over columns 144 (w / 2) .. +143 (2 x 9 MFMA tiles: each B fragment read from LDS
// serves two channel tiles) across its block's tiles, and waves 0-1 the bias sums; the block writes them to
// part[block][64 * 576 + 64] (weights, then bias), which k_dw_reduce adds to the gradient.
constexpr int kDwPart = 64 * kPatchK + 64;

__host__ __device__ inline int dw_plane(const PatchGeom &g) { return g.n_img_t * g.PR * g.Ho + 8; }
// registers of 16 bytes per thread staging the input rows (64 pixels each) and the dZ slice, by stride
__host__ __device__ constexpr int dw_xregs(int S) { return S == 1 ? 5 : 9; }
__host__ __device__ constexpr int dw_zregs(int S) { return 4 / S; }  // 64 x (256 / S) bf16 over 512 threads
// LDS bf16 of the kernel: the three transposed copies, then the dZ slice [64][256 / S + 8]
__host__ __device__ inline size_t dw_lds_elems(const PatchGeom &g, int S) {
    return (size_t)3 * 64 * dw_plane(g) + (size_t)64 * (256 / S + 8);
}

template <int S>
__global__ void __launch_bounds__(kPatchThreads) k_conv_dw_patch(const bf16 *__restrict__ X, int n, PatchGeom g,
                                                                 const bf16 *__restrict__ dZt,
                                                                 float *__restrict__ part) {
    extern __shared__ uint4 s_dyn[];
    constexpr int kTP = 256 / S, kZP = kTP + 8;   // positions per tile; dZ slice pitch
    constexpr int kXR = dw_xregs(S), kZR = dw_zregs(S);
    bf16 *s_c = reinterpret_cast<bf16 *>(s_dyn);  // [3][64][CP]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, gq = lane >> 4;
    const int cp = wave & 1, kq = wave >> 1;  // channels 32 cp .. +31, columns 144 kq .. +143 (9 MFMA tiles)
    const int H = g.H, Ho = g.Ho, HoHo = Ho * Ho, CP = dw_plane(g);
    const int rows = g.n_img_t * g.PR;  // staged rows of a plane
    bf16 *s_z = s_c + (size_t)3 * 64 * CP;  // [64][kZP]: dZt[c][m0 + q], q < kTP
    const int M = n * HoHo;  // below 2^31: the host caps the batch
    const int n_tiles = (M + g.TP - 1) / g.TP;
    // entries no staged pixel reaches: column ox = 0 of copy 0 (input column -1), and for S = 1 column Ho - 1 of
    // copy 2 (input column H)
    for (int q = tid; q < 64 * rows; q += kPatchThreads) {
        const int ci = q / rows, rr = q - ci * rows;
        s_c[(size_t)ci * CP + rr * Ho] = 0;
        if (S == 1) s_c[(size_t)(128 + ci) * CP + rr * Ho + Ho - 1] = 0;
    }
    auto tile_origin = [&](int t, int &img0, int &oy0) {
        const int m0 = t * g.TP;
        img0 = (int)(m0 / HoHo);
        oy0 = (int)((m0 - img0 * HoHo) / Ho);
    };
    uint4 pre[kXR], zpre[kZR];
    // unit u: channels 8 c8 .. + 7, c8 = (u / 64) % 8, of staged pixel (u / 512) * 64 + u % 64 (image it, staged row,
    // column ix; column fastest), so that the 64 lanes of a wave store one plane at consecutive columns; staged row r
    // of an image is input row oy0 * S - 1 + r. dZ unit u: channel u / (kTP / 8), positions 8 (u % (kTP / 8)) .. + 7.
    const int npix = rows * H;
    auto load_patch = [&](int t) {
        int img0, oy0;
        tile_origin(t, img0, oy0);
#pragma unroll
        for (int r = 0; r < kZR; ++r) {
            const int u = tid + r * kPatchThreads;
            const int c = u / (kTP / 8), q8 = u - c * (kTP / 8);
            const int m = t * g.TP + 8 * q8;
            zpre[r] = m < M ? *reinterpret_cast<const uint4 *>(dZt + (size_t)c * M + m) : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (int r = 0; r < kXR; ++r) {
            const int u = tid + r * kPatchThreads;
            const int pix = (u >> 9) * 64 + (u & 63), c8 = (u >> 6) & 7;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (pix < npix) {
                const int rr = pix / H, ix = pix - rr * H;
                const int it = rr / g.PR, row = rr - it * g.PR;
                const int img = img0 + it, iy = oy0 * S - 1 + row;
                if (img < n && iy >= 0 && iy < H)
                    v = *reinterpret_cast<const uint4 *>(X + (((size_t)img * H + iy) * H + ix) * 64 + 8 * c8);
            }
            pre[r] = v;
        }
    };
    floatx4_t acc[2][9];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 9; ++j) acc[i][j] = floatx4_t{0.f, 0.f, 0.f, 0.f};
    float bsum[2] = {0.f, 0.f};
    if ((int)blockIdx.x < n_tiles) load_patch(blockIdx.x);
    for (int t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        __syncthreads();  // zero columns stored (first tile); the previous tile's reads done
#pragma unroll
        for (int r = 0; r < kZR; ++r) {
            const int u = tid + r * kPatchThreads;
            const int c = u / (kTP / 8), q8 = u - c * (kTP / 8);
            *reinterpret_cast<uint4 *>(s_z + c * kZP + 8 * q8) = zpre[r];
        }
#pragma unroll
        for (int r = 0; r < kXR; ++r) {
            const int u = tid + r * kPatchThreads;
            const int pix = (u >> 9) * 64 + (u & 63), c8 = (u >> 6) & 7;
            if (pix < npix) {
                const int rr = pix / H, ix = pix - rr * H;
                const bf16 *e = reinterpret_cast<const bf16 *>(&pre[r]);
#pragma unroll
                for (int kx = 0; kx < 3; ++kx) {
                    const int tt = ix + 1 - kx;
                    if (tt < 0 || (S == 2 && (tt & 1))) continue;
                    const int ox = tt / S;
                    if (ox >= Ho) continue;
                    bf16 *d = s_c + (size_t)(kx * 64 + 8 * c8) * CP + rr * Ho + ox;
#pragma unroll
                    for (int c = 0; c < 8; ++c) d[(size_t)c * CP] = e[c];
                }
            }
        }
        __syncthreads();
        if (t + gridDim.x < n_tiles) load_patch(t + gridDim.x);
        int img0, oy0;
        tile_origin(t, img0, oy0);
        const int m0 = t * g.TP;
        const int row0 = m0 / Ho;  // global output row of the tile's first position
        for (int s = 0; s < g.TP / 32; ++s) {
            const int q0 = 32 * s + 8 * gq;  // this lane's 8 positions q0 .. q0 + 7 of the tile
            bf16x8_t a[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const uint4 av = *reinterpret_cast<const uint4 *>(s_z + (32 * cp + 16 * i + r16) * kZP + q0);
                if (kq == 0) {
                    const bf16 *e = reinterpret_cast<const bf16 *>(&av);
#pragma unroll
                    for (int c = 0; c < 8; ++c) bsum[i] += from_bf16(e[c]);
                }
                a[i] = __builtin_bit_cast(bf16x8_t, av);
            }
            const int R = row0 + q0 / Ho;
            const int img = (int)(R / Ho), oy = (int)(R - img * Ho);
            const int pbase = ((img - img0) * g.PR + (oy - oy0) * S) * Ho + (q0 % Ho);
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const int kt = 9 * kq + j, tap = kt >> 2, ci0 = 16 * (kt & 3);
                const int ky = tap / 3, kx = tap - 3 * ky;
                const bf16x8_t b = __builtin_bit_cast(
                    bf16x8_t, *reinterpret_cast<const uint4 *>(s_c + (size_t)(kx * 64 + ci0 + r16) * CP + pbase + ky * Ho));
#pragma unroll
                for (int i = 0; i < 2; ++i) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b, acc[i][j], 0, 0, 0);
            }
        }
    }
    // D row 4 gq + r of channel tile 2 cp + i, column r16 of MFMA tile kt
    float *o = part + (size_t)blockIdx.x * kDwPart;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            const int kt = 9 * kq + j;
#pragma unroll
            for (int r = 0; r < 4; ++r)
                o[(size_t)(32 * cp + 16 * i + 4 * gq + r) * kPatchK + 16 * kt + r16] = acc[i][j][r];
        }
    if (kq == 0) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            float b = bsum[i];
            b += __shfl_xor(b, 16);
            b += __shfl_xor(b, 32);
            if (gq == 0) o[64 * kPatchK + 32 * cp + 16 * i + r16] = b;
        }
    }
}

// out[e] += sum over blocks b of part[b][e], e < kDwPart (four consecutive entries per thread); blockIdx.y sums
// the blocks [y nb / gy, (y + 1) nb / gy) and adds its sums atomically.
__global__ void __launch_bounds__(kB) k_dw_reduce(const float *__restrict__ part, int nb, float *__restrict__ out) {
    const int e4 = blockIdx.x * kB + threadIdx.x;
    if (4 * e4 >= kDwPart) return;
    const int b0 = (int)((long long)blockIdx.y * nb / gridDim.y), b1 = (int)((long long)(blockIdx.y + 1) * nb / gridDim.y);
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll 8
    for (int b = b0; b < b1; ++b) {
        const float4 v = reinterpret_cast<const float4 *>(part + (size_t)b * kDwPart)[e4];
        s.x += v.x;
        s.y += v.y;
        s.z += v.z;
        s.w += v.w;
    }
    atomicAdd(&out[4 * e4], s.x);
    atomicAdd(&out[4 * e4 + 1], s.y);
    atomicAdd(&out[4 * e4 + 2], s.z);
    atomicAdd(&out[4 * e4 + 3], s.w);
}

// --------------------------------------------------------------------------------------------------
// Loss and prediction over NB bins: one block per image.
// --------------------------------------------------------------------------------------------------
__device__ __forceinline__ float block_max(float v, float *red) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = red[0];
    for (int w = 1; w < kB / 64; ++w) r = fmaxf(r, red[w]);
    return r;
}
__device__ __forceinline__ double block_sum(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.;
    for (int w = 0; w < kB / 64; ++w) r += red[w];
    return r;
}

// loss[b] = logsumexp(z) - z[label]; dz = (softmax(z) - onehot) * scale (bf16)
__global__ void __launch_bounds__(kB) k_softmax_xent(const float *__restrict__ z, const int *__restrict__ labels,
                                                     int NB, float scale, double *__restrict__ loss,
                                                     bf16 *__restrict__ dz) {
    __shared__ float redf[kB / 64];
    __shared__ double redd[kB / 64];
    const int b = blockIdx.x, tid = threadIdx.x;
    const float *zb = z + (size_t)b * NB;
    float mx = -3.0e38f;
    for (int j = tid; j < NB; j += kB) mx = fmaxf(mx, zb[j]);
    mx = block_max(mx, redf);
    double s = 0.;
    for (int j = tid; j < NB; j += kB) s += exp((double)(zb[j] - mx));
    s = block_sum(s, redd);
    const int lab = labels[b];
    if (tid == 0) loss[b] = log(s) + (double)mx - (double)zb[lab];
    const double inv = 1. / s;
    for (int j = tid; j < NB; j += kB) {
        const float p = (float)(exp((double)(zb[j] - mx)) * inv);
        dz[(size_t)b * NB + j] = to_bf16(((j == lab) ? p - 1.f : p) * scale);
    }
}

constexpr int kMaxTopK = 64;

// top-k bins of softmax(z) per image (descending probability, ties to the lower bin)
__global__ void __launch_bounds__(kB) k_topk(const float *__restrict__ z, int NB, int topk, int *__restrict__ bins,
                                             float *__restrict__ probs) {
    extern __shared__ float zs[];
    __shared__ float redf[kB / 64];
    __shared__ double redd[kB / 64];
    __shared__ int redi[kB / 64];
    const int b = blockIdx.x, tid = threadIdx.x;
    const float *zb = z + (size_t)b * NB;
    float mx = -3.0e38f;
    for (int j = tid; j < NB; j += kB) {
        zs[j] = zb[j];
        mx = fmaxf(mx, zs[j]);
    }
    mx = block_max(mx, redf);
    double s = 0.;
    for (int j = tid; j < NB; j += kB) s += exp((double)(zs[j] - mx));
    s = block_sum(s, redd);
    for (int t = 0; t < topk; ++t) {
        float bv = -3.0e38f;
        int bi = 0x7fffffff;
        for (int j = tid; j < NB; j += kB)
            if (zs[j] > bv || (zs[j] == bv && j < bi)) {
                bv = zs[j];
                bi = j;
            }
        for (int o = 32; o > 0; o >>= 1) {
            const float ov = __shfl_xor(bv, o);
            const int oi = __shfl_xor(bi, o);
            if (ov > bv || (ov == bv && oi < bi)) {
                bv = ov;
                bi = oi;
            }
        }
        __syncthreads();
        if ((tid & 63) == 0) {
            redf[tid >> 6] = bv;
            redi[tid >> 6] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            for (int w = 1; w < kB / 64; ++w)
                if (redf[w] > bv || (redf[w] == bv && redi[w] < bi)) {
                    bv = redf[w];
                    bi = redi[w];
                }
            bins[(size_t)b * topk + t] = bi;
            probs[(size_t)b * topk + t] = (float)(exp((double)(bv - mx)) / s);
            zs[bi] = -3.0e38f;
        }
        __syncthreads();
    }
}

// Adam (Kingma & Ba 2015) on FP32 master weights, refreshing the BF16 copies; grads are cleared.
__global__ void __launch_bounds__(kB) k_adam(float *__restrict__ w, float *__restrict__ grad, float *__restrict__ m1,
                                             float *__restrict__ m2, bf16 *__restrict__ wb, long long n, float lr,
                                             float b1, float b2, float eps, float c1, float c2) {
    const long long i = (long long)blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const float gr = grad[i];
    const float a = b1 * m1[i] + (1.f - b1) * gr;
    const float v = b2 * m2[i] + (1.f - b2) * gr * gr;
    m1[i] = a;
    m2[i] = v;
    const float wi = w[i] - lr * (a * c1) / (sqrtf(v * c2) + eps);
    w[i] = wi;
    wb[i] = to_bf16(wi);
    grad[i] = 0.f;
}

__global__ void __launch_bounds__(kB) k_to_bf16(const float *__restrict__ x, long long n, bf16 *__restrict__ y) {
    const long long i = (long long)blockIdx.x * kB + threadIdx.x;
    if (i < n) y[i] = to_bf16(x[i]);
}

// gather a batch of stored images: dst[i] = src[idx[i]] (W*W bf16 each)
__global__ void __launch_bounds__(kB) k_gather_images(const bf16 *__restrict__ src, const int *__restrict__ idx,
                                                      int WW, bf16 *__restrict__ dst) {
    const int b = blockIdx.y;
    const int j = blockIdx.x * kB + threadIdx.x;
    if (j < WW) dst[(size_t)b * WW + j] = src[(size_t)idx[b] * WW + j];
}

__global__ void __launch_bounds__(kB) k_gather_labels(const int *__restrict__ src, const int *__restrict__ idx, int n,
                                                      int *__restrict__ dst) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

}  // namespace rlnp
