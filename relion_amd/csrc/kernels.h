// Device kernels of the plain-FP32 E-step for CDNA4 (gfx950). Included by estep.hip only.
//
// Image geometry: an s x (s/2+1) half-complex image, pixel p = i*(s/2+1) + x, frequency
// (x, y) with y = i for i <= s/2 else i - s (FOR_ALL_ELEMENTS_IN_FFTW_TRANSFORM, src/fftw.h:99-102).
// Volume geometry: projector layout of src/projector.cpp:59-89, interleaved float2 [z][y][x].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.h"

namespace rlnk {

// Builds the corner-packed volume from the plain [z][y][x] float2 volume. One thread per base voxel.
__global__ void __launch_bounds__(kBlock) k_pack_volume(const float2 *__restrict__ plain, VolGeom g,
                                                        float4 *__restrict__ packed) {
    const size_t n = (size_t)(g.xdim - 1) * (g.ydim - 1) * (g.zdim - 1);
    const size_t idx = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (idx >= n) return;
    const int bx = g.xdim - 1, by = g.ydim - 1;
    const int x0 = (int)(idx % bx);
    const size_t zy = idx / bx;
    const int y0 = (int)(zy % by), z0 = (int)(zy / by);
    const size_t row = (size_t)g.xdim, slab = row * (size_t)g.ydim;
    const float2 *b = plain + (size_t)z0 * slab + (size_t)y0 * row + x0;
    const float2 d000 = b[0], d001 = b[1], d010 = b[row], d011 = b[row + 1];
    const float2 d100 = b[slab], d101 = b[slab + 1], d110 = b[slab + row], d111 = b[slab + row + 1];
    float4 *o = packed + 4 * packed_index(g, x0, y0, z0);
    o[0] = make_float4(d000.x, d000.y, d001.x, d001.y);
    o[1] = make_float4(d010.x, d010.y, d011.x, d011.y);
    o[2] = make_float4(d100.x, d100.y, d101.x, d101.y);
    o[3] = make_float4(d110.x, d110.y, d111.x, d111.y);
}

// Backprojection accumulators in corner slots: per base voxel 32 floats = one 128-byte line,
// [re c0..c7 | im c0..c7 | weight c0..c7 | unused], corner c = 4*dz + 2*dy + dx. All 24 atomics of one
// trilinear scatter land in one line, so a wave's atomic instruction can be served line by line.
// k_fold_slots sums the slots into plain volumes.
constexpr int kSlot = 32;

// Plain data/weight volumes from the slots: v(z,y,x) = sum over corners c (in order) of slot[base(z,y,x) - c][c].
// One block per 4 x 4 x 4 brick of output voxels: the slots of the 5 x 5 x 5 bases it needs (the brick's own and
// the lower neighbours' faces) are read into LDS, 24 floats per base, then 192 threads form the 64 voxels' re, im
// and weight. Bricks from layer bz0 on (z from 4 bz0): the outputs hold planes z - 4 bz0, so a slab of layers can be
// folded into a smaller buffer. grid (ceil(xdim/4) * ceil(ydim/4) * layers), at most ceil(zdim/4) - bz0 layers.
__global__ void __launch_bounds__(kBlock) k_fold_slots(const float *__restrict__ slots, VolGeom g, int bz0,
                                                       float *__restrict__ re, float *__restrict__ im,
                                                       float *__restrict__ w) {
    __shared__ float s[125][25];  // [base][re c0..7 | im c0..7 | weight c0..7], padded against bank conflicts
    const int obx = (g.xdim + 3) >> 2, oby = (g.ydim + 3) >> 2;
    const int bx = blockIdx.x % obx, by = (blockIdx.x / obx) % oby, bz = bz0 + (int)(blockIdx.x / (obx * oby));
    for (int q = threadIdx.x; q < 125 * 24; q += kBlock) {
        const int bi = q / 24, f = q - bi * 24;
        const int x0 = 4 * bx - 1 + bi % 5, y0 = 4 * by - 1 + (bi / 5) % 5, z0 = 4 * bz - 1 + bi / 25;
        float v = 0.f;
        if (x0 >= 0 && y0 >= 0 && z0 >= 0 && x0 < g.xdim - 1 && y0 < g.ydim - 1 && z0 < g.zdim - 1)
            v = slots[kSlot * packed_index(g, x0, y0, z0) + f];
        s[bi][f] = v;
    }
    __syncthreads();
    if (threadIdx.x >= 192) return;
    const int comp = threadIdx.x >> 6, vi = threadIdx.x & 63;  // comp: 0 re, 1 im, 2 weight
    const int lx = vi & 3, ly = (vi >> 2) & 3, lz = vi >> 4;
    const int x = 4 * bx + lx, y = 4 * by + ly, z = 4 * bz + lz;
    if (x >= g.xdim || y >= g.ydim || z >= g.zdim) return;
    float sum = 0.f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {  // base (x - dx, y - dy, z - dz) is local (lx + 1 - dx, ...)
        const int bi = (lx + 1 - (c & 1)) + 5 * ((ly + 1 - ((c >> 1) & 1)) + 5 * (lz + 1 - (c >> 2)));
        sum += s[bi][8 * comp + c];
    }
    float *out = comp == 0 ? re : (comp == 1 ? im : w);
    out[((size_t)(z - 4 * bz0) * g.ydim + y) * g.xdim + x] = sum;
}

// Blocks of k_fold_slots for `layers` brick layers (all of them by default).
inline unsigned fold_blocks(const VolGeom &g, int layers = -1) {
    if (layers < 0) layers = (g.zdim + 3) / 4;
    return (unsigned)(((g.xdim + 3) / 4) * ((g.ydim + 3) / 4) * layers);
}


// --------------------------------------------------------------------------------------------------
// Coarse pass
// --------------------------------------------------------------------------------------------------

// Projections of every coarse orientation of every class at the coarse size:
// proj[(k*n_or + o)*P + p] (class-major). One thread per pixel, grid (K*n_or, ceil(P/256)).
__global__ void __launch_bounds__(kBlock) k_project_grid(const float4 *const *__restrict__ vols, VolGeom vg,
                                                         const Frame *__restrict__ frames, int n_or, ImgGeom ig,
                                                         float2 *__restrict__ proj) {
    const int p = blockIdx.y * kBlock + threadIdx.x;
    const int ko = blockIdx.x;
    if (p >= ig.P) return;
    const int k = ko / n_or, o = ko - k * n_or;
    int x, y;
    pixel_xy(p, ig.s, ig.xs, x, y);
    float2 v = make_float2(0.f, 0.f);
    if (x * x + y * y <= ig.proj_r2_max) v = project_pixel(vols[k], vg, frames[o], x, y);
    proj[(size_t)ko * ig.P + p] = v;
}

// Windowed, sqrt(Minvsigma2/2)-scaled images of the coarse pass for B particles:
//   ctfq[b][p] = q_p * ctf_b(p) * scale_b  and  yq[b][t][p] = q_p * X_b(p) * exp(-2 pi i (x tx + y ty)/N),
// with the phase factored as ex[t][x] * ey[t][i]. X and ctf are read from the fine-size input (window of
// src/fftw.h:809-857). grid (ceil(P/256), NT, B).
__global__ void __launch_bounds__(kBlock) k_prep_coarse(const float2 *__restrict__ fimg, const float *__restrict__ fctf,
                                                        const float *__restrict__ scale, int sf, int Pf, ImgGeom ig,
                                                        const float *__restrict__ q, const float2 *__restrict__ ex,
                                                        const float2 *__restrict__ ey, int NT, int do_ctf,
                                                        float *__restrict__ ctfq, float2 *__restrict__ yq) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    const int t = blockIdx.y, b = blockIdx.z;
    if (p >= ig.P) return;
    const int i = p / ig.xs, x = p - i * ig.xs;
    const int y = (i < ig.xs) ? i : i - ig.s;
    const int isrc = (y < 0) ? y + sf : y;
    const size_t src = (size_t)b * Pf + (size_t)isrc * (sf / 2 + 1) + x;
    const float qp = q[p];
    const float2 X = fimg[src];
    const float2 ph = cmul(ex[(size_t)t * ig.xs + x], ey[(size_t)t * ig.s + i]);
    const float2 v = cmul(make_float2(X.x * qp, X.y * qp), ph);
    yq[((size_t)b * NT + t) * ig.P + p] = v;
    if (t == 0) ctfq[(size_t)b * ig.P + p] = qp * (do_ctf ? fctf[src] : 1.f) * scale[b];
}

// Tile of the coarse squared differences: TO orientations x TT translations of one particle.
constexpr int kCoTO = 64, kCoTT = 64, kCoCP = 32;

struct CoarseTile {
    int b;        // particle within the batch
    int o_begin;  // first orientation (within the particle's list)
};

// diff2[b][o][t] = highres_Xi2/2 + sum_p |ctfq_b(p) * P_o(p) - yq_b,t(p)|^2 (getAllSquaredDifferences,
// src/ml_optimiser.cpp:7430-7442). Orientations whose prior or class prior is zero are marked -999 (never
// computed, as in the reference's exp_Mweight initialisation, :7155). grid (n_tiles, ceil(NT/TT)).
__global__ void __launch_bounds__(kBlock) k_coarse_diff2(const CoarseTile *__restrict__ tiles,
                                                         const int *__restrict__ or_off, const int *__restrict__ or_cnt,
                                                         const int *__restrict__ or_proj, const float2 *__restrict__ proj,
                                                         const float *__restrict__ ctfq, const float2 *__restrict__ yq,
                                                         const float *__restrict__ xi2half, const long long *__restrict__ d2_off,
                                                         int NT, int P, float *__restrict__ diff2) {
    __shared__ float2 As[kCoCP][kCoTO];
    __shared__ float2 Ys[kCoCP][kCoTT];
    __shared__ int og_s[kCoTO];
    const CoarseTile tile = tiles[blockIdx.x];
    const int b = tile.b;
    const int n_o = or_cnt[b];
    const int t_begin = blockIdx.y * kCoTT;
    const int tid = threadIdx.x;
    const int to = tid >> 4, tt = tid & 15;
    if (tid < kCoTO) {
        const int o = tile.o_begin + tid;
        og_s[tid] = (o < n_o) ? or_proj[or_off[b] + o] : -1;
    }
    __syncthreads();
    const float *cq = ctfq + (size_t)b * P;
    const float2 *yb = yq + (size_t)b * NT * P;
    float tot[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) tot[i][j] = 0.f;
    for (int p0 = 0; p0 < P; p0 += kCoCP) {
        // stage c*P for TO orientations (coalesced along p)
        for (int idx = tid; idx < kCoCP * kCoTO; idx += kBlock) {
            const int oo = idx / kCoCP, pp = idx - oo * kCoCP;
            const int p = p0 + pp;
            float2 v = make_float2(0.f, 0.f);
            const int og = og_s[oo];
            if (og >= 0 && p < P) {
                const float2 pr = proj[(size_t)og * P + p];
                const float c = cq[p];
                v = make_float2(c * pr.x, c * pr.y);
            }
            As[pp][oo] = v;
        }
        for (int idx = tid; idx < kCoCP * kCoTT; idx += kBlock) {
            const int tq = idx / kCoCP, pp = idx - tq * kCoCP;
            const int p = p0 + pp, t = t_begin + tq;
            Ys[pp][tq] = (t < NT && p < P) ? yb[(size_t)t * P + p] : make_float2(0.f, 0.f);
        }
        __syncthreads();
        float acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
#pragma unroll 4
        for (int pp = 0; pp < kCoCP; ++pp) {
            float2 a[4], y[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[pp][to + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) y[j] = Ys[pp][tt + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float dr = a[i].x - y[j].x, di = a[i].y - y[j].y;
                    acc[i][j] = fmaf(dr, dr, acc[i][j]);
                    acc[i][j] = fmaf(di, di, acc[i][j]);
                }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) tot[i][j] += acc[i][j];
        __syncthreads();
    }
    const float base = xi2half[b];
    float *out = diff2 + d2_off[b];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int oo = to + 16 * i;
        const int o = tile.o_begin + oo;
        if (o >= n_o) continue;
        const bool skip = og_s[oo] < 0 || or_proj[or_off[b] + o] < 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int t = t_begin + tt + 16 * j;
            if (t < NT) out[(size_t)o * NT + t] = skip ? -999.f : base + tot[i][j];
        }
    }
}

// --------------------------------------------------------------------------------------------------
// Split-BF16 coarse pass on MFMA.
// diff2(o,t) = Xi2/2 + sum_p c_p^2 |P_o(p)|^2 + sum_p |y_t(p)|^2 - 2 sum_p Re(P_o(p) conj(c_p y_t(p))),
// with y_t(p) = q_p X(p) exp(-2 pi i (x tx + y ty)/N) (|y_t|^2 does not depend on t) and c_p = q_p ctf_p s.
// The cross term is a real contraction over k = 2p + {re, im} of A[o][k] (the projection, shared by all
// particles) and B[t][k] = c_p y_t(p); both are split x = hi + lo in BF16 and the products hi*hi + hi*lo +
// lo*hi are accumulated in FP32 by v_mfma_f32_16x16x32_bf16.
// --------------------------------------------------------------------------------------------------
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float floatx2_t __attribute__((ext_vector_type(2)));

// round-to-nearest-even FP32 -> BF16 bit pattern
__device__ __forceinline__ unsigned short f2bf(float x) {
    unsigned int u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }
__device__ __forceinline__ void split_bf16(float x, unsigned short &hi, unsigned short &lo) {
    hi = f2bf(x);
    lo = f2bf(x - bf2f(hi));
}
__device__ __forceinline__ bf16x8_t load_bf16x8(const unsigned short *p) {
    return __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const uint4 *>(p));
}

// Lane layout of v_mfma_f32_16x16x32_bf16 used below (checked by k_selftest_mfma): lane l supplies row l%16 of
// A and column l%16 of B, each with the 8 consecutive k = 8*(l/16) .. 8*(l/16)+7; it receives D rows
// 4*(l/16) .. 4*(l/16)+3 of column l%16.
__global__ void k_selftest_mfma(const float *__restrict__ A, const float *__restrict__ B, float *__restrict__ D) {
    const int l = threadIdx.x;
    unsigned short a[8], b[8];
    for (int j = 0; j < 8; ++j) {
        a[j] = f2bf(A[(l % 16) * 32 + 8 * (l / 16) + j]);
        b[j] = f2bf(B[(8 * (l / 16) + j) * 16 + l % 16]);
    }
    floatx4_t acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(load_bf16x8(a), load_bf16x8(b), acc, 0, 0, 0);
    for (int j = 0; j < 4; ++j) D[(4 * (l / 16) + j) * 16 + l % 16] = acc[j];
}

// Contraction layout of the split-BF16 pass: pixel groups of 4 consecutive x in one row. Group g = (row i, x0)
// holds pixels (i, x0 .. x0+3) at k = 8g + 2j + {re, im}; the groups cover, row by row, x = 0 up to the last
// pixel whose Minvsigma2 is non-zero, so rows and corners that carry no weight are not contracted. Kp = 8 *
// (groups rounded up to a multiple of 4); slots past the last group and pixels with x >= xs hold zeros.

// hi + lo split of two floats with the hardware round-to-nearest-even conversion; hi and lo are packed pairs
__device__ __forceinline__ void split2_bf16(float a, float b, unsigned int &hi, unsigned int &lo) {
    const bf16x2_t h = __builtin_convertvector((floatx2_t){a, b}, bf16x2_t);
    hi = __builtin_bit_cast(unsigned int, h);
    const float ra = a - __uint_as_float(hi << 16), rb = b - __uint_as_float(hi & 0xffff0000u);
    lo = __builtin_bit_cast(unsigned int, __builtin_convertvector((floatx2_t){ra, rb}, bf16x2_t));
}

// Projection tables of the split pass in the group layout: FP32 complex pf [K*n_or][Kp/2] (the coarse kernel
// multiplies them by the particle and splits them in BF16). grid (K*n_or, ceil(Kp/2/256)).
__global__ void __launch_bounds__(kBlock) k_project_grid_split(const float4 *const *__restrict__ vols, VolGeom vg,
                                                               const Frame *__restrict__ frames, int n_or, ImgGeom ig,
                                                               const int2 *__restrict__ grp, int G, int Kp,
                                                               float2 *__restrict__ pf) {
    const int idx = blockIdx.y * kBlock + threadIdx.x;  // complex slot 4g + j
    const int ko = blockIdx.x;
    if (2 * idx >= Kp) return;
    const int g = idx >> 2, j = idx & 3;
    float2 v = make_float2(0.f, 0.f);
    if (g < G) {
        const int2 gr = grp[g];
        const int x = gr.y + j, i = gr.x;
        const int y = (i < ig.xs) ? i : i - ig.s;
        if (x < ig.xs && x * x + y * y <= ig.proj_r2_max) {
            const int k = ko / n_or, o = ko - k * n_or;
            v = project_pixel(vols[k], vg, frames[o], x, y);
        }
    }
    pf[(size_t)ko * (Kp / 2) + idx] = v;
}

// Particle side of the split pass, in the group layout: w[b][4g+j] = c_p q_p X(p) with c_p = q_p ctf_p s (the
// translation phase is the coarse kernel's B operand), ctf2[b][4g+j] = c_p^2, and ynorm[b] = sum_p |q_p X(p)|^2
// over all pixels (FP64). grid (ceil(max(Kp/2, P)/256), B).
__global__ void __launch_bounds__(kBlock) k_prep_coarse_split(const float2 *__restrict__ fimg,
                                                              const float *__restrict__ fctf,
                                                              const float *__restrict__ scale, int sf, int Pf,
                                                              ImgGeom ig, const float *__restrict__ q,
                                                              const int2 *__restrict__ grp, int G, int Kp, int do_ctf,
                                                              float2 *__restrict__ w, float *__restrict__ ctf2,
                                                              double *__restrict__ ynorm) {
    __shared__ double red[kBlock / 64];
    const int idx = blockIdx.x * kBlock + threadIdx.x;
    const int b = blockIdx.y;
    const float sc = scale[b];
    auto src_of = [&](int i, int x) {
        const int y = (i < ig.xs) ? i : i - ig.s;
        const int isrc = (y < 0) ? y + sf : y;
        return (size_t)b * Pf + (size_t)isrc * (sf / 2 + 1) + x;
    };
    if (idx < Kp / 2) {
        const int g = idx >> 2, j = idx & 3;
        float2 v = make_float2(0.f, 0.f);
        float c2 = 0.f;
        if (g < G) {
            const int2 gr = grp[g];
            const int x = gr.y + j, i = gr.x;
            if (x < ig.xs) {
                const size_t src = src_of(i, x);
                const float qp = q[i * ig.xs + x];
                const float c = qp * (do_ctf ? fctf[src] : 1.f) * sc;
                const float2 X = fimg[src];
                v = make_float2(c * qp * X.x, c * qp * X.y);
                c2 = c * c;
            }
        }
        w[(size_t)b * (Kp / 2) + idx] = v;
        ctf2[(size_t)b * (Kp / 2) + idx] = c2;
    }
    float yn = 0.f;
    if (idx < ig.P) {
        const int i = idx / ig.xs, x = idx - i * ig.xs;
        const float qp = q[idx];
        const float2 X = fimg[src_of(i, x)];
        yn = qp * qp * (X.x * X.x + X.y * X.y);
    }
    double s = yn;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double t = red[0] + red[1] + red[2] + red[3];
        if (t != 0.) atomicAdd(&ynorm[b], t);
    }
}

// Translation operand of the split pass, the same for every particle: the phase E[t][4g + m] = ex[t][x0 + m] *
// ey[t][row] of every complex slot of pixel group g = (row, x0) (rows of exg padded with zeros past xs), split
// into BF16 hi and lo (re, im) pairs. grid (ceil(NT * Kc / 256)).
__global__ void __launch_bounds__(kBlock) k_phase_split(const float2 *__restrict__ exg, int xs4,
                                                        const float2 *__restrict__ ey, int s,
                                                        const int2 *__restrict__ grp, int Kc, int NT,
                                                        unsigned int *__restrict__ ehi, unsigned int *__restrict__ elo) {
    const long long idx = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (idx >= (long long)NT * Kc) return;
    const int t = (int)(idx / Kc), c = (int)(idx - (long long)t * Kc);
    const int2 gr = grp[c >> 2];
    const float2 v = cmul(exg[(size_t)t * xs4 + gr.y + (c & 3)], ey[(size_t)t * s + gr.x]);
    unsigned int h, l;
    split2_bf16(v.x, v.y, h, l);
    ehi[idx] = h;
    elo[idx] = l;
}

// Per-particle state of the coarse weights and significance selection (k_wpart_init .. k_wsel_stats); the
// split coarse kernel already folds its minimum diff2 into mn_bits.
struct WPart {
    unsigned int mn_bits;   // min diff2 (non-negative float bits)
    int status;
    double sum;             // sum of weights
    double target;          // adaptive_fraction * sum
    unsigned long long prefix, mask;  // radix selection state
    double above_sum;       // weights in digit bins above the crossing bin
    long long above_cnt;
    int gather;             // 1: crossing bin small enough, elements gathered
    int ncand;
    double thr;             // significance threshold
    long long nr_sign;
    long long n_above;      // entries >= thr
};

constexpr int kMfO = 64;    // orientations per tile (4 MFMA row tiles)
constexpr int kMfG = 1;     // tiles per block, each with its own 4 waves
constexpr int kMfBlock = 256 * kMfG;

// One block per two coarse tiles (64 orientations of one particle each, possibly different particles) and all
// translations. The cross term sum_p Re(conj(P_o(p)) w(p) E_t(p)) is a real dot product of (re, im) pairs
// between A[o][p] = P_o(p) conj(w(p)), formed while a group stages its tile's projection rows (64 rows a K step)
// and split in BF16, and B[t][p] = E_t(p), the particle-independent phase table split once per model
// (k_phase_split) and read straight into the MFMA operands: the two groups read the same B fragments, the second
// from the cache. Wave w of a group owns translation tiles w, w+4, w+8, ... and the group's four orientation
// tiles; row tiles past the tile's orientations are skipped. grid (ceil(n_tiles / 2)).
template <int NTT>  // translation tiles of 16 in the block (16*NTT >= NT)
__global__ void __launch_bounds__(kMfBlock, 3) k_coarse_mfma(const CoarseTile *__restrict__ tiles, int n_tiles,
                                                          const int *__restrict__ or_off,
                                                          const int *__restrict__ or_cnt,
                                                          const int *__restrict__ or_proj,
                                                          const float2 *__restrict__ pf,
                                                          const float2 *__restrict__ w,
                                                          const float *__restrict__ ctf2,
                                                          const double *__restrict__ ynorm,
                                                          const unsigned int *__restrict__ ehi,
                                                          const unsigned int *__restrict__ elo,
                                                          const float *__restrict__ xi2half,
                                                          const long long *__restrict__ d2_off, int NT, int Kp,
                                                          const float2 *__restrict__ zero_row,
                                                          float *__restrict__ diff2, WPart *__restrict__ part) {
    constexpr int kTW = (NTT + 3) / 4;  // translation tiles per wave
    constexpr int kAStride = 40;        // bf16 per staged A row (32 + 8 padding against bank conflicts)
    __shared__ int og_s[kMfG][kMfO];
    __shared__ float a_norm[kMfG][kMfO];
    // A operand of one K step (hi and lo planes of the block's 128 rows), double buffered
    __shared__ __align__(16) unsigned short s_a[2][2][kMfG * kMfO * kAStride];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int grp = wave >> 2, wv = wave & 3, gtid = tid & 255;
    const int ti = kMfG * blockIdx.x + grp;
    const bool has = ti < n_tiles;
    const CoarseTile tile = has ? tiles[ti] : CoarseTile{0, 0};
    const int b = tile.b;
    const int n_o = has ? or_cnt[b] : 0;
    const int n_rows = has ? min(kMfO, n_o - tile.o_begin) : 0;  // valid orientations of the tile
    const int Kc = Kp / 2;  // complex slots per row
    if (gtid < kMfO) og_s[grp][gtid] = gtid < n_rows ? or_proj[or_off[b] + tile.o_begin + gtid] : -1;
    __syncthreads();
    floatx4_t acc[4][kTW];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < kTW; ++j) acc[i][j] = floatx4_t{0.f, 0.f, 0.f, 0.f};
    const int r16 = lane & 15, gq = lane >> 4;
    const int row_tiles = (n_rows + 15) >> 4;  // row tiles of this wave's group with orientations (uniform)
    // staging: thread gtid of a group forms the 4 complex slots 16 ks + 4 sq of its tile's row gtid / 4 in K step
    // ks (8 bf16 of hi and of lo), its projection values loaded one K step ahead
    const int srow = gtid >> 2, sq = gtid & 3, arow = grp * kMfO + srow;
    const float4 *st_p;
    {
        const int og = og_s[grp][srow];
        st_p = reinterpret_cast<const float4 *>(og >= 0 ? pf + (size_t)og * Kc : zero_row);
    }
    const float4 *wrow = reinterpret_cast<const float4 *>(w + (size_t)b * Kc);
    const float4 *c2row = reinterpret_cast<const float4 *>(ctf2 + (size_t)b * Kc);
    float an = 0.f;  // this thread's part of sum_p c_p^2 |P_o(p)|^2 over its row's slots
    auto stage = [&](float4 p0, float4 p1, int ks, int to) {
        const float4 w0 = wrow[8 * ks + 2 * sq], w1 = wrow[8 * ks + 2 * sq + 1];
        const float4 c2 = c2row[4 * ks + sq];
        an = fmaf(p0.x * p0.x + p0.y * p0.y, c2.x, an);
        an = fmaf(p0.z * p0.z + p0.w * p0.w, c2.y, an);
        an = fmaf(p1.x * p1.x + p1.y * p1.y, c2.z, an);
        an = fmaf(p1.z * p1.z + p1.w * p1.w, c2.w, an);
        const float2 pv[4] = {make_float2(p0.x, p0.y), make_float2(p0.z, p0.w), make_float2(p1.x, p1.y),
                              make_float2(p1.z, p1.w)};
        const float2 wq[4] = {make_float2(w0.x, w0.y), make_float2(w0.z, w0.w), make_float2(w1.x, w1.y),
                              make_float2(w1.z, w1.w)};
        unsigned int hh[4], ll[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {  // P conj(w)
            const float re = pv[m].x * wq[m].x + pv[m].y * wq[m].y;
            const float im = pv[m].y * wq[m].x - pv[m].x * wq[m].y;
            split2_bf16(re, im, hh[m], ll[m]);
        }
        *reinterpret_cast<uint4 *>(&s_a[to][0][arow * kAStride + 8 * sq]) = make_uint4(hh[0], hh[1], hh[2], hh[3]);
        *reinterpret_cast<uint4 *>(&s_a[to][1][arow * kAStride + 8 * sq]) = make_uint4(ll[0], ll[1], ll[2], ll[3]);
    };
    int tt[kTW];
#pragma unroll
    for (int j = 0; j < kTW; ++j) {
        const int t = 16 * (wv + 4 * j) + r16;
        tt[j] = ((wv + 4 * j) < NTT && t < NT) ? t : -1;
    }
    // projection values of kPre K steps in flight: register set d holds step ks with ks % kPre == d
    constexpr int kPre = 3;
    const int nK = Kp / 32;
    float4 pa[kPre][2];
#pragma unroll
    for (int d = 0; d < kPre; ++d) {
        pa[d][0] = d < nK ? st_p[8 * d + 2 * sq] : make_float4(0.f, 0.f, 0.f, 0.f);
        pa[d][1] = d < nK ? st_p[8 * d + 2 * sq + 1] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    stage(pa[0][0], pa[0][1], 0, 0);
    __syncthreads();
    for (int ks0 = 0; ks0 < nK; ks0 += kPre) {
#pragma unroll
        for (int d = 0; d < kPre; ++d) {
            const int ks = ks0 + d, buf = ks & 1;
            if (ks >= nK) break;    // uniform over the block
            if (ks + kPre < nK) {  // set d was staged at the end of step ks - 1: reload it for step ks + kPre
                pa[d][0] = st_p[8 * (ks + kPre) + 2 * sq];
                pa[d][1] = st_p[8 * (ks + kPre) + 2 * sq + 1];
            }

            if (row_tiles > 0) {
                bf16x8_t ah[4], al[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = grp * kMfO + 16 * i + r16;
                    ah[i] = __builtin_bit_cast(
                        bf16x8_t, *reinterpret_cast<const uint4 *>(&s_a[buf][0][row * kAStride + 8 * gq]));
                    al[i] = __builtin_bit_cast(
                        bf16x8_t, *reinterpret_cast<const uint4 *>(&s_a[buf][1][row * kAStride + 8 * gq]));
                }
                const int slot = 16 * ks + 4 * gq;  // this lane's 4 complex slots of the B fragment
#pragma unroll
                for (int j = 0; j < kTW; ++j) {
                    if (wv + 4 * j >= NTT) continue;
                    uint4 hv = make_uint4(0u, 0u, 0u, 0u), lv = hv;
                    if (tt[j] >= 0) {
                        hv = *reinterpret_cast<const uint4 *>(ehi + (size_t)tt[j] * Kc + slot);
                        lv = *reinterpret_cast<const uint4 *>(elo + (size_t)tt[j] * Kc + slot);
                    }
                    const bf16x8_t bh = __builtin_bit_cast(bf16x8_t, hv);
                    const bf16x8_t bl = __builtin_bit_cast(bf16x8_t, lv);
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        if (i >= row_tiles) continue;
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh, acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl, acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh, acc[i][j], 0, 0, 0);
                    }
                }
            }
            // the other buffer was last read in the previous step, before the barrier that ended it
            if (ks + 1 < nK) stage(pa[(d + 1) % kPre][0], pa[(d + 1) % kPre][1], ks + 1, buf ^ 1);
            __syncthreads();
        }
    }
    an += __shfl_xor(an, 1);  // the 4 threads of a row are consecutive lanes
    an += __shfl_xor(an, 2);
    if (sq == 0) a_norm[grp][srow] = an;
    __syncthreads();
    if (!has) return;
    const float base = xi2half[b] + (float)ynorm[b];
    float *out = diff2 + d2_off[b];
    float mn = 3.0e38f;  // minimum of the non-negative values written, for the weights' exponent offset
#pragma unroll
    for (int j = 0; j < kTW; ++j) {
        if (tt[j] < 0) continue;
        const int t = tt[j];  // D column
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int oo = 16 * i + gq * 4 + r;  // D row within the tile
                if (oo >= n_rows) continue;
                const float v = og_s[grp][oo] < 0 ? -999.f : base + a_norm[grp][oo] - 2.f * acc[i][j][r];
                out[(size_t)(tile.o_begin + oo) * NT + t] = v;
                if (v >= 0.f) mn = fminf(mn, v);
            }
    }
    for (int o = 32; o > 0; o >>= 1) mn = fminf(mn, __shfl_xor(mn, o));
    // sign bit cleared: -0 orders as +0 among the unsigned bit patterns of non-negative floats
    if (lane == 0 && mn < 3.0e38f) atomicMin(&part[b].mn_bits, __float_as_uint(mn) & 0x7fffffffu);
}

// --------------------------------------------------------------------------------------------------
// Weights and significance (convertAllSquaredDifferencesToWeights, src/ml_optimiser.cpp:7704-8239)
// --------------------------------------------------------------------------------------------------

struct PassStats {
    double min_diff2;
    double sum_weight;
    double threshold;      // exp_significant_weight
    double max_weight;
    long long nr_sign;     // my_nr_significant_coarse_samples
    long long n_above;     // entries with weight >= threshold
    long long best_key;    // tie-break key of the max weight (wsum loop order)
    long long best_index;  // entry index of the max weight
    int status;            // 0 ok, RLN_ERR_* otherwise
    int pad_;
};

constexpr int kWBlock = 1024;
constexpr int kCand = 2048;  // threshold candidates sorted in LDS

__device__ __forceinline__ double block_sum_d(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0.;
    if (threadIdx.x < 64) {
        t = (threadIdx.x < (blockDim.x >> 6)) ? red[threadIdx.x] : 0.;
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
        if (threadIdx.x == 0) red[0] = t;
    }
    __syncthreads();
    const double r = red[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ float block_min_f(float v, float *red) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x < 64) {
        float t = (threadIdx.x < (blockDim.x >> 6)) ? red[threadIdx.x] : 3.0e38f;
        for (int o = 32; o > 0; o >>= 1) t = fminf(t, __shfl_xor(t, o));
        if (threadIdx.x == 0) red[0] = t;
    }
    __syncthreads();
    const float r = red[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ double block_min_d(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x < 64) {
        double t = (threadIdx.x < (blockDim.x >> 6)) ? red[threadIdx.x] : 1.0e300;
        for (int o = 32; o > 0; o >>= 1) t = fmin(t, __shfl_xor(t, o));
        if (threadIdx.x == 0) red[0] = t;
    }
    __syncthreads();
    const double r = red[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ long long block_sum_ll(long long v, long long *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x < 64) {
        long long t = (threadIdx.x < (blockDim.x >> 6)) ? red[threadIdx.x] : 0;
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
        if (threadIdx.x == 0) red[0] = t;
    }
    __syncthreads();
    const long long r = red[0];
    __syncthreads();
    return r;
}

// Device-side header of a chunk's fine pass, written by k_fine_plan / k_fine_items and read by the persistent
// kernels, so that the fine pass needs no host round trip. When a count exceeds the capacity the host sized
// the buffers for, overflow is set, every later kernel of the pass returns at once, and the host grows the
// buffers and runs the pass again.
struct FinePlan {
    long long n_pairs;   // significant coarse poses of the chunk
    int n_jobs;          // (particle, orientation) jobs
    int n_small, n_wide; // work items of k_fine_diff2<2> and k_fine_diff2_wide
    int overflow;
    int B;               // particles in the chunk
    int next_fine;       // next work item of k_fine_diff2 (dynamic fetch; zeroed by k_fine_plan)
    int next_wsum;       // next work item of k_wsum_bp
    int next_fine_x[8];  // next work unit of each XCD's share (k_fine_diff2 with per-XCD queues)
    int next_wsum_x[8];  // the same for k_wsum_bp
    int pad_;
};

// Per-chunk input images (device pointers), read by the fine-pass kernels through one indirection so that a
// captured graph of the pass serves every chunk.
struct ChunkPtrs {
    const float2 *fimg;
    const float2 *fimg_nomask;
    const float *fctf;
    int B;  // particles in the chunk
    int pad_;
};

// Mode of k_weights: coarse entries are e = o*NT + t; fine entries are e = s*32 + r*4 + u where pair s
// holds the coarse index (o*NT + t) in pairs[].
struct WeightArgs {
    const float *diff2;          // coarse mode: per particle at d2_off[b]; entries < 0 were not evaluated
    const double *diff2d;        // fine mode: per particle at d2_off[b], every entry evaluated
    const long long *d2_off;
    const long long *n_entries;  // per particle
    const double *or_prior;      // normalised pdf_orientation per orientation entry (or_off based)
    const int *or_off;
    const int *oc_off;          // per-particle offsets of per-orientation state (ocnt)
    const double *tr_prior;      // normalised pdf_offset [b][NT]
    const int *pairs;            // fine mode: coarse index per pair, at pair_off[b]
    const long long *pair_off;
    int NT;
    int fine;                    // 0 coarse, 1 fine
    double adaptive_fraction;
    long long maxsig;            // <= 0: none (coarse pass only)
    double *weight;              // out, same indexing as diff2
    unsigned char *mask;         // out (coarse mode), weight >= threshold
    PassStats *stats;            // out [b]
    const FinePlan *plan;        // fine mode: chunk header (overflow: skip; blocks >= plan->B: idle); else null
};

__device__ __forceinline__ double entry_prior(const WeightArgs &a, int b, long long e) {
    long long c;
    if (a.fine) {
        c = a.pairs[a.pair_off[b] + (e >> 5)];
    } else {
        c = e;
    }
    const long long o = c / a.NT, t = c - o * a.NT;
    return a.or_prior[a.or_off[b] + o] * a.tr_prior[(size_t)b * a.NT + t];
}

// Key that orders entries like storeWeightedSums visits them (class, direction, psi, iover_rot, itrans,
// iover_trans; src/ml_optimiser.cpp:8329-8470), used to break ties of the max weight as the reference does.
__device__ __forceinline__ long long wsum_order_key(const WeightArgs &a, int b, long long e) {
    if (!a.fine) return e;
    const long long s = e >> 5;
    const int r = (int)((e >> 2) & 7), u = (int)(e & 3);
    const long long c = a.pairs[a.pair_off[b] + s];
    const long long o = c / a.NT, t = c - o * a.NT;
    return ((o * 8 + r) * a.NT + t) * 4 + u;
}

// One 1024-thread block per particle: min diff2, weights in FP64, their sum, the significance threshold by
// radix selection over the weights' bit patterns (positive doubles order like their uint64 patterns), the
// coarse significance mask and the max weight. The threshold is the weight at which the descending
// cumulative sum first exceeds adaptive_fraction * sum (src/ml_optimiser.cpp:8145-8168); maxsig caps the
// count as in :8147-8157.
__global__ void __launch_bounds__(kWBlock) k_weights(WeightArgs a) {
    __shared__ double red_d[16];
    __shared__ float red_f[16];
    __shared__ long long red_l[16];
    __shared__ unsigned int hcnt[256];
    __shared__ double hsum[256];
    __shared__ unsigned long long s_prefix, s_mask;
    __shared__ double s_cand[kCand];
    __shared__ int s_ncand, s_cand_ok;
    __shared__ double s_cand_thr;
    __shared__ long long s_cand_nsig;
    __shared__ double s_above_sum;
    __shared__ long long s_above_cnt;
    __shared__ int s_done;
    const int b = blockIdx.x;
    if (a.plan && (a.plan->overflow || b >= a.plan->B)) return;
    const long long n = a.n_entries[b];
    double *w = a.weight + a.d2_off[b];
    const int tid = threadIdx.x;

    // fine entries are FP64 and all evaluated; coarse entries FP32, < 0 where not evaluated
    const float *d2 = a.fine ? nullptr : a.diff2 + a.d2_off[b];
    const double *d2d = a.fine ? a.diff2d + a.d2_off[b] : nullptr;
    double mind;
    if (a.fine) {
        double mnd = 1.0e300;
        for (long long e = tid; e < n; e += kWBlock) mnd = fmin(mnd, d2d[e]);
        mind = block_min_d(mnd, red_d);
    } else {
        float mn = 3.0e38f;
        for (long long e = tid; e < n; e += kWBlock) {
            const float v = d2[e];
            if (v >= 0.f) mn = fminf(mn, v);
        }
        mind = (double)block_min_f(mn, red_f);
    }
    double sum = 0.;
    for (long long e = tid; e < n; e += kWBlock) {
        const double v = a.fine ? d2d[e] : (double)d2[e];
        double wt = 0.;
        if (a.fine || v >= 0.) {
            const double diff = v - mind;
            wt = entry_prior(a, b, e);
            wt = (diff > 700.) ? 0. : wt * exp(-diff);
        }
        w[e] = wt;
        sum += wt;
    }
    sum = block_sum_d(sum, red_d);
    PassStats st;
    st.min_diff2 = mind;
    st.sum_weight = sum;
    st.threshold = 0.;
    st.max_weight = 0.;
    st.nr_sign = 0;
    st.n_above = 0;
    st.best_key = -1;
    st.best_index = -1;
    st.status = 0;
    st.pad_ = 0;
    if (!(sum > 0.)) {  // zero or NaN (src/ml_optimiser.cpp:8054-8104)
        if (tid == 0) {
            st.status = -6;
            a.stats[b] = st;
        }
        return;
    }

    const double target = a.adaptive_fraction * sum;
    double thr = 0.;
    long long nr_sign = 0;
    bool have_thr = false;
    // Candidates: only weights >= (1 - f) * sum / n can be the threshold, since all smaller weights together
    // hold less than the (1 - f) tail. When they are few, sort them in LDS and accumulate in descending order
    // exactly as the reference's loop does (src/ml_optimiser.cpp:8145-8168).
    {
        const double xmin = (1. - a.adaptive_fraction) * sum / (double)(n > 0 ? n : 1);
        if (tid == 0) s_ncand = 0;
        __syncthreads();
        for (long long e = tid; e < n; e += kWBlock) {
            const double v = w[e];
            if (v > 0. && v >= xmin) {
                const int slot = atomicAdd(&s_ncand, 1);
                if (slot < kCand) s_cand[slot] = v;
            }
        }
        __syncthreads();
        const int nc = s_ncand;
        if (nc <= kCand) {
            // sort size: the next power of two holding the candidates (uniform), padded with -1 below every weight
            int n2 = 2;
            while (n2 < nc) n2 <<= 1;
            for (int i = tid; i < n2; i += kWBlock)
                if (i >= nc) s_cand[i] = -1.;
            __syncthreads();
            // bitonic sort, ascending
            for (int ks = 2; ks <= n2; ks <<= 1)
                for (int j = ks >> 1; j > 0; j >>= 1) {
                    for (int i = tid; i < n2; i += kWBlock) {
                        const int l = i ^ j;
                        if (l > i) {
                            const double x = s_cand[i], y = s_cand[l];
                            if (((i & ks) == 0) ? (x > y) : (x < y)) {
                                s_cand[i] = y;
                                s_cand[l] = x;
                            }
                        }
                    }
                    __syncthreads();
                }
            // descending cumulative sum by wave 0, 64 candidates per step (inclusive lane scan)
            if (tid < 64) {
                double base = 0.;
                int crossed = 0;
                long long rank = -1;
                for (int r0 = 0; r0 < nc; r0 += 64) {
                    const int r = r0 + tid;
                    const double v = (r < nc) ? s_cand[n2 - 1 - r] : 0.;
                    double x = v;
                    for (int o = 1; o < 64; o <<= 1) {
                        const double y = __shfl_up(x, o);
                        if (tid >= o) x += y;
                    }
                    const double cum = base + x;
                    const unsigned long long hit = __ballot(r < nc && cum > target);
                    if (hit) {
                        crossed = 1;
                        rank = r0 + (__ffsll((long long)hit) - 1);
                        break;
                    }
                    base = __shfl(cum, 63);
                }
                if (tid == 0) {
                    long long nsig = rank + 1;
                    if (crossed && a.maxsig > 0 && !a.fine && nsig > a.maxsig) nsig = a.maxsig;
                    s_cand_ok = crossed;
                    s_cand_nsig = nsig;
                    s_cand_thr = crossed ? s_cand[n2 - nsig] : 0.;
                }
            }
            __syncthreads();
            if (s_cand_ok) {
                have_thr = true;
                thr = s_cand_thr;
                nr_sign = s_cand_nsig;
            }
        }
    }
    if (!have_thr) {
    // radix selection of the crossing element: 8 digits of 8 bits from the most significant
    if (tid == 0) {
        s_prefix = 0ULL;
        s_mask = 0ULL;
        s_above_sum = 0.;
        s_above_cnt = 0;
        s_done = 0;
    }
    __syncthreads();
    for (int shift = 56; shift >= 0; shift -= 8) {
        for (int i = tid; i < 256; i += kWBlock) {
            hcnt[i] = 0;
            hsum[i] = 0.;
        }
        __syncthreads();
        const unsigned long long pre = s_prefix, msk = s_mask;
        for (long long e = tid; e < n; e += kWBlock) {
            const double v = w[e];
            if (!(v > 0.)) continue;
            const unsigned long long key = (unsigned long long)__double_as_longlong(v);
            if ((key & msk) != pre) continue;
            const int bin = (int)((key >> shift) & 255ULL);
            atomicAdd(&hcnt[bin], 1u);
            atomicAdd(&hsum[bin], v);
        }
        __syncthreads();
        if (tid == 0) {
            double acc = s_above_sum;
            long long cnt = s_above_cnt;
            int bin = -1;
            for (int k = 255; k >= 0; --k) {
                if (hcnt[k] == 0) continue;
                if (acc + hsum[k] > target) {
                    bin = k;
                    break;
                }
                acc += hsum[k];
                cnt += hcnt[k];
            }
            if (bin < 0) {  // rounding: the total never exceeds the target; take the smallest present bin
                for (int k = 0; k < 256; ++k)
                    if (hcnt[k]) {
                        bin = k;
                        break;
                    }
                // undo its accumulation
                acc -= hsum[bin];
                cnt -= hcnt[bin];
            }
            s_above_sum = acc;
            s_above_cnt = cnt;
            s_prefix |= ((unsigned long long)bin << shift);
            s_mask |= (255ULL << shift);
        }
        __syncthreads();
    }
    thr = __longlong_as_double((long long)s_prefix);
    // elements equal to thr needed to cross the target
    {
        const double rem = target - s_above_sum;
        long long m = (long long)floor(rem / thr) + 1;
        if (m < 1) m = 1;
        nr_sign = s_above_cnt + m;
    }
    if (a.maxsig > 0 && !a.fine && nr_sign > a.maxsig) {
        // rank selection of the maxsig-th largest weight
        if (tid == 0) {
            s_prefix = 0ULL;
            s_mask = 0ULL;
            s_above_cnt = 0;
        }
        __syncthreads();
        for (int shift = 56; shift >= 0; shift -= 8) {
            for (int i = tid; i < 256; i += kWBlock) hcnt[i] = 0;
            __syncthreads();
            const unsigned long long pre = s_prefix, msk = s_mask;
            for (long long e = tid; e < n; e += kWBlock) {
                const double v = w[e];
                if (!(v > 0.)) continue;
                const unsigned long long key = (unsigned long long)__double_as_longlong(v);
                if ((key & msk) != pre) continue;
                atomicAdd(&hcnt[(int)((key >> shift) & 255ULL)], 1u);
            }
            __syncthreads();
            if (tid == 0) {
                long long cnt = s_above_cnt;
                int bin = 0;
                for (int k = 255; k >= 0; --k) {
                    if (cnt + (long long)hcnt[k] >= a.maxsig) {
                        bin = k;
                        break;
                    }
                    cnt += hcnt[k];
                }
                s_above_cnt = cnt;
                s_prefix |= ((unsigned long long)bin << shift);
                s_mask |= (255ULL << shift);
            }
            __syncthreads();
        }
        thr = __longlong_as_double((long long)s_prefix);
        nr_sign = a.maxsig;
    }
    }  // radix selection
    // mask, count above threshold, max weight with the reference's loop-order tie break
    long long above = 0;
    double bw = -1.;
    long long bkey = 0x7fffffffffffffffLL, bidx = -1;
    for (long long e = tid; e < n; e += kWBlock) {
        const double v = w[e];
        const bool sig = v >= thr && v > 0.;
        if (!a.fine) a.mask[a.d2_off[b] + e] = sig ? 1 : 0;
        above += sig ? 1 : 0;
        if (v > 0.) {
            const long long key = wsum_order_key(a, b, e);
            if (v > bw || (v == bw && key < bkey)) {
                bw = v;
                bkey = key;
                bidx = e;
            }
        }
    }
    above = block_sum_ll(above, red_l);
    // arg-max across the block
    for (int o = 32; o > 0; o >>= 1) {
        const double ow = __shfl_xor(bw, o);
        const long long ok = __shfl_xor(bkey, o), oi = __shfl_xor(bidx, o);
        if (ow > bw || (ow == bw && ok < bkey)) {
            bw = ow;
            bkey = ok;
            bidx = oi;
        }
    }
    __shared__ double s_bw[16];
    __shared__ long long s_bk[16], s_bi[16];
    if ((tid & 63) == 0) {
        s_bw[tid >> 6] = bw;
        s_bk[tid >> 6] = bkey;
        s_bi[tid >> 6] = bidx;
    }
    __syncthreads();
    if (tid == 0) {
        for (int i = 1; i < (kWBlock >> 6); ++i)
            if (s_bw[i] > bw || (s_bw[i] == bw && s_bk[i] < bkey)) {
                bw = s_bw[i];
                bkey = s_bk[i];
                bidx = s_bi[i];
            }
        st.threshold = thr;
        st.nr_sign = nr_sign;
        st.n_above = above;
        st.max_weight = bw;
        st.best_key = bkey;
        st.best_index = bidx;
        if (!a.fine && nr_sign == 0) st.status = -5;
        a.stats[b] = st;
    }
}

// --------------------------------------------------------------------------------------------------
// Coarse weights over segments: the same results as k_weights (min, FP64 weights, sum, significance
// threshold, mask, max weight), with every pass over the entries split into blocks of kSegEntries, so that a
// particle with millions of coarse poses (an exhaustive search) is spread over the whole GPU. The threshold is
// found by a multi-block radix selection over the weights' bit patterns (8-bit digits from the top; positive
// doubles order like their uint64 patterns) until the crossing digit bin holds at most kGather elements, which
// are then sorted in LDS and scanned exactly.
// --------------------------------------------------------------------------------------------------
constexpr long long kSegEntries = 32768;  // entries per segment block
constexpr int kGather = 4096;             // crossing-bin elements sorted exactly

struct WSeg {
    int b;             // particle
    int pad_;
    long long begin;   // entry range [begin, end) within the particle
    long long end;
};


struct WSegBest {
    double w;
    long long key, idx;
};

__global__ void __launch_bounds__(kBlock) k_wpart_init(WPart *__restrict__ part, int B) {
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= B) return;
    WPart p{};
    p.mn_bits = 0x7f800000u;  // +inf
    part[b] = p;
}

__global__ void __launch_bounds__(kBlock) k_wseg_min(WeightArgs a, const WSeg *__restrict__ segs,
                                                     WPart *__restrict__ part) {
    __shared__ float red[kBlock / 64];
    const WSeg sg = segs[blockIdx.x];
    const float *d2 = a.diff2 + a.d2_off[sg.b];
    float mn = 3.0e38f;
    for (long long e = sg.begin + threadIdx.x; e < sg.end; e += kBlock) {
        const float v = d2[e];
        if (v >= 0.f) mn = fminf(mn, v);
    }
    for (int o = 32; o > 0; o >>= 1) mn = fminf(mn, __shfl_xor(mn, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mn;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) mn = fminf(mn, red[w]);
        if (mn < 3.0e38f) atomicMin(&part[sg.b].mn_bits, __float_as_uint(mn) & 0x7fffffffu);
    }
}

// Weights of one segment, their sum, and the first radix pass of the selection (the top byte of every positive
// weight: the pass before any prefix is fixed), accumulated here instead of in a separate read of the weights.
__global__ void __launch_bounds__(kBlock) k_wseg_exp(WeightArgs a, const WSeg *__restrict__ segs,
                                                     WPart *__restrict__ part, unsigned int *__restrict__ hcnt,
                                                     double *__restrict__ hsum) {
    __shared__ double red[kBlock / 64];
    __shared__ unsigned int c[256];
    __shared__ double sm[256];
    c[threadIdx.x] = 0;
    sm[threadIdx.x] = 0.;
    const WSeg sg = segs[blockIdx.x];
    const int b = sg.b;
    const float *d2 = a.diff2 + a.d2_off[b];
    double *w = a.weight + a.d2_off[b];
    const double mind = (double)__uint_as_float(part[b].mn_bits);
    __syncthreads();
    double sum = 0.;
    for (long long e = sg.begin + threadIdx.x; e < sg.end; e += kBlock) {
        const float v = d2[e];
        double wt = 0.;
        if (v >= 0.f) {
            const double diff = (double)v - mind;
            wt = entry_prior(a, b, e);
            wt = (diff > 700.) ? 0. : wt * exp(-diff);
        }
        w[e] = wt;
        sum += wt;
        if (wt > 0.) {
            const int bin = (int)((unsigned long long)__double_as_longlong(wt) >> 56);
            atomicAdd(&c[bin], 1u);
            atomicAdd(&sm[bin], wt);
        }
    }
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kBlock / 64; ++i) sum += red[i];
        atomicAdd(&part[b].sum, sum);
    }
    if (c[threadIdx.x]) {
        atomicAdd(&hcnt[(size_t)b * 256 + threadIdx.x], c[threadIdx.x]);
        atomicAdd(&hsum[(size_t)b * 256 + threadIdx.x], sm[threadIdx.x]);
    }
}

// per particle: status and selection start
__global__ void __launch_bounds__(kBlock) k_wsel_start(WeightArgs a, WPart *__restrict__ part, int B) {
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= B) return;
    WPart &p = part[b];
    if (!(p.sum > 0.)) {
        p.status = -6;  // zero or NaN sum (src/ml_optimiser.cpp:8054-8104)
        p.gather = 2;   // no selection
        return;
    }
    p.target = a.adaptive_fraction * p.sum;
}

// one radix pass: per-particle digit histograms (count, sum) of the weights that match the prefix
__global__ void __launch_bounds__(kBlock) k_wseg_hist(WeightArgs a, const WSeg *__restrict__ segs,
                                                      const WPart *__restrict__ part, int shift,
                                                      unsigned int *__restrict__ hcnt, double *__restrict__ hsum) {
    __shared__ unsigned int c[256];
    __shared__ double sm[256];
    const WSeg sg = segs[blockIdx.x];
    const WPart p = part[sg.b];
    if (p.gather) return;
    c[threadIdx.x] = 0;
    sm[threadIdx.x] = 0.;
    __syncthreads();
    const double *w = a.weight + a.d2_off[sg.b];
    for (long long e = sg.begin + threadIdx.x; e < sg.end; e += kBlock) {
        const double v = w[e];
        if (!(v > 0.)) continue;
        const unsigned long long key = (unsigned long long)__double_as_longlong(v);
        if ((key & p.mask) != p.prefix) continue;
        const int bin = (int)((key >> shift) & 255ULL);
        atomicAdd(&c[bin], 1u);
        atomicAdd(&sm[bin], v);
    }
    __syncthreads();
    if (c[threadIdx.x]) {
        atomicAdd(&hcnt[(size_t)sg.b * 256 + threadIdx.x], c[threadIdx.x]);
        atomicAdd(&hsum[(size_t)sg.b * 256 + threadIdx.x], sm[threadIdx.x]);
    }
}

// per particle: the digit bin where the descending cumulative sum crosses the target; histograms cleared
__global__ void __launch_bounds__(kBlock) k_wsel_pick(WPart *__restrict__ part, int B, int shift,
                                                      unsigned int *__restrict__ hcnt, double *__restrict__ hsum) {
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= B) return;
    WPart &p = part[b];
    unsigned int *hc = hcnt + (size_t)b * 256;
    double *hs = hsum + (size_t)b * 256;
    if (!p.gather) {
        double acc = p.above_sum;
        long long cnt = p.above_cnt;
        int bin = -1;
        for (int k = 255; k >= 0; --k) {
            if (hc[k] == 0) continue;
            if (acc + hs[k] > p.target) {
                bin = k;
                break;
            }
            acc += hs[k];
            cnt += hc[k];
        }
        if (bin < 0) {  // rounding: the total never exceeds the target; take the smallest present bin
            for (int k = 0; k < 256; ++k)
                if (hc[k]) {
                    bin = k;
                    break;
                }
            acc -= hs[bin];
            cnt -= hc[bin];
        }
        p.above_sum = acc;
        p.above_cnt = cnt;
        p.prefix |= ((unsigned long long)bin << shift);
        p.mask |= (255ULL << shift);
        if (hc[bin] <= (unsigned)kGather || shift == 0) p.gather = 1;
    }
    for (int k = 0; k < 256; ++k) {
        hc[k] = 0;
        hs[k] = 0.;
    }
}

// elements of the crossing bin (weights matching the prefix) into the per-particle candidate buffer
__global__ void __launch_bounds__(kBlock) k_wseg_gather(WeightArgs a, const WSeg *__restrict__ segs,
                                                        WPart *__restrict__ part, double *__restrict__ cand) {
    const WSeg sg = segs[blockIdx.x];
    const WPart p = part[sg.b];
    if (p.gather != 1) return;
    const double *w = a.weight + a.d2_off[sg.b];
    for (long long e = sg.begin + threadIdx.x; e < sg.end; e += kBlock) {
        const double v = w[e];
        if (!(v > 0.)) continue;
        const unsigned long long key = (unsigned long long)__double_as_longlong(v);
        if ((key & p.mask) != p.prefix) continue;
        const int slot = atomicAdd(&part[sg.b].ncand, 1);
        if (slot < kGather) cand[(size_t)sg.b * kGather + slot] = v;
    }
}

// per particle (one block): exact threshold from the sorted crossing-bin elements, maxsig cap by rank
__global__ void __launch_bounds__(kWBlock) k_wsel_final(WeightArgs a, WPart *__restrict__ part,
                                                        const double *__restrict__ cand) {
    __shared__ double s_c[kGather];
    __shared__ long long s_rank;
    __shared__ int s_ok;
    const int b = blockIdx.x, tid = threadIdx.x;
    WPart &p = part[b];
    if (p.gather == 2) return;
    const int nc = min(p.ncand, kGather);
    // sort size: the next power of two holding the candidates (uniform over the block), padded with -1 below
    // every weight
    int n2 = 2;
    while (n2 < nc) n2 <<= 1;
    for (int i = tid; i < n2; i += kWBlock) s_c[i] = (i < nc) ? cand[(size_t)b * kGather + i] : -1.;
    __syncthreads();
    for (int ks = 2; ks <= n2; ks <<= 1)  // bitonic sort, ascending
        for (int j = ks >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < n2; i += kWBlock) {
                const int l = i ^ j;
                if (l > i) {
                    const double x = s_c[i], y = s_c[l];
                    if (((i & ks) == 0) ? (x > y) : (x < y)) {
                        s_c[i] = y;
                        s_c[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    // descending cumulative sum from above_sum, 64 elements per step (wave 0)
    if (tid < 64) {
        double base = p.above_sum;
        int crossed = 0;
        long long rank = nc - 1;
        for (int r0 = 0; r0 < nc; r0 += 64) {
            const int r = r0 + tid;
            double x = (r < nc) ? s_c[n2 - 1 - r] : 0.;
            for (int o = 1; o < 64; o <<= 1) {
                const double y = __shfl_up(x, o);
                if (tid >= o) x += y;
            }
            const double cum = base + x;
            const unsigned long long hit = __ballot(r < nc && cum > p.target);
            if (hit) {
                crossed = 1;
                rank = r0 + (__ffsll((long long)hit) - 1);
                break;
            }
            base = __shfl(cum, 63);
        }
        if (tid == 0) {
            s_rank = rank;
            s_ok = crossed;
        }
    }
    __syncthreads();
    if (tid != 0) return;
    double thr;
    long long nr_sign;
    if (p.ncand <= kGather && nc > 0) {
        // rank within the bin; without a crossing (rounding) the smallest element, as the reference's loop ends
        const long long rank = s_ok ? s_rank : (long long)nc - 1;
        thr = s_c[n2 - 1 - rank];
        nr_sign = p.above_cnt + rank + 1;
    } else {
        // all 64 bits fixed and more than kGather equal elements: the bin is one value
        thr = __longlong_as_double((long long)p.prefix);
        long long m = (long long)floor((p.target - p.above_sum) / thr) + 1;
        if (m < 1) m = 1;
        nr_sign = p.above_cnt + m;
    }
    p.thr = thr;
    p.nr_sign = nr_sign;
}

// maxsig cap (src/ml_optimiser.cpp:8147-8157): the maxsig-th largest weight becomes the threshold. One block
// per capped particle, radix selection by counts over the particle's entries.
__global__ void __launch_bounds__(kWBlock) k_wsel_maxsig(WeightArgs a, WPart *__restrict__ part) {
    __shared__ unsigned int hc[256];
    __shared__ unsigned long long s_prefix, s_mask;
    __shared__ long long s_cnt;
    const int b = blockIdx.x, tid = threadIdx.x;
    WPart &p = part[b];
    if (p.gather == 2 || a.maxsig <= 0 || p.nr_sign <= a.maxsig) return;
    const long long n = a.n_entries[b];
    const double *w = a.weight + a.d2_off[b];
    if (tid == 0) {
        s_prefix = 0ULL;
        s_mask = 0ULL;
        s_cnt = 0;
    }
    __syncthreads();
    for (int shift = 56; shift >= 0; shift -= 8) {
        for (int i = tid; i < 256; i += kWBlock) hc[i] = 0;
        __syncthreads();
        const unsigned long long pre = s_prefix, msk = s_mask;
        for (long long e = tid; e < n; e += kWBlock) {
            const double v = w[e];
            if (!(v > 0.)) continue;
            const unsigned long long key = (unsigned long long)__double_as_longlong(v);
            if ((key & msk) != pre) continue;
            atomicAdd(&hc[(int)((key >> shift) & 255ULL)], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            long long cnt = s_cnt;
            int bin = 0;
            for (int k = 255; k >= 0; --k) {
                if (cnt + (long long)hc[k] >= a.maxsig) {
                    bin = k;
                    break;
                }
                cnt += hc[k];
            }
            s_cnt = cnt;
            s_prefix |= ((unsigned long long)bin << shift);
            s_mask |= (255ULL << shift);
        }
        __syncthreads();
    }
    if (tid == 0) {
        p.thr = __longlong_as_double((long long)s_prefix);
        p.nr_sign = a.maxsig;
    }
}

// mask, count above threshold, and each segment's max weight with the wsum-order tie break
// Also counts the significant translations of every orientation entry (ocnt) and, per particle, the orientations
// with at least one (njobs): the inputs of the fine-pass job lists.
__global__ void __launch_bounds__(kBlock) k_wseg_mask(WeightArgs a, const WSeg *__restrict__ segs,
                                                      WPart *__restrict__ part, WSegBest *__restrict__ best,
                                                      int *__restrict__ ocnt, int *__restrict__ njobs) {
    __shared__ double s_bw[kBlock / 64];
    __shared__ long long s_bk[kBlock / 64], s_bi[kBlock / 64], s_ab[kBlock / 64];
    const WSeg sg = segs[blockIdx.x];
    const int b = sg.b;
    const WPart p = part[b];
    const double thr = p.thr;
    const bool ok = p.gather != 2;
    const double *w = a.weight + a.d2_off[b];
    long long above = 0;
    double bw = -1.;
    long long bkey = 0x7fffffffffffffffLL, bidx = -1;
    for (long long e = sg.begin + threadIdx.x; e < sg.end; e += kBlock) {
        const double v = w[e];
        const bool sig = ok && v >= thr && v > 0.;
        a.mask[a.d2_off[b] + e] = sig ? 1 : 0;
        above += sig ? 1 : 0;
        if (sig) {
            const int o = (int)(e / a.NT);
            if (atomicAdd(&ocnt[a.oc_off[b] + o], 1) == 0) atomicAdd(&njobs[b], 1);
        }
        if (v > 0.) {
            const long long key = wsum_order_key(a, b, e);
            if (v > bw || (v == bw && key < bkey)) {
                bw = v;
                bkey = key;
                bidx = e;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        above += __shfl_xor(above, o);
        const double ow = __shfl_xor(bw, o);
        const long long ok2 = __shfl_xor(bkey, o), oi = __shfl_xor(bidx, o);
        if (ow > bw || (ow == bw && ok2 < bkey)) {
            bw = ow;
            bkey = ok2;
            bidx = oi;
        }
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_bw[wv] = bw;
        s_bk[wv] = bkey;
        s_bi[wv] = bidx;
        s_ab[wv] = above;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kBlock / 64; ++i) {
            above += s_ab[i];
            if (s_bw[i] > bw || (s_bw[i] == bw && s_bk[i] < bkey)) {
                bw = s_bw[i];
                bkey = s_bk[i];
                bidx = s_bi[i];
            }
        }
        if (above) atomicAdd((unsigned long long *)&part[b].n_above, (unsigned long long)above);
        best[blockIdx.x] = WSegBest{bw, bkey, bidx};
    }
}

// per particle: PassStats from the particle state and its segments' maxima
__global__ void __launch_bounds__(kBlock) k_wsel_stats(const WPart *__restrict__ part, const int *__restrict__ seg_first,
                                                       const WSegBest *__restrict__ best, int B, int fine,
                                                       PassStats *__restrict__ stats, int *__restrict__ npairs) {
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= B) return;
    const WPart p = part[b];
    npairs[b] = (int)p.n_above;
    PassStats st;
    st.min_diff2 = (double)__uint_as_float(p.mn_bits);
    st.sum_weight = p.sum;
    st.threshold = p.thr;
    st.nr_sign = p.nr_sign;
    st.n_above = p.n_above;
    st.status = p.status;
    st.pad_ = 0;
    double bw = -1.;
    long long bkey = 0x7fffffffffffffffLL, bidx = -1;
    for (int s = seg_first[b]; s < seg_first[b + 1]; ++s) {
        const WSegBest q = best[s];
        if (q.w > bw || (q.w == bw && q.key < bkey)) {
            bw = q.w;
            bkey = q.key;
            bidx = q.idx;
        }
    }
    st.max_weight = p.status ? 0. : bw;
    st.best_key = p.status ? -1 : bkey;
    st.best_index = p.status ? -1 : bidx;
    if (!p.status) {
        if (!fine && p.nr_sign == 0) st.status = -5;
    } else {
        st.threshold = 0.;
        st.nr_sign = 0;
        st.n_above = 0;
    }
    stats[b] = st;
}

// --------------------------------------------------------------------------------------------------
// Fine-pass job lists from the coarse mask (the reference builds them on the host:
// makeJobsForDiff2Fine, src/acc/acc_helper_functions_impl.h:26-102).
// A job is one coarse orientation with >= 1 significant translation; its pairs are consecutive in
// ihidden order. pairs[] holds the particle-local coarse index o*NT + t of every significant pair.
// --------------------------------------------------------------------------------------------------

struct FineJob {
    int b;          // particle
    int o;          // orientation entry within the particle's list
    int s_begin;    // first pair (particle-local)
    int s_count;    // number of significant translations
};


__device__ __forceinline__ int block_exclusive_scan_i(int v, int *tmp, int &total) {
    // wave-level inclusive scan
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    __syncthreads();
    if (lane == 63) tmp[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
            const int t = tmp[i];
            tmp[i] = acc;
            acc += t;
        }
        tmp[blockDim.x >> 6] = acc;
    }
    __syncthreads();
    const int r = x - v + tmp[w];
    total = tmp[blockDim.x >> 6];
    __syncthreads();
    return r;
}


// Offsets of the chunk's fine pass from the per-particle counts: exclusive scans over particles of pairs and
// jobs, fine-weight entry counts and offsets, totals and the capacity check. One block of 1024.
__global__ void __launch_bounds__(kWBlock) k_fine_plan(const int *__restrict__ n_pairs_b, const int *__restrict__ n_jobs_b,
                                                       const ChunkPtrs *__restrict__ cp, long long cap_pairs, int cap_jobs,
                                                       long long *__restrict__ pair_off, int *__restrict__ job_off,
                                                       long long *__restrict__ fine_n, long long *__restrict__ fine_off,
                                                       FinePlan *__restrict__ plan) {
    __shared__ int tmp[24];
    const int B = cp->B;
    long long pbase = 0;
    int jbase = 0;
    for (int b0 = 0; b0 < B; b0 += kWBlock) {
        const int b = b0 + threadIdx.x;
        const int np = b < B ? n_pairs_b[b] : 0, nj = b < B ? n_jobs_b[b] : 0;
        int tp, tj;
        const int ps = block_exclusive_scan_i(np, tmp, tp);
        const int js = block_exclusive_scan_i(nj, tmp, tj);
        if (b < B) {
            pair_off[b] = pbase + ps;
            job_off[b] = jbase + js;
            fine_n[b] = 32LL * np;
            fine_off[b] = 32LL * (pbase + ps);
        }
        pbase += tp;
        jbase += tj;
    }
    if (threadIdx.x == 0) {
        plan->n_pairs = pbase;
        plan->n_jobs = jbase;
        plan->n_small = plan->n_wide = 0;
        plan->next_fine = plan->next_wsum = 0;
        for (int x = 0; x < 8; ++x) plan->next_fine_x[x] = plan->next_wsum_x[x] = 0;
        plan->B = B;
        plan->overflow = (pbase > cap_pairs || jbase > cap_jobs) ? 1 : 0;
    }
}

// Writes pairs and jobs at the offsets computed by k_fine_plan from the per-orientation significant counts
// (ocnt, counted by k_wseg_mask). grid (B), block 1024.
__global__ void __launch_bounds__(kWBlock) k_write_jobs(const unsigned char *__restrict__ mask,
                                                        const long long *__restrict__ d2_off,
                                                        const int *__restrict__ or_cnt, const int *__restrict__ oc_off,
                                                        const int *__restrict__ ocnt, int NT,
                                                        const long long *__restrict__ pair_off,
                                                        const int *__restrict__ job_off, int *__restrict__ pairs,
                                                        FineJob *__restrict__ jobs, const FinePlan *__restrict__ plan) {
    __shared__ int tmp[24];
    const int b = blockIdx.x;
    if (plan && (plan->overflow || b >= plan->B)) return;
    const unsigned char *m = mask + d2_off[b];
    int pair_base = 0, job_base = 0;
    const int n_o = or_cnt[b];
    for (int o0 = 0; o0 < n_o; o0 += kWBlock) {
        const int o = o0 + threadIdx.x;
        const int c = (o < n_o) ? ocnt[oc_off[b] + o] : 0;  // significant translations of orientation o
        int tot_p, tot_j;
        const int ps = block_exclusive_scan_i(c, tmp, tot_p);
        const int js = block_exclusive_scan_i(c > 0 ? 1 : 0, tmp, tot_j);
        if (o < n_o && c > 0) {
            int s = pair_base + ps;
            FineJob j;
            j.b = b;
            j.o = o;
            j.s_begin = s;
            j.s_count = c;
            jobs[job_off[b] + job_base + js] = j;
            for (int t = 0; t < NT; ++t)
                if (m[(size_t)o * NT + t]) pairs[pair_off[b] + (s++)] = o * NT + t;
        }
        pair_base += tot_p;
        job_base += tot_j;
    }
}

// --------------------------------------------------------------------------------------------------
// Fine pass: one block per (job, chunk of TC coarse translations); each thread owns pixels, projects the
// 8 oversampled orientations once per pixel and accumulates the 8 x 4*TC squared differences.
// --------------------------------------------------------------------------------------------------

// Jobs are keyed for the fine pass's order by class and coarse orientation (direction-major), key >> key_shift so that
// the keys fit the LDS histogram of k_fine_items.
constexpr int kJobKeys = 4096;

// Work items of the fine pass: jobs with at most small_max significant translations are one item of
// the small path of k_fine_diff2; the others one wide item per wide_tc translations. With `sorted`, the jobs are first
// put in the order of their class and coarse orientation (a counting sort over kJobKeys keys; the order within a key
// is arbitrary), so that consecutive items are jobs of nearby orientations, often of different particles: jobs of one
// coarse direction sample the same four fine-direction planes of the reference, and items processed together then
// share its corner blocks in the caches. job_order receives the order (the weighted sums follow it too); without
// `sorted` it is the identity (particle order). One block of 1024.
__global__ void __launch_bounds__(kWBlock) k_fine_items(const FineJob *__restrict__ jobs, const int *__restrict__ or_off,
                                                        const int *__restrict__ or_frame, const int *__restrict__ or_class,
                                                        int n_or_grid, int key_shift, int sorted, int small_max,
                                                        int wide_tc, int cap_small, int cap_wide,
                                                        int *__restrict__ job_order, int2 *__restrict__ small_items,
                                                        int2 *__restrict__ wide_items, FinePlan *__restrict__ plan) {
    __shared__ int tmp[24];
    __shared__ int hist[kJobKeys];
    if (plan->overflow) return;
    const int n_jobs = plan->n_jobs;
    if (sorted) {
        for (int q = threadIdx.x; q < kJobKeys; q += kWBlock) hist[q] = 0;
        __syncthreads();
        auto key_of = [&](int j) {
            const FineJob jb = jobs[j];
            const int oe = or_off[jb.b] + jb.o;
            const long long g = (long long)or_class[oe] * n_or_grid + (or_frame[oe] >> 3);
            return (int)min((long long)(kJobKeys - 1), g >> key_shift);
        };
        for (int j = threadIdx.x; j < n_jobs; j += kWBlock) atomicAdd(&hist[key_of(j)], 1);
        __syncthreads();
        // exclusive scan of the histogram, kJobKeys / kWBlock consecutive keys per thread
        constexpr int kPer = kJobKeys / kWBlock;
        int loc[kPer], sum = 0;
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            loc[q] = hist[threadIdx.x * kPer + q];
            sum += loc[q];
        }
        int tot;
        int run = block_exclusive_scan_i(sum, tmp, tot);
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            hist[threadIdx.x * kPer + q] = run;
            run += loc[q];
        }
        __syncthreads();
        for (int j = threadIdx.x; j < n_jobs; j += kWBlock) job_order[atomicAdd(&hist[key_of(j)], 1)] = j;
    } else {
        for (int j = threadIdx.x; j < n_jobs; j += kWBlock) job_order[j] = j;
    }
    __syncthreads();
    int sbase = 0, wbase = 0;
    for (int j0 = 0; j0 < n_jobs; j0 += kWBlock) {
        const int jj = j0 + threadIdx.x;
        int ns = 0, nw = 0, sc = 0, j = 0;
        if (jj < n_jobs) {
            j = job_order[jj];
            sc = jobs[j].s_count;
            if (sc <= small_max) ns = 1;
            else nw = (sc + wide_tc - 1) / wide_tc;
        }
        int ts, tw;
        const int ss = block_exclusive_scan_i(ns, tmp, ts);
        const int ws = block_exclusive_scan_i(nw, tmp, tw);
        if (ns && sbase + ss < cap_small) small_items[sbase + ss] = make_int2(j, 0);
        for (int q = 0; q < nw; ++q)
            if (wbase + ws + q < cap_wide) wide_items[wbase + ws + q] = make_int2(j, q * wide_tc);
        sbase += ts;
        wbase += tw;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        plan->n_small = sbase;
        plan->n_wide = wbase;
        if (sbase > cap_small || wbase > cap_wide) plan->overflow = 1;
    }
}

struct FineArgs {
    const FineJob *jobs;
    const int2 *items_w;           // (job, first job-local pair) per wide work item (plan->n_wide)
    const int2 *items_s;           // per small work item (plan->n_small)
    FinePlan *plan;                // header; next_fine is the dynamic item counter
    const ChunkPtrs *cp;           // this chunk's images
    const int *pairs;
    const long long *pair_off;
    const int *or_off;
    const int *or_frame;           // fine frame base (coarse orientation * 8) per orientation entry
    const int *or_class;
    const float4 *const *vols;
    VolGeom vg;
    const Frame *fine_frames;      // [n_dir * n_psi * 8]
    ImgGeom ig;
    const float *scale;
    const float *q;                // sqrt(Minvsigma2/2) of the fine size
    const float2 *ex;              // [n_xs][xs] phase factors of the fine x shifts
    const float2 *ey;              // [n_ys][s]
    const int *tr_ix;              // [NT*4] x-shift index of fine translation t*4+u
    const int *tr_iy;
    const int2 *pix;               // sampled pixels (Minvsigma2 > 0, inside the projection radius) as (p, x | row
                                   // << 16), in n_bands bands of equal count by radius, pixel order within a band
    int n_pix;
    int n_bands;                   // band b: pix[b n_pix / n_bands .. (b + 1) n_pix / n_bands)
    int item_major;                // unit order: 0 band-major (all items of band 0, then band 1, ...), 1 item-major
    int xcd_queues;                // 1: each XCD takes its own eighth of the items (band-major), then helps the others
    int NT;
    int do_ctf;
    double *diff2;                 // [pair_off[b]*32 + s*32 + r*4 + u]: k_fine_init's base, plus each band's sum
};

// Pose-independent part of a particle's fine diff2, in FP64: highres_Xi2 / 2 + sum over the pixels with
// Minvsigma2 > 0 of |q X|^2 (q = sqrt(Minvsigma2 / 2)). The fine kernels accumulate only the pose-dependent
// remainder |A|^2 - 2 Re(A conj(Y)) of |A - Y|^2 (A the scaled projection, Y = q X shifted; |Y| = |q X|), whose
// magnitude is that of the signal, not of the noise power, so that FP32 partial sums keep the differences between
// poses that the weights exponentiate (the reference CPU path sums diff2 in double, src/ml_optimiser.cpp:7436-7442).
// grid (capacity of the chunk), block kBlock; blocks beyond the chunk's particle count return.
__global__ void __launch_bounds__(kBlock) k_fine_base(const ChunkPtrs *__restrict__ cp, const float *__restrict__ q,
                                                      int P, const double *__restrict__ xi2half,
                                                      double *__restrict__ fbase) {
    __shared__ double red[kBlock / 64];
    const int b = blockIdx.x;
    if (b >= cp->B) return;
    const float2 *X = cp->fimg + (size_t)b * P;
    double acc = 0.;
    for (int p = threadIdx.x; p < P; p += kBlock) {
        const float qp = q[p];
        if (qp == 0.f) continue;
        const float2 x = X[p];
        const double xr = (double)qp * x.x, xi = (double)qp * x.y;
        acc += xr * xr + xi * xi;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = xi2half[b];
        for (int w = 0; w < kBlock / 64; ++w) t += red[w];
        fbase[b] = t;
    }
}

// Fine diff2 of the chunk's evaluated poses set to their particle's pose-independent part; the bands of the
// fine pass add their pose-dependent sums to it. grid (capacity of the chunk), block kBlock.
__global__ void __launch_bounds__(kBlock) k_fine_init(const int *__restrict__ n_pairs_b,
                                                      const long long *__restrict__ pair_off,
                                                      const double *__restrict__ fbase, const FinePlan *__restrict__ plan,
                                                      double *__restrict__ diff2) {
    const int b = blockIdx.x;
    if (plan->overflow || b >= plan->B) return;
    const long long n = 32LL * n_pairs_b[b];
    double *d = diff2 + 32LL * pair_off[b];
    const double v = fbase[b];
    for (long long e = threadIdx.x; e < n; e += kBlock) d[e] = v;
}

// Pixel range [i0, i1) of band `band` in FineArgs::pix.
__device__ __forceinline__ void band_range(const FineArgs &a, int band, int &i0, int &i1) {
    i0 = band * a.n_pix / a.n_bands;
    i1 = (band + 1) * a.n_pix / a.n_bands;
}

// Work-item shapes of the fine pass (chosen per engine by the tuner): a job with at most TC significant coarse
// translations is one small item; a larger job is split into wide items of 4 * CW coarse translations, CW per wave.
struct FineShape {
    int tc, cw;
};
constexpr FineShape kFineShapes[] = {{2, 2}, {1, 2}, {2, 1}, {1, 1}};
constexpr int kNumFineShapes = (int)(sizeof(kFineShapes) / sizeof(kFineShapes[0]));

// One work item on one band of pixels: a job with at most TC significant coarse translations. Each thread owns
// pixels, projects the 8 oversampled orientations once per pixel and accumulates the 8 x 4*TC squared differences.
template <int TC>
struct FineSmallLds {
    float red[kBlock / 64][8 * 4 * TC];
    int tix[4 * TC], tiy[4 * TC];
    int t[TC];
};

template <int TC>
__device__ __forceinline__ void fine_small_item(const FineArgs &a, int job_id, int first, int band,
                                                FineSmallLds<TC> &L) {
    constexpr int kFineSmallTC = TC;
    constexpr int kFineNF = 4 * kFineSmallTC;  // fine translations per block
    const FineJob job = a.jobs[job_id];
    const int b = job.b;
    const int nc = min(kFineSmallTC, job.s_count - first);
    const int tid = threadIdx.x;
    if (tid < kFineSmallTC) {
        const int s = job.s_begin + first + tid;
        L.t[tid] = (tid < nc) ? a.pairs[a.pair_off[b] + s] % a.NT : 0;
    }
    __syncthreads();
    if (tid < kFineNF) {
        const int c = tid >> 2, u = tid & 3;
        const int t = L.t[c];
        L.tix[tid] = a.tr_ix[t * 4 + u];
        L.tiy[tid] = a.tr_iy[t * 4 + u];
    }
    __syncthreads();
    const int oe = a.or_off[b] + job.o;
    const int k = a.or_class[oe];
    const Frame *fr = a.fine_frames + a.or_frame[oe];
    const float4 *vol = a.vols[k];
    const float2 *X = a.cp->fimg + (size_t)b * a.ig.P;
    const float *C = a.cp->fctf + (size_t)b * a.ig.P;
    const float sc = a.scale[b];
    float acc[8][kFineNF];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int j = 0; j < kFineNF; ++j) acc[r][j] = 0.f;
    int i0, i1;
    band_range(a, band, i0, i1);
    // pixels with Minvsigma2 == 0 contribute nothing, those outside the projection radius only |Y|^2 (in fbase)
    for (int ip = i0 + tid; ip < i1; ip += kBlock) {
        const int2 pe = a.pix[ip];
        const int p = pe.x, x = pe.y & 0xffff, i = pe.y >> 16;
        const int y = (i < a.ig.xs) ? i : i - a.ig.s;
        const float qp = a.q[p];
        const float c = qp * (a.do_ctf ? C[p] : 1.f) * sc;
        float2 A[8];
        // every listed pixel is inside the projection radius; testing it anyway keeps the compiler from hoisting
        // all 32 corner-block loads of the 8 orientations at once (which spills registers at 3 waves per SIMD)
        const bool inside = x * x + y * y <= a.ig.proj_r2_max;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            float2 v = make_float2(0.f, 0.f);
            if (inside) v = project_pixel(vol, a.vg, fr[r], x, y);
            A[r] = make_float2(c * v.x, c * v.y);
        }
        const float2 Xp = X[p];
        const float2 Xq2 = make_float2(2.f * qp * Xp.x, 2.f * qp * Xp.y);
#pragma unroll
        for (int j = 0; j < kFineNF; ++j) {
            const float2 ph = cmul(a.ex[(size_t)L.tix[j] * a.ig.xs + x], a.ey[(size_t)L.tiy[j] * a.ig.s + i]);
            const float2 Y2 = cmul(Xq2, ph);  // 2 Y
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                // |A|^2 - 2 Re(A conj(Y)) = A.x (A.x - 2 Y.x) + A.y (A.y - 2 Y.y)
                acc[r][j] = fmaf(A[r].x, A[r].x - Y2.x, acc[r][j]);
                acc[r][j] = fmaf(A[r].y, A[r].y - Y2.y, acc[r][j]);
            }
        }
    }
    // block reduction: wave shuffles, then across the 4 waves
    const int lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int j = 0; j < kFineNF; ++j) {
            float v = acc[r][j];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            if (lane == 0) L.red[w][r * kFineNF + j] = v;
        }
    __syncthreads();
    if (tid < 8 * kFineNF) {
        const int r = tid / kFineNF, j = tid - r * kFineNF;
        const int c = j >> 2, u = j & 3;
        if (c < nc) {
            float v = 0.f;
            for (int ww = 0; ww < kBlock / 64; ++ww) v += L.red[ww][tid];
            const long long s = job.s_begin + first + c;
            atomicAdd(&a.diff2[(a.pair_off[b] + s) * 32 + r * 4 + u], (double)v);
        }
    }
}

// One work item of a job with more than TC significant translations: up to 4 * kWideCW coarse
// translations. Per tile of 64 pixels the four waves first project the 8 fine orientations together (two per
// thread) into LDS; then wave w accumulates the squared differences of its kWideCW coarse translations (4 * kWideCW
// fine translations) over the tile's pixels. A job with up to kWideTC significant translations reads each
// trilinear sample once.
template <int CW>
struct FineWideLds {
    float2 A[8][64];
    float2 X[64];
    int p[64];
    int tix[4][4 * CW], tiy[4][4 * CW];  // per wave
};

template <int CW>
__device__ __forceinline__ void fine_wide_item(const FineArgs &a, int job_id, int first, int band,
                                               FineWideLds<CW> &L) {
    constexpr int kWideCW = CW;            // coarse translations per wave
    constexpr int kWideTC = 4 * kWideCW;   // coarse translations per block
    constexpr int kNF = 4 * kWideCW;  // fine translations per wave
    const FineJob job = a.jobs[job_id];
    const int b = job.b;
    const int nc = min(kWideTC, job.s_count - first);  // coarse translations of this block
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < 4 * kNF) {
        const int w = tid / kNF, j = tid % kNF;  // wave w, fine translation j = 4c + u
        const int c = kWideCW * w + (j >> 2), u = j & 3;
        int t = 0;
        if (c < nc) t = a.pairs[a.pair_off[b] + job.s_begin + first + c] % a.NT;
        L.tix[w][j] = a.tr_ix[t * 4 + u];
        L.tiy[w][j] = a.tr_iy[t * 4 + u];
    }
    const int oe = a.or_off[b] + job.o;
    const int k = a.or_class[oe];
    const Frame *fr = a.fine_frames + a.or_frame[oe];
    const float4 *vol = a.vols[k];
    const float2 *X = a.cp->fimg + (size_t)b * a.ig.P;
    const float *C = a.cp->fctf + (size_t)b * a.ig.P;
    const float sc = a.scale[b];
    const bool active = kWideCW * wave < nc;  // this wave has translations
    float acc[8][kNF];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int j = 0; j < kNF; ++j) acc[r][j] = 0.f;
    int i0, i1;
    band_range(a, band, i0, i1);
    for (int t0 = i0; t0 < i1; t0 += 64) {
        __syncthreads();  // previous tile consumed (and the translation tables written, first time)
        // phase A: 8 projections of 64 sampled pixels, 2 per thread
        {
            const int ip = t0 + lane;
            float2 v0 = make_float2(0.f, 0.f), v1 = make_float2(0.f, 0.f);
            float c = 0.f;
            if (ip < i1) {
                const int2 pe = a.pix[ip];
                const int p = pe.x, x = pe.y & 0xffff, i = pe.y >> 16;
                const int y = (i < a.ig.xs) ? i : i - a.ig.s;
                const float qp = a.q[p];
                c = qp * (a.do_ctf ? C[p] : 1.f) * sc;
                v0 = project_pixel(vol, a.vg, fr[2 * wave], x, y);
                v1 = project_pixel(vol, a.vg, fr[2 * wave + 1], x, y);
                if (wave == 0) {
                    const float2 Xp = X[p];
                    L.X[lane] = make_float2(2.f * qp * Xp.x, 2.f * qp * Xp.y);  // 2 q X
                    L.p[lane] = p;
                }
            } else if (wave == 0) {
                L.p[lane] = -1;
            }
            L.A[2 * wave][lane] = make_float2(c * v0.x, c * v0.y);
            L.A[2 * wave + 1][lane] = make_float2(c * v1.x, c * v1.y);
        }
        __syncthreads();
        // phase B: this wave's fine translations over the tile
        if (active) {
            const int p = L.p[lane];
            if (p >= 0) {
                const int i = p / a.ig.xs, x = p - i * a.ig.xs;
                float2 A[8];
#pragma unroll
                for (int r = 0; r < 8; ++r) A[r] = L.A[r][lane];
                const float2 Xq2 = L.X[lane];
#pragma unroll
                for (int j = 0; j < kNF; ++j) {
                    const float2 ph = cmul(a.ex[(size_t)L.tix[wave][j] * a.ig.xs + x],
                                           a.ey[(size_t)L.tiy[wave][j] * a.ig.s + i]);
                    const float2 Y2 = cmul(Xq2, ph);  // 2 Y
#pragma unroll
                    for (int r = 0; r < 8; ++r) {
                        // pose-dependent part of |A - Y|^2 (|Y|^2 is in fbase)
                        acc[r][j] = fmaf(A[r].x, A[r].x - Y2.x, acc[r][j]);
                        acc[r][j] = fmaf(A[r].y, A[r].y - Y2.y, acc[r][j]);
                    }
                }
            }
        }
    }
    if (active) {
        // wave reduction; lanes 0 .. kWideCW-1 write the wave's translations
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
            for (int j = 0; j < kNF; ++j) {
                float v = acc[r][j];
                for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
                acc[r][j] = v;
            }
        if (lane < kWideCW) {
            const int c = kWideCW * wave + lane;  // one coarse translation per lane
            if (c < nc) {
                const long long s = job.s_begin + first + c;
                double *out = a.diff2 + (a.pair_off[b] + s) * 32;
#pragma unroll
                for (int r = 0; r < 8; ++r)
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        // acc index j = 4 * lane + u, selected without dynamic register indexing
                        float v = acc[r][u];
#pragma unroll
                        for (int q = 1; q < kWideCW; ++q)
                            if (lane == q) v = acc[r][4 * q + u];
                        atomicAdd(&out[r * 4 + u], (double)v);
                    }
            }
        }
    }
}

// Fine pass, persistent: blocks fetch work units from plan->next_fine until none is left, so uneven items spread
// over the GPU; every block exits once the counter passes the last unit. A unit is one work item on one band of
// pixels, either band-major (all items of band 0, then band 1, ...; within a band the wide items first, the heaviest,
// then the small ones), so that the units in flight sample one or two radial shells of the reference, or item-major
// (the bands of an item one after the other), so that with items in orientation order (k_fine_items) the units in
// flight sample the planes of few coarse directions. Each unit adds its partial
// sums to the FP64 diff2 (k_fine_init wrote the pose-independent part). Three waves per SIMD, as the two item
// paths have when compiled alone (at most 168 VGPRs). kTuning only separates the tuning launches from the production
// ones in kernel traces.
template <int kTuning, int TC, int CW>
__global__ void __launch_bounds__(kBlock, 3) k_fine_diff2(FineArgs a) {
    __shared__ union {
        FineSmallLds<TC> s;
        FineWideLds<CW> w;
    } lds;
    __shared__ int s_item;
    if (a.plan->overflow) return;
    const int n_wide = a.plan->n_wide;
    const int n_items = n_wide + a.plan->n_small;
    const int n_units = n_items * a.n_bands;
    if (a.xcd_queues) {
        // Workgroups are dispatched to the 8 XCDs in turn, so block b runs on XCD b % 8. XCD x owns the x-th eighth of
        // the wide and of the small items (contiguous in orientation order: a subset of the directions for its L2)
        // and takes their units band-major; a block whose XCD's share is exhausted takes units of the next XCDs'
        // shares, so the pass ends when every share is. Each unit index comes from one atomic increment of its
        // share's counter, so every unit runs exactly once whichever XCD the blocks actually run on.
        __shared__ int s_queue;
        const int n_small = a.plan->n_small;
        const int home = blockIdx.x & 7;
        for (;;) {
            __syncthreads();
            if (threadIdx.x == 0) {
                int got = -1, q = 0;
                for (int k = 0; k < 8 && got < 0; ++k) {
                    const int x = (home + k) & 7;
                    const int nx = ((x + 1) * n_wide / 8 - x * n_wide / 8) + ((x + 1) * n_small / 8 - x * n_small / 8);
                    const int lim = nx * a.n_bands;
                    if (*(volatile int *)&a.plan->next_fine_x[x] >= lim) continue;  // exhausted (a stale read only
                                                                                    // costs the increment below)
                    const int u = atomicAdd(&a.plan->next_fine_x[x], 1);
                    if (u < lim) {
                        got = u;
                        q = x;
                    }
                }
                s_item = got;
                s_queue = q;
            }
            __syncthreads();
            const int unit = __builtin_amdgcn_readfirstlane(s_item);
            if (unit < 0) break;
            const int x = __builtin_amdgcn_readfirstlane(s_queue);
            const int wlo = x * n_wide / 8, nxw = (x + 1) * n_wide / 8 - wlo;
            const int slo = x * n_small / 8, nx = nxw + (x + 1) * n_small / 8 - slo;
            const int band = unit / nx, idx = unit - band * nx;
            if (idx < nxw) {
                const int2 it = a.items_w[wlo + idx];
                fine_wide_item<CW>(a, it.x, it.y, band, lds.w);
            } else {
                const int2 it = a.items_s[slo + idx - nxw];
                fine_small_item<TC>(a, it.x, it.y, band, lds.s);
            }
        }
        return;
    }
    for (;;) {
        __syncthreads();  // previous unit's LDS and s_item consumed
        if (threadIdx.x == 0) s_item = atomicAdd(&a.plan->next_fine, 1);
        __syncthreads();
        // one value per block: readfirstlane makes the branches below provably uniform, so the two item paths
        // are not linearised and do not hold their registers at the same time
        const int unit = __builtin_amdgcn_readfirstlane(s_item);
        if (unit >= n_units) break;
        int band, item;
        if (a.item_major) {
            item = unit / a.n_bands;
            band = unit - item * a.n_bands;
        } else {
            band = unit / n_items;
            item = unit - band * n_items;
        }
        if (item < n_wide) {
            const int2 it = a.items_w[item];
            fine_wide_item<CW>(a, it.x, it.y, band, lds.w);
        } else {
            const int2 it = a.items_s[item - n_wide];
            fine_small_item<TC>(a, it.x, it.y, band, lds.s);
        }
    }
}

// Resets the dynamic item counters (tuning launches of the fine pass reuse one plan).
__global__ void k_reset_items(FinePlan *plan) {
    if (threadIdx.x == 0) {
        plan->next_fine = 0;
        plan->next_wsum = 0;
        for (int x = 0; x < 8; ++x) plan->next_fine_x[x] = plan->next_wsum_x[x] = 0;
    }
}

// Zero fill used inside captured graphs instead of memset nodes.
// Sum over the B particles of each shell's weighted sigma2_noise contribution: block s sums column s of the
// [B][nshell] per-particle accumulators in a fixed order (strided partial sums, then a tree). grid (nshell).
__global__ void __launch_bounds__(kBlock) k_shell_total(const double *__restrict__ acc, int B, int nshell,
                                                        double *__restrict__ out) {
    __shared__ double red[kBlock];
    const int s = blockIdx.x, tid = threadIdx.x;
    double v = 0.;
    for (int b = tid; b < B; b += kBlock) v += acc[(size_t)b * nshell + s];
    red[tid] = v;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    if (tid == 0) out[s] = red[0];
}

__global__ void __launch_bounds__(kBlock) k_zero_d(double *__restrict__ p, long long n) {
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock) p[i] = 0.;
}

// --------------------------------------------------------------------------------------------------
// Weighted sums and backprojection (storeWeightedSums, src/ml_optimiser.cpp:8241-9156, with
// BackProjector::backproject2Dto3D, src/backprojector.cpp:55-357). Persistent blocks over (job, pixel band) units.
// --------------------------------------------------------------------------------------------------
constexpr int kMaxSigT = 1024;  // fine translations of one orientation (4 per coarse translation); host checks 4*NT

struct WsumArgs {
    const FineJob *jobs;
    const int *pairs;
    const long long *pair_off;
    const int *or_off;
    const int *or_frame;
    const int *or_class;
    const int *or_dir;              // global direction index per orientation entry
    const float4 *const *vols;
    VolGeom vg;
    const Frame *fine_frames;
    ImgGeom ig;
    const ChunkPtrs *cp;            // this chunk's images
    FinePlan *plan;                 // items: plan->n_jobs * 8, fetched through plan->next_wsum
    const int *job_order;           // jobs in the fine pass's order (k_fine_items)
    const float *scale_w;           // max(scale, 0.001) (src/ml_optimiser.cpp:8402-8416)
    const float *minv_w;            // Minvsigma2 with the origin restored
    const int *mresol;              // resolution shell per pixel, -1 excluded
    const int2 *wpix;               // the pixels that contribute (a shell >= 0 or backprojected) as (p, x | row << 16),
                                    // n_bands bands of equal count by radius, pixel order within a band
    int n_wpix;
    int n_bands;
    int unit_order;                 // 0 band-major, 1 item-major (the bands of a job in a row), 2 band-major in per-XCD
                                    // queues
    const unsigned char *xa_shell;  // [K][nshell] data_vs_prior > 3 (scale correction) or null
    const float2 *ex;
    const float2 *ey;
    const int *tr_ix;
    const int *tr_iy;
    const double *over_x;           // [NT*4] fine shifts in pixels
    const double *over_y;
    const double *weight;           // fine weights [pair_off*32 + ...]
    const PassStats *stats;         // fine-pass stats per particle
    const double *old_off;          // [b][2]
    const double *prior_off;        // [b][2]
    double pixel_size;
    int NT;
    int nshell;
    int n_dir;
    int do_ctf;
    int do_scale;
    float *const *bp_slots;         // [K] corner-slot accumulators of the backprojectors
    double *shell_acc;              // [b][nshell] wsum sigma2_noise per particle
    double *part_acc;               // [b][8]: sumw, sigma2_offset, norm_corr, XA, AA, blocks, translations
    double *pdf_class;              // [b][K]
    double *pdf_direction;          // [K][n_dir] (all particles)
    int K;
};

__global__ void __launch_bounds__(kBlock, 5) k_wsum_bp(WsumArgs a) {
    __shared__ int s_ix[kMaxSigT], s_iy[kMaxSigT];
    __shared__ float s_w[kMaxSigT];
    __shared__ int s_tmp[24];
    __shared__ int s_cnt[8];
    __shared__ float shell[512];
    __shared__ double s_red[16];
    __shared__ long long sx_base[kBlock];
    __shared__ float sx_val[kBlock][3];
    __shared__ float sx_w[kBlock][8];
    __shared__ int s_item;
    if (a.plan->overflow) return;
    const int n_items = a.plan->n_jobs;
    const int n_units = n_items * a.n_bands;
    const int tid = threadIdx.x;
    // persistent: units fetched dynamically. A unit is one job (a coarse orientation of a particle with its
    // significant coarse translations) on one radial band of pixels; it visits the job's 8 oversampled orientations
    // and skips those without a significant fine pose (most of them). Band-major or item-major in one queue, or
    // band-major in per-XCD queues (XCD b % 8 first takes its eighth of the jobs, contiguous in the fine pass's
    // orientation order, then helps the others), in the fine pass's job order.
    __shared__ int s_queue;
    for (;;) {
    __syncthreads();  // previous unit's LDS and s_item consumed
    if (tid == 0) {
        if (a.unit_order == 2) {
            const int home = blockIdx.x & 7;
            int got = -1, q = 0;
            for (int k = 0; k < 8 && got < 0; ++k) {
                const int x = (home + k) & 7;
                const int lim = ((x + 1) * n_items / 8 - x * n_items / 8) * a.n_bands;
                if (*(volatile int *)&a.plan->next_wsum_x[x] >= lim) continue;
                const int u = atomicAdd(&a.plan->next_wsum_x[x], 1);
                if (u < lim) {
                    got = u;
                    q = x;
                }
            }
            s_item = got;
            s_queue = q;
        } else {
            s_item = atomicAdd(&a.plan->next_wsum, 1);
        }
    }
    __syncthreads();
    const int unit = s_item;
    if (unit >= n_units || unit < 0) break;
    int band, item;
    if (a.unit_order == 2) {
        const int x = s_queue;
        const int lo = x * n_items / 8, nx = (x + 1) * n_items / 8 - lo;
        band = unit / nx;
        item = lo + unit - band * nx;
    } else if (a.unit_order == 1) {
        item = unit / a.n_bands;
        band = unit - item * a.n_bands;
    } else {
        band = unit / n_items;
        item = unit - band * n_items;
    }
    const int job_id = a.job_order[item];
    const FineJob job = a.jobs[job_id];
    const int b = job.b;
    // a particle whose fine weights failed (zero or NaN sum) contributes nothing: its weights are not
    // normalisable and would turn the shared accumulators into NaN (the run returns the particle's status)
    if (a.stats[b].status != 0) continue;
    const double thr = a.stats[b].threshold;
    const double sumw = a.stats[b].sum_weight;
    for (int i = tid; i < a.nshell; i += kBlock) shell[i] = 0.f;
    const int oe = a.or_off[b] + job.o;
    const int k = a.or_class[oe];
    const float4 *vol = a.vols[k];
    const float2 *X = a.cp->fimg + (size_t)b * a.ig.P;
    const float2 *Xn = a.cp->fimg_nomask + (size_t)b * a.ig.P;
    const float *C = a.cp->fctf + (size_t)b * a.ig.P;
    const float scw = a.do_scale ? a.scale_w[b] : 1.f;
    float *slots = a.bp_slots[k];
    const unsigned char *xash = a.xa_shell ? a.xa_shell + (size_t)k * a.nshell : nullptr;
    double xa = 0., aa = 0., nc = 0.;
    const int sh = a.ig.xs;
    const long long base = (a.pair_off[b] + job.s_begin) * 32;
    const int n4 = 4 * job.s_count;  // fine translations of one oversampled orientation
    const int w0 = band * a.n_wpix / a.n_bands, w1 = (band + 1) * a.n_wpix / a.n_bands;
    const int n_chunks = (w1 - w0 + kBlock - 1) / kBlock;
    // significant fine poses per oversampled orientation, in one coalesced pass over the job's 32 weights per
    // significant coarse translation (entry sidx * 32 + r * 4 + u)
    if (tid < 8) s_cnt[tid] = 0;
    __syncthreads();
    for (int e = tid; e < 32 * job.s_count; e += kBlock)
        if (a.weight[base + e] >= thr) atomicAdd(&s_cnt[(e >> 2) & 7], 1);
    __syncthreads();
    for (int r = 0; r < 8; ++r) {
        if (s_cnt[r] == 0) continue;  // uniform
        // significant fine translations of orientation r, compacted in (coarse translation, sub-shift) order; their
        // normalised weights' sums (pdf_class, pdf_direction, sigma2_offset) once per job and orientation, with band 0
        int cnt = 0;
        double sw = 0., so = 0.;
        const double px2 = a.pixel_size * a.pixel_size;
        for (int e0 = 0; e0 < n4; e0 += kBlock) {
            const int e = e0 + tid;
            bool sig = false;
            double wn = 0.;
            int t = 0, u = 0;
            if (e < n4) {
                const int sidx = e >> 2;
                u = e & 3;
                const double wt = a.weight[base + (long long)sidx * 32 + r * 4 + u];
                sig = wt >= thr;
                if (sig) {
                    t = a.pairs[a.pair_off[b] + job.s_begin + sidx] % a.NT;
                    wn = wt / sumw;
                }
            }
            int tot;
            const int pos = cnt + block_exclusive_scan_i(sig ? 1 : 0, s_tmp, tot);
            if (sig) {
                if (pos < kMaxSigT) {
                    s_ix[pos] = a.tr_ix[t * 4 + u];
                    s_iy[pos] = a.tr_iy[t * 4 + u];
                    s_w[pos] = (float)wn;
                }
                if (band == 0) {
                    sw += wn;
                    const double dx = a.prior_off[2 * b] - a.old_off[2 * b] - a.over_x[t * 4 + u];
                    const double dy = a.prior_off[2 * b + 1] - a.old_off[2 * b + 1] - a.over_y[t * 4 + u];
                    so += wn * px2 * (dx * dx + dy * dy);
                }
            }
            cnt += tot;
        }
        if (cnt == 0) continue;  // uniform: cnt comes from the block scan
        if (band == 0) {
            sw = block_sum_d(sw, s_red);
            so = block_sum_d(so, s_red);
            if (tid == 0) {
                atomicAdd(&a.part_acc[8 * b + 0], sw);
                atomicAdd(&a.part_acc[8 * b + 1], so);
                atomicAdd(&a.part_acc[8 * b + 5], 1.0);
                atomicAdd(&a.part_acc[8 * b + 6], (double)cnt);
                atomicAdd(&a.pdf_class[(size_t)b * a.K + k], sw);
                atomicAdd(&a.pdf_direction[(size_t)k * a.n_dir + a.or_dir[oe]], sw);
            }
        }
        __syncthreads();  // compacted lists visible
        const int nsig = min(cnt, kMaxSigT);
        const Frame fr = a.fine_frames[a.or_frame[oe] + r];
        // the band's pixels in chunks of kBlock; each block starts at a different chunk so that concurrent blocks do
        // not all scatter into the same voxels at the same time
        const int c0 = n_chunks > 0 ? (int)(((unsigned)(8 * item + r) * 2654435761u) % (unsigned)n_chunks) : 0;
        for (int cc = 0; cc < n_chunks; ++cc) {
            int chunk = c0 + cc;
            if (chunk >= n_chunks) chunk -= n_chunks;
            const int ip = w0 + chunk * kBlock + tid;
            long long sc_base = -1;  // trilinear base voxel of this pixel's scatter (-1: none)
            float sc_val[3] = {0.f, 0.f, 0.f};
            float sc_w[8];
            if (ip < w1) {
                const int2 pe = a.wpix[ip];
                const int p = pe.x, x = pe.y & 0xffff, i = pe.y >> 16;
                const int y = (i < a.ig.xs) ? i : i - a.ig.s;
                const float ctf = (a.do_ctf ? C[p] : 1.f) * scw;
                float2 ref = make_float2(0.f, 0.f);
                if (x * x + y * y <= a.ig.proj_r2_max) {
                    const float2 v = project_pixel(vol, a.vg, fr, x, y);
                    ref = make_float2(ctf * v.x, ctf * v.y);
                }
                const int ires = a.mresol[p];
                const float minv = a.minv_w[p];
                const bool do_xa = a.do_scale && xash && ires > -1 && xash[ires];
                float wd2 = 0.f, fre = 0.f, fim = 0.f, fw = 0.f, pxa = 0.f, paa = 0.f;
                const float2 Xp = X[p], Xnp = Xn[p];
                for (int j = 0; j < nsig; ++j) {
                    const float wn = s_w[j];
                    const float2 ph = cmul(a.ex[(size_t)s_ix[j] * a.ig.xs + x], a.ey[(size_t)s_iy[j] * a.ig.s + i]);
                    if (ires > -1) {
                        const float2 Y = cmul(Xp, ph);
                        const float dr = ref.x - Y.x, di = ref.y - Y.y;
                        wd2 = fmaf(wn, dr * dr + di * di, wd2);
                        if (do_xa) {
                            pxa = fmaf(wn, ref.x * Y.x + ref.y * Y.y, pxa);
                            paa = fmaf(wn, ref.x * ref.x + ref.y * ref.y, paa);
                        }
                    }
                    const float2 Z = cmul(Xnp, ph);
                    const float wx = wn * ctf * minv;
                    fre = fmaf(Z.x, wx, fre);
                    fim = fmaf(Z.y, wx, fim);
                    fw = fmaf(wx, ctf, fw);
                }
                if (ires > -1) {
                    atomicAdd(&shell[ires], wd2);
                    nc += wd2;
                    xa += pxa;
                    aa += paa;
                }
                // backprojection of (fre, fim) with weight fw (backproject2Dto3D, src/backprojector.cpp:124-296)
                if (fw > 0.f && x * x + y * y <= a.ig.bp_r2_max && !(i >= sh && x == 0)) {
                    const float fxp = (float)x, fyp = (float)y;
                    float xp = fr.e0x * fxp + fr.e1x * fyp;
                    float yp = fr.e0y * fxp + fr.e1y * fyp;
                    float zp = fr.e0z * fxp + fr.e1z * fyp;
                    float vim = fim;
                    if (xp < 0.f) {
                        xp = -xp;
                        yp = -yp;
                        zp = -zp;
                        vim = -vim;
                    }
                    const float fx0 = floorf(xp), fy0 = floorf(yp), fz0 = floorf(zp);
                    const float fx = xp - fx0, fy = yp - fy0, fz = zp - fz0;
                    const int x0 = (int)fx0, y0 = (int)fy0 - a.vg.starty, z0 = (int)fz0 - a.vg.startz;
                    if (!(x0 < 0 || x0 + 1 >= a.vg.xdim || y0 < 0 || y0 + 1 >= a.vg.ydim || z0 < 0 ||
                          z0 + 1 >= a.vg.zdim)) {
                        const float mfx = 1.f - fx, mfy = 1.f - fy, mfz = 1.f - fz;
                        sc_base = (long long)packed_index(a.vg, x0, y0, z0);
                        sc_val[0] = fre;
                        sc_val[1] = vim;
                        sc_val[2] = fw;
                        sc_w[0] = mfz * mfy * mfx;
                        sc_w[1] = mfz * mfy * fx;
                        sc_w[2] = mfz * fy * mfx;
                        sc_w[3] = mfz * fy * fx;
                        sc_w[4] = fz * mfy * mfx;
                        sc_w[5] = fz * mfy * fx;
                        sc_w[6] = fz * fy * mfx;
                        sc_w[7] = fz * fy * fx;
                    }
                }
            }
            // Scatter of the block's pixels into the corner slots. The 24 values of one pixel share one 128-byte
            // slot line; lanes are regrouped through LDS so that each atomic instruction covers two pixels'
            // lines (32 lanes each) instead of 64 lines, which is what bounds atomic throughput.
            __syncthreads();
            sx_base[tid] = sc_base;
#pragma unroll
            for (int c = 0; c < 3; ++c) sx_val[tid][c] = sc_val[c];
#pragma unroll
            for (int c = 0; c < 8; ++c) sx_w[tid][c] = (sc_base >= 0) ? sc_w[c] : 0.f;
            __syncthreads();
            {
                const int lane = tid & 63, wave0 = tid & ~63;
                const int f = lane & 31;
                for (int k2 = 0; k2 < 32; ++k2) {
                    const int src = wave0 + 2 * k2 + (lane >> 5);
                    const long long sb = sx_base[src];
                    if (sb >= 0 && f < 24) atomicAdd(slots + kSlot * sb + f, sx_w[src][f & 7] * sx_val[src][f >> 3]);
                }
            }
        }
        __syncthreads();  // this orientation's lists consumed before the next one's are written
    }  // r
    nc = block_sum_d(nc, s_red);
    xa = block_sum_d(xa, s_red);
    aa = block_sum_d(aa, s_red);
    __syncthreads();
    double *sa = a.shell_acc + (size_t)b * a.nshell;
    for (int s2 = tid; s2 < a.nshell; s2 += kBlock)
        if (shell[s2] != 0.f) atomicAdd(&sa[s2], (double)shell[s2]);
    if (tid == 0 && (nc != 0. || xa != 0. || aa != 0.)) {
        atomicAdd(&a.part_acc[8 * b + 2], nc);
        atomicAdd(&a.part_acc[8 * b + 3], xa);
        atomicAdd(&a.part_acc[8 * b + 4], aa);
    }
    }  // units
}

}  // namespace rlnk
