// Reconstruction on the GPU: BackProjector::reconstruct (src/backprojector.cpp:1379-1780) for a 3-D reference, 2-D
// data and trilinear interpolation, the default path (window to the original size, no blob downsampling). The
// backprojector's accumulated data and weight never leave the device:
//   k_rc_init        decenter data and weight into the half-complex layout of the pad_size^3 box, MAP
//                    regularisation of the weight per shell, division by normalise, Fnewweight = 1 in the sphere
//   without gridding iterations (skip_gridding, relion_refine's default): k_rc_radavg (shell sums of the weight),
//                    k_rc_skip (data / max(weight, shell average / 1000))
//   gridding iterations (max_iter_preweight): k_rc_product (Fconv = Fnewweight * Fweight), Z2D, k_rc_blob (times the
//                    tabulated blob transform in real space), D2Z, k_rc_update (Fnewweight /= max(1e-6, |Fconv|))
//   k_rc_window      Fconv = data * Fnewweight, window to the padded box, CenterFFTbySign; Z2D
//   k_rc_bg, k_rc_out  real-space window to the original box, 1 / (pad^3 ori), soft mask with the average background,
//                    griddingCorrect
// Everything in FP64 (hipFFT D2Z / Z2D), as the reference's default CPU build (DoublePrec_CPU, CMakeLists.txt:135)
// runs it: the MAP term 1 / (0.001 weight) of shells with a zero tau2 spans many decades, beyond FP32 transforms.
// Every kernel streams its volumes once (HBM-bound).
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/rln_estep.h"
#include "host_common.h"
#include "map_workspace.h"

using namespace rln;

namespace rlnc {

constexpr int kThreads = 256;
constexpr int kBlobTable = 10000;
constexpr int kMaxShells = 1024;  // TabFtBlob entries over [0, 0.5) (src/tabfuncs.cpp:95-112)

struct Geo {
    int S, h;          // pad_size and its half
    long long max_r2;  // ROUND(pad r_max)^2
    double pad;
};

__device__ __forceinline__ int fftw_k(int k, int n, int ncols) { return k < ncols ? k : k - n; }

// Decenter (src/projector.h:249-259), MAP regularisation (src/backprojector.cpp:1440-1473), normalise, initial
// Fnewweight. One thread per element of the half-complex S^3 layout [S][S][h+1].
__global__ void __launch_bounds__(kThreads) k_rc_init(const float *__restrict__ re, const float *__restrict__ im,
                                                      const float *__restrict__ w, Geo g, int do_map,
                                                      const double *__restrict__ tau2, int ntau2, double inv_ovs_fudge,
                                                      int minres_map, double normalise, double *__restrict__ fweight,
                                                      double *__restrict__ fnew, double2 *__restrict__ fdata) {
    const int S = g.S, xs = g.h + 1;
    const size_t n = (size_t)S * S * xs;
    const size_t e = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (e >= n) return;
    const int j = (int)(e % xs);
    const size_t zy = e / xs;
    const int i = (int)(zy % S), k = (int)(zy / S);
    const int kp = fftw_k(k, S, xs), ip = fftw_k(i, S, xs), jp = j;
    const long long r2 = (long long)kp * kp + (long long)ip * ip + (long long)jp * jp;
    double fw = 0.;
    double2 fd = make_double2(0., 0.);
    if (r2 <= g.max_r2) {
        const size_t src = ((size_t)(kp + g.h) * S + (ip + g.h)) * xs + jp;
        fw = (double)w[src];
        fd = make_double2((double)re[src], (double)im[src]);
    }
    if (do_map && r2 < g.max_r2) {
        const int ires = (int)floor(sqrt((double)r2) / g.pad + 0.5);
        const double t = tau2[min(ires, ntau2 - 1)];
        double inv = 0.;
        if (t > 0.)
            inv = inv_ovs_fudge / t;
        else if (fw > 1e-20)
            inv = 1. / (0.001 * fw);
        if (ires >= minres_map) fw += inv;
    }
    fweight[e] = fw / normalise;
    fnew[e] = r2 < g.max_r2 ? 1. : 0.;
    fdata[e] = make_double2(fd.x / normalise, fd.y / normalise);
}

// Radial sums of the regularised weight for the path without gridding iterations (src/backprojector.cpp:1513-1532):
// shells FLOOR(r / pad) over r^2 < ROUND((r_max pad)^2), FP64 shared-memory bins; grid-stride over a bounded
// grid so that each global bin takes few atomics.
__global__ void __launch_bounds__(kThreads) k_rc_radavg(const double *__restrict__ fweight, Geo g, long long round_max_r2,
                                                        int r_max, double *__restrict__ acc) {
    __shared__ double s_sum[kMaxShells], s_cnt[kMaxShells];
    for (int t = threadIdx.x; t < r_max; t += kThreads) s_sum[t] = s_cnt[t] = 0.;
    __syncthreads();
    const int S = g.S, xs = g.h + 1;
    const size_t n = (size_t)S * S * xs;
    for (size_t e = (size_t)blockIdx.x * kThreads + threadIdx.x; e < n; e += (size_t)gridDim.x * kThreads) {
        const int j = (int)(e % xs);
        const size_t zy = e / xs;
        const int i = (int)(zy % S), k = (int)(zy / S);
        const int kp = fftw_k(k, S, xs), ip = fftw_k(i, S, xs);
        const long long r2 = (long long)kp * kp + (long long)ip * ip + (long long)j * j;
        if (r2 < round_max_r2) {
            const int ires = min((int)floor(sqrt((double)r2) / g.pad), r_max - 1);
            atomicAdd(&s_sum[ires], fweight[e]);
            atomicAdd(&s_cnt[ires], 1.);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < r_max; t += kThreads)
        if (s_cnt[t] != 0.) {
            atomicAdd(&acc[t], s_sum[t]);
            atomicAdd(&acc[r_max + t], s_cnt[t]);
        }
}

// data / max(weight, radial average / 1000), no division where that is 0 (src/backprojector.cpp:1535-1571); the
// gridding weight becomes 1 everywhere so that the window keeps every decentred element.
__global__ void __launch_bounds__(kThreads) k_rc_skip(const double *__restrict__ fweight, Geo g, int r_max,
                                                      const double *__restrict__ acc, double2 *__restrict__ fdata,
                                                      double *__restrict__ fnew) {
    const int S = g.S, xs = g.h + 1;
    const size_t n = (size_t)S * S * xs;
    const size_t e = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (e >= n) return;
    const int j = (int)(e % xs);
    const size_t zy = e / xs;
    const int i = (int)(zy % S), k = (int)(zy / S);
    const int kp = fftw_k(k, S, xs), ip = fftw_k(i, S, xs);
    const long long r2 = (long long)kp * kp + (long long)ip * ip + (long long)j * j;
    const int ires = min((int)floor(sqrt((double)r2) / g.pad), r_max - 1);
    const double c = acc[r_max + ires];
    const double avg = c > 0. ? acc[ires] / (1000. * c) : 0.;
    const double w = fmax(fweight[e], avg);
    if (w != 0.) {
        const double2 d = fdata[e];
        fdata[e] = make_double2(d.x / w, d.y / w);
    }
    fnew[e] = 1.;
}

__global__ void __launch_bounds__(kThreads) k_rc_product(const double *__restrict__ fnew, const double *__restrict__ fweight,
                                                         size_t n, double2 *__restrict__ fconv) {
    const size_t e = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (e < n) fconv[e] = make_double2(fnew[e] * fweight[e], 0.);
}

// Real-space multiplication by the normalised blob transform (convoluteBlobRealSpace, src/backprojector.cpp:2483-2528):
// radius from the wrapped indices (k < S/2 ? k : k - S) over ori * pad, table entry at floor(rval / (0.5 / n)).
__global__ void __launch_bounds__(kThreads) k_rc_blob(double *__restrict__ m, int S, double inv_oripad,
                                                      const double *__restrict__ table) {
    const size_t n = (size_t)S * S * S;
    const size_t v = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (v >= n) return;
    const int x = (int)(v % S);
    const size_t zy = v / S;
    const int y = (int)(zy % S), z = (int)(zy / S);
    const int hh = S / 2;
    const int kp = z < hh ? z : z - S, ip = y < hh ? y : y - S, jp = x < hh ? x : x - S;
    const double rval = sqrt((double)kp * kp + (double)ip * ip + (double)jp * jp) * inv_oripad;
    const int idx = (int)(rval / (0.5 / kBlobTable));
    const double f = idx < kBlobTable ? table[idx] / table[0] : 0.;
    m[v] *= f;
}

__global__ void __launch_bounds__(kThreads) k_rc_update(const double2 *__restrict__ fconv, Geo g, double inv_n,
                                                        double *__restrict__ fnew) {
    const int S = g.S, xs = g.h + 1;
    const size_t n = (size_t)S * S * xs;
    const size_t e = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (e >= n) return;
    const int j = (int)(e % xs);
    const size_t zy = e / xs;
    const int i = (int)(zy % S), k = (int)(zy / S);
    const int kp = fftw_k(k, S, xs), ip = fftw_k(i, S, xs);
    const long long r2 = (long long)kp * kp + (long long)ip * ip + (long long)j * j;
    if (r2 < g.max_r2) {
        const double2 c = fconv[e];
        const double a = sqrt(c.x * c.x + c.y * c.y) * inv_n;
        fnew[e] /= fmax(1e-6, a);
    }
}

// Fconv = data * Fnewweight windowed from the S^3 to the P^3 half-complex layout (windowFourierTransform,
// src/fftw.h:809-860), with CenterFFTbySign on the P array. One thread per P element.
__global__ void __launch_bounds__(kThreads) k_rc_window(const double2 *__restrict__ fdata, const double *__restrict__ fnew,
                                                        Geo g, int P, double2 *__restrict__ out) {
    const int xp = P / 2 + 1, S = g.S, xs = g.h + 1;
    const size_t n = (size_t)P * P * xp;
    const size_t e = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (e >= n) return;
    const int j = (int)(e % xp);
    const size_t zy = e / xp;
    const int i = (int)(zy % P), k = (int)(zy / P);
    const int kp = fftw_k(k, P, xp), ip = fftw_k(i, P, xp), jp = j;
    double2 v = make_double2(0., 0.);
    bool take;
    if (xp <= xs) {
        take = true;  // shrink: every element of the smaller transform from the larger one
    } else {          // grow: elements of the smaller one within (S/2)^2
        take = kp >= -g.h && kp <= g.h && ip >= -g.h && ip <= g.h && jp <= g.h &&
               (long long)kp * kp + (long long)ip * ip + (long long)jp * jp <= (long long)g.h * g.h;
    }
    if (take) {
        const int sk = kp < 0 ? kp + S : kp, si = ip < 0 ? ip + S : ip;
        const size_t src = ((size_t)sk * S + si) * xs + jp;
        const double f = fnew[src];
        const double2 d = fdata[src];
        const double sg = ((k + i + j) & 1) ? -f : f;
        v = make_double2(sg * d.x, sg * d.y);
    }
    out[e] = v;
}

// Average background of the soft mask over the original-size window (src/mask.cpp:53-74): sum of w and w * v,
// grid-stride over a bounded grid (two FP64 atomics per block).
__global__ void __launch_bounds__(kThreads) k_rc_bg(const double *__restrict__ m, int P, int ori, double scale,
                                                    double radius, double width, double *__restrict__ acc) {
    __shared__ double red[2][kThreads / 64];
    const size_t n = (size_t)ori * ori * ori;
    double sw = 0., swv = 0.;
    for (size_t v = (size_t)blockIdx.x * kThreads + threadIdx.x; v < n; v += (size_t)gridDim.x * kThreads) {
        const int x = (int)(v % ori);
        const size_t zy = v / ori;
        const int y = (int)(zy % ori), z = (int)(zy / ori);
        const int h = ori / 2, o = P / 2 - h;
        const double r = sqrt((double)(x - h) * (x - h) + (double)(y - h) * (y - h) + (double)(z - h) * (z - h));
        if (r >= radius) {
            const double val = m[((size_t)(z + o) * P + (y + o)) * P + (x + o)] * scale;
            const double w = r > radius + width ? 1. : 0.5 + 0.5 * cos(3.14159265358979323846 * (radius + width - r) / width);
            sw += w;
            swv += w * val;
        }
    }
    for (int s = 32; s > 0; s >>= 1) {
        sw += __shfl_xor(sw, s);
        swv += __shfl_xor(swv, s);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = sw;
        red[1][threadIdx.x >> 6] = swv;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0., b = 0.;
        for (int q = 0; q < kThreads / 64; ++q) {
            a += red[0][q];
            b += red[1][q];
        }
        if (a != 0.) {
            atomicAdd(&acc[0], a);
            atomicAdd(&acc[1], b);
        }
    }
}

// Window, scale, soft mask (softMaskOutsideMap, src/mask.cpp:76-95) and gridding correction (src/projector.cpp:595-628).
__global__ void __launch_bounds__(kThreads) k_rc_out(const double *__restrict__ m, int P, int ori, double scale,
                                                     double radius, double width, const double *__restrict__ acc,
                                                     double inv_oripad, float *__restrict__ vol) {
    const size_t n = (size_t)ori * ori * ori;
    const size_t v = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (v >= n) return;
    const int x = (int)(v % ori);
    const size_t zy = v / ori;
    const int y = (int)(zy % ori), z = (int)(zy / ori);
    const int h = ori / 2, o = P / 2 - h;
    double val = m[((size_t)(z + o) * P + (y + o)) * P + (x + o)] * scale;
    const double r = sqrt((double)(x - h) * (x - h) + (double)(y - h) * (y - h) + (double)(z - h) * (z - h));
    if (r >= radius) {
        const double bg = acc[0] != 0. ? acc[1] / acc[0] : 0.;
        if (r > radius + width) {
            val = bg;
        } else {
            const double rc = 0.5 + 0.5 * cos(3.14159265358979323846 * (radius + width - r) / width);
            val = (1. - rc) * val + rc * bg;
        }
    }
    if (r > 0.) {
        const double rv = 3.14159265358979323846 * r * inv_oripad;
        const double sinc = sin(rv) / rv;
        val /= sinc * sinc;
    }
    vol[v] = (float)val;
}

// I_0 by its power series, I_{3/2} and J_{3/2} by their closed forms (half-integer orders).
double bessel_i0(double x) {
    double sum = 1., term = 1.;
    for (int k = 1; k < 200; ++k) {
        term *= (x / (2. * k)) * (x / (2. * k));
        sum += term;
        if (term < 1e-17 * sum) break;
    }
    return sum;
}

// Lewitt's transform of the order-0 Kaiser-Bessel blob (radius a, taper alpha) at frequency w.
double kaiser_ft(double w, double a, double alpha) {
    const double pi = 3.14159265358979323846;
    const double t = 2. * pi * a * w;
    const double s = sqrt(fabs(alpha * alpha - t * t));
    const double c = pow(2. * pi, 1.5) * a * a * a / bessel_i0(alpha);
    double f;
    if (s < 1e-3) {
        f = sqrt(2. / pi) / 3. * (t > alpha ? 1. - s * s / 10. : 1. + s * s / 10.);
    } else if (t > alpha) {
        f = sqrt(2. / (pi * s)) * (sin(s) / s - cos(s)) / pow(s, 1.5);
    } else {
        f = sqrt(2. / (pi * s)) * (cosh(s) - sinh(s) / s) / pow(s, 1.5);
    }
    return c * f;
}

}  // namespace rlnc

extern "C" int rln_backprojector_reconstruct(rln_backprojector *b, int ori_size, int skip_gridding, int max_iter_preweight,
                                             int do_map,
                                             const double *tau2, int n_tau2, double tau2_fudge, double normalise,
                                             int minres_map, float *vol_out, int out_on_device) {
    if (!b || !b->slots.ptr || !vol_out) RLN_FAIL(RLN_ERR_STATE, "backprojector not dimensioned or null output");
    if (ori_size < 4 || (ori_size & 1)) RLN_FAIL(RLN_ERR_ARG, "ori_size %d must be even and at least 4", ori_size);
    if (do_map && (!tau2 || n_tau2 < 1)) RLN_FAIL(RLN_ERR_ARG, "MAP reconstruction needs the tau2 spectrum");
    if (!(normalise > 0.)) RLN_FAIL(RLN_ERR_ARG, "invalid normalise %g", normalise);
    const bool gridding = !skip_gridding && max_iter_preweight > 0;  // src/backprojector.cpp:1507
    const MdlDims &d = b->dims;
    const double pad = d.pad;
    const int P0 = ROUND(pad * ori_size);
    if (std::fabs(pad * ori_size - P0) > 1e-9 || (P0 & 1))
        RLN_FAIL(RLN_ERR_ARG, "padding_factor * ori_size must be an even integer (%g * %d)", pad, ori_size);
    rlnc::Geo g;
    g.S = d.ydim;
    g.h = g.S / 2;
    g.pad = pad;
    const long long rr = ROUND(d.maxr * pad);
    g.max_r2 = rr * rr;
    const int S = g.S, P = P0, r_max = d.maxr;
    if (r_max < 1 || r_max > rlnc::kMaxShells) RLN_FAIL(RLN_ERR_STATE, "backprojector r_max %d out of range", r_max);
    if (do_map && n_tau2 < r_max + 1)
        RLN_FAIL(RLN_ERR_ARG, "tau2 holds %d shells, the MAP term needs %d", n_tau2, r_max + 1);
    const size_t nS = (size_t)S * S * (g.h + 1), nR = (size_t)S * S * S, nP = (size_t)P * P * (P / 2 + 1),
                 nPr = (size_t)P * P * P, nv = d.voxels(), nvol = (size_t)ori_size * ori_size * ori_size;
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipDeviceSynchronize());  // the E-step's scatter into the slots has finished
    rln::MapWorkspace &ws = rln::map_workspace(b->device);
    std::lock_guard<std::mutex> lk(ws.mtx);
    hipStream_t st;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {  // drains the stream before the workspace is released to the next call
        hipStream_t s;
        ~StreamGuard() {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    } guard{st};
    DevBuf &d_plain = ws.buf[0], &d_fw = ws.buf[1], &d_fn = ws.buf[2], &d_fd = ws.buf[3], &d_fc = ws.buf[4],
           &d_m = ws.buf[5], &d_table = ws.buf[6], &d_tau = ws.buf[7], &d_acc = ws.buf[8], &d_out = ws.buf[9];
    RLN_TRY(d_plain.ensure(3 * sizeof(float) * nv));
    RLN_TRY(d_fw.ensure(sizeof(double) * nS));
    RLN_TRY(d_fn.ensure(sizeof(double) * nS));
    RLN_TRY(d_fd.ensure(sizeof(double2) * nS));
    RLN_TRY(d_fc.ensure(sizeof(double2) * std::max(nS, nP)));
    RLN_TRY(d_m.ensure(sizeof(double) * std::max(nR, nPr)));
    RLN_TRY(d_table.ensure(sizeof(double) * rlnc::kBlobTable));
    RLN_TRY(d_tau.ensure(sizeof(double) * std::max(1, n_tau2)));
    RLN_TRY(d_acc.ensure(sizeof(double) * 2 * r_max));
    float *re = d_plain.as<float>(), *im = re + nv, *wt = im + nv;
    RLN_TRY(rln::fold_backprojector_device(b, re, im, wt, st));
    // blob transform table (radius 2 * 1.9, alpha 15, order 0: src/backprojector.h:80, 113-115)
    std::vector<double> table(rlnc::kBlobTable);
    for (int i = 0; i < rlnc::kBlobTable; ++i) table[i] = rlnc::kaiser_ft(i * (0.5 / rlnc::kBlobTable), 2. * 1.9, 15.);
    HIP_TRY(hipMemcpyAsync(d_table.ptr, table.data(), sizeof(double) * table.size(), hipMemcpyHostToDevice, st));
    if (do_map) HIP_TRY(hipMemcpyAsync(d_tau.ptr, tau2, sizeof(double) * n_tau2, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(d_acc.ptr, 0, sizeof(double) * 2 * r_max, st));
    auto grid = [](size_t n) { return dim3((unsigned)((n + rlnc::kThreads - 1) / rlnc::kThreads)); };
    // reductions: at most 2048 blocks (8 per CU), each looping over its share
    auto reduce_grid = [](size_t n) {
        return dim3((unsigned)std::min<size_t>(2048, (n + rlnc::kThreads - 1) / rlnc::kThreads));
    };
    const double ovs = pad * pad * pad;
    hipLaunchKernelGGL(rlnc::k_rc_init, grid(nS), dim3(rlnc::kThreads), 0, st, re, im, wt, g, do_map ? 1 : 0,
                       d_tau.as<const double>(), std::max(1, n_tau2), 1. / (ovs * tau2_fudge), minres_map,
                       gridding ? normalise : 1., d_fw.as<double>(), d_fn.as<double>(), d_fd.as<double2>());
    HIP_TRY(hipGetLastError());
    if (!gridding) {
        const double rm = (double)r_max * pad;
        hipLaunchKernelGGL(rlnc::k_rc_radavg, reduce_grid(nS), dim3(rlnc::kThreads), 0, st, d_fw.as<const double>(), g,
                           (long long)ROUND(rm * rm), r_max, d_acc.as<double>());
        hipLaunchKernelGGL(rlnc::k_rc_skip, grid(nS), dim3(rlnc::kThreads), 0, st, d_fw.as<const double>(), g, r_max,
                           d_acc.as<const double>(), d_fd.as<double2>(), d_fn.as<double>());
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemsetAsync(d_acc.ptr, 0, sizeof(double) * 2, st));  // reused by the soft mask's background sums
    }
    hipfftHandle c2r_s = 0, r2c_s = 0, c2r_p = 0;
    if (gridding) {
        RLN_TRY(ws.plan3d(S, HIPFFT_Z2D, st, &c2r_s));
        RLN_TRY(ws.plan3d(S, HIPFFT_D2Z, st, &r2c_s));
    }
    for (int it = 0; gridding && it < max_iter_preweight; ++it) {
        hipLaunchKernelGGL(rlnc::k_rc_product, grid(nS), dim3(rlnc::kThreads), 0, st, d_fn.as<const double>(),
                           d_fw.as<const double>(), nS, d_fc.as<double2>());
        if (hipfftExecZ2D(c2r_s, d_fc.as<hipfftDoubleComplex>(), d_m.as<hipfftDoubleReal>()) != HIPFFT_SUCCESS)
            RLN_FAIL(RLN_ERR_HIP, "Z2D failed");
        hipLaunchKernelGGL(rlnc::k_rc_blob, grid(nR), dim3(rlnc::kThreads), 0, st, d_m.as<double>(), S,
                           1. / (ori_size * pad), d_table.as<const double>());
        if (hipfftExecD2Z(r2c_s, d_m.as<hipfftDoubleReal>(), d_fc.as<hipfftDoubleComplex>()) != HIPFFT_SUCCESS)
            RLN_FAIL(RLN_ERR_HIP, "D2Z failed");
        hipLaunchKernelGGL(rlnc::k_rc_update, grid(nS), dim3(rlnc::kThreads), 0, st, d_fc.as<const double2>(), g,
                           1. / (double)nR, d_fn.as<double>());
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(rlnc::k_rc_window, grid(nP), dim3(rlnc::kThreads), 0, st, d_fd.as<const double2>(),
                       d_fn.as<const double>(), g, P, d_fc.as<double2>());
    HIP_TRY(hipGetLastError());
    RLN_TRY(ws.plan3d(P, HIPFFT_Z2D, st, &c2r_p));
    if (hipfftExecZ2D(c2r_p, d_fc.as<hipfftDoubleComplex>(), d_m.as<hipfftDoubleReal>()) != HIPFFT_SUCCESS)
        RLN_FAIL(RLN_ERR_HIP, "Z2D of the padded box failed");
    const double scale = 1. / (ovs * ori_size), radius = ori_size / 2., width = 3.;  // softMaskOutsideMap defaults
    hipLaunchKernelGGL(rlnc::k_rc_bg, reduce_grid(nvol), dim3(rlnc::kThreads), 0, st, d_m.as<const double>(), P, ori_size, scale,
                       radius, width, d_acc.as<double>());
    float *dst = vol_out;
    if (!out_on_device) {
        RLN_TRY(d_out.ensure(sizeof(float) * nvol));
        dst = d_out.as<float>();
    }
    hipLaunchKernelGGL(rlnc::k_rc_out, grid(nvol), dim3(rlnc::kThreads), 0, st, d_m.as<const double>(), P, ori_size, scale,
                       radius, width, d_acc.as<const double>(), 1. / (ori_size * pad), dst);
    HIP_TRY(hipGetLastError());
    if (!out_on_device) HIP_TRY(hipMemcpyAsync(vol_out, dst, sizeof(float) * nvol, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return RLN_OK;
}
