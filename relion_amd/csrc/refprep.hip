// Reference preparation on the GPU: Projector::computeFourierTransformMap (src/projector.cpp:116-593) for a 3-D
// reference and 2-D data, with the trilinear gridding correction (Projector::griddingCorrect, :595-628). The real-space
// map becomes the projector's padded Fourier volume without leaving device memory:
//   k_grid_pad        gridding correction (divide by sinc^2(r / padoridim)) and zero padding to padoridim^3
//   hipFFT R2C        3-D forward transform of the padded map
//   k_fourier_window  1/padoridim^3 and pad^3 * ori scaling, CenterFFTbySign, the r <= ROUND(pad r_max) window into
//                     the projector's plain array, per-shell power (FP64 shared-memory bins), low-resolution cut
//   k_pack_volume     the plain array into the bricked corner layout of the projector
// All three kernels stream through their volumes once (HBM-bound); the FFT is the library's.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/rln_estep.h"
#include "host_common.h"
#include "map_workspace.h"

using namespace rln;

namespace rlnr {

constexpr int kThreads = 256;
constexpr int kMaxShells = 1025;

// Gridding correction and padding (src/projector.cpp:595-628, :377-392): out[z][y][x] of the padded box P^3 holds
// vol / sinc^2(r / P) inside the original box (both centred at half their size), 0 outside.
__global__ void __launch_bounds__(kThreads) k_grid_pad(const float *__restrict__ vol, int ori, int P, int do_gridding,
                                                       float *__restrict__ out) {
    const size_t n = (size_t)P * P * P;
    const size_t v = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (v >= n) return;
    const int x = (int)(v % P);
    const size_t zy = v / P;
    const int y = (int)(zy % P), z = (int)(zy / P);
    const int k = z - P / 2, i = y - P / 2, j = x - P / 2;  // logical coordinates
    const int sz = k + ori / 2, sy = i + ori / 2, sx = j + ori / 2;
    float val = 0.f;
    if (sz >= 0 && sz < ori && sy >= 0 && sy < ori && sx >= 0 && sx < ori) {
        double f = (double)vol[((size_t)sz * ori + sy) * ori + sx];
        if (do_gridding) {
            const double r = sqrt((double)(k * k + i * i + j * j));
            if (r > 0.) {
                const double rv = 3.14159265358979323846 * r / (double)P;
                const double sinc = sin(rv) / rv;
                f /= sinc * sinc;
            }
        }
        val = (float)f;
    }
    out[v] = val;
}

// Window of the centred transform into the projector's plain array [S][S][S/2+1] (S = pad_size, z and y from -S/2,
// x from 0) and the per-shell power (src/projector.cpp:506-530). F: unnormalised R2C output [P][P][P/2+1].
// Grid-stride over a bounded grid, so that each global shell bin takes few FP64 atomics.
__global__ void __launch_bounds__(kThreads) k_fourier_window(const float2 *__restrict__ F, int P, int S, long long max_r2,
                                                             long long min_r2, double scale, double pad, int nshell,
                                                             float2 *__restrict__ data, double *__restrict__ power,
                                                             double *__restrict__ count) {
    __shared__ double s_pow[kMaxShells], s_cnt[kMaxShells];
    for (int t = threadIdx.x; t < nshell; t += kThreads) s_pow[t] = s_cnt[t] = 0.;
    __syncthreads();
    const int h = S / 2, xs = h + 1;
    const size_t n = (size_t)S * S * xs;
    for (size_t q = (size_t)blockIdx.x * kThreads + threadIdx.x; q < n; q += (size_t)gridDim.x * kThreads) {
        const int ix = (int)(q % xs);
        const size_t zy = q / xs;
        const int iy = (int)(zy % S), iz = (int)(zy / S);
        const int kz = iz - h, ky = iy - h, kx = ix;
        const long long r2 = (long long)kz * kz + (long long)ky * ky + (long long)kx * kx;
        float2 out = make_float2(0.f, 0.f);
        // the transform holds z and y frequencies -(P/2 - 1) .. P/2: -P/2 has no element and stays zero
        if (r2 <= max_r2 && kz > -P / 2 && ky > -P / 2) {
            const int pz = kz < 0 ? kz + P : kz, py = ky < 0 ? ky + P : ky;
            const float2 f = F[((size_t)pz * P + py) * (P / 2 + 1) + kx];
            const double sg = ((pz + py + kx) & 1) ? -scale : scale;  // CenterFFTbySign on the transform's indices
            const double re = sg * (double)f.x, im = sg * (double)f.y;
            const int ires = (int)floor(sqrt((double)r2) / pad + 0.5);
            if (ires < nshell) {
                atomicAdd(&s_pow[ires], (re * re + im * im) / 2.);
                atomicAdd(&s_cnt[ires], 1.);
            }
            if (r2 > min_r2) out = make_float2((float)re, (float)im);
        }
        data[q] = out;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nshell; t += kThreads)
        if (s_cnt[t] != 0.) {
            atomicAdd(&power[t], s_pow[t]);
            atomicAdd(&count[t], s_cnt[t]);
        }
}

}  // namespace rlnr

extern "C" int rln_projector_compute_fourier_map(rln_projector *p, const float *vol, int on_device, int ori_size,
                                                 int current_size, double padding_factor, int do_gridding, int min_ires,
                                                 double *power_spectrum, double *data_out) {
    if (!p || !vol) RLN_FAIL(RLN_ERR_ARG, "rln_projector_compute_fourier_map: null argument");
    if (ori_size < 4 || (ori_size & 1)) RLN_FAIL(RLN_ERR_ARG, "ori_size %d must be even and at least 4", ori_size);
    if (!(padding_factor >= 1.)) RLN_FAIL(RLN_ERR_ARG, "padding_factor %g below 1", padding_factor);
    // sizes (src/projector.cpp:59-89, :131-134)
    int P = ROUND(padding_factor * ori_size);
    P += P % 2;
    const double pad = (double)P / ori_size;
    const int r_max = std::min(current_size > 0 ? current_size : ori_size, ori_size) / 2;
    const int rr = ROUND(pad * r_max);
    const int S = 2 * (rr + 1) + 1, h = S / 2;
    const long long max_r2 = (long long)rr * rr;
    const int mr = min_ires > 0 ? ROUND(min_ires * pad) : -1;
    const long long min_r2 = min_ires > 0 ? (long long)mr * mr : -1;
    const int nshell = ori_size / 2 + 1;
    if (nshell > rlnr::kMaxShells) RLN_FAIL(RLN_ERR_ARG, "ori_size %d above the supported %d", ori_size, 2 * (rlnr::kMaxShells - 1));
    const double normfft = pad * pad * pad * ori_size;  // 3-D reference, 2-D data
    RLN_TRY(rln_projector_set_mdl_dim(p, h + 1, S, S, -h, -h, r_max, pad, nullptr));
    HIP_TRY(hipSetDevice(p->device));
    rln::MapWorkspace &ws = rln::map_workspace(p->device);
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
    const size_t nvol = (size_t)ori_size * ori_size * ori_size, npad = (size_t)P * P * P;
    const size_t nF = (size_t)P * P * (P / 2 + 1), nd = (size_t)S * S * (h + 1);
    DevBuf &d_vol = ws.buf[0], &d_pad = ws.buf[1], &d_F = ws.buf[2], &d_data = ws.buf[3], &d_pow = ws.buf[4];
    const float *src = vol;
    if (!on_device) {
        RLN_TRY(d_vol.ensure(sizeof(float) * nvol));
        HIP_TRY(hipMemcpyAsync(d_vol.ptr, vol, sizeof(float) * nvol, hipMemcpyHostToDevice, st));
        src = d_vol.as<const float>();
    }
    RLN_TRY(d_pad.ensure(sizeof(float) * npad));
    RLN_TRY(d_F.ensure(sizeof(float2) * nF));
    RLN_TRY(d_data.ensure(sizeof(float2) * nd));
    RLN_TRY(d_pow.ensure(sizeof(double) * 2 * nshell));
    HIP_TRY(hipMemsetAsync(d_pow.ptr, 0, sizeof(double) * 2 * nshell, st));
    hipLaunchKernelGGL(rlnr::k_grid_pad, dim3((unsigned)((npad + rlnr::kThreads - 1) / rlnr::kThreads)),
                       dim3(rlnr::kThreads), 0, st, src, ori_size, P, do_gridding ? 1 : 0, d_pad.as<float>());
    HIP_TRY(hipGetLastError());
    hipfftHandle plan = 0;
    RLN_TRY(ws.plan3d(P, HIPFFT_R2C, st, &plan));
    if (hipfftExecR2C(plan, d_pad.as<hipfftReal>(), d_F.as<hipfftComplex>()) != HIPFFT_SUCCESS)
        RLN_FAIL(RLN_ERR_HIP, "hipfftExecR2C failed");
    hipLaunchKernelGGL(rlnr::k_fourier_window,
                       dim3((unsigned)std::min<size_t>(2048, (nd + rlnr::kThreads - 1) / rlnr::kThreads)),
                       dim3(rlnr::kThreads), 0, st, d_F.as<const float2>(), P, S, max_r2, min_r2,
                       normfft / (double)npad, pad, nshell, d_data.as<float2>(), d_pow.as<double>(),
                       d_pow.as<double>() + nshell);
    HIP_TRY(hipGetLastError());
    RLN_TRY(rln::pack_device_into_projector(p, d_data.as<const float2>(), st));
    std::vector<double> pw(2 * (size_t)nshell);
    HIP_TRY(hipMemcpyAsync(pw.data(), d_pow.ptr, sizeof(double) * 2 * nshell, hipMemcpyDeviceToHost, st));
    std::vector<float2> hd;
    if (data_out) {
        hd.resize(nd);
        HIP_TRY(hipMemcpyAsync(hd.data(), d_data.ptr, sizeof(float2) * nd, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    if (power_spectrum)
        for (int i = 0; i < nshell; ++i) power_spectrum[i] = pw[nshell + i] < 1. ? 0. : pw[i] / pw[nshell + i];
    if (data_out)
        for (size_t i = 0; i < nd; ++i) {
            data_out[2 * i] = hd[i].x;
            data_out[2 * i + 1] = hd[i].y;
        }
    return RLN_OK;
}
