// Particle image preparation on the GPU (include/rln_imgprep.h): the image part of
// MlOptimiser::getFourierTransformsAndCtfs (src/ml_optimiser.cpp:5840-6531) for 2-D single-image particles.
//
// One chunk of particles goes through three kernels and the library FFT:
//   k_prep_real     one workgroup per image: norm correction, integer shift with zero fill, soft mask
//                   (average background, or the noise image) -> masked and unmasked real images
//   hipFFT R2C      one batched plan over the 2 x chunk real images
//   k_prep_fourier  one workgroup per image: 1/N^2 scale, power spectrum and high-resolution power of the masked
//                   transform, window to current_size, CenterFFTbySign, CTF image
// With a noise background, k_noise_fourier and a batched C2R plan make the noise images first.
// All three kernels are bandwidth-bound streams over the images (about 1 MB per particle at N = 256).
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/rln_imgprep.h"
#include "host_common.h"

using namespace rln;

namespace rlnp {

constexpr int kThreads = 256;
constexpr int kMaxHalf = 1025;  // ori_size / 2 + 1 for boxes up to 2048

// Per-image parameters, prepared on the host.
struct PrepImage {
    double axx, axy, ayy;  // astigmatic defocus bilinear form (src/ctf.cpp:134-261)
    double k4, k5;         // -bfactor / 4, phase shift in radians
    double ctf_scale;      // kfactor
    double norm;           // avg_norm_correction / norm_correction (1 without norm correction)
    int dx, dy;            // rounded old offset
    unsigned long long seed;
};

// Constants of one run.
struct PrepConst {
    int N, s;
    double radius, radius_p, width;
    int zero_mask;
    double inv_box_A;  // 1 / (ori_size * pixel_size)
    double K1, K2, K3;
    int do_ctf, do_abs, flip, intact, premult;
};

// splitmix64 finaliser
__host__ __device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Two N(0,1) deviates for element e of the image with seed `seed` (Box-Muller in FP64).
__device__ __forceinline__ double2 noise_pair(unsigned long long seed, unsigned long long e) {
    const unsigned long long h1 = mix64(mix64(seed) + e);
    const unsigned long long h2 = mix64(h1);
    const double u1 = ((double)(h1 >> 11) + 0.5) * 0x1.0p-53;
    const double u2 = ((double)(h2 >> 11) + 0.5) * 0x1.0p-53;
    const double r = sqrt(-2.0 * log(u1));
    double sn, cs;
    sincos(6.283185307179586 * u2, &sn, &cs);
    return make_double2(r * cs, r * sn);
}

__device__ __forceinline__ double block_sum(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0.;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
    __syncthreads();
    if (threadIdx.x == 0) red[0] = t;
    __syncthreads();
    const double r = red[0];
    __syncthreads();
    return r;
}

// Background noise in Fourier space (src/ml_optimiser.cpp:6300-6335): real and imaginary parts N(0, sigma),
// sigma = sqrt(fudge * sigma2[ires]), ires = ROUND(|k|) < N/2 + 1, zero beyond. The x = 0 and x = N/2 columns are
// made Hermitian (the value at -y is the conjugate of the one at y; y = 0 and y = N/2 are real) so that every
// inverse real FFT returns the same image. grid (ceil(N*(N/2+1)/256), chunk); out [chunk][N][N/2+1].
__global__ void __launch_bounds__(kThreads) k_noise_fourier(const double *__restrict__ sigma2, double fudge,
                                                            const PrepImage *__restrict__ img, int N,
                                                            float2 *__restrict__ out) {
    const int xh = N / 2 + 1;
    const int e = blockIdx.x * kThreads + threadIdx.x;
    if (e >= N * xh) return;
    const int b = blockIdx.y;
    const int i = e / xh, j = e - i * xh;
    const int ip = (i < xh) ? i : i - N;
    int gi = i;  // element whose deviates are used
    bool conj = false, real_only = false;
    if (j == 0 || j == N / 2) {
        if (ip < 0) {
            gi = -ip;
            conj = true;
        } else if (ip == 0 || ip == N / 2) {
            real_only = true;
        }
    }
    const int ires = (int)rint(sqrt((double)(ip * ip + j * j)));
    float2 v = make_float2(0.f, 0.f);
    if (ires < xh) {
        const double sg = sqrt(fudge * sigma2[ires]);
        const double2 z = noise_pair(img[b].seed, (unsigned long long)gi * xh + j);
        v.x = (float)(sg * z.x);
        v.y = real_only ? 0.f : (float)((conj ? -sg : sg) * z.y);
    }
    out[(size_t)b * N * xh + e] = v;
}

// Real-space stage (src/ml_optimiser.cpp:6225-6247, src/mask.cpp:43-96). One workgroup per image.
// in [chunk][N][N]; noise [chunk][N][N] or null; out [2][chunk][N][N]: masked images, then unmasked ones.
__global__ void __launch_bounds__(kThreads) k_prep_real(const float *__restrict__ in, const float *__restrict__ noise,
                                                        const PrepImage *__restrict__ img, PrepConst c, int chunk,
                                                        float *__restrict__ out) {
    __shared__ double red[kThreads / 64];
    const int b = blockIdx.x;
    const int N = c.N, h = N / 2;
    const PrepImage pi = img[b];
    const float *src = in + (size_t)b * N * N;
    const float *nz = noise ? noise + (size_t)b * N * N : nullptr;
    float *om = out + (size_t)b * N * N;
    float *ou = out + ((size_t)chunk + b) * N * N;
    const double pi_over_w = 3.14159265358979323846 / c.width;
    auto value = [&](int p) -> double {
        const int i = p / N, j = p - (p / N) * N;
        const int si = i - pi.dy, sj = j - pi.dx;
        if (si < 0 || si >= N || sj < 0 || sj >= N) return 0.;
        return pi.norm * (double)src[si * N + sj];
    };
    double bg = 0.;
    if (c.zero_mask) {
        // average background value, weighted by the raised cosine inside the edge
        double sw = 0., swv = 0.;
        for (int p = threadIdx.x; p < N * N; p += kThreads) {
            const int y = p / N - h, x = p - (p / N) * N - h;
            const double r = sqrt((double)(x * x + y * y));
            if (r < c.radius) continue;
            const double w = (r > c.radius_p) ? 1. : 0.5 + 0.5 * cos(pi_over_w * (c.radius_p - r));
            sw += w;
            swv += w * value(p);
        }
        sw = block_sum(sw, red);
        swv = block_sum(swv, red);
        bg = sw > 0. ? swv / sw : 0.;
    }
    for (int p = threadIdx.x; p < N * N; p += kThreads) {
        const int y = p / N - h, x = p - (p / N) * N - h;
        const double v = value(p);
        ou[p] = (float)v;
        const double r = sqrt((double)(x * x + y * y));
        double m = v;
        if (r >= c.radius) {
            const double add = nz ? (double)nz[p] : bg;
            if (r > c.radius_p) {
                m = add;
            } else {
                const double rc = 0.5 + 0.5 * cos(pi_over_w * (c.radius_p - r));
                m = (1. - rc) * v + rc * add;
            }
        }
        om[p] = (float)m;
    }
}

// CTF::getCTF (src/ctf.h:184-240) with do_damping and no dose model, at spatial frequency (X, Y) in 1/A, FP64.
__device__ __forceinline__ double ctf_value(const PrepConst &c, const PrepImage &pi, double X, double Y) {
    const double u2 = X * X + Y * Y;
    const double gamma = c.K1 * (pi.axx * X * X + 2.0 * pi.axy * X * Y + pi.ayy * Y * Y) + c.K2 * u2 * u2 - pi.k5 - c.K3;
    double r = (c.intact && fabs(gamma) < 1.5707963267948966) ? 1.0 : -sin(gamma);
    r *= exp(pi.k4 * u2);
    if (c.do_abs)
        r = fabs(r);
    else if (c.flip)
        r = (r < 0.) ? -1. : 1.;
    r *= pi.ctf_scale;
    if (fabs(r) < 1e-8) r = (r >= 0. ? 1. : -1.) * 1e-8;
    return r;
}

// Fourier-space stage (src/ml_optimiser.cpp:6278-6289, 6379-6420, 6483-6531). One workgroup per image.
// F [2][chunk][N][N/2+1] unnormalised forward transforms (masked, then unmasked).
// Outputs [chunk][s][s/2+1]: fimg, fimg_nomask (float2), fctf (float); power [chunk][N/2+1], hx [chunk] (FP64).
__global__ void __launch_bounds__(kThreads) k_prep_fourier(const float2 *__restrict__ F,
                                                           const PrepImage *__restrict__ img, PrepConst c, int chunk,
                                                           float2 *__restrict__ fimg, float2 *__restrict__ fimg_nm,
                                                           float *__restrict__ fctf, double *__restrict__ power,
                                                           double *__restrict__ hx_out) {
    __shared__ double spec[kMaxHalf];
    __shared__ double red[kThreads / 64];
    const int b = blockIdx.x;
    const int N = c.N, xh = N / 2 + 1, s = c.s, xs = s / 2 + 1;
    const float inv = 1.f / ((float)N * (float)N);
    const PrepImage pi = img[b];
    for (int k = threadIdx.x; k < xh; k += kThreads) spec[k] = 0.;
    __syncthreads();
    const float2 *Fm = F + (size_t)b * N * xh;
    const float2 *Fu = F + ((size_t)chunk + b) * N * xh;
    float2 *om = fimg + (size_t)b * s * xs;
    float2 *ou = fimg_nm + (size_t)b * s * xs;
    float *oc = fctf + (size_t)b * s * xs;
    const int hires_from = s / 2 + 1;
    double hx = 0.;
    // masked transform: spectrum over the full transform (Hermitian half of the x = 0 column skipped), window
    for (int e = threadIdx.x; e < N * xh; e += kThreads) {
        const int i = e / xh, j = e - i * xh;
        const int ip = (i < xh) ? i : i - N;
        const float2 f = make_float2(Fm[e].x * inv, Fm[e].y * inv);
        const int ires = (int)rintf(sqrtf((float)(ip * ip + j * j)));
        if (ires > 0 && ires < xh && !(j == 0 && ip < 0)) {
            const double p2 = (double)f.x * f.x + (double)f.y * f.y;
            atomicAdd(&spec[ires], p2);
            if (ires >= hires_from) hx += p2;
        }
        if (j < xs && ip <= s / 2 && ip > -s / 2) {
            const int io = ip >= 0 ? ip : ip + s;
            const float sg = ((io + j) & 1) ? -1.f : 1.f;
            om[io * xs + j] = make_float2(sg * f.x, sg * f.y);
        }
    }
    // unmasked transform and CTF on the window
    for (int q = threadIdx.x; q < s * xs; q += kThreads) {
        const int io = q / xs, j = q - io * xs;
        const int ip = (io < xs) ? io : io - s;
        const int i = ip >= 0 ? ip : ip + N;
        const float2 f = Fu[i * xh + j];
        const float sg = ((io + j) & 1) ? -inv : inv;
        ou[q] = make_float2(sg * f.x, sg * f.y);
        float cv = 1.f;
        if (c.do_ctf) {
            double v = ctf_value(c, pi, j * c.inv_box_A, ip * c.inv_box_A);
            if (c.premult) v *= v;
            cv = (float)v;
        }
        oc[q] = cv;
    }
    hx = block_sum(hx, red);
    __syncthreads();
    for (int k = threadIdx.x; k < xh; k += kThreads) power[(size_t)b * xh + k] = spec[k];
    if (threadIdx.x == 0) hx_out[b] = (s < N) ? hx : 0.;
}

}  // namespace rlnp

using namespace rlnp;

#define FFT_TRY(expr)                                                                        \
    do {                                                                                     \
        hipfftResult _r = (expr);                                                            \
        if (_r != HIPFFT_SUCCESS) RLN_FAIL(RLN_ERR_HIP, "%s failed: hipfft status %d", #expr, (int)_r); \
    } while (0)

struct rln_imgprep {
    int device = 0;
    rln_imgprep_config cfg{};
    hipStream_t st = nullptr;
    std::mutex mu;
    int chunk = 0;
    DevBuf d_in, d_real, d_cplx, d_noise_c, d_noise_r, d_img, d_sigma, d_power, d_hx, d_out;
    std::vector<double> sigma2;
    std::map<int, hipfftHandle> r2c, c2r;
    hipEvent_t ev[6] = {};
    double ms[5] = {};
};

namespace {

int plan_for(std::map<int, hipfftHandle> &cache, hipfftType type, int N, int batch, hipStream_t st,
             hipfftHandle *out) {
    auto it = cache.find(batch);
    if (it != cache.end()) {
        *out = it->second;
        return RLN_OK;
    }
    if (cache.size() >= 4) {
        for (auto &kv : cache) hipfftDestroy(kv.second);
        cache.clear();
    }
    hipfftHandle h;
    int n2[2] = {N, N};
    FFT_TRY(hipfftPlanMany(&h, 2, n2, nullptr, 1, 0, nullptr, 1, 0, type, batch));
    if (hipfftSetStream(h, st) != HIPFFT_SUCCESS) {
        hipfftDestroy(h);
        RLN_FAIL(RLN_ERR_HIP, "hipfftSetStream failed");
    }
    cache[batch] = h;
    *out = h;
    return RLN_OK;
}

size_t per_image_bytes(const rln_imgprep_config &c) {
    const size_t N = c.ori_size, s = c.current_size;
    size_t b = N * N * 4 /* input */ + 2 * N * N * 4 /* real */ + 2 * N * (N / 2 + 1) * 8 /* transforms */;
    if (!c.zero_mask) b += N * (N / 2 + 1) * 8 + N * N * 4;
    b += s * (s / 2 + 1) * 20;  // staged outputs
    return b;
}

}  // namespace

extern "C" int rln_imgprep_create(int device, const rln_imgprep_config *cfg, rln_imgprep **out) {
    if (!cfg || !out) RLN_FAIL(RLN_ERR_ARG, "null argument");
    *out = nullptr;
    const int N = cfg->ori_size, s = cfg->current_size;
    if (N < 4 || (N & 1) || N > 2 * (kMaxHalf - 1))
        RLN_FAIL(RLN_ERR_ARG, "ori_size %d must be even and in [4, %d]", N, 2 * (kMaxHalf - 1));
    if (s < 2 || (s & 1) || s > N) RLN_FAIL(RLN_ERR_ARG, "current_size %d must be even and in [2, ori_size]", s);
    if (!(cfg->pixel_size > 0.)) RLN_FAIL(RLN_ERR_ARG, "pixel_size must be positive");
    if (!(cfg->width_mask_edge > 0.)) RLN_FAIL(RLN_ERR_ARG, "width_mask_edge must be positive");
    if (cfg->do_ctf_correction && (!(cfg->voltage > 0.) || !(cfg->q0 >= 0. && cfg->q0 < 1.)))
        RLN_FAIL(RLN_ERR_ARG, "invalid optics: voltage %g kV, Q0 %g", cfg->voltage, cfg->q0);
    HIP_TRY(hipSetDevice(device));
    auto *p = new rln_imgprep();
    p->device = device;
    p->cfg = *cfg;
    const size_t budget = (size_t)6 << 30;
    p->chunk = (int)std::max<size_t>(1, std::min<size_t>(4096, budget / per_image_bytes(*cfg)));
    if (hipStreamCreateWithFlags(&p->st, hipStreamNonBlocking) != hipSuccess) {
        delete p;
        RLN_FAIL(RLN_ERR_HIP, "stream creation failed");
    }
    for (auto &e : p->ev) (void)hipEventCreate(&e);
    *out = p;
    return RLN_OK;
}

extern "C" void rln_imgprep_destroy(rln_imgprep *p) {
    if (!p) return;
    (void)hipSetDevice(p->device);
    (void)hipStreamSynchronize(p->st);
    for (auto &kv : p->r2c) hipfftDestroy(kv.second);
    for (auto &kv : p->c2r) hipfftDestroy(kv.second);
    for (auto &e : p->ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(p->st);
    delete p;
}

extern "C" int rln_imgprep_set_sigma2_noise(rln_imgprep *p, const double *sigma2, int n) {
    if (!p || (!sigma2 && n > 0) || n < 0) RLN_FAIL(RLN_ERR_ARG, "invalid noise spectrum");
    std::lock_guard<std::mutex> lk(p->mu);
    const int xh = p->cfg.ori_size / 2 + 1;
    // shells beyond the image's Nyquist are dropped, missing ones are zero (remapping of equal boxes, :6309-6317)
    p->sigma2.assign(xh, 0.);
    for (int i = 0; i < std::min(n, xh); ++i) p->sigma2[i] = sigma2[i];
    return RLN_OK;
}

extern "C" int rln_imgprep_timings(rln_imgprep *p, double *ms5) {
    if (!p || !ms5) RLN_FAIL(RLN_ERR_ARG, "null argument");
    for (int i = 0; i < 5; ++i) ms5[i] = p->ms[i];
    return RLN_OK;
}

extern "C" int rln_imgprep_run(rln_imgprep *p, int n, const float *images, int in_on_device,
                               const rln_image_meta *meta, float *fimg, float *fimg_nomask, float *fctf,
                               int out_on_device, double *highres_Xi2, double *power, double *old_offset) {
    if (!p) RLN_FAIL(RLN_ERR_ARG, "null engine");
    if (n < 0) RLN_FAIL(RLN_ERR_ARG, "negative particle count %d", n);
    if (n == 0) return RLN_OK;
    if (!images || !meta || !fimg || !fimg_nomask || !fctf) RLN_FAIL(RLN_ERR_ARG, "null input or output");
    std::lock_guard<std::mutex> lk(p->mu);
    const rln_imgprep_config &cf = p->cfg;
    if (!cf.zero_mask && (int)p->sigma2.size() != cf.ori_size / 2 + 1)
        RLN_FAIL(RLN_ERR_STATE, "noise background needs rln_imgprep_set_sigma2_noise first");
    const bool do_norm = cf.avg_norm_correction > 0.;
    for (int b = 0; b < n; ++b)
        if (do_norm && !(meta[b].norm_correction > 0.))
            RLN_FAIL(RLN_ERR_ARG, "particle %d: norm correction %g is not positive", b, meta[b].norm_correction);
    HIP_TRY(hipSetDevice(p->device));
    const int N = cf.ori_size, s = cf.current_size, xh = N / 2 + 1, xs = s / 2 + 1;
    const int C = std::min(n, p->chunk);
    const size_t npix = (size_t)N * N, nf = (size_t)N * xh, no = (size_t)s * xs;

    PrepConst c{};
    c.N = N;
    c.s = s;
    c.radius = cf.particle_diameter / (2. * cf.pixel_size);
    if (c.radius < 0.) c.radius = N / 2.;  // src/mask.cpp:48-49
    c.width = cf.width_mask_edge;
    c.radius_p = c.radius + c.width;
    c.zero_mask = cf.zero_mask ? 1 : 0;
    c.inv_box_A = 1. / (N * cf.pixel_size);
    // CTF::initialise (src/ctf.cpp:134-261): kV -> V, mm -> A, relativistic wavelength
    const double kV = cf.voltage * 1e3, lambda = 12.2643247 / sqrt(kV * (1. + kV * 0.978466e-6));
    c.K1 = kPi / 2 * 2 * lambda;
    c.K2 = kPi / 2 * (cf.cs * 1e7) * lambda * lambda * lambda;
    c.K3 = atan(cf.q0 / sqrt(1 - cf.q0 * cf.q0));
    c.do_ctf = cf.do_ctf_correction ? 1 : 0;
    c.do_abs = cf.ctf_phase_flipped ? 1 : 0;
    c.flip = cf.only_flip_phases ? 1 : 0;
    c.intact = cf.intact_ctf_first_peak ? 1 : 0;
    c.premult = cf.ctf_premultiplied ? 1 : 0;

    RLN_TRY(p->d_real.ensure(sizeof(float) * 2 * npix * C));
    RLN_TRY(p->d_cplx.ensure(sizeof(float2) * 2 * nf * C));
    RLN_TRY(p->d_img.ensure(sizeof(PrepImage) * C));
    RLN_TRY(p->d_power.ensure(sizeof(double) * xh * C));
    RLN_TRY(p->d_hx.ensure(sizeof(double) * C));
    if (!in_on_device) RLN_TRY(p->d_in.ensure(sizeof(float) * npix * C));
    if (!out_on_device) RLN_TRY(p->d_out.ensure(sizeof(float) * 5 * no * C));
    if (!cf.zero_mask) {
        RLN_TRY(p->d_noise_c.ensure(sizeof(float2) * nf * C));
        RLN_TRY(p->d_noise_r.ensure(sizeof(float) * npix * C));
        RLN_TRY(p->d_sigma.ensure(sizeof(double) * xh));
        HIP_TRY(hipMemcpyAsync(p->d_sigma.ptr, p->sigma2.data(), sizeof(double) * xh, hipMemcpyHostToDevice, p->st));
    }
    std::vector<PrepImage> hp(C);
    std::vector<double> hpow((size_t)xh * C), hhx(C);
    for (int i = 0; i < 5; ++i) p->ms[i] = 0.;
    HIP_TRY(hipEventRecord(p->ev[5], p->st));
    for (int b0 = 0; b0 < n; b0 += C) {
        const int m = std::min(C, n - b0);
        for (int k = 0; k < m; ++k) {
            const rln_image_meta &mt = meta[b0 + k];
            PrepImage &q = hp[k];
            q.dx = ROUND(mt.old_offset_x);
            q.dy = ROUND(mt.old_offset_y);
            q.norm = do_norm ? cf.avg_norm_correction / mt.norm_correction : 1.;
            const double az = mt.defocus_angle * kPi / 180., ca = cos(az), sa = sin(az);
            const double D0 = -mt.defocus_u, D1 = -mt.defocus_v;
            q.axx = ca * D0 * ca + sa * D1 * sa;
            q.axy = ca * D0 * sa - sa * D1 * ca;
            q.ayy = sa * D0 * sa + ca * D1 * ca;
            q.k4 = -mt.bfactor / 4.;
            q.k5 = mt.phase_shift * kPi / 180.;
            q.ctf_scale = mt.kfactor;
            q.seed = cf.random_seed + (unsigned long long)mt.part_id;
            if (old_offset) {
                old_offset[2 * (b0 + k)] = q.dx;
                old_offset[2 * (b0 + k) + 1] = q.dy;
            }
        }
        HIP_TRY(hipMemcpyAsync(p->d_img.ptr, hp.data(), sizeof(PrepImage) * m, hipMemcpyHostToDevice, p->st));
        const float *src = images + (size_t)b0 * npix;
        if (!in_on_device) {
            HIP_TRY(hipMemcpyAsync(p->d_in.ptr, src, sizeof(float) * npix * m, hipMemcpyHostToDevice, p->st));
            src = p->d_in.as<float>();
        }
        const float *noise = nullptr;
        HIP_TRY(hipEventRecord(p->ev[0], p->st));
        if (!cf.zero_mask) {
            hipfftHandle hc;
            RLN_TRY(plan_for(p->c2r, HIPFFT_C2R, N, m, p->st, &hc));
            dim3 g((unsigned)((nf + kThreads - 1) / kThreads), (unsigned)m);
            hipLaunchKernelGGL(k_noise_fourier, g, dim3(kThreads), 0, p->st, p->d_sigma.as<double>(), cf.sigma2_fudge,
                               p->d_img.as<PrepImage>(), N, p->d_noise_c.as<float2>());
            HIP_TRY(hipGetLastError());
            FFT_TRY(hipfftExecC2R(hc, (hipfftComplex *)p->d_noise_c.ptr, (hipfftReal *)p->d_noise_r.ptr));
            noise = p->d_noise_r.as<float>();
        }
        HIP_TRY(hipEventRecord(p->ev[1], p->st));
        hipLaunchKernelGGL(k_prep_real, dim3(m), dim3(kThreads), 0, p->st, src, noise, p->d_img.as<PrepImage>(), c, m,
                           p->d_real.as<float>());
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(p->ev[2], p->st));
        hipfftHandle hr;
        RLN_TRY(plan_for(p->r2c, HIPFFT_R2C, N, 2 * m, p->st, &hr));
        FFT_TRY(hipfftExecR2C(hr, (hipfftReal *)p->d_real.ptr, (hipfftComplex *)p->d_cplx.ptr));
        HIP_TRY(hipEventRecord(p->ev[3], p->st));
        float2 *o_m, *o_u;
        float *o_c;
        if (out_on_device) {
            o_m = reinterpret_cast<float2 *>(fimg) + (size_t)b0 * no;
            o_u = reinterpret_cast<float2 *>(fimg_nomask) + (size_t)b0 * no;
            o_c = fctf + (size_t)b0 * no;
        } else {
            o_m = p->d_out.as<float2>();
            o_u = o_m + no * m;
            o_c = reinterpret_cast<float *>(o_u + no * m);
        }
        hipLaunchKernelGGL(k_prep_fourier, dim3(m), dim3(kThreads), 0, p->st, p->d_cplx.as<float2>(),
                           p->d_img.as<PrepImage>(), c, m, o_m, o_u, o_c, p->d_power.as<double>(),
                           p->d_hx.as<double>());
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(p->ev[4], p->st));
        if (!out_on_device) {
            HIP_TRY(hipMemcpyAsync(fimg + 2 * (size_t)b0 * no, o_m, sizeof(float2) * no * m, hipMemcpyDeviceToHost, p->st));
            HIP_TRY(hipMemcpyAsync(fimg_nomask + 2 * (size_t)b0 * no, o_u, sizeof(float2) * no * m,
                                   hipMemcpyDeviceToHost, p->st));
            HIP_TRY(hipMemcpyAsync(fctf + (size_t)b0 * no, o_c, sizeof(float) * no * m, hipMemcpyDeviceToHost, p->st));
        }
        HIP_TRY(hipMemcpyAsync(hpow.data(), p->d_power.ptr, sizeof(double) * xh * m, hipMemcpyDeviceToHost, p->st));
        HIP_TRY(hipMemcpyAsync(hhx.data(), p->d_hx.ptr, sizeof(double) * m, hipMemcpyDeviceToHost, p->st));
        HIP_TRY(hipStreamSynchronize(p->st));
        float t;
        if (hipEventElapsedTime(&t, p->ev[1], p->ev[2]) == hipSuccess) p->ms[0] += t;
        if (hipEventElapsedTime(&t, p->ev[2], p->ev[3]) == hipSuccess) p->ms[1] += t;
        if (hipEventElapsedTime(&t, p->ev[3], p->ev[4]) == hipSuccess) p->ms[2] += t;
        if (hipEventElapsedTime(&t, p->ev[0], p->ev[1]) == hipSuccess) p->ms[3] += t;
        if (power) std::copy(hpow.begin(), hpow.begin() + (size_t)xh * m, power + (size_t)b0 * xh);
        if (highres_Xi2) std::copy(hhx.begin(), hhx.begin() + m, highres_Xi2 + b0);
    }
    HIP_TRY(hipEventRecord(p->ev[0], p->st));
    HIP_TRY(hipStreamSynchronize(p->st));
    float t;
    if (hipEventElapsedTime(&t, p->ev[5], p->ev[0]) == hipSuccess) p->ms[4] = t;
    (void)hipGetLastError();
    return RLN_OK;
}
