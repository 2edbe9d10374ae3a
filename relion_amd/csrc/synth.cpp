// Seeded synthetic cryo-EM data for the E-step: Gaussian-blob maps, their padded Fourier references, and
// particles = CTF-modulated central slices of the analytic 3-D Fourier transform, shifted, plus white noise.
//
// Everything is analytic and separable: the 3-D FT of a blob a*exp(-|r-c|^2/(2 s^2)) is
// a (2 pi s^2)^{3/2} exp(-2 pi^2 s^2 |f|^2) exp(-2 pi i f.c), which factorises over the axes, and the
// central slice at Euler matrix A samples it at f3 = A^T f2, whose phase f2.(A c) factorises over the
// two image axes. The projector convention is Projector::project (src/projector.cpp:630-797); the
// reference's own simulator is src/apps/project.cpp:66-106.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

#include "rln_common.h"
#include "sampling.h"

namespace rln {

namespace {

using cd = std::complex<double>;

// Counter-based generator: splitmix64 finaliser over (seed, stream, counter).
inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

struct Stream {
    uint64_t key;
    uint64_t ctr = 0;
    Stream(uint64_t seed, uint64_t stream) : key(mix64(seed ^ mix64(stream * 0x632be59bd9b4e019ULL))) {}
    double uniform() { return (double)(mix64(key + (ctr++) * 0xd1b54a32d192ed03ULL) >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
    double normal() {
        double u1 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        const double u2 = uniform();
        return sqrt(-2. * log(u1)) * cos(2. * kPi * u2);
    }
};

struct Blob {
    double x, y, z, sigma, amp;
};

enum : uint64_t { kStreamShared = 0x5348u, kStreamClass = 0x434c0000u, kStreamParticle = 0x50410000000ULL };

std::vector<Blob> blobs_for_class(const rln_synth_spec &s, int iclass, double share_fraction) {
    const int n_shared = (int)std::lround(share_fraction * s.nr_blobs);
    std::vector<Blob> out;
    auto draw = [&](Stream &st, int count) {
        for (int i = 0; i < count; ++i) {
            Blob b;
            // uniform in a ball: rejection from the cube
            do {
                b.x = st.uniform(-1., 1.);
                b.y = st.uniform(-1., 1.);
                b.z = st.uniform(-1., 1.);
            } while (b.x * b.x + b.y * b.y + b.z * b.z > 1.);
            b.x *= s.blob_radius;
            b.y *= s.blob_radius;
            b.z *= s.blob_radius;
            b.sigma = st.uniform(s.sigma_min, s.sigma_max);
            b.amp = st.uniform(s.amp_min, s.amp_max);
            out.push_back(b);
        }
    };
    Stream shared(s.seed, kStreamShared);
    draw(shared, n_shared);
    Stream own(s.seed, kStreamClass + (uint64_t)iclass);
    draw(own, s.nr_blobs - n_shared);
    return out;
}

// 1-D factor of a blob's transform along one axis at integer index k of an n-point grid (f = k/n).
inline cd axis_factor(double sigma, double c, double k, double n) {
    const double f = k / n;
    const double env = exp(-2. * kPi * kPi * sigma * sigma * f * f);
    const double ph = -2. * kPi * f * c;
    return cd(env * cos(ph), env * sin(ph));
}

}  // namespace

// CTF::initialise + CTF::getCTF (src/ctf.cpp:207-261, src/ctf.h:184-240); x, y in 1/Angstrom.
double ctf_value(double defU, double defV, double defAngle, double voltage, double cs, double q0, double bfactor,
                 double scale, double phase_shift, double x, double y) {
    const double kv = voltage * 1e3;
    const double lambda = 12.2643247 / sqrt(kv * (1. + kv * 0.978466e-6));
    const double K1 = kPi * lambda;
    const double K2 = kPi / 2. * cs * 1e7 * lambda * lambda * lambda;
    const double K3 = atan(q0 / sqrt(1. - q0 * q0));
    const double K4 = -bfactor / 4.;
    const double K5 = phase_shift * kPi / 180.;
    const double az = defAngle * kPi / 180.;
    const double c = cos(az), s = sin(az);
    // defocus as the quadratic form Q^T diag(-defU, -defV) Q with Q the rotation by the azimuth
    const double axx = -defU * c * c - defV * s * s;
    const double axy = -defU * c * s + defV * s * c;
    const double ayy = -defU * s * s - defV * c * c;
    const double u2 = x * x + y * y;
    const double gamma = K1 * (axx * x * x + 2. * axy * x * y + ayy * y * y) + K2 * u2 * u2 - K5 - K3;
    return scale * (-sin(gamma)) * exp(K4 * u2);
}

}  // namespace rln

extern "C" double rln_ctf_value(double defU, double defV, double defAngle, double voltage, double cs, double q0,
                                double bfactor, double scale, double phase_shift, double x, double y) {
    return rln::ctf_value(defU, defV, defAngle, voltage, cs, q0, bfactor, scale, phase_shift, x, y);
}

extern "C" int rln_synth_real_map(const rln_synth_spec *spec, int iclass, double share_fraction, float *out) {
    using namespace rln;
    if (!spec || !out || spec->ori_size <= 0) RLN_FAIL(RLN_ERR_ARG, "rln_synth_real_map: invalid argument");
    const int n = spec->ori_size, h = n / 2;
    const std::vector<Blob> blobs = blobs_for_class(*spec, iclass, share_fraction);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < n; ++z) {
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) {
                double v = 0.;
                for (const Blob &B : blobs) {
                    const double dx = x - h - B.x, dy = y - h - B.y, dz = z - h - B.z;
                    v += B.amp * exp(-(dx * dx + dy * dy + dz * dz) / (2. * B.sigma * B.sigma));
                }
                out[((size_t)z * n + y) * n + x] = (float)v;
            }
    }
    return RLN_OK;
}

extern "C" int rln_synth_reference(const rln_synth_spec *spec, int iclass, double share_fraction, int current_size,
                                   double padding_factor, double *out) {
    using namespace rln;
    if (!spec || !out || current_size <= 0 || current_size > spec->ori_size || !(padding_factor >= 1.))
        RLN_FAIL(RLN_ERR_ARG, "rln_synth_reference: invalid argument");
    const int r_max = current_size / 2;
    const int pad_size = 2 * (ROUND(padding_factor * r_max) + 1) + 1;
    const int xdim = pad_size / 2 + 1, ydim = pad_size, zdim = pad_size;
    const int h = pad_size / 2;  // y,z logical range [-h, h]
    const long long max_r2 = (long long)ROUND(r_max * padding_factor) * ROUND(r_max * padding_factor);
    const double n = padding_factor * spec->ori_size;  // padded grid length (f = k / n cycles per pixel)
    const std::vector<Blob> blobs = blobs_for_class(*spec, iclass, share_fraction);
    const int nb = (int)blobs.size();
    // Image normalisation 1/N^2 (2-D DFT of an N x N image divided by its pixel count).
    const double norm = 1. / ((double)spec->ori_size * spec->ori_size);
    std::vector<cd> fx((size_t)nb * xdim), fy((size_t)nb * ydim), fz((size_t)nb * zdim);
    for (int b = 0; b < nb; ++b) {
        const Blob &B = blobs[b];
        const double amp = B.amp * std::pow(2. * kPi * B.sigma * B.sigma, 1.5) * norm;
        for (int x = 0; x < xdim; ++x) fx[(size_t)b * xdim + x] = amp * axis_factor(B.sigma, B.x, x, n);
        for (int i = 0; i < ydim; ++i) fy[(size_t)b * ydim + i] = axis_factor(B.sigma, B.y, i - h, n);
        for (int i = 0; i < zdim; ++i) fz[(size_t)b * zdim + i] = axis_factor(B.sigma, B.z, i - h, n);
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (int iz = 0; iz < zdim; ++iz) {
        std::vector<cd> row(xdim);
        std::vector<cd> yz(nb);
        const long long kz = iz - h;
        for (int iy = 0; iy < ydim; ++iy) {
            const long long ky = iy - h;
            double *dst = out + 2 * ((size_t)iz * ydim + iy) * xdim;
            const long long rem = max_r2 - kz * kz - ky * ky;
            if (rem < 0) {
                std::fill(dst, dst + 2 * xdim, 0.);
                continue;
            }
            const int xmax = std::min<long long>(xdim - 1, (long long)floor(sqrt((double)rem)));
            std::fill(row.begin(), row.end(), cd(0., 0.));
            for (int b = 0; b < nb; ++b) {
                const cd t = fy[(size_t)b * ydim + iy] * fz[(size_t)b * zdim + iz];
                const cd *ax = &fx[(size_t)b * xdim];
                for (int x = 0; x <= xmax; ++x) row[x] += t * ax[x];
            }
            for (int x = 0; x < xdim; ++x) {
                const bool inside = x <= xmax && (long long)x * x <= rem;
                dst[2 * x] = inside ? row[x].real() : 0.;
                dst[2 * x + 1] = inside ? row[x].imag() : 0.;
            }
        }
    }
    return RLN_OK;
}

extern "C" int rln_synth_particles(const rln_synth_spec *spec, int nr_classes, double share_fraction,
                                   int current_size, long long first, int n, float *fimg, float *fctf,
                                   rln_particle_meta *meta, double *true_pose, double *noise_var) {
    using namespace rln;
    if (!spec || n < 0 || current_size <= 0 || (current_size & 1) || nr_classes <= 0)
        RLN_FAIL(RLN_ERR_ARG, "rln_synth_particles: invalid argument");
    const int s = current_size, xs = s / 2 + 1;
    const size_t P = (size_t)s * xs;
    const double N = spec->ori_size;
    std::vector<std::vector<Blob>> cls(nr_classes);
    for (int k = 0; k < nr_classes; ++k) cls[k] = blobs_for_class(*spec, k, share_fraction);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n; ++i) {
        const long long id = first + i;
        Stream st(spec->seed, kStreamParticle + (uint64_t)id);
        const int k = nr_classes > 1 ? std::min(nr_classes - 1, (int)(st.uniform() * nr_classes)) : 0;
        // uniform rotation from a uniform unit quaternion (Shoemake 1992)
        const double u1 = st.uniform(), u2 = st.uniform(), u3 = st.uniform();
        const double qw = sqrt(1. - u1) * sin(2. * kPi * u2), qx = sqrt(1. - u1) * cos(2. * kPi * u2);
        const double qy = sqrt(u1) * sin(2. * kPi * u3), qz = sqrt(u1) * cos(2. * kPi * u3);
        const double R[9] = {1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw),
                             2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw),
                             2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)};
        double rot, tilt, psi;
        matrix_to_angles(R, rot, tilt, psi);
        double A[9];
        euler_angles2matrix(rot, tilt, psi, A);
        const double dx = st.uniform(-spec->offset_max, spec->offset_max);
        const double dy = st.uniform(-spec->offset_max, spec->offset_max);
        const double defU = st.uniform(spec->defocus_min, spec->defocus_max);
        const double defV = defU - st.uniform(0., spec->max_astigmatism);
        const double defAng = st.uniform(0., 180.);
        double snr = spec->snr_min;
        if (spec->snr_max > spec->snr_min)
            snr = exp(st.uniform(log(spec->snr_min), log(spec->snr_max)));
        // prior angles: true pose perturbed by Gaussian angle noise
        const double ps = spec->prior_sigma_deg;
        double prot = rot + ps * st.normal(), ptilt = tilt + ps * st.normal(), ppsi = psi + ps * st.normal();
        if (ptilt < 0.) {
            ptilt = -ptilt;
            prot += 180.;
        }
        if (ptilt > 180.) {
            ptilt = 360. - ptilt;
            prot += 180.;
        }
        prot = prot - 360. * floor((prot + 180.) / 360.);
        ppsi = ppsi - 360. * floor(ppsi / 360.);
        // slice of the analytic transform: phase k.(A c), per-axis factors
        const std::vector<Blob> &bl = cls[k];
        const int nb = (int)bl.size();
        std::vector<cd> ex((size_t)nb * xs), ey((size_t)nb * s);
        const double norm = 1. / (N * N);
        for (int b = 0; b < nb; ++b) {
            const Blob &B = bl[b];
            const double u = A[0] * B.x + A[1] * B.y + A[2] * B.z;
            const double v = A[3] * B.x + A[4] * B.y + A[5] * B.z;
            const double amp = B.amp * std::pow(2. * kPi * B.sigma * B.sigma, 1.5) * norm;
            for (int x = 0; x < xs; ++x) ex[(size_t)b * xs + x] = amp * axis_factor(B.sigma, u, x, N);
            for (int r = 0; r < s; ++r) {
                const int y = (r < xs) ? r : r - s;
                ey[(size_t)b * s + r] = axis_factor(B.sigma, v, y, N);
            }
        }
        std::vector<cd> img(P);
        std::vector<double> ctf(P);
        double sig_pow = 0.;
        long cnt = 0;
        for (int r = 0; r < s; ++r) {
            const int y = (r < xs) ? r : r - s;
            for (int x = 0; x < xs; ++x) {
                cd acc(0., 0.);
                for (int b = 0; b < nb; ++b) acc += ex[(size_t)b * xs + x] * ey[(size_t)b * s + r];
                // shift by -d so that translating by +d recovers the slice (shiftImageInFourierTransform sign)
                const double ph = 2. * kPi * (x * dx + y * dy) / N;
                acc *= cd(cos(ph), sin(ph));
                const double c = ctf_value(defU, defV, defAng, spec->voltage, spec->cs, spec->q0, spec->bfactor, 1.,
                                           0., x / (N * spec->pixel_size), y / (N * spec->pixel_size));
                acc *= c;
                ctf[(size_t)r * xs + x] = c;
                img[(size_t)r * xs + x] = acc;
                if (x * x + y * y <= (s / 2) * (s / 2)) {
                    sig_pow += std::norm(acc);
                    ++cnt;
                }
            }
        }
        const double nv = (sig_pow / (double)std::max(1L, cnt)) / (2. * snr);  // per-component variance
        for (size_t p = 0; p < P; ++p) {
            const double nr = sqrt(nv) * st.normal(), ni = sqrt(nv) * st.normal();
            if (fimg) {
                fimg[2 * (P * i + p)] = (float)(img[p].real() + nr);
                fimg[2 * (P * i + p) + 1] = (float)(img[p].imag() + ni);
            }
            if (fctf) fctf[P * i + p] = (float)ctf[p];
        }
        if (meta) {
            rln_particle_meta &m = meta[i];
            m.old_offset_x = m.old_offset_y = 0.;
            m.prior_offset_x = m.prior_offset_y = 0.;
            m.prior_rot = prot;
            m.prior_tilt = ptilt;
            m.prior_psi = ppsi;
            m.scale = 1.;
            m.highres_Xi2 = 0.;
        }
        if (true_pose) {
            double *tp = true_pose + 6 * (size_t)i;
            tp[0] = rot;
            tp[1] = tilt;
            tp[2] = psi;
            tp[3] = dx;
            tp[4] = dy;
            tp[5] = k + 1;
        }
        if (noise_var) noise_var[i] = nv;
    }
    return RLN_OK;
}
