/*
 * TEST INFRASTRUCTURE ONLY -- the parity oracle. Never linked into, loaded by or called from the
 * product library (relion_amd/); only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg use it, to check the product's results.
 *
 * An FP64 restatement, written from the reference's behaviour, of RELION's CPU expectation step:
 *   MlOptimiser::getAllSquaredDifferences            src/ml_optimiser.cpp:7098-7701
 *   MlOptimiser::convertAllSquaredDifferencesToWeights src/ml_optimiser.cpp:7704-8239
 *   MlOptimiser::storeWeightedSums                   src/ml_optimiser.cpp:8241-9156
 *   MlOptimiser::precalculateShiftedImagesCtfsAndInvSigma2s src/ml_optimiser.cpp:6755-7073
 *   Projector::project                               src/projector.cpp:630-797
 *   BackProjector::backproject2Dto3D                 src/backprojector.cpp:55-357
 *   HealpixSampling (grid, oversampling, priors)     src/healpix_sampling.cpp
 *   Euler_angles2matrix                              src/euler.cpp:52-91
 *   CTF::getCTF / CTF::initialise                    src/ctf.h:184-240, src/ctf.cpp:207-261
 *
 * Scope: 2-D particles of one image, 3-D references, C1 symmetry, one body, no helical/tomo,
 * squared-difference (non-CC) target, trilinear interpolation, no random perturbation of the
 * sampling. Exact-arithmetic choices where the reference's result depends on the last bit of a
 * rotation (pixels exactly on the r_max circle) are stated at the functions concerned.
 */
#include "estep_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double PI = 3.14159265358979323846;
const double LARGE_NUMBER = 99e99;  // src/macros.h:78 (double build)

// src/macros.h:197,212,226
inline int ROUND(double x) { return (x > 0) ? (int)(x + 0.5) : (int)(x - 0.5); }
inline int CEIL(double x) { return (x == (int)x) ? (int)x : ((x > 0) ? (int)(x + 1) : (int)x); }
inline int FLOOR(double x) { return (x == (int)x) ? (int)x : ((x > 0) ? (int)x : (int)(x - 1)); }
inline double DEG2RAD(double d) { return d * PI / 180.; }
inline double RAD2DEG(double r) { return r * 180. / PI; }

// ------------------------------------------------------------------------------------------------
// HEALPix NEST scheme (Gorski et al. 2005, ApJ 622:759), as Healpix_Base implements it in
// src/Healpix_2.15a/healpix_base.cc (nest2xyf, xyf2nest, pix2ang_z_phi).
// ------------------------------------------------------------------------------------------------
const int JRLL[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
const int JPLL[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};

long long spread_bits(int v) {
    long long r = 0;
    for (int b = 0; b < 31; ++b)
        if (v & (1 << b)) r |= (1LL << (2 * b));
    return r;
}
int compress_bits(long long v) {
    int r = 0;
    for (int b = 0; b < 31; ++b)
        if (v & (1LL << (2 * b))) r |= (1 << b);
    return r;
}

}  // namespace

extern "C" void orc_hp_nest2xyf(int order, long long pix, int *ix, int *iy, int *face) {
    const long long npface = 1LL << (2 * order);
    *face = (int)(pix >> (2 * order));
    const long long ipf = pix & (npface - 1);
    *ix = compress_bits(ipf);
    *iy = compress_bits(ipf >> 1);
}

extern "C" long long orc_hp_xyf2nest(int order, int ix, int iy, int face) {
    return ((long long)face << (2 * order)) + spread_bits(ix) + (spread_bits(iy) << 1);
}

extern "C" void orc_hp_pix2ang_z_phi(int order, long long pix, double *z, double *phi) {
    const long long nside = 1LL << order;
    const long long npix = 12 * nside * nside;
    const long long nl4 = 4 * nside;
    const double fact2 = 4. / (double)npix;
    const double fact1 = (double)(nside << 1) * fact2;
    int ix, iy, face;
    orc_hp_nest2xyf(order, pix, &ix, &iy, &face);
    const long long jr = ((long long)JRLL[face] << order) - ix - iy - 1;
    long long nr, kshift;
    if (jr < nside) {
        nr = jr;
        *z = 1. - (double)(nr * nr) * fact2;
        kshift = 0;
    } else if (jr > 3 * nside) {
        nr = nl4 - jr;
        *z = (double)(nr * nr) * fact2 - 1.;
        kshift = 0;
    } else {
        nr = nside;
        *z = (double)(2 * nside - jr) * fact1;
        kshift = (jr - nside) & 1;
    }
    long long jp = (JPLL[face] * nr + ix - iy + 1 + kshift) / 2;
    if (jp > nl4) jp -= nl4;
    if (jp < 1) jp += nl4;
    *phi = ((double)jp - (double)(kshift + 1) * 0.5) * ((PI / 2.) / (double)nr);
}

// src/euler.cpp:52-91
extern "C" void orc_euler_angles2matrix(double alpha, double beta, double gamma, double *A) {
    alpha = DEG2RAD(alpha);
    beta = DEG2RAD(beta);
    gamma = DEG2RAD(gamma);
    const double ca = cos(alpha), cb = cos(beta), cg = cos(gamma);
    const double sa = sin(alpha), sb = sin(beta), sg = sin(gamma);
    const double cc = cb * ca, cs = cb * sa, sc = sb * ca, ss = sb * sa;
    A[0] = cg * cc - sg * sa;
    A[1] = cg * cs + sg * ca;
    A[2] = -cg * sb;
    A[3] = -sg * cc - cg * sa;
    A[4] = -sg * cs + cg * ca;
    A[5] = sg * sb;
    A[6] = sc;
    A[7] = ss;
    A[8] = cb;
}

// CTF::initialise (src/ctf.cpp:207-261) followed by CTF::getCTF (src/ctf.h:184-240) with
// do_damping, no magnification matrix, no dose model.
extern "C" double orc_ctf_value(double defU, double defV, double defAngle, double voltage, double cs,
                                double q0, double bfactor, double scale, double phase_shift, double X,
                                double Y) {
    const double local_Cs = cs * 1e7;
    const double local_kV = voltage * 1e3;
    const double rad_azimuth = DEG2RAD(defAngle);
    const double lambda = 12.2643247 / sqrt(local_kV * (1. + local_kV * 0.978466e-6));
    const double K1 = PI / 2 * 2 * lambda;
    const double K2 = PI / 2 * local_Cs * lambda * lambda * lambda;
    const double K3 = atan(q0 / sqrt(1 - q0 * q0));
    const double K4 = -bfactor / 4.;
    const double K5 = DEG2RAD(phase_shift);
    // astigmatism as the bilinear form A = Q^T D Q
    const double sin_az = sin(rad_azimuth), cos_az = cos(rad_azimuth);
    // Q = [[c, s], [-s, c]], Qt = [[c, -s], [s, c]], D = diag(-defU, -defV)
    const double D0 = -defU, D1 = -defV;
    const double Axx = cos_az * D0 * cos_az + (-sin_az) * D1 * (-sin_az);
    const double Axy = cos_az * D0 * sin_az + (-sin_az) * D1 * cos_az;
    const double Ayy = sin_az * D0 * sin_az + cos_az * D1 * cos_az;
    const double u2 = X * X + Y * Y;
    const double u4 = u2 * u2;
    const double gamma = K1 * (Axx * X * X + 2.0 * Axy * X * Y + Ayy * Y * Y) + K2 * u4 - K5 - K3;
    double retval = -sin(gamma);
    retval *= exp(K4 * u2);
    return scale * retval;
}

namespace {

struct Grid {
    int order = 0;
    int n_dir = 0, n_psi = 0, n_trans = 0;
    double psi_step = 0;
    std::vector<double> rot, tilt, psi, tx, ty;  // translations in Angstrom
};

// HealpixSampling::checkDirection (src/healpix_sampling.cpp:1541-1564)
void check_direction(double &rot, double &tilt) {
    if (tilt < 0.) {
        tilt = -tilt;
        rot += 180.;
    }
    while (true) {
        if (rot > 180.)
            rot -= 360.;
        else if (rot < -180.)
            rot += 360.;
        else
            break;
    }
}

// getDirectionFromHealPix (src/healpix_sampling.cpp:1566-1577)
void direction_from_healpix(int order, long long ipix, double &rot, double &tilt) {
    double z, phi;
    orc_hp_pix2ang_z_phi(order, ipix, &z, &phi);
    rot = RAD2DEG(phi);
    tilt = acos(z) * 180. / PI;  // ACOSD
    check_direction(rot, tilt);
}

// initialise / setTranslations / setOrientations (src/healpix_sampling.cpp:51-138, 291-450, 477-556)
// for C1 (no symmetry-equivalent removal), 2-D translations, no tilt limits.
bool build_grid(const rln_estep_config *cfg, Grid &g) {
    if (cfg->healpix_order < 0 || cfg->offset_step <= 0. || cfg->offset_range < 0.) return false;
    g.order = cfg->healpix_order;
    double psi_step = cfg->psi_step;
    if (psi_step <= 0.) psi_step = 360. / (6 * ROUND(std::pow(2., g.order)));
    const long long npix = 12LL << (2 * g.order);
    g.rot.resize(npix);
    g.tilt.resize(npix);
    for (long long ipix = 0; ipix < npix; ++ipix) direction_from_healpix(g.order, ipix, g.rot[ipix], g.tilt[ipix]);
    g.n_dir = (int)npix;
    const int nr_psi = CEIL(360. / psi_step);
    g.psi_step = 360. / (double)nr_psi;
    g.psi.resize(nr_psi);
    for (int i = 0; i < nr_psi; ++i) g.psi[i] = i * g.psi_step;
    g.n_psi = nr_psi;
    const int maxp = CEIL(cfg->offset_range / cfg->offset_step);
    g.tx.clear();
    g.ty.clear();
    for (long ix = -maxp; ix <= maxp; ++ix) {
        const double xoff = ix * cfg->offset_step;
        for (long iy = -maxp; iy <= maxp; ++iy) {
            const double yoff = iy * cfg->offset_step;
            const double max2 = xoff * xoff + yoff * yoff;
            if (max2 < cfg->offset_range * cfg->offset_range + 0.001) {
                g.tx.push_back(xoff);
                g.ty.push_back(yoff);
            }
        }
    }
    g.n_trans = (int)g.tx.size();
    return true;
}

// getOrientations at oversampling order 1 (src/healpix_sampling.cpp:1832-1939) with
// pushbackOversampledPsiAngles (:1941-1966): directions of the 4 order+1 sub-pixels (j outer, i inner),
// 2 psi angles each (innermost).
void oversampled_orientations(const Grid &g, int idir, int ipsi, double *rot, double *tilt, double *psi) {
    const int over_order = g.order + 1;
    const int fact = 2;
    int x, y, face;
    orc_hp_nest2xyf(g.order, idir, &x, &y, &face);
    int k = 0;
    for (int j = fact * y; j < fact * (y + 1); ++j)
        for (int i = fact * x; i < fact * (x + 1); ++i) {
            const long long overpix = orc_hp_xyf2nest(over_order, i, j, face);
            double r, t;
            direction_from_healpix(over_order, overpix, r, t);
            for (int ipo = 0; ipo < 2; ++ipo) {
                rot[k] = r;
                tilt[k] = t;
                psi[k] = g.psi[ipsi] - 0.5 * g.psi_step + (0.5 + ipo) * g.psi_step / 2.;
                ++k;
            }
        }
}

// getTranslationsInPixel at oversampling order 1 (src/healpix_sampling.cpp:1724-1830)
void oversampled_translations(const rln_estep_config *cfg, const Grid &g, int itrans, double *x, double *y) {
    const double step = cfg->offset_step;
    int k = 0;
    for (int ox = 0; ox < 2; ++ox) {
        const double over_x = g.tx[itrans] - 0.5 * step + (0.5 + ox) * step / 2.;
        for (int oy = 0; oy < 2; ++oy) {
            const double over_y = g.ty[itrans] - 0.5 * step + (0.5 + oy) * step / 2.;
            x[k] = over_x / cfg->pixel_size;
            y[k] = over_y / cfg->pixel_size;
            ++k;
        }
    }
}

// gaussian1D (src/funcs.cpp:356-360)
double gaussian1D(double x, double sigma, double mu) {
    x -= mu;
    return 1 / sqrt(2 * PI * sigma * sigma) * exp(-0.5 * ((x / sigma) * (x / sigma)));
}

// selectOrientationsWithNonZeroPriorProbability (src/healpix_sampling.cpp:695-1076), 3-D, C1,
// sigma_rot and sigma_tilt both > 0, no bimodal psi, no sigma_tilt_from_ninety / sigma_psi_from_zero.
void local_priors(const rln_estep_config *cfg, const Grid &g, double prior_rot, double prior_tilt,
                  double prior_psi, std::vector<int> &pdir, std::vector<double> &dprior, std::vector<int> &ppsi,
                  std::vector<double> &pprior) {
    pdir.clear();
    dprior.clear();
    ppsi.clear();
    pprior.clear();
    const double cutoff = cfg->sigma_cutoff > 0 ? cfg->sigma_cutoff : 3.;
    // Euler_angles2direction (src/euler.cpp:94-113)
    auto direction = [](double a, double b, double *v) {
        a = DEG2RAD(a);
        b = DEG2RAD(b);
        v[0] = sin(b) * cos(a);
        v[1] = sin(b) * sin(a);
        v[2] = cos(b);
    };
    double pd[3];
    direction(prior_rot, prior_tilt, pd);
    double sumprior = 0., best_ang = 9999.;
    long best_idir = -999;
    const double biggest_sigma = std::max(cfg->sigma_rot, cfg->sigma_tilt);
    for (int idir = 0; idir < g.n_dir; ++idir) {
        if (cfg->sigma_rot > 0. && cfg->sigma_tilt > 0.) {
            double md[3];
            direction(g.rot[idir], g.tilt[idir], md);
            double diffang = acos(md[0] * pd[0] + md[1] * pd[1] + md[2] * pd[2]) * 180. / PI;
            if (diffang > 180.) diffang = std::fabs(diffang - 360.);
            if (diffang < cutoff * biggest_sigma) {
                pdir.push_back(idir);
                const double prior = gaussian1D(diffang, biggest_sigma, 0.);
                sumprior += prior;
                dprior.push_back(prior);
            }
            if (diffang < best_ang) {
                best_idir = idir;
                best_ang = diffang;
            }
        } else {
            pdir.push_back(idir);
            dprior.push_back(1.);
            sumprior += 1.;
        }
    }
    for (size_t i = 0; i < dprior.size(); ++i) dprior[i] /= sumprior;
    if (dprior.empty()) {
        pdir.push_back((int)best_idir);
        dprior.push_back(1.);
    }
    sumprior = 0.;
    double best_diff = 9999.;
    long best_ipsi = -999;
    for (int ipsi = 0; ipsi < g.n_psi; ++ipsi) {
        if (cfg->sigma_psi > 0.) {
            double diffpsi = std::fabs(g.psi[ipsi] - prior_psi);
            if (diffpsi > 180.) diffpsi = std::fabs(diffpsi - 360.);
            if (diffpsi < cutoff * cfg->sigma_psi) {
                const double prior = gaussian1D(diffpsi, cfg->sigma_psi, 0.);
                ppsi.push_back(ipsi);
                pprior.push_back(prior);
                sumprior += prior;
            }
            if (diffpsi < best_diff) {
                best_ipsi = ipsi;
                best_diff = diffpsi;
            }
        } else {
            ppsi.push_back(ipsi);
            pprior.push_back(1.);
            sumprior += 1.;
        }
    }
    for (size_t i = 0; i < pprior.size(); ++i) pprior[i] /= sumprior;
    if (pprior.empty()) {
        ppsi.push_back((int)best_ipsi);
        pprior.push_back(1.);
    }
}

// Mresol (src/ml_optimiser.cpp:5782-5812): resolution shell per pixel of an s x (s/2+1) image,
// -1 beyond s/2 and on the duplicated half column (x == 0, y < 0).
std::vector<int> mresol(int s) {
    const int xs = s / 2 + 1;
    std::vector<int> m((size_t)s * xs, -1);
    for (int i = 0; i < s; ++i) {
        const int ip = (i < xs) ? i : i - s;
        for (int j = 0; j < xs; ++j) {
            const int ires = ROUND(sqrt((double)(ip * ip + j * j)));
            if (ires < s / 2 + 1 && !(j == 0 && ip < 0)) m[(size_t)i * xs + j] = ires;
        }
    }
    return m;
}

// Minvsigma2 (src/ml_optimiser.cpp:6868-6883), model and image at the same pixel size and box.
std::vector<double> minvsigma2(int s, const rln_estep_config *cfg, const double *sigma2_noise) {
    const std::vector<int> m = mresol(s);
    const int nsig = cfg->ori_size / 2 + 1;
    std::vector<double> out(m.size(), 0.);
    for (size_t n = 0; n < m.size(); ++n) {
        const int ires = m[n];
        if (ires > 0 && ires < nsig) out[n] = 1. / (cfg->sigma2_fudge * sigma2_noise[ires]);
    }
    return out;
}

// windowFourierTransform (src/fftw.h:809-857) of an s_in x (s_in/2+1) array to s_out <= s_in.
template <int NC>
std::vector<double> window(const double *in, int s_in, int s_out) {
    const int xin = s_in / 2 + 1, xout = s_out / 2 + 1;
    std::vector<double> out((size_t)s_out * xout * NC, 0.);
    for (int i = 0; i < s_out; ++i) {
        const int ip = (i < xout) ? i : i - s_out;
        const int iin = (ip < 0) ? ip + s_in : ip;
        for (int j = 0; j < xout; ++j)
            for (int c = 0; c < NC; ++c)
                out[((size_t)i * xout + j) * NC + c] = in[((size_t)iin * xin + j) * NC + c];
    }
    return out;
}

}  // namespace

// shiftImageInFourierTransform, 2-D case (src/fftw.cpp:842-960): out = in * exp(-2 pi i (x dx + y dy)/N).
extern "C" void orc_shift_image(const double *in, int s, double oridim, double xshift, double yshift,
                                double *out) {
    const int xs = s / 2 + 1;
    xshift /= -oridim;
    yshift /= -oridim;
    if (std::fabs(xshift) < 1e-6 && std::fabs(yshift) < 1e-6) {  // XMIPP_EQUAL_ACCURACY
        std::memcpy(out, in, sizeof(double) * 2 * (size_t)s * xs);
        return;
    }
    for (int i = 0; i < s; ++i) {
        const double y = (i < xs) ? i : i - s;
        for (int j = 0; j < xs; ++j) {
            const double x = j;
            const double dotp = 2 * PI * (x * xshift + y * yshift);
            const double a = cos(dotp), b = sin(dotp);
            const double c = in[2 * ((size_t)i * xs + j)], d = in[2 * ((size_t)i * xs + j) + 1];
            const double ac = a * c, bd = b * d, ab_cd = (a + b) * (c + d);
            out[2 * ((size_t)i * xs + j)] = ac - bd;
            out[2 * ((size_t)i * xs + j) + 1] = ab_cd - ac - bd;
        }
    }
}

// Projector::project (src/projector.cpp:630-797), TRILINEAR. Uses Ainv = A^T (A is a rotation) scaled
// by padding_factor. The reference additionally skips pixels with |Ainv (x,y,0)|^2 > (r_max*pad)^2;
// for pixels inside the output circle (x^2+y^2 <= (s/2)^2 <= r_max^2) that test can only fire through
// rounding of the rotation, so it is evaluated in exact integer arithmetic: pad^2 (x^2+y^2) > r_max_ref^2.
extern "C" void orc_project(const double *vol, int xdim, int ydim, int zdim, int r_max, double padding_factor,
                            const double *A, int s, double *img) {
    const int xs = s / 2 + 1;
    std::memset(img, 0, sizeof(double) * 2 * (size_t)s * xs);
    double Ainv[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Ainv[r * 3 + c] = A[c * 3 + r] * padding_factor;
    const int r_max_out = xs - 1;
    const int r_max_out_2 = r_max_out * r_max_out;
    const int r_max_ref = (int)(r_max * padding_factor);
    const long long r_max_ref_2 = (long long)r_max_ref * r_max_ref;
    const int starty = -(int)((float)ydim / 2.0), startz = -(int)((float)zdim / 2.0);
    const double pad2 = padding_factor * padding_factor;
    for (int i = 0; i < s; ++i) {
        const int y = (i <= r_max_out) ? i : i - s;
        const int y2 = y * y;
        const int x_max = FLOOR(sqrt((double)(r_max_out_2 - y2)));
        for (int x = 0; x <= x_max; ++x) {
            double xp = Ainv[0] * x + Ainv[1] * y;
            double yp = Ainv[3] * x + Ainv[4] * y;
            double zp = Ainv[6] * x + Ainv[7] * y;
            if (pad2 * (double)(x * x + y2) > (double)r_max_ref_2) continue;
            const bool is_neg_x = xp < 0;
            if (is_neg_x) {
                xp = -xp;
                yp = -yp;
                zp = -zp;
            }
            const int x0 = FLOOR(xp);
            const double fx = xp - x0;
            int y0 = FLOOR(yp);
            const double fy = yp - y0;
            y0 -= starty;
            int z0 = FLOOR(zp);
            const double fz = zp - z0;
            z0 -= startz;
            if (x0 < 0 || x0 + 1 >= xdim || y0 < 0 || y0 + 1 >= ydim || z0 < 0 || z0 + 1 >= zdim) continue;
            double res[2];
            for (int c = 0; c < 2; ++c) {
                auto V = [&](int zz, int yy, int xx) { return vol[2 * (((size_t)zz * ydim + yy) * xdim + xx) + c]; };
                const double d000 = V(z0, y0, x0), d001 = V(z0, y0, x0 + 1);
                const double d010 = V(z0, y0 + 1, x0), d011 = V(z0, y0 + 1, x0 + 1);
                const double d100 = V(z0 + 1, y0, x0), d101 = V(z0 + 1, y0, x0 + 1);
                const double d110 = V(z0 + 1, y0 + 1, x0), d111 = V(z0 + 1, y0 + 1, x0 + 1);
                // LIN_INTERP(a, l, h) = l + (h - l) * a
                const double dx00 = d000 + (d001 - d000) * fx;
                const double dx01 = d100 + (d101 - d100) * fx;
                const double dx10 = d010 + (d011 - d010) * fx;
                const double dx11 = d110 + (d111 - d110) * fx;
                const double dxy0 = dx00 + (dx10 - dx00) * fy;
                const double dxy1 = dx01 + (dx11 - dx01) * fy;
                res[c] = dxy0 + (dxy1 - dxy0) * fz;
            }
            img[2 * ((size_t)i * xs + x)] = res[0];
            img[2 * ((size_t)i * xs + x) + 1] = is_neg_x ? -res[1] : res[1];
        }
    }
}

// BackProjector::backproject2Dto3D (src/backprojector.cpp:55-357), TRILINEAR, no Ewald sphere, no
// magnification matrix. The reference selects, per row, the x range inside the ellipse |Ainv (x,y)|^2 <=
// max_r2 and then re-tests r2_3D > max_r2; for a rotation both reduce to x^2+y^2 <= max_r2/pad^2, which
// is evaluated exactly here (the reference's FP rounding decides the pixels exactly on that circle).
extern "C" void orc_backproject(double *vol, double *weight, int xdim, int ydim, int zdim, int r_max,
                                double padding_factor, const double *A, int s, const double *img,
                                const double *mweight) {
    const int sh = s / 2 + 1;
    double Ainv[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Ainv[r * 3 + c] = A[c * 3 + r] * padding_factor;
    const long long rr = ROUND(r_max * padding_factor);
    const long long max_r2 = rr * rr;
    const int starty = -(int)((float)ydim / 2.0), startz = -(int)((float)zdim / 2.0);
    const double pad2 = padding_factor * padding_factor;
    for (int i = 0; i < s; ++i) {
        int y, first_allowed_x;
        if (i < sh) {
            y = i;
            first_allowed_x = 0;
        } else {
            y = i - s;
            first_allowed_x = 1;  // the x == 0 column is stored twice in the FFTW layout
        }
        for (int x = first_allowed_x; x <= sh - 1; ++x) {
            if (pad2 * (double)(x * x + y * y) > (double)max_r2) continue;
            const size_t n = (size_t)i * sh + x;
            const double my_weight = mweight ? mweight[n] : 1.;
            if (my_weight <= 0.) continue;
            double vre = img[2 * n], vim = img[2 * n + 1];
            double xp = Ainv[0] * x + Ainv[1] * y;
            double yp = Ainv[3] * x + Ainv[4] * y;
            double zp = Ainv[6] * x + Ainv[7] * y;
            bool is_neg_x = false;
            if (xp < 0) {
                xp = -xp;
                yp = -yp;
                zp = -zp;
                is_neg_x = true;
            }
            const int x0 = FLOOR(xp);
            const double fx = xp - x0;
            int y0 = FLOOR(yp);
            const double fy = yp - y0;
            y0 -= starty;
            int z0 = FLOOR(zp);
            const double fz = zp - z0;
            z0 -= startz;
            if (x0 < 0 || x0 + 1 >= xdim || y0 < 0 || y0 + 1 >= ydim || z0 < 0 || z0 + 1 >= zdim) continue;
            const double mfx = 1. - fx, mfy = 1. - fy, mfz = 1. - fz;
            const double dd[8] = {mfz * mfy * mfx, mfz * mfy * fx, mfz * fy * mfx, mfz * fy * fx,
                                  fz * mfy * mfx,  fz * mfy * fx,  fz * fy * mfx,  fz * fy * fx};
            if (is_neg_x) vim = -vim;
            const int zz[8] = {z0, z0, z0, z0, z0 + 1, z0 + 1, z0 + 1, z0 + 1};
            const int yy[8] = {y0, y0, y0 + 1, y0 + 1, y0, y0, y0 + 1, y0 + 1};
            const int xx[8] = {x0, x0 + 1, x0, x0 + 1, x0, x0 + 1, x0, x0 + 1};
            for (int c = 0; c < 8; ++c) {
                const size_t v = ((size_t)zz[c] * ydim + yy[c]) * xdim + xx[c];
                vol[2 * v] += dd[c] * vre;
                vol[2 * v + 1] += dd[c] * vim;
                weight[v] += dd[c] * my_weight;
            }
        }
    }
}

extern "C" int orc_sampling_counts(const rln_estep_config *cfg, int *n_dir, int *n_psi, int *n_trans) {
    Grid g;
    if (!build_grid(cfg, g)) return RLN_ERR_ARG;
    *n_dir = g.n_dir;
    *n_psi = g.n_psi;
    *n_trans = g.n_trans;
    return RLN_OK;
}

extern "C" int orc_sampling_grid(const rln_estep_config *cfg, double *rot, double *tilt, double *psi,
                                 double *trans_x, double *trans_y) {
    Grid g;
    if (!build_grid(cfg, g)) return RLN_ERR_ARG;
    std::copy(g.rot.begin(), g.rot.end(), rot);
    std::copy(g.tilt.begin(), g.tilt.end(), tilt);
    std::copy(g.psi.begin(), g.psi.end(), psi);
    std::copy(g.tx.begin(), g.tx.end(), trans_x);
    std::copy(g.ty.begin(), g.ty.end(), trans_y);
    return RLN_OK;
}

extern "C" int orc_oversampled_orientations(const rln_estep_config *cfg, int idir, int ipsi, double *rot,
                                            double *tilt, double *psi) {
    Grid g;
    if (!build_grid(cfg, g) || idir < 0 || idir >= g.n_dir || ipsi < 0 || ipsi >= g.n_psi) return RLN_ERR_ARG;
    oversampled_orientations(g, idir, ipsi, rot, tilt, psi);
    return RLN_OK;
}

extern "C" int orc_oversampled_translations(const rln_estep_config *cfg, int itrans, double *x, double *y) {
    Grid g;
    if (!build_grid(cfg, g) || itrans < 0 || itrans >= g.n_trans) return RLN_ERR_ARG;
    oversampled_translations(cfg, g, itrans, x, y);
    return RLN_OK;
}

extern "C" int orc_sampling_local(const rln_estep_config *cfg, double prior_rot, double prior_tilt,
                                  double prior_psi, int *dir_index, double *dir_prior, int max_dir, int *n_dir,
                                  int *psi_index, double *psi_prior, int max_psi, int *n_psi) {
    Grid g;
    if (!build_grid(cfg, g)) return RLN_ERR_ARG;
    std::vector<int> pd, pp;
    std::vector<double> dp, ppr;
    local_priors(cfg, g, prior_rot, prior_tilt, prior_psi, pd, dp, pp, ppr);
    *n_dir = (int)pd.size();
    *n_psi = (int)pp.size();
    if ((int)pd.size() > max_dir || (int)pp.size() > max_psi) return RLN_ERR_ARG;
    for (size_t i = 0; i < pd.size(); ++i) {
        dir_index[i] = pd[i];
        dir_prior[i] = dp[i];
    }
    for (size_t i = 0; i < pp.size(); ++i) {
        psi_index[i] = pp[i];
        psi_prior[i] = ppr[i];
    }
    return RLN_OK;
}

namespace {

struct WsumAcc {
    std::vector<double> sigma2_noise, pdf_class, pdf_direction;
    double sigma2_offset = 0, sumw_group = 0, XA = 0, AA = 0, LL = 0, Pmax = 0;
};

struct ParticleCtx {
    const rln_estep_config *cfg;
    const Grid *g;
    const orc_model *model;
    int xdim, ydim, zdim, r_max;
};

// The significance rule of convertAllSquaredDifferencesToWeights (src/ml_optimiser.cpp:8114-8168):
// descending accumulation of the positive weights until the running sum exceeds adaptive_fraction*sum.
int significance(const rln_estep_config *cfg, const std::vector<double> &w, double sum, bool pass0,
                 double &sig_weight, long &nr_sig) {
    std::vector<double> sorted;
    sorted.reserve(w.size());
    for (double v : w)
        if (v > 0.) sorted.push_back(v);
    if (sorted.empty()) return RLN_ERR_FILTERED_ZERO;
    std::sort(sorted.begin(), sorted.end());
    double frac = 0.;
    nr_sig = 0;
    double my_sig = 0.;
    for (long i = (long)sorted.size() - 1; i >= 0; --i) {
        if (cfg->maximum_significants > 0) {
            if (nr_sig < cfg->maximum_significants) {
                if (pass0) nr_sig++;
                my_sig = sorted[i];
            }
        } else {
            if (pass0) nr_sig++;
            my_sig = sorted[i];
        }
        frac += sorted[i];
        if (frac > cfg->adaptive_fraction * sum) break;
    }
    if (pass0 && nr_sig == 0) return RLN_ERR_NO_SIGNIFICANT;
    sig_weight = my_sig;
    return RLN_OK;
}

int particle_estep(const ParticleCtx &ctx, const rln_particle_meta &meta, const double *fimg,
                   const double *fimg_nomask, const double *fctf, rln_particle_result &res, WsumAcc &acc,
                   double *bp_data, double *bp_weight, std::vector<std::mutex> &bp_locks, orc_debug *dbg) {
    const rln_estep_config *cfg = ctx.cfg;
    const Grid &g = *ctx.g;
    const orc_model *model = ctx.model;
    const int K = cfg->nr_classes;
    const double px = cfg->pixel_size;
    const int NT = g.n_trans;
    const size_t vox = (size_t)ctx.xdim * ctx.ydim * ctx.zdim;

    // ---- orientations with non-zero prior (getFourierTransformsAndCtfs, src/ml_optimiser.cpp:5955-6030)
    std::vector<int> pdir, ppsi;
    std::vector<double> dprior, pprior;
    if (cfg->local_search) {
        local_priors(cfg, g, meta.prior_rot, meta.prior_tilt, meta.prior_psi, pdir, dprior, ppsi, pprior);
    } else {
        pdir.resize(g.n_dir);
        ppsi.resize(g.n_psi);
        for (int i = 0; i < g.n_dir; ++i) pdir[i] = i;
        for (int i = 0; i < g.n_psi; ++i) ppsi[i] = i;
    }
    const int ND = (int)pdir.size(), NP = (int)ppsi.size();
    const long n_coarse = (long)K * ND * NP * NT;
    auto pdf_or = [&](int k, int id, int ip) -> double {
        if (cfg->local_search) return dprior[id] * pprior[ip];
        return model->pdf_direction[(size_t)k * g.n_dir + id];
    };
    const double scale = cfg->do_scale_correction ? meta.scale : 1.;

    // ---- offset priors (src/ml_optimiser.cpp:7736-7749, 7790-7872)
    const double s2off = cfg->sigma2_offset;
    const double offset_norm = 2. * PI * sqrt(s2off) * sqrt(s2off);
    std::vector<double> pdf_off(NT);
    double pdf_offset_mean = 0., pdf_orientation_mean = 0.;
    for (int t = 0; t < NT; ++t) {
        const double ox = meta.old_offset_x + g.tx[t];  // pixels + Angstrom, as in the reference
        const double oy = meta.old_offset_y + g.ty[t];
        double tdiff2 = (ox - meta.prior_offset_x) * (ox - meta.prior_offset_x) / (-2. * s2off);
        tdiff2 += (oy - meta.prior_offset_y) * (oy - meta.prior_offset_y) / (-2. * s2off);
        tdiff2 *= px * px;
        pdf_off[t] = (s2off < 0.0001) ? ((tdiff2 > 0.) ? 0. : 1.) : exp(tdiff2) / offset_norm;
    }
    for (int k = 0; k < K; ++k) {
        for (int id = 0; id < ND; ++id)
            for (int ip = 0; ip < NP; ++ip) pdf_orientation_mean += pdf_or(k, id, ip);
        for (int t = 0; t < NT; ++t) pdf_offset_mean += pdf_off[t];
    }
    pdf_orientation_mean /= (double)((long)K * ND * NP);
    pdf_offset_mean /= (double)((long)K * NT);

    // ---- pass 0: coarse (getAllSquaredDifferences with exp_ipass = 0)
    const int sc = cfg->coarse_size, scx = sc / 2 + 1;
    const size_t Pc = (size_t)sc * scx;
    const int sf = cfg->current_size, sfx = sf / 2 + 1;
    const size_t Pf = (size_t)sf * sfx;
    std::vector<double> Fimg_c = window<2>(fimg, sf, sc);
    std::vector<double> ctf_c = window<1>(fctf, sf, sc);
    std::vector<double> minv_c = minvsigma2(sc, cfg, model->sigma2_noise);
    std::vector<std::vector<double>> shifted_c(NT, std::vector<double>(2 * Pc));
    for (int t = 0; t < NT; ++t)
        orc_shift_image(Fimg_c.data(), sc, cfg->ori_size, g.tx[t] / px, g.ty[t] / px, shifted_c[t].data());

    std::vector<double> Mw0(n_coarse, -999.);
    double min0 = LARGE_NUMBER;
    std::vector<double> Fref(2 * Pf), Frefctf(2 * Pf);
    double A[9];
    for (int k = 0; k < K; ++k) {
        if (!(model->pdf_class[k] > 0.)) continue;
        for (int id = 0; id < ND; ++id)
            for (int ip = 0; ip < NP; ++ip) {
                const long iorientclass = (long)k * ND * NP + (long)id * NP + ip;
                if (!(pdf_or(k, id, ip) > 0.)) continue;
                orc_euler_angles2matrix(g.rot[pdir[id]], g.tilt[pdir[id]], g.psi[ppsi[ip]], A);
                orc_project(model->volumes[k], ctx.xdim, ctx.ydim, ctx.zdim, ctx.r_max, cfg->padding_factor, A,
                            sc, Fref.data());
                for (size_t n = 0; n < Pc; ++n) {
                    const double c = cfg->do_ctf_correction ? ctf_c[n] : 1.;
                    Frefctf[2 * n] = Fref[2 * n] * c * scale;
                    Frefctf[2 * n + 1] = Fref[2 * n + 1] * c * scale;
                }
                for (int t = 0; t < NT; ++t) {
                    const double *fs = shifted_c[t].data();
                    double diff2 = meta.highres_Xi2 / 2.;
                    for (size_t n = 0; n < Pc; ++n) {
                        const double dr = Frefctf[2 * n] - fs[2 * n];
                        const double di = Frefctf[2 * n + 1] - fs[2 * n + 1];
                        diff2 += (dr * dr + di * di) * 0.5 * minv_c[n];
                    }
                    const long ihidden = iorientclass * NT + t;
                    Mw0[ihidden] = diff2;
                    if (diff2 < min0) min0 = diff2;
                }
            }
    }
    if (dbg && dbg->coarse_diff2) std::copy(Mw0.begin(), Mw0.end(), dbg->coarse_diff2);

    // ---- convertAllSquaredDifferencesToWeights, pass 0 (src/ml_optimiser.cpp:7874-8050)
    auto convert = [&](std::vector<double> &Mw, int nover, double mind, double &sum) {
        sum = 0.;
        for (int k = 0; k < K; ++k)
            for (int id = 0; id < ND; ++id)
                for (int ip = 0; ip < NP; ++ip) {
                    const long iorientclass = (long)k * ND * NP + (long)id * NP + ip;
                    double po = pdf_or(k, id, ip);
                    if (pdf_orientation_mean != 0.) po /= pdf_orientation_mean;
                    for (int t = 0; t < NT; ++t) {
                        double pf = pdf_off[t];
                        if (pdf_offset_mean > 0.) pf /= pdf_offset_mean;
                        const long ihidden = iorientclass * NT + t;
                        for (int o = 0; o < nover; ++o) {
                            double &m = Mw[ihidden * nover + o];
                            if (m < 0.) {
                                m = 0.;
                            } else {
                                double weight = po * pf;
                                const double diff2 = m - mind;
                                if (diff2 > 700.)
                                    weight = 0.;
                                else
                                    weight *= exp(-diff2);
                                m = weight;
                                sum += weight;
                            }
                        }
                    }
                }
    };
    double sum0;
    convert(Mw0, 1, min0, sum0);
    if (sum0 == 0. || std::isnan(sum0)) return RLN_ERR_SUM_WEIGHT_ZERO;
    double sig0;
    long nsig0;
    int rc = significance(cfg, Mw0, sum0, true, sig0, nsig0);
    if (rc != RLN_OK) return rc;
    std::vector<unsigned char> mask(n_coarse);
    for (long i = 0; i < n_coarse; ++i) mask[i] = (Mw0[i] >= sig0) ? 1 : 0;
    if (dbg && dbg->coarse_weight) std::copy(Mw0.begin(), Mw0.end(), dbg->coarse_weight);
    if (dbg && dbg->coarse_significant) std::copy(mask.begin(), mask.end(), dbg->coarse_significant);
    res.nr_significant = (double)nsig0;
    res.coarse_sum_weight = sum0;
    res.coarse_min_diff2 = min0;
    res.coarse_significant_weight = sig0;
    auto any_sig = [&](long iorientclass) {
        for (int t = 0; t < NT; ++t)
            if (mask[iorientclass * NT + t]) return true;
        return false;
    };

    // ---- pass 1: fine (getAllSquaredDifferences with exp_ipass = 1, adaptive_oversampling = 1)
    const std::vector<double> minv_f = minvsigma2(sf, cfg, model->sigma2_noise);
    std::vector<std::vector<double>> over_x(NT, std::vector<double>(4)), over_y(NT, std::vector<double>(4));
    for (int t = 0; t < NT; ++t) oversampled_translations(cfg, g, t, over_x[t].data(), over_y[t].data());
    // shifted images on demand, index t*4+u
    std::vector<std::vector<double>> shifted_f(NT * 4), shifted_fn(NT * 4);
    auto get_shift = [&](std::vector<std::vector<double>> &cache, const double *src, int t, int u) -> const double * {
        std::vector<double> &v = cache[t * 4 + u];
        if (v.empty()) {
            v.resize(2 * Pf);
            orc_shift_image(src, sf, cfg->ori_size, over_x[t][u], over_y[t][u], v.data());
        }
        return v.data();
    };
    std::vector<double> Mw1((size_t)n_coarse * 32, -999.);
    double min1 = LARGE_NUMBER;
    double orot[8], otilt[8], opsi[8];
    int n_fine = 0;
    for (int k = 0; k < K; ++k) {
        if (!(model->pdf_class[k] > 0.)) continue;
        for (int id = 0; id < ND; ++id)
            for (int ip = 0; ip < NP; ++ip) {
                const long iorientclass = (long)k * ND * NP + (long)id * NP + ip;
                if (!any_sig(iorientclass) || !(pdf_or(k, id, ip) > 0.)) continue;
                oversampled_orientations(g, pdir[id], ppsi[ip], orot, otilt, opsi);
                for (int r = 0; r < 8; ++r) {
                    orc_euler_angles2matrix(orot[r], otilt[r], opsi[r], A);
                    orc_project(model->volumes[k], ctx.xdim, ctx.ydim, ctx.zdim, ctx.r_max, cfg->padding_factor,
                                A, sf, Fref.data());
                    for (size_t n = 0; n < Pf; ++n) {
                        const double c = cfg->do_ctf_correction ? fctf[n] : 1.;
                        Frefctf[2 * n] = Fref[2 * n] * c * scale;
                        Frefctf[2 * n + 1] = Fref[2 * n + 1] * c * scale;
                    }
                    for (int t = 0; t < NT; ++t) {
                        const long ihidden = iorientclass * NT + t;
                        if (!mask[ihidden]) continue;
                        for (int u = 0; u < 4; ++u) {
                            const double *fs = get_shift(shifted_f, fimg, t, u);
                            double diff2 = meta.highres_Xi2 / 2.;
                            for (size_t n = 0; n < Pf; ++n) {
                                const double dr = Frefctf[2 * n] - fs[2 * n];
                                const double di = Frefctf[2 * n + 1] - fs[2 * n + 1];
                                diff2 += (dr * dr + di * di) * 0.5 * minv_f[n];
                            }
                            const long ihidden_over = (ihidden * 8 + r) * 4 + u;
                            Mw1[ihidden_over] = diff2;
                            if (diff2 < min1) min1 = diff2;
                            ++n_fine;
                        }
                    }
                }
            }
    }
    if (dbg && dbg->fine_diff2) std::copy(Mw1.begin(), Mw1.end(), dbg->fine_diff2);
    double sum1;
    convert(Mw1, 32, min1, sum1);
    if (sum1 == 0. || std::isnan(sum1)) return RLN_ERR_SUM_WEIGHT_ZERO;
    double sig1;
    long nsig1;
    rc = significance(cfg, Mw1, sum1, false, sig1, nsig1);
    if (rc != RLN_OK) return rc;
    if (dbg && dbg->fine_weight) std::copy(Mw1.begin(), Mw1.end(), dbg->fine_weight);
    res.sum_weight = sum1;
    res.min_diff2 = min1;
    res.significant_weight = sig1;
    res.nr_fine_evaluated = n_fine;

    // ---- storeWeightedSums (src/ml_optimiser.cpp:8241-9156)
    std::vector<double> minv_w = minv_f;
    minv_w[0] = 1. / (cfg->sigma2_fudge * model->sigma2_noise[0]);  // origin restored (:8307)
    const std::vector<int> mres = mresol(sf);
    double myscale = 1.;
    if (cfg->do_scale_correction) {
        myscale = meta.scale;
        if (myscale > 10000.) return RLN_ERR_ARG;
        if (myscale < 0.001) myscale = 0.001;
    }
    std::vector<double> Mctf(Pf), Fimg_acc(2 * Pf), Fweight(Pf);
    double max_weight = -1.;
    double norm_corr = 0., XA = 0., AA = 0.;
    for (int k = 0; k < K; ++k)
        for (int id = 0; id < ND; ++id)
            for (int ip = 0; ip < NP; ++ip) {
                const long iorientclass = (long)k * ND * NP + (long)id * NP + ip;
                if (!any_sig(iorientclass)) continue;
                oversampled_orientations(g, pdir[id], ppsi[ip], orot, otilt, opsi);
                for (int r = 0; r < 8; ++r) {
                    orc_euler_angles2matrix(orot[r], otilt[r], opsi[r], A);
                    orc_project(model->volumes[k], ctx.xdim, ctx.ydim, ctx.zdim, ctx.r_max, cfg->padding_factor,
                                A, sf, Fref.data());
                    std::fill(Fimg_acc.begin(), Fimg_acc.end(), 0.);
                    std::fill(Fweight.begin(), Fweight.end(), 0.);
                    for (size_t n = 0; n < Pf; ++n) {
                        Mctf[n] = cfg->do_ctf_correction ? fctf[n] : 1.;
                        Frefctf[2 * n] = Fref[2 * n] * Mctf[n];
                        Frefctf[2 * n + 1] = Fref[2 * n + 1] * Mctf[n];
                    }
                    if (cfg->do_scale_correction)
                        for (size_t n = 0; n < Pf; ++n) {
                            Frefctf[2 * n] *= myscale;
                            Frefctf[2 * n + 1] *= myscale;
                            Mctf[n] *= myscale;
                        }
                    bool any = false;
                    for (int t = 0; t < NT; ++t) {
                        const long ihidden = iorientclass * NT + t;
                        for (int u = 0; u < 4; ++u) {
                            const long ihidden_over = ihidden * 32 + r * 4 + u;
                            double weight = Mw1[ihidden_over];
                            if (!(weight >= sig1)) continue;
                            weight /= sum1;
                            any = true;
                            const double *fs = get_shift(shifted_f, fimg, t, u);
                            const double *fn = get_shift(shifted_fn, fimg_nomask, t, u);
                            for (size_t n = 0; n < Pf; ++n) {
                                const int ires = mres[n];
                                if (ires > -1) {
                                    const double dr = Frefctf[2 * n] - fs[2 * n];
                                    const double di = Frefctf[2 * n + 1] - fs[2 * n + 1];
                                    const double wdiff2 = weight * (dr * dr + di * di);
                                    acc.sigma2_noise[ires] += wdiff2;
                                    norm_corr += wdiff2;
                                    if (cfg->do_scale_correction && model->data_vs_prior &&
                                        model->data_vs_prior[(size_t)k * (cfg->ori_size / 2 + 1) + ires] > 3.) {
                                        XA += weight * (Frefctf[2 * n] * fs[2 * n] + Frefctf[2 * n + 1] * fs[2 * n + 1]);
                                        AA += weight * (Frefctf[2 * n] * Frefctf[2 * n] +
                                                        Frefctf[2 * n + 1] * Frefctf[2 * n + 1]);
                                    }
                                }
                            }
                            acc.sumw_group += weight;
                            acc.pdf_class[k] += weight;
                            const double diffx = meta.prior_offset_x - meta.old_offset_x - over_x[t][u];
                            const double diffy = meta.prior_offset_y - meta.old_offset_y - over_y[t][u];
                            acc.sigma2_offset += weight * px * px * (diffx * diffx + diffy * diffy);
                            acc.pdf_direction[(size_t)k * g.n_dir + pdir[id]] += weight;
                            for (size_t n = 0; n < Pf; ++n) {
                                const double myctf = Mctf[n];
                                const double wx = weight * myctf * minv_w[n];
                                Fimg_acc[2 * n] += fn[2 * n] * wx;
                                Fimg_acc[2 * n + 1] += fn[2 * n + 1] * wx;
                                Fweight[n] += wx * myctf;
                            }
                            if (weight > max_weight) {
                                max_weight = weight;
                                res.rot = orot[r];
                                res.tilt = otilt[r];
                                res.psi = opsi[r];
                                res.xoff = meta.old_offset_x + over_x[t][u];
                                res.yoff = meta.old_offset_y + over_y[t][u];
                                res.iclass = k + 1;
                                res.pmax = max_weight;
                                res.best_ihidden_over = ihidden_over;
                            }
                        }
                    }
                    if (any && bp_data) {
                        std::lock_guard<std::mutex> lk(bp_locks[k]);
                        orc_backproject(bp_data + 2 * vox * k, bp_weight + vox * k, ctx.xdim, ctx.ydim, ctx.zdim,
                                        ctx.r_max, cfg->padding_factor, A, sf, Fimg_acc.data(), Fweight.data());
                    }
                }
            }
    if (cfg->do_scale_correction) {
        XA /= meta.scale;
        AA /= meta.scale * meta.scale;
    }
    res.signal_product = XA;
    res.reference_power = AA;
    res.wsum_norm_correction = norm_corr;
    double logsigma2 = 0.;
    for (size_t n = 0; n < Pf; ++n) {
        const int ires = mres[n];
        if (ires > 0 && ires < cfg->ori_size / 2 + 1) logsigma2 += log(2. * PI * model->sigma2_noise[ires]);
    }
    res.dLL = log(sum1) - min1 - logsigma2;
    acc.XA += XA;
    acc.AA += AA;
    acc.LL += res.dLL;
    acc.Pmax += res.pmax;
    return RLN_OK;
}

}  // namespace

extern "C" int orc_estep(const rln_estep_config *cfg, const orc_model *model, int n, const rln_particle_meta *meta,
                         const double *fimg, const double *fimg_nomask, const double *fctf,
                         rln_particle_result *results, rln_wsum *wsum, double *bp_data, double *bp_weight,
                         orc_debug *dbg, int nr_threads) {
    if (!cfg || !model || n < 0 || cfg->adaptive_oversampling != 1 || cfg->coarse_size > cfg->current_size ||
        cfg->current_size > cfg->ori_size || (cfg->coarse_size & 1) || (cfg->current_size & 1))
        return RLN_ERR_ARG;
    Grid g;
    if (!build_grid(cfg, g)) return RLN_ERR_ARG;
    ParticleCtx ctx;
    ctx.cfg = cfg;
    ctx.g = &g;
    ctx.model = model;
    ctx.r_max = std::min(cfg->current_size / 2, cfg->ori_size / 2);
    const int pad_size = 2 * (ROUND(cfg->padding_factor * ctx.r_max) + 1) + 1;
    ctx.xdim = pad_size / 2 + 1;
    ctx.ydim = ctx.zdim = pad_size;
    const int K = cfg->nr_classes;
    const int nsig = cfg->ori_size / 2 + 1;
    const size_t Pf = (size_t)cfg->current_size * (cfg->current_size / 2 + 1);
    std::vector<std::mutex> bp_locks(K);
    int status = RLN_OK;
    std::mutex merge;
#ifdef _OPENMP
    if (nr_threads <= 0) nr_threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nr_threads)
#endif
    for (int i = 0; i < n; ++i) {
        WsumAcc acc;
        acc.sigma2_noise.assign(nsig, 0.);
        acc.pdf_class.assign(K, 0.);
        acc.pdf_direction.assign((size_t)K * g.n_dir, 0.);
        rln_particle_result r;
        std::memset(&r, 0, sizeof(r));
        const int rc = particle_estep(ctx, meta[i], fimg + 2 * Pf * i, fimg_nomask + 2 * Pf * i, fctf + Pf * i, r,
                                      acc, bp_data, bp_weight, bp_locks, dbg ? &dbg[i] : nullptr);
        std::lock_guard<std::mutex> lk(merge);
        if (rc != RLN_OK) {
            status = rc;
            continue;
        }
        if (results) results[i] = r;
        if (wsum) {
            if (wsum->sigma2_noise)
                for (int s = 0; s < nsig; ++s) wsum->sigma2_noise[s] += acc.sigma2_noise[s];
            if (wsum->pdf_class)
                for (int k = 0; k < K; ++k) wsum->pdf_class[k] += acc.pdf_class[k];
            if (wsum->pdf_direction)
                for (size_t d = 0; d < acc.pdf_direction.size(); ++d) wsum->pdf_direction[d] += acc.pdf_direction[d];
            wsum->sigma2_offset += acc.sigma2_offset;
            wsum->sumw_group += acc.sumw_group;
            wsum->signal_product += acc.XA;
            wsum->reference_power += acc.AA;
            wsum->LL += acc.LL;
            wsum->ave_Pmax += acc.Pmax;
        }
    }
    return status;
}
