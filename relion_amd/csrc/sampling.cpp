// Pose grid of the E-step on the host (HealpixSampling, src/healpix_sampling.cpp), C1 symmetry.
#include "sampling.h"

#include <algorithm>
#include <cmath>

namespace rln {

namespace {

// Morton (bit-interleave) helpers for the NEST pixel index within a face.
inline unsigned long long part1by1(unsigned long long v) {
    v &= 0x00000000ffffffffULL;
    v = (v | (v << 16)) & 0x0000ffff0000ffffULL;
    v = (v | (v << 8)) & 0x00ff00ff00ff00ffULL;
    v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0fULL;
    v = (v | (v << 2)) & 0x3333333333333333ULL;
    v = (v | (v << 1)) & 0x5555555555555555ULL;
    return v;
}
inline unsigned long long compact1by1(unsigned long long v) {
    v &= 0x5555555555555555ULL;
    v = (v | (v >> 1)) & 0x3333333333333333ULL;
    v = (v | (v >> 2)) & 0x0f0f0f0f0f0f0f0fULL;
    v = (v | (v >> 4)) & 0x00ff00ff00ff00ffULL;
    v = (v | (v >> 8)) & 0x0000ffff0000ffffULL;
    v = (v | (v >> 16)) & 0x00000000ffffffffULL;
    return v;
}

// Ring index offset and longitude offset of each of the 12 base faces (Gorski et al. 2005, Table 1).
const int kFaceRing[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
const int kFacePhi[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};

inline double deg(double rad) { return rad * 180. / kPi; }
inline double rad(double d) { return d * kPi / 180.; }

// HealpixSampling::checkDirection (src/healpix_sampling.cpp:1541-1564)
void wrap_direction(double &rot, double &tilt) {
    if (tilt < 0.) {
        tilt = -tilt;
        rot += 180.;
    }
    for (;;) {
        if (rot > 180.)
            rot -= 360.;
        else if (rot < -180.)
            rot += 360.;
        else
            break;
    }
}

void healpix_direction(int order, long long pix, double &rot, double &tilt) {
    double z, phi;
    hp_pix2ang_z_phi(order, pix, z, phi);
    rot = deg(phi);
    tilt = acos(z) * 180. / kPi;
    wrap_direction(rot, tilt);
}

// gaussian1D (src/funcs.cpp:356-360)
inline double normal_pdf(double x, double sigma) {
    return 1 / sqrt(2 * kPi * sigma * sigma) * exp(-0.5 * ((x / sigma) * (x / sigma)));
}

}  // namespace

void hp_nest2xyf(int order, long long pix, int &ix, int &iy, int &face) {
    const unsigned long long within = (unsigned long long)pix & ((1ULL << (2 * order)) - 1);
    face = (int)(pix >> (2 * order));
    ix = (int)compact1by1(within);
    iy = (int)compact1by1(within >> 1);
}

long long hp_xyf2nest(int order, int ix, int iy, int face) {
    return ((long long)face << (2 * order)) + (long long)(part1by1(ix) | (part1by1(iy) << 1));
}

void hp_pix2ang_z_phi(int order, long long pix, double &z, double &phi) {
    const long long nside = 1LL << order;
    const double npix = 12. * (double)nside * (double)nside;
    const double f2 = 4. / npix;
    const double f1 = (double)(2 * nside) * f2;
    int ix, iy, face;
    hp_nest2xyf(order, pix, ix, iy, face);
    const long long ring = ((long long)kFaceRing[face] << order) - ix - iy - 1;  // 1-based ring number
    long long nr;
    int shift;
    if (ring < nside) {  // north polar cap
        nr = ring;
        z = 1. - (double)(nr * nr) * f2;
        shift = 0;
    } else if (ring > 3 * nside) {  // south polar cap
        nr = 4 * nside - ring;
        z = (double)(nr * nr) * f2 - 1.;
        shift = 0;
    } else {  // equatorial belt
        nr = nside;
        z = (double)(2 * nside - ring) * f1;
        shift = (int)((ring - nside) & 1);
    }
    long long jp = (kFacePhi[face] * nr + ix - iy + 1 + shift) / 2;
    if (jp > 4 * nside) jp -= 4 * nside;
    if (jp < 1) jp += 4 * nside;
    phi = ((double)jp - (double)(shift + 1) * 0.5) * ((kPi / 2.) / (double)nr);
}

// HEALPix NEST pixel containing direction (z = cos(theta), phi): the ang2pix of Gorski et al. 2005 (ApJ 622,
// 759), as Healpix_Base::ang2pix_z_phi (src/Healpix_2.15a/healpix_base.cc:368).
long long hp_ang2pix_nest(int order, double z, double phi) {
    const int nside = 1 << order;
    const double za = std::fabs(z);
    double tt = std::fmod(phi, 2. * kPi);
    if (tt < 0.) tt += 2. * kPi;
    tt *= 2. / kPi;  // in [0, 4)
    int face, ix, iy;
    if (za <= 2. / 3.) {
        const double t1 = nside * (0.5 + tt), t2 = nside * z * 0.75;
        const int jp = (int)(t1 - t2), jm = (int)(t1 + t2);
        const int ifp = jp / nside, ifm = jm / nside;
        face = (ifp == ifm) ? (ifp | 4) : ((ifp < ifm) ? ifp : ifm + 8);
        ix = jm & (nside - 1);
        iy = nside - (jp & (nside - 1)) - 1;
    } else {
        int ntt = (int)tt;
        if (ntt >= 4) ntt = 3;
        const double tp = tt - ntt;
        const double tmp = nside * std::sqrt(3. * (1. - za));
        int jp = (int)(tp * tmp), jm = (int)((1. - tp) * tmp);
        jp = std::min(jp, nside - 1);
        jm = std::min(jm, nside - 1);
        if (z >= 0.) {
            face = ntt;
            ix = nside - jm - 1;
            iy = nside - jp - 1;
        } else {
            face = ntt + 8;
            ix = jp;
            iy = jm;
        }
    }
    return hp_xyf2nest(order, ix, iy, face);
}


void euler_angles2matrix(double rot, double tilt, double psi, double *A) {
    const double a = rad(rot), b = rad(tilt), g = rad(psi);
    const double ca = cos(a), sa = sin(a), cb = cos(b), sb = sin(b), cg = cos(g), sg = sin(g);
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

// ZYZ angles of a rotation matrix in the convention of Euler_angles2matrix (src/euler.cpp:52-91).
void matrix_to_angles(const double *A, double &rot, double &tilt, double &psi) {
    const double sb = std::sqrt(A[2] * A[2] + A[5] * A[5]);
    if (sb > 1e-9) {
        psi = std::atan2(A[5], -A[2]);
        rot = std::atan2(A[7], A[6]);
        tilt = std::atan2(sb, A[8]);
    } else {
        rot = 0.;
        tilt = A[8] > 0 ? 0. : kPi;
        psi = std::atan2(-A[3], A[0]);
    }
    rot *= 180. / kPi;
    tilt *= 180. / kPi;
    psi *= 180. / kPi;
}

bool PoseGrid::build(const rln_estep_config &cfg) {
    if (cfg.healpix_order < 0 || cfg.healpix_order > 12 || !(cfg.offset_step > 0.) || cfg.offset_range < 0. ||
        !(cfg.pixel_size > 0.))
        return false;
    order = cfg.healpix_order;
    offset_step = cfg.offset_step;
    pixel_size = cfg.pixel_size;
    // directions (setOrientations, src/healpix_sampling.cpp:477-556)
    n_dir = 12 << (2 * order);
    rot.resize(n_dir);
    tilt.resize(n_dir);
    dir_unit.resize(3 * (size_t)n_dir);
    for (int p = 0; p < n_dir; ++p) {
        healpix_direction(order, p, rot[p], tilt[p]);
        const double a = rad(rot[p]), b = rad(tilt[p]);
        dir_unit[3 * p] = sin(b) * cos(a);
        dir_unit[3 * p + 1] = sin(b) * sin(a);
        dir_unit[3 * p + 2] = cos(b);
    }
    // psi (default step: src/healpix_sampling.cpp:73-74)
    double step = cfg.psi_step;
    if (step <= 0.) step = 360. / (6 * ROUND(std::pow(2., order)));
    n_psi = CEIL(360. / step);
    psi_step = 360. / (double)n_psi;
    psi.resize(n_psi);
    for (int i = 0; i < n_psi; ++i) psi[i] = i * psi_step;
    // translations (setTranslations, src/healpix_sampling.cpp:291-450): disc of radius offset_range
    const int m = CEIL(cfg.offset_range / cfg.offset_step);
    const double lim = cfg.offset_range * cfg.offset_range + 0.001;
    tx.clear();
    ty.clear();
    for (int ix = -m; ix <= m; ++ix)
        for (int iy = -m; iy <= m; ++iy) {
            const double x = ix * cfg.offset_step, y = iy * cfg.offset_step;
            if (x * x + y * y < lim) {
                tx.push_back(x);
                ty.push_back(y);
            }
        }
    n_trans = (int)tx.size();
    return true;
}

void PoseGrid::oversampled_orientations(int idir, int ipsi, double *rot8, double *tilt8, double *psi8) const {
    int x, y, face;
    hp_nest2xyf(order, idir, x, y, face);
    int k = 0;
    for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
            double r, t;
            healpix_direction(order + 1, hp_xyf2nest(order + 1, 2 * x + di, 2 * y + dj, face), r, t);
            for (int q = 0; q < 2; ++q, ++k) {
                rot8[k] = r;
                tilt8[k] = t;
                psi8[k] = psi[ipsi] - 0.5 * psi_step + (0.5 + q) * psi_step / 2.;
            }
        }
}

void PoseGrid::oversampled_orientation(int idir, int ipsi, int iover_rot, double &r, double &t, double &p) const {
    int x, y, face;
    hp_nest2xyf(order, idir, x, y, face);
    const int sub = iover_rot >> 1, q = iover_rot & 1;
    const int dj = sub >> 1, di = sub & 1;
    healpix_direction(order + 1, hp_xyf2nest(order + 1, 2 * x + di, 2 * y + dj, face), r, t);
    p = psi[ipsi] - 0.5 * psi_step + (0.5 + q) * psi_step / 2.;
}

void PoseGrid::oversampled_translations(int itrans, double *x4, double *y4) const {
    int k = 0;
    for (int ox = 0; ox < 2; ++ox)
        for (int oy = 0; oy < 2; ++oy, ++k) {
            x4[k] = (tx[itrans] - 0.5 * offset_step + (0.5 + ox) * offset_step / 2.) / pixel_size;
            y4[k] = (ty[itrans] - 0.5 * offset_step + (0.5 + oy) * offset_step / 2.) / pixel_size;
        }
}

void PoseGrid::local_priors(const rln_estep_config &cfg, double prior_rot, double prior_tilt, double prior_psi,
                            std::vector<int> &dir_idx, std::vector<double> &dir_prior, std::vector<int> &psi_idx,
                            std::vector<double> &psi_prior) const {
    dir_idx.clear();
    dir_prior.clear();
    psi_idx.clear();
    psi_prior.clear();
    const double cutoff = cfg.sigma_cutoff > 0. ? cfg.sigma_cutoff : 3.;
    auto unit = [](double a, double b, double *v) {  // Euler_angles2direction (src/euler.cpp:94-113)
        a = rad(a);
        b = rad(b);
        v[0] = sin(b) * cos(a);
        v[1] = sin(b) * sin(a);
        v[2] = cos(b);
    };
    double pv[3];
    unit(prior_rot, prior_tilt, pv);
    const bool use_dir = cfg.sigma_rot > 0. && cfg.sigma_tilt > 0.;
    const double sig = std::max(cfg.sigma_rot, cfg.sigma_tilt);
    double total = 0., best = 9999.;
    int best_dir = -1;
    // Directions whose dot product with the prior is clearly below cos(cutoff*sigma) cannot pass the angular
    // test below; they are skipped without the acos (the test itself is unchanged).
    const double lim = cutoff * sig;
    const double dot_min = (lim < 179.) ? cos((lim + 1.) * kPi / 180.) : -2.;
    for (int pass = 0; pass < 2; ++pass) {
        for (int d = 0; d < n_dir; ++d) {
            if (!use_dir) {
                dir_idx.push_back(d);
                dir_prior.push_back(1.);
                total += 1.;
                continue;
            }
            const double *v = &dir_unit[3 * (size_t)d];
            const double dot = v[0] * pv[0] + v[1] * pv[1] + v[2] * pv[2];
            if (pass == 0 && dot < dot_min) continue;
            double ang = acos(dot) * 180. / kPi;
            if (ang > 180.) ang = std::fabs(ang - 360.);
            if (ang < lim) {
                const double p = normal_pdf(ang, sig);
                dir_idx.push_back(d);
                dir_prior.push_back(p);
                total += p;
            }
            if (ang < best) {
                best = ang;
                best_dir = d;
            }
        }
        // the fallback to the nearest direction needs every direction's angle
        if (!dir_idx.empty() || !use_dir) break;
        best = 9999.;
        best_dir = -1;
    }
    for (double &p : dir_prior) p /= total;
    if (dir_idx.empty()) {
        dir_idx.push_back(best_dir);
        dir_prior.push_back(1.);
    }
    total = 0.;
    best = 9999.;
    int best_psi = -1;
    for (int i = 0; i < n_psi; ++i) {
        if (!(cfg.sigma_psi > 0.)) {
            psi_idx.push_back(i);
            psi_prior.push_back(1.);
            total += 1.;
            continue;
        }
        double dpsi = std::fabs(psi[i] - prior_psi);
        if (dpsi > 180.) dpsi = std::fabs(dpsi - 360.);
        if (dpsi < cutoff * cfg.sigma_psi) {
            const double p = normal_pdf(dpsi, cfg.sigma_psi);
            psi_idx.push_back(i);
            psi_prior.push_back(p);
            total += p;
        }
        if (dpsi < best) {
            best = dpsi;
            best_psi = i;
        }
    }
    for (double &p : psi_prior) p /= total;
    if (psi_idx.empty()) {
        psi_idx.push_back(best_psi);
        psi_prior.push_back(1.);
    }
}

void local_priors_batch(const PoseGrid &g, const rln_estep_config &cfg, const rln_particle_meta *meta, int n,
                        std::vector<LocalSet> &out) {
    out.resize(n);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i)
        g.local_priors(cfg, meta[i].prior_rot, meta[i].prior_tilt, meta[i].prior_psi, out[i].dir_idx,
                       out[i].dir_prior, out[i].psi_idx, out[i].psi_prior);
}

}  // namespace rln

// ------------------------------------------------------------------------------------------------
// C ABI sampling helpers
// ------------------------------------------------------------------------------------------------
extern "C" int rln_coarse_image_size(int ori_size, double pixel_size, int healpix_order, double particle_diameter,
                                     int current_size) {
    // src/ml_optimiser.cpp:5757-5779 with adaptive_oversampling > 0, no strict_highres_exp, no max_coarse_size,
    // one optics group at the model's pixel size and box.
    if (ori_size <= 0 || !(pixel_size > 0.) || !(particle_diameter > 0.) || healpix_order < 0) {
        rln::set_error("rln_coarse_image_size: invalid argument");
        return RLN_ERR_ARG;
    }
    const double angular = 360. / (6 * rln::ROUND(std::pow(2., healpix_order)));
    const double rotated_distance = (angular / 360.) * rln::kPi * particle_diameter;
    const double coarse_resolution = rotated_distance / 1.2;  // keepsafe factor for 3-D references
    int coarse = 2 * rln::CEIL(pixel_size * ori_size / coarse_resolution);
    coarse = std::min(coarse, ori_size);
    coarse = std::min(coarse, current_size);
    return coarse;
}

extern "C" int rln_sampling_counts(const rln_estep_config *cfg, int *n_dir, int *n_psi, int *n_trans) {
    rln::PoseGrid g;
    if (!cfg || !g.build(*cfg)) RLN_FAIL(RLN_ERR_ARG, "rln_sampling_counts: invalid sampling options");
    *n_dir = g.n_dir;
    *n_psi = g.n_psi;
    *n_trans = g.n_trans;
    return RLN_OK;
}

extern "C" int rln_sampling_grid(const rln_estep_config *cfg, double *rot, double *tilt, double *psi,
                                 double *trans_x, double *trans_y) {
    rln::PoseGrid g;
    if (!cfg || !g.build(*cfg)) RLN_FAIL(RLN_ERR_ARG, "rln_sampling_grid: invalid sampling options");
    std::copy(g.rot.begin(), g.rot.end(), rot);
    std::copy(g.tilt.begin(), g.tilt.end(), tilt);
    std::copy(g.psi.begin(), g.psi.end(), psi);
    std::copy(g.tx.begin(), g.tx.end(), trans_x);
    std::copy(g.ty.begin(), g.ty.end(), trans_y);
    return RLN_OK;
}

extern "C" int rln_sampling_local(const rln_estep_config *cfg, double prior_rot, double prior_tilt, double prior_psi,
                                  int *dir_index, double *dir_prior, int max_dir, int *n_dir, int *psi_index,
                                  double *psi_prior, int max_psi, int *n_psi) {
    rln::PoseGrid g;
    if (!cfg || !g.build(*cfg)) RLN_FAIL(RLN_ERR_ARG, "rln_sampling_local: invalid sampling options");
    std::vector<int> di, pi;
    std::vector<double> dp, pp;
    g.local_priors(*cfg, prior_rot, prior_tilt, prior_psi, di, dp, pi, pp);
    *n_dir = (int)di.size();
    *n_psi = (int)pi.size();
    if ((int)di.size() > max_dir || (int)pi.size() > max_psi)
        RLN_FAIL(RLN_ERR_ARG, "rln_sampling_local: output capacity too small (%d dirs, %d psi)", *n_dir, *n_psi);
    std::copy(di.begin(), di.end(), dir_index);
    std::copy(dp.begin(), dp.end(), dir_prior);
    std::copy(pi.begin(), pi.end(), psi_index);
    std::copy(pp.begin(), pp.end(), psi_prior);
    return RLN_OK;
}

extern "C" int rln_sampling_ang2pix(int order, double rot, double tilt, long long *pix) {
    if (order < 0 || order > 12 || !pix) {
        rln::set_error("rln_sampling_ang2pix: invalid argument");
        return RLN_ERR_ARG;
    }
    const double a = rot * rln::kPi / 180., b = tilt * rln::kPi / 180.;
    *pix = rln::hp_ang2pix_nest(order, std::cos(b), std::atan2(std::sin(b) * std::sin(a), std::sin(b) * std::cos(a)));
    return RLN_OK;
}
