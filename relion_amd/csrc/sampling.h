// Host-side pose grid of the E-step: HEALPix NEST directions, psi angles, 2-D translations, order-1
// oversampling and local-search priors (the reference's HealpixSampling, src/healpix_sampling.cpp).
#pragma once

#include <vector>

#include "rln_common.h"

namespace rln {

struct PoseGrid {
    int order = 0;
    int n_dir = 0, n_psi = 0, n_trans = 0;
    double psi_step = 0., offset_step = 0., pixel_size = 1.;
    std::vector<double> rot, tilt;   // per direction, degrees
    std::vector<double> psi;         // per psi index, degrees
    std::vector<double> tx, ty;      // translations, Angstrom
    std::vector<double> dir_unit;    // per direction: Euler_angles2direction (x, y, z)

    // Builds the C1 grid of cfg (HealpixSampling::initialise, src/healpix_sampling.cpp:51-138).
    bool build(const rln_estep_config &cfg);
    // getOrientations at oversampling order 1: 8 (rot, tilt, psi), src/healpix_sampling.cpp:1832-1966.
    void oversampled_orientations(int idir, int ipsi, double *rot8, double *tilt8, double *psi8) const;
    // one of them: iover_rot = (2*dj + di)*2 + q (sub-pixel row dj, column di, psi half q)
    void oversampled_orientation(int idir, int ipsi, int iover_rot, double &rot, double &tilt, double &psi) const;
    // getTranslationsInPixel at oversampling order 1: 4 shifts in pixels, src/healpix_sampling.cpp:1724-1830.
    void oversampled_translations(int itrans, double *x4, double *y4) const;
    // selectOrientationsWithNonZeroPriorProbability (src/healpix_sampling.cpp:695-1076), C1, 3-D.
    void local_priors(const rln_estep_config &cfg, double prior_rot, double prior_tilt, double prior_psi,
                      std::vector<int> &dir_idx, std::vector<double> &dir_prior, std::vector<int> &psi_idx,
                      std::vector<double> &psi_prior) const;
};

// Non-zero-prior orientations of one particle.
struct LocalSet {
    std::vector<int> dir_idx, psi_idx;
    std::vector<double> dir_prior, psi_prior;
};

// local_priors for n particles in parallel (OpenMP); out[i] for particle i.
void local_priors_batch(const PoseGrid &g, const rln_estep_config &cfg, const rln_particle_meta *meta, int n,
                        std::vector<LocalSet> &out);

// HEALPix NEST primitives (Gorski et al. 2005), as in src/Healpix_2.15a/healpix_base.cc.
void hp_nest2xyf(int order, long long pix, int &ix, int &iy, int &face);
long long hp_xyf2nest(int order, int ix, int iy, int face);
void hp_pix2ang_z_phi(int order, long long pix, double &z, double &phi);
long long hp_ang2pix_nest(int order, double z, double phi);

// Euler_angles2matrix (src/euler.cpp:52-91), row-major.
void euler_angles2matrix(double rot, double tilt, double psi, double *A);
void matrix_to_angles(const double *A, double &rot, double &tilt, double &psi);

}  // namespace rln
