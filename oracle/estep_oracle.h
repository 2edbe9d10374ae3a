/*
 * TEST INFRASTRUCTURE ONLY -- never linked into or called by the product library.
 *
 * FP64 CPU restatement of RELION's CPU expectation step, used as the parity oracle
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * Every function cites the reference file:line it restates.
 */
#ifndef RLN_ESTEP_ORACLE_H
#define RLN_ESTEP_ORACLE_H

#include "../include/rln_estep.h"

#ifdef __cplusplus
extern "C" {
#endif

/* HEALPix NEST helpers (src/Healpix_2.15a/healpix_base.cc). */
void orc_hp_pix2ang_z_phi(int order, long long pix, double *z, double *phi);
void orc_hp_nest2xyf(int order, long long pix, int *ix, int *iy, int *face);
long long orc_hp_xyf2nest(int order, int ix, int iy, int face);

/* Euler_angles2matrix (src/euler.cpp:52-91), row-major 3x3. */
void orc_euler_angles2matrix(double rot, double tilt, double psi, double *A);

/* CTF::getCTF after CTF::setValues (src/ctf.h:184-240, src/ctf.cpp:134-261), x,y in 1/Angstrom. */
double orc_ctf_value(double defU, double defV, double defAngle, double voltage, double cs, double q0,
                     double bfactor, double scale, double phase_shift, double x, double y);

/* Coarse sampling grid (HealpixSampling::setTranslations / setOrientations, C1). */
int orc_sampling_counts(const rln_estep_config *cfg, int *n_dir, int *n_psi, int *n_trans);
int orc_sampling_grid(const rln_estep_config *cfg, double *rot, double *tilt, double *psi,
                      double *trans_x, double *trans_y);
/* HealpixSampling::getOrientations / getTranslationsInPixel at oversampling order 1 for one coarse
 * (direction, psi) / translation: 8 (rot,tilt,psi) triplets, 4 (x,y) pixel shifts. */
int orc_oversampled_orientations(const rln_estep_config *cfg, int idir, int ipsi, double *rot, double *tilt,
                                 double *psi);
int orc_oversampled_translations(const rln_estep_config *cfg, int itrans, double *x, double *y);
int orc_sampling_local(const rln_estep_config *cfg, double prior_rot, double prior_tilt, double prior_psi,
                       int *dir_index, double *dir_prior, int max_dir, int *n_dir,
                       int *psi_index, double *psi_prior, int max_psi, int *n_psi);

/* Projector::project (src/projector.cpp:630-797): volume in the projector layout (interleaved
 * double complex), image of size s x (s/2+1) (interleaved), zeroed first. */
void orc_project(const double *vol, int xdim, int ydim, int zdim, int r_max, double padding_factor,
                 const double *A, int s, double *img);
/* BackProjector::backproject2Dto3D (src/backprojector.cpp:55-357), accumulating. */
void orc_backproject(double *vol, double *weight, int xdim, int ydim, int zdim, int r_max,
                     double padding_factor, const double *A, int s, const double *img, const double *mweight);
/* shiftImageInFourierTransform (src/fftw.cpp:842-960), 2-D. */
void orc_shift_image(const double *in, int s, double oridim, double xshift, double yshift, double *out);

/* Model of one E-step: K padded references + per-iteration spectra. */
typedef struct orc_model {
    const double *const *volumes;   /* [K] interleaved complex, projector layout */
    const double *pdf_class;        /* [K] */
    const double *pdf_direction;    /* [K][n_dir] */
    const double *sigma2_noise;     /* [ori_size/2+1] */
    const double *data_vs_prior;    /* [K][ori_size/2+1] or NULL */
} orc_model;

/* Optional per-particle intermediates (dense arrays in ihidden / ihidden_over order). */
typedef struct orc_debug {
    double *coarse_diff2;     /* [K*n_dir*n_psi*n_trans], -999 where not computed */
    double *coarse_weight;    /* same size */
    unsigned char *coarse_significant;
    double *fine_diff2;       /* [K*n_dir*n_psi*n_trans*32], -999 where not computed */
    double *fine_weight;
} orc_debug;

/* Full CPU E-step of n particles (expectationOneParticle, src/ml_optimiser.cpp:4376-4728), OpenMP
 * over particles. bp_data [K][z][y][x] interleaved complex and bp_weight [K][z][y][x] accumulate.
 * dbg may be NULL or an array of n orc_debug (entries with NULL pointers are skipped). */
int orc_estep(const rln_estep_config *cfg, const orc_model *model, int n,
              const rln_particle_meta *meta, const double *fimg, const double *fimg_nomask,
              const double *fctf, rln_particle_result *results, rln_wsum *wsum,
              double *bp_data, double *bp_weight, orc_debug *dbg, int nr_threads);

#ifdef __cplusplus
}
#endif
#endif
