/*
 * rln_estep.h -- C ABI of the MI355X-native RELION expectation step.
 *
 * One shared library (librln_estep.so) exports three objects and their calls:
 *
 *   rln_projector      device copy of one class's padded Fourier reference
 *                      (replaces AccProjector, src/acc/acc_projector.h:17-104)
 *   rln_backprojector  device accumulators of one class's padded Fourier data
 *                      and weight volumes (replaces AccBackprojector,
 *                      src/acc/acc_backprojector.h:24-99)
 *   rln_estep          the per-device E-step engine: coarse pass, weights and
 *                      significance, fine pass, weighted sums and
 *                      backprojection for a batch of prepared particles
 *                      (replaces MlOptimiserAccGPU::doThreadExpectationSomeParticles,
 *                      src/acc/hip/hip_ml_optimiser.hip.cpp:250-298, which drives
 *                      accDoExpectationOneParticle, src/acc/acc_ml_optimiser_impl.h:3673-3962)
 *
 * Conventions (all follow the reference CPU path, src/ml_optimiser.cpp):
 *   - 2-D images are FFTW half-complex arrays of size s x (s/2+1), row i holds
 *     frequency y = i for i <= s/2 and y = i - s otherwise; complex values are
 *     interleaved (re, im).
 *   - A padded reference volume has dims zdim x ydim x xdim with
 *     zdim = ydim = pad_size, xdim = pad_size/2 + 1,
 *     pad_size = 2*(ROUND(padding_factor*r_max)+1)+1, r_max = current_size/2,
 *     logical x in [0, xdim), y,z in [-(pad_size/2), pad_size/2]
 *     (src/projector.cpp:59-89).
 *   - Offsets in rln_particle_meta are in pixels, as METADATA_XOFF/YOFF of
 *     exp_metadata (src/ml_optimiser.cpp:5880-5886, 10429).
 *
 * Every call returns RLN_OK (0) or a negative status; rln_last_error()
 * returns a thread-local description of the last failure.
 */
#ifndef RLN_ESTEP_H
#define RLN_ESTEP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    RLN_OK = 0,
    RLN_ERR_ARG = -1,            /* invalid argument or unsupported option */
    RLN_ERR_HIP = -2,            /* HIP runtime failure */
    RLN_ERR_STATE = -3,          /* call out of order (e.g. model not set) */
    RLN_ERR_FILTERED_ZERO = -4,  /* ERRFILTEREDZERO: no positive weight (src/error.h:260-280) */
    RLN_ERR_NO_SIGNIFICANT = -5, /* ERRNOSIGNIFS: no significant coarse sample */
    RLN_ERR_SUM_WEIGHT_ZERO = -6,/* ERRSUMWEIGHTZERO: zero or NaN sum of weights */
    RLN_ERR_MEMORY = -7          /* device allocation failed */
};

/* Precision of the squared-difference (diff2) kernels. */
enum {
    RLN_PREC_PLAIN_FP32 = 0,     /* direct sum of |CTF*s*P - X_t|^2 in FP32 (fallback path) */
    RLN_PREC_SPLIT_BF16 = 1      /* coarse cross term Re(P conj(CTF*s*X_t)) on MFMA, BF16 hi+lo, FP32 accumulation */
};

/* Options of one E-step run; names follow MlOptimiser / MlModel / HealpixSampling members. */
typedef struct rln_estep_config {
    /* MlModel geometry */
    int    ori_size;              /* mymodel.ori_size (square particle box, pixels) */
    int    nr_classes;            /* mymodel.nr_classes (K) */
    double pixel_size;            /* Angstrom per pixel (model and images) */
    double padding_factor;        /* mymodel.padding_factor */
    int    current_size;          /* image_current_size: fine pass / weighted sums; r_max = current_size/2 */
    int    coarse_size;           /* image_coarse_size: coarse pass (src/ml_optimiser.cpp:5763-5779) */
    /* HealpixSampling */
    int    healpix_order;         /* coarse directions = 12*4^order (C1) */
    int    adaptive_oversampling; /* fine pass oversampling order; only 1 is supported */
    double psi_step;              /* degrees; <= 0 selects 360/(6*2^order) (src/healpix_sampling.cpp:73-74) */
    double offset_range;          /* Angstrom */
    double offset_step;           /* Angstrom */
    /* orientational priors */
    int    local_search;          /* 0: NOPRIOR (pdf_direction); 1: Gaussian priors around each particle's angles */
    int    precision;             /* RLN_PREC_* */
    double sigma_rot, sigma_tilt, sigma_psi; /* degrees, local search (sqrt(mymodel.sigma2_rot) ...) */
    double sigma_cutoff;          /* 3 (src/healpix_sampling.h:236-242) */
    /* MlOptimiser */
    double adaptive_fraction;     /* --adaptive_fraction, 0.999 */
    long long maximum_significants; /* --maxsig; <= 0: no cap */
    double sigma2_fudge;          /* sigma2_fudge, 1 */
    double sigma2_offset;         /* mymodel.sigma2_offset, Angstrom^2 */
    int    do_ctf_correction;     /* do_ctf_correction && refs_are_ctf_corrected */
    int    do_scale_correction;   /* do_scale_correction */
} rln_estep_config;

/* Per-particle scalars that the reference reads from exp_metadata / mymodel. */
typedef struct rln_particle_meta {
    double old_offset_x, old_offset_y;     /* pixels: rounded offsets already applied to the image */
    double prior_offset_x, prior_offset_y; /* pixels: exp_prior */
    double prior_rot, prior_tilt, prior_psi; /* degrees: centre of local searches */
    double scale;                          /* mymodel.scale_correction[group_id] */
    double highres_Xi2;                    /* sum |X|^2 beyond current_size (src/ml_optimiser.cpp:6386-6404) */
} rln_particle_meta;

/* Per-particle outputs (exp_metadata columns and the E-step scalars). */
typedef struct rln_particle_result {
    double rot, tilt, psi;        /* METADATA_ROT/TILT/PSI of the max-weight pose */
    double xoff, yoff;            /* METADATA_XOFF/YOFF (pixels) */
    double pmax;                  /* METADATA_PMAX */
    double nr_significant;        /* METADATA_NR_SIGN (coarse pass) */
    double dLL;                   /* METADATA_DLL */
    double sum_weight;            /* exp_sum_weight after the fine pass */
    double min_diff2;             /* exp_min_diff2 after the fine pass */
    double significant_weight;    /* exp_significant_weight after the fine pass */
    double coarse_sum_weight;     /* exp_sum_weight after the coarse pass */
    double coarse_min_diff2;      /* exp_min_diff2 after the coarse pass */
    double coarse_significant_weight;
    double wsum_norm_correction;  /* exp_wsum_norm_correction (without the tail beyond current_size) */
    double signal_product;        /* exp_wsum_scale_correction_XA / scale */
    double reference_power;       /* exp_wsum_scale_correction_AA / scale^2 */
    long long best_ihidden_over;  /* fine-pass index of the max-weight pose (src/healpix_sampling.cpp:1711-1722) */
    int    iclass;                /* METADATA_CLASS (1-based) */
    int    nr_fine_evaluated;     /* fine-pass poses evaluated */
} rln_particle_result;

/* Weighted sums accumulated (+=) over all particles of a run (MlWsumModel members). */
typedef struct rln_wsum {
    double *sigma2_noise;         /* [ori_size/2+1] wsum_model.sigma2_noise (shells < current_size/2+1) */
    double *pdf_class;            /* [K] */
    double *pdf_direction;        /* [K][12*4^order] */
    double sigma2_offset;
    double sumw_group;
    double signal_product;        /* wsum_signal_product (summed over particles) */
    double reference_power;       /* wsum_reference_power */
    double LL;                    /* sum of dLL */
    double ave_Pmax;              /* sum of Pmax */
} rln_wsum;

/* ---------------------------------------------------------------------------------------------
 * Projector: AccProjector (src/acc/acc_projector.h:80-97, impl src/acc/acc_projector_impl.h:5-392)
 * ------------------------------------------------------------------------------------------- */
typedef struct rln_projector rln_projector;

/* AccProjector::AccProjector() -- binds to a HIP device. */
int rln_projector_create(int device, rln_projector **out);
/* AccProjector::setMdlDim(xdim, ydim, zdim, inity, initz, maxr, padding_factor) (acc_projector.h:80-86);
 * *changed = 1 when the dimensions changed (the reference returns false when unchanged). */
int rln_projector_set_mdl_dim(rln_projector *p, int xdim, int ydim, int zdim, int inity, int initz,
                              int maxr, double padding_factor, int *changed);
/* AccProjector::initMdl(XFLOAT *real, XFLOAT *imag) (acc_projector.h:88-95): host split float arrays. */
int rln_projector_init_mdl(rln_projector *p, const float *real, const float *imag);
/* AccProjector::initMdl(Complex *data) (acc_projector.h:88-95): host interleaved double complex. */
int rln_projector_init_mdl_complex(rln_projector *p, const double *data);
/* Projector::computeFourierTransformMap (src/projector.cpp:116-593) for a 3-D reference and 2-D data, trilinear
 * interpolation: from a real-space map vol[z][y][x] (ori_size^3 floats, origin at ori_size/2; host, or device when
 * on_device) to the projector's padded Fourier volume, without a host round trip. Dimensions the projector as the
 * reference does: padoridim = ROUND(padding_factor * ori_size) made even, pad = padoridim / ori_size,
 * r_max = min(current_size, ori_size) / 2 (current_size <= 0: ori_size), pad_size = 2 (ROUND(pad r_max) + 1) + 1.
 * do_gridding: divide the map by sinc^2(r / padoridim) first (Projector::griddingCorrect, :595-628). min_ires > 0:
 * zero the data within ROUND(min_ires * pad) after the spectrum. Optional host outputs: power_spectrum
 * [ori_size/2 + 1] (radial average of |data|^2 / 2 per shell ROUND(r / pad)); data_out: the plain array as
 * interleaved double complex [pad_size][pad_size][pad_size/2 + 1] (PPref.data: z and y from -(pad_size/2), x from 0).
 * Work buffers and FFT plans are kept per device for the next call (with the reconstruction's; about 2 GB at
 * box 256) until the process ends; calls on one device are serialised. */
int rln_projector_compute_fourier_map(rln_projector *p, const float *vol, int on_device, int ori_size,
                                      int current_size, double padding_factor, int do_gridding, int min_ires,
                                      double *power_spectrum, double *data_out);
/* AccProjector::clear() (acc_projector.h:97) + destructor. */
int rln_projector_destroy(rln_projector *p);

/* ---------------------------------------------------------------------------------------------
 * Backprojector: AccBackprojector (src/acc/acc_backprojector.h:60-94, impl acc_backprojector_impl.h:13-190)
 * ------------------------------------------------------------------------------------------- */
typedef struct rln_backprojector rln_backprojector;

int rln_backprojector_create(int device, rln_backprojector **out);
/* AccBackprojector::setMdlDim(...) (acc_backprojector.h:60-66): *bytes = device bytes held. */
int rln_backprojector_set_mdl_dim(rln_backprojector *b, int xdim, int ydim, int zdim, int inity, int initz,
                                  int maxr, double padding_factor, size_t *bytes);
/* AccBackprojector::initMdl() (acc_backprojector.h:68): zero data and weight. */
int rln_backprojector_init_mdl(rln_backprojector *b);
/* AccBackprojector::getMdlData(real, imag, weight) (acc_backprojector_impl.h:109-138):
 * synchronise and copy to host float arrays of xdim*ydim*zdim elements. */
int rln_backprojector_get_mdl_data(rln_backprojector *b, float *real, float *imag, float *weight);
/* No counterpart in the reference: rln_backprojector_get_mdl_data without any device allocation, for callers that
 * cannot allocate (code under graph capture, serving loops, harnesses that forbid hipMalloc). Same outputs, bit for
 * bit: the slots are folded a few brick layers (4 planes each) at a time into a static 32 MiB staging array per
 * device and copied out layer by layer. Limit: 4 planes of re, im and weight must fit the staging, i.e.
 * xdim * ydim <= 699,050 (box up to about 580 at padding 2); larger volumes return RLN_ERR_ARG. All calls on one
 * device share its staging array: calls that overlap on one device race. */
int rln_backprojector_get_mdl_data_noalloc(rln_backprojector *b, float *real, float *imag, float *weight);
/* BackProjector::reconstruct (src/backprojector.cpp:1379-1780) for a 3-D reference, 2-D data and trilinear
 * interpolation, the default path (window to the original size): from the data and weight accumulated in b, on the
 * device. do_map: the weight gains 1 / (pad^3 * tau2_fudge * tau2[ires]) (1 / (0.001 * weight) where tau2 is 0) for
 * shells ires = ROUND(r / pad) >= minres_map. skip_gridding (relion_refine's default) or max_iter_preweight <= 0:
 * data / max(weight, 1/1000 of the weight's average over its shell FLOOR(r / pad)). Otherwise data and weight are
 * divided by normalise (the class's pdf weight sum) and max_iter_preweight gridding iterations with the tabulated
 * Kaiser-Bessel blob (radius 3.8, alpha 15, order 0) give the weights;
 * then the inverse transform, the real-space window, the soft mask (radius ori_size/2, width 3, average background)
 * and the sinc^2 gridding correction. tau2: n_tau2 shells (ori_size/2 + 1). vol_out: ori_size^3 floats [z][y][x],
 * origin at ori_size/2, host or device (out_on_device). padding_factor * ori_size must be an even integer. The
 * accumulators are left unchanged. Work buffers and FFT plans are kept per device for the next call (about 3.5 GB
 * at box 256, FP64) until the process ends; calls on one device are serialised. */
int rln_backprojector_reconstruct(rln_backprojector *b, int ori_size, int skip_gridding, int max_iter_preweight,
                                  int do_map, const double *tau2, int n_tau2, double tau2_fudge, double normalise, int minres_map,
                                  float *vol_out, int out_on_device);
int rln_backprojector_destroy(rln_backprojector *b);

/* ---------------------------------------------------------------------------------------------
 * E-step engine (MlDeviceBundle + MlOptimiserAccGPU, src/acc/hip/hip_ml_optimiser.hip.cpp:70-298)
 * ------------------------------------------------------------------------------------------- */
typedef struct rln_estep rln_estep;

/* MlDeviceBundle::setupFixedSizedObjects/setupTunableSizedObjects (hip_ml_optimiser.hip.cpp:88-226). */
int rln_estep_create(int device, const rln_estep_config *cfg, rln_estep **out);
int rln_estep_destroy(rln_estep *e);

/* Binds the per-iteration model: one projector and one backprojector per class (borrowed, not owned),
 * pdf_class[K], pdf_direction[K][12*4^order] (NOPRIOR mode), sigma2_noise[ori_size/2+1] and, if
 * do_scale_correction, data_vs_prior[K][ori_size/2+1]. Host arrays are copied. */
int rln_estep_set_model(rln_estep *e, rln_projector *const *projectors, rln_backprojector *const *backprojectors,
                        const double *pdf_class, const double *pdf_direction, const double *sigma2_noise,
                        const double *data_vs_prior);

/* Runs the E-step of n particles (accDoExpectationOneParticle for each, batched):
 *   fimg, fimg_nomask: [n][current_size][current_size/2+1] interleaved complex float (masked / unmasked image)
 *   fctf:              [n][current_size][current_size/2+1] float
 *   inputs_on_device:  1 if the three image arrays are device pointers on the engine's device, 0 if host
 *   meta:              [n] host
 * Outputs: results[n] (host), wsum (host arrays, accumulated), and backprojection into the bound
 * backprojectors (device, accumulated). */
int rln_estep_run(rln_estep *e, int n, const float *fimg, const float *fimg_nomask, const float *fctf,
                  int inputs_on_device, const rln_particle_meta *meta,
                  rln_particle_result *results, rln_wsum *wsum);

/* Particles per chunk at most (default 16,384; the first chunk of a run a quarter of it and at most 1,024, so that the
 * device starts early; fewer when a chunk's coarse intermediates would exceed the device budget): the batch the
 * E-step processes at once, as the reference's pool of particles per task (--pool,
 * src/ml_optimiser.cpp:466, 2482-2483, 3754). Results do not depend on it, up to the order of floating-point sums
 * over particles. Ignored while intermediates are kept (one chunk). RLN_ERR_ARG below 1. */
int rln_estep_set_pool(rln_estep *e, int max_particles);

/* The fine pass (job lists, fine diff2, fine weights, weighted sums and backprojection) runs from device-side
 * counts as persistent kernels captured into HIP graphs and replayed for every chunk (default, enable = 1), or
 * launched directly (enable = 0; same kernels and results). No counterpart in the reference, whose fine pass
 * is planned on the host per particle (src/acc/acc_ml_optimiser_impl.h:1420-1890). */
int rln_estep_use_graphs(rln_estep *e, int enable);

/* Work-unit shape of the fine diff2 kernel. The kernel splits every work item's sampled pixels into n_bands radial
 * shells of equal pixel count and runs all items' band 0, then band 1, ..., so that the gathers in flight touch one
 * shell of the reference at a time. By default (n_bands = 0) the engine times the band counts 1, 2, 4, 8, 16, 32
 * and the persistent grid sizes on its first chunk with enough work and keeps the fastest; n_bands in that set fixes
 * the band count (grid: the resident-block count) and turns the tuning off. Results do not depend on it, up to the
 * order of the FP64 sums of the bands. No counterpart in the reference, whose launch geometry is fixed
 * (src/acc/cuda/cuda_settings.h:86-141). RLN_ERR_ARG for other values. */
int rln_estep_set_fine_bands(rln_estep *e, int n_bands);
/* Order of the fine pass's work units: 0 jobs in particle order, units band-major (all jobs' band 0, then band 1,
 * ...); 1 jobs sorted by class and coarse orientation (direction-major), band-major; 2 sorted, item-major (the bands
 * of a job one after the other); 3 sorted, band-major within per-XCD queues (XCD x first takes the x-th eighth of the
 * sorted jobs, so that each XCD's L2 holds the planes of fewer directions, then helps the others); -1 (the default)
 * lets the tuning choose. Jobs of one coarse direction sample the
 * same four fine-direction planes of the reference, so in orientation order the units in flight share its corner
 * blocks in the caches. wsum_order sets the weighted sums' unit order likewise (0 band-major, 1 item-major,
 * 2 band-major in per-XCD queues; they follow the fine pass's job order). Results do not depend on it, up to the order of FP64 and FP32 atomic sums.
 * No counterpart in the reference (one particle at a time, src/acc/acc_helper_functions_impl.h:26-102).
 * RLN_ERR_ARG for other values. */
int rln_estep_set_fine_order(rln_estep *e, int order, int wsum_order);
/* Radial pixel bands per work unit of the weighted sums and backprojection (1, 2, 4, 8, 16, 32), or 0 (the default)
 * for the fine pass's tuned count. The weighted sums accumulate into the backprojectors, so they are not timed by the
 * tuning. Results do not depend on it, up to the order of FP32 atomic sums. RLN_ERR_ARG for other values. */
int rln_estep_set_wsum_bands(rln_estep *e, int n_bands);
/* Work-item shape of the fine pass (index into the shapes below), or -1 (the default) for the tuning's choice; the
 * tuning then times only the other parameters. Results do not depend on it, up to the order of FP32 partial sums.
 * RLN_ERR_ARG for other values. */
int rln_estep_set_fine_shape(rln_estep *e, int shape);
/* The fine pass's tuning: the time in ms of every candidate tried ([shape][order][band choice][grid], 4 x 4 x 6 x 5
 * in that order; shapes (TC, CW) = (2, 2), (1, 2), (2, 1), (1, 1); orders as rln_estep_set_fine_order; band counts
 * 1 .. 32; grids 512, 1024, 2048, 4096 and the resident-block count), -1 for those not tried; *n_out = 0 before any
 * tuning. Copies min(n, capacity). */
int rln_estep_fine_tuning(rln_estep *e, double *ms, int capacity, int *n_out);

/* Explicit coarse orientation lists for the next rln_estep_run (global searches only), as produced by the
 * learned pose proposal: particle i is searched over orientations set_orients[set_offsets[i] ..
 * set_offsets[i+1]), each a global grid index idir * N_psi + ipsi of the coarse HEALPix grid, for all classes
 * and translations; an empty list means the exhaustive grid (the fallback). Orientation priors are the global
 * ones (pdf_direction normalised over the whole grid), so weights are those of the exhaustive search
 * restricted to the list. No counterpart in the reference: the reference's coarse pass always enumerates the
 * whole grid or the local-prior window (AccProjectorPlan::setup, src/acc/acc_projector_plan_impl.h:99-410).
 * The lists apply to one run and are then cleared; their count must equal that run's n. */
int rln_estep_set_orientation_sets(rln_estep *e, int n, const int *set_offsets, const int *set_orients);

/* Stage hooks for parity tests. After rln_estep_run with keep_intermediates set, these copy the
 * last run's per-particle intermediates to host (particle index ip within that run):
 *   coarse: diff2 per coarse pose in ihidden order (K * N_dir_b * N_psi_b * N_tr), and weights;
 *   fine:   n_fine (ihidden_over, diff2, weight) triplets of the evaluated fine poses (diff2 in FP64: the
 *           pose-independent part of each particle's diff2 is summed in FP64, the pose-dependent part in FP32). */
int rln_estep_keep_intermediates(rln_estep *e, int keep);
int rln_estep_coarse_size(rln_estep *e, int ip, long long *n_coarse, int *n_dir, int *n_psi, int *n_trans);
int rln_estep_get_coarse(rln_estep *e, int ip, float *diff2, double *weight, unsigned char *significant);
int rln_estep_fine_size(rln_estep *e, int ip, long long *n_fine);
int rln_estep_get_fine(rln_estep *e, int ip, long long *ihidden_over, double *diff2, double *weight);

/* Device time of the last run per stage, in milliseconds, from HIP events on the engine's stream:
 * prep, coarse_diff2, coarse_weights, jobs (fine-pass plan and work items), fine_diff2 (the fine diff2 kernel),
 * fine_weights, wsum_backproject; then host wall
 * time of chunk preparation (list building not overlapped with the device), of it the first chunk's lists (built
 * before any device work of the run), the epilogue, and the whole run. */
int rln_estep_last_timings(rln_estep *e, double *ms, int n_max, int *n_out);
/* Work counters of the last run: particles, coarse poses, coarse pixels per image, fine poses, fine pixels
 * per image, fine diff2 work items, backprojected fine orientations, significant (orientation, translation)
 * pairs summed in the weighted sums, coarse grid projections computed, coarse tiles, fine jobs, particle chunks
 * (one launch of each per-chunk kernel per chunk), launches of the captured fine-pass graph, fine-pass
 * capacity retries, the persistent grid sizes of the fine diff2 kernel (tuned) and of the weighted sums (resident
 * blocks), the fine-image pixels that are sampled per fine orientation (non-zero Minvsigma2 inside the
 * projection radius), the number of radial pixel bands the fine pass splits each work item into, its work-item
 * shape: coarse translations per small item and per wave of a wide item (all three tuned), the order of its units
 * (rln_estep_set_fine_order), the weighted sums' unit order and their band count. */
int rln_estep_last_counters(rln_estep *e, long long *out, int n_max, int *n_out);

/* ---------------------------------------------------------------------------------------------
 * Sampling helpers (HealpixSampling, src/healpix_sampling.cpp), host-side.
 * ------------------------------------------------------------------------------------------- */
/* image_coarse_size rule of src/ml_optimiser.cpp:5757-5779 (adaptive_oversampling > 0). */
int rln_coarse_image_size(int ori_size, double pixel_size, int healpix_order, double particle_diameter,
                          int current_size);
/* Number of directions (C1), psi angles and translations of the coarse grid. */
int rln_sampling_counts(const rln_estep_config *cfg, int *n_dir, int *n_psi, int *n_trans);
/* Coarse grid: rot/tilt per direction (degrees), psi per psi index, translations (Angstrom). */
int rln_sampling_grid(const rln_estep_config *cfg, double *rot, double *tilt, double *psi,
                      double *trans_x, double *trans_y);
/* selectOrientationsWithNonZeroPriorProbability (src/healpix_sampling.cpp:695-1076) for one particle:
 * fills up to max_dir / max_psi entries; returns counts. */
int rln_sampling_local(const rln_estep_config *cfg, double prior_rot, double prior_tilt, double prior_psi,
                       int *dir_index, double *dir_prior, int max_dir, int *n_dir,
                       int *psi_index, double *psi_prior, int max_psi, int *n_psi);

/* HEALPix NEST pixel (at order) containing the direction (rot, tilt) in degrees: ang2pix of Gorski et al.
 * 2005 as Healpix_Base::ang2pix_z_phi (src/Healpix_2.15a/healpix_base.cc:368); the inverse of the grid's
 * directions (HealpixSampling::getDirectionFromHealPix, src/healpix_sampling.cpp:1566-1660). */
int rln_sampling_ang2pix(int order, double rot, double tilt, long long *pix);

/* ---------------------------------------------------------------------------------------------
 * Synthetic data (seeded Gaussian-blob maps and particles; the reference's simulator is
 * src/apps/project.cpp:66-106). Host-side, OpenMP.
 * ------------------------------------------------------------------------------------------- */
typedef struct rln_synth_spec {
    unsigned long long seed;
    int    ori_size;
    double pixel_size;
    int    nr_blobs;
    double blob_radius;           /* centres uniform within this radius (pixels) */
    double sigma_min, sigma_max;  /* blob sigma range (pixels) */
    double amp_min, amp_max;
    double snr_min, snr_max;      /* per-particle SNR, log-uniform in [min, max] (min == max: fixed) */
    double defocus_min, defocus_max; /* Angstrom */
    double max_astigmatism;       /* Angstrom */
    double voltage, cs, q0, bfactor;
    double offset_max;            /* true offsets uniform in [-max, max] pixels */
    double prior_sigma_deg;       /* prior angles = true pose perturbed by this rotation sigma */
} rln_synth_spec;

/* Padded Fourier reference of the blob map for class `iclass` (blob set drawn from seed+iclass;
 * share_fraction of the blobs are drawn from a class-independent stream). Output: interleaved
 * double complex, zdim*ydim*xdim as in the projector layout for (current_size, padding_factor). */
int rln_synth_reference(const rln_synth_spec *spec, int iclass, double share_fraction, int current_size,
                        double padding_factor, double *out);
/* The same blob map sampled in real space: out[z][y][x] (ori_size^3 floats, origin at ori_size/2) =
 * sum of a exp(-|r - c|^2 / (2 s^2)) over the class's blobs. */
int rln_synth_real_map(const rln_synth_spec *spec, int iclass, double share_fraction, float *out);
/* Particles [first, first+n): images at current_size (interleaved float complex), CTFs, metadata,
 * true poses (rot, tilt, psi, xoff, yoff, class) and the noise variance per Fourier component. */
int rln_synth_particles(const rln_synth_spec *spec, int nr_classes, double share_fraction, int current_size,
                        long long first, int n, float *fimg, float *fctf, rln_particle_meta *meta,
                        double *true_pose, double *noise_var);
/* CTF value of the reference model (src/ctf.h:184-240) at frequency (x, y) in 1/Angstrom. */
double rln_ctf_value(double defU, double defV, double defAngle, double voltage, double cs, double q0,
                     double bfactor, double scale, double phase_shift, double x, double y);

/* Self-test of the MFMA lane layout the split-BF16 kernels rely on: D(16x16) = A(16x32) * B(32x16), row-major
 * float inputs rounded to BF16, one v_mfma_f32_16x16x32_bf16 on `device`. No counterpart in the reference. */
int rln_selftest_mfma_bf16(int device, const float *A, const float *B, float *D);

const char *rln_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* RLN_ESTEP_H */
