/*
 * rln_refine.h -- the relion_refine command line for the E-step's drop-in.
 *
 * relion_refine's options (MlOptimiser::parseInitial, src/ml_optimiser.cpp:613-1088) parsed with the reference's
 * names, defaults and unit conventions, and turned into the E-step's configuration (include/rln_estep.h) with the
 * reference's size rules (MlOptimiser::updateCurrentResolution and updateImageSizeAndResolutionPointers,
 * src/ml_optimiser.cpp:5579-5780). A caller that drove relion_refine hands the same argument vector here.
 *
 * Options are handled in three ways:
 *   - those that set the E-step or its inputs are parsed into rln_refine_args;
 *   - those that only concern the run around it (threads, MPI and disc I/O, scratch, verbosity, GPU memory) are
 *     accepted and ignored, as they change no result;
 *   - those selecting branches this E-step does not implement (helical, subtomogram, multi-body, SGD, SOM,
 *     cross-correlation target, symmetry other than C1, nearest-neighbour interpolation, oversampling other than 1,
 *     tilt-limited grids, Fourier masks, relaxed or local symmetry, skip_align / skip_rotate, bimodal psi, CTF
 *     variants (intact first peak, phase flipping), a low-resolution limit of the probabilities, continuation and
 *     models or samplings read from STAR files) fail with RLN_ERR_ARG and a message naming the option. An option relion_refine does not know fails as its own parser fails
 *     (IOParser::checkForUnknownArguments, src/args.cpp:333-360).
 * --perturb (default 0.5) is parsed; the E-step samples the unperturbed grid, so rln_refine_args.perturbation_factor
 * is reported for the caller to check (0 reproduces the reference exactly; the reference's own perturbation is a
 * random draw per iteration, src/healpix_sampling.cpp:166-174).
 */
#ifndef RLN_REFINE_H
#define RLN_REFINE_H

#include "rln_estep.h"

#ifdef __cplusplus
extern "C" {
#endif

#define RLN_REFINE_PATH 1024

/* The options of one relion_refine run that the E-step and its inputs depend on (names of MlOptimiser /
 * MlModel / HealpixSampling members; defaults of parseInitial). */
typedef struct rln_refine_args {
    /* General */
    char   fn_data[RLN_REFINE_PATH];   /* --i: particles STAR file */
    char   fn_ref[RLN_REFINE_PATH];    /* --ref: reference map(s) ("None") */
    char   fn_out[RLN_REFINE_PATH];    /* --o: output rootname */
    char   fn_mask[RLN_REFINE_PATH];   /* --solvent_mask ("None") */
    int    nr_iter;                    /* --iter (-1) */
    int    nr_classes;                 /* --K (1) */
    double tau2_fudge;                 /* --tau2_fudge (1 when not given) */
    double particle_diameter;          /* --particle_diameter, Angstrom (-1: (ori_size - maskedge) * pixel size) */
    int    do_zero_mask;               /* --zero_mask */
    int    do_solvent;                 /* --flatten_solvent */
    int    do_split_random_halves;     /* --split_random_halves */
    /* Initialisation */
    double sigma_offset;               /* --offset, Angstrom (10); sigma2_offset = sigma_offset^2 */
    double ini_high;                   /* --ini_high, Angstrom (-1) */
    /* Orientations */
    int    adaptive_oversampling;      /* --oversampling (1) */
    int    healpix_order;              /* --healpix_order (2) */
    double psi_step;                   /* --psi_step, degrees (-1: from the HEALPix order) */
    double limit_tilt;                 /* --limit_tilt (-91: no limit) */
    char   sym[32];                    /* --sym ("c1") */
    double offset_range_px;            /* --offset_range, pixels (6) */
    double offset_step_px;             /* --offset_step, pixels (2) */
    double perturbation_factor;        /* --perturb (0.5) */
    int    do_auto_refine;             /* --auto_refine */
    int    auto_local_healpix_order;   /* --auto_local_healpix_order (4) */
    double sigma_rot, sigma_tilt, sigma_psi; /* --sigma_ang / --sigma_rot / --sigma_tilt / --sigma_psi, degrees (0) */
    int    local_search;               /* 1 when any of those is > 0 (PRIOR_ROTTILT_PSI), else NOPRIOR */
    /* Corrections */
    int    do_ctf_correction;          /* --ctf */
    int    refs_are_ctf_corrected;     /* not --ctf_uncorrected_ref */
    int    do_norm_correction;         /* --norm, unless --no_norm */
    int    do_scale_correction;        /* --scale, unless --no_scale */
    /* Computation and expert options */
    int    nr_pool;                    /* --pool (1) */
    int    nr_threads;                 /* --j (1) */
    int    do_gpu;                     /* --gpu */
    char   gpu_ids[256];               /* --gpu's device list ("default") */
    double padding_factor;             /* --pad (2) */
    int    maximum_significants;       /* --maxsig (-1) */
    double adaptive_fraction;          /* --adaptive_fraction (0.999) */
    int    max_coarse_size;            /* --coarse_size (-1) */
    int    width_mask_edge;            /* --maskedge (5) */
    int    incr_size;                  /* --incr_size (10) */
    double strict_highres_exp;         /* --strict_highres_exp, Angstrom (-1) */
    double strict_lowres_exp;          /* --strict_lowres_exp, Angstrom (-1) */
    int    fix_sigma_noise;            /* --fix_sigma_noise */
    int    fix_sigma_offset;           /* --fix_sigma_offset */
    int    skip_gridding;              /* not --dont_skip_gridding */
    int    random_seed;                /* --random_seed (-1) */
    int    do_map;                     /* not --no_map (MAP estimation of the reference, 1) */
    int    minres_map;                 /* --minres_map (5) */
    double sigma2_fudge;               /* --sigma2_fudge (1) */
    int    do_use_all_data;            /* --use_all_data: current size = ori_size */
} rln_refine_args;

/* Parses relion_refine's argument vector (argv[0], the program name, is skipped) into *a, defaults first.
 * RLN_ERR_ARG, with rln_last_error() naming the option, for an option relion_refine does not know, a missing or
 * malformed value, or a branch this E-step does not implement. */
int rln_refine_parse_args(int argc, const char *const *argv, rln_refine_args *a);

/* Image sizes of one optics group for the next E-step (updateCurrentResolution, src/ml_optimiser.cpp:5579-5677, and
 * updateImageSizeAndResolutionPointers, :5684-5780): the model's ori_size and pixel_size (model_pixel_size), the
 * optics group's image size and pixel size, the iteration about to run (iter 1 is the first), the resolution shell
 * at which data_vs_prior falls below 1 (max over classes; ignored in the first iteration when --ini_high is given),
 * the previous iteration's ave_Pmax and whether its FSC was high at the limit. Outputs image_current_size and
 * image_coarse_size of the group. */
int rln_refine_image_sizes(const rln_refine_args *a, int model_ori_size, double model_pixel_size, int image_size,
                           double image_pixel_size, int iter, int maxres_shell, double ave_pmax,
                           int has_high_fsc_at_limit, int *current_size, int *coarse_size);

/* The E-step configuration of one optics group: the parsed options with the offsets converted from pixels to
 * Angstrom by the model's pixel size (src/ml_optimiser.cpp:2346-2355), the particle diameter defaulted from the mask
 * edge (:2369-2370), sigma2_offset = --offset^2, do_ctf_correction = --ctf and the references CTF-corrected, and the
 * sizes from rln_refine_image_sizes. precision selects the diff2 kernels (RLN_PREC_*). */
int rln_refine_estep_config(const rln_refine_args *a, int model_ori_size, double model_pixel_size,
                            double image_pixel_size, int current_size, int coarse_size, int precision,
                            rln_estep_config *cfg);

#ifdef __cplusplus
}
#endif

#endif /* RLN_REFINE_H */
