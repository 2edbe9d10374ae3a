// relion_refine's command line for the E-step's drop-in (include/rln_refine.h): the options of
// MlOptimiser::parseInitial (src/ml_optimiser.cpp:613-1088) with their defaults, and the size rules of
// updateCurrentResolution / updateImageSizeAndResolutionPointers (:5579-5780).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <string>

#include "../../include/rln_refine.h"
#include "rln_common.h"

namespace {

using Setter = std::function<int(rln_refine_args &, const std::string &)>;

struct Option {
    bool takes_value;
    Setter set;  // null: accepted and ignored
};

int copy_path(char *dst, const std::string &v, const char *name) {
    if (v.size() >= RLN_REFINE_PATH) RLN_FAIL(RLN_ERR_ARG, "%s: path longer than %d characters", name, RLN_REFINE_PATH - 1);
    std::memcpy(dst, v.c_str(), v.size() + 1);
    return RLN_OK;
}

// textToFloat / textToInteger (src/strings.cpp): the whole token must be a number.
int to_double(const std::string &v, const char *name, double &out) {
    char *end = nullptr;
    out = std::strtod(v.c_str(), &end);
    if (v.empty() || !end || *end != '\0') RLN_FAIL(RLN_ERR_ARG, "%s: '%s' is not a number", name, v.c_str());
    return RLN_OK;
}

int to_int(const std::string &v, const char *name, int &out) {
    double d = 0.;
    if (to_double(v, name, d) != RLN_OK) return RLN_ERR_ARG;
    out = (int)d;  // textToInteger reads through a float (src/strings.cpp:211-226)
    return RLN_OK;
}

Setter dbl(double rln_refine_args::*field, const char *name) {
    return [field, name](rln_refine_args &a, const std::string &v) -> int { return to_double(v, name, a.*field); };
}

Setter integer(int rln_refine_args::*field, const char *name) {
    return [field, name](rln_refine_args &a, const std::string &v) -> int { return to_int(v, name, a.*field); };
}

Setter path(char (rln_refine_args::*field)[RLN_REFINE_PATH], const char *name) {
    return [field, name](rln_refine_args &a, const std::string &v) -> int { return copy_path(a.*field, v, name); };
}

Setter flag(int rln_refine_args::*field, int value) {
    return [field, value](rln_refine_args &a, const std::string &) -> int {
        a.*field = value;
        return RLN_OK;
    };
}

// An option whose use selects a branch this E-step does not implement.
Setter refuse(const char *name, const char *what) {
    return [name, what](rln_refine_args &, const std::string &) -> int {
        RLN_FAIL(RLN_ERR_ARG, "relion_refine option %s (%s) is not supported by this E-step", name, what);
    };
}

// A value option that is refused unless it keeps the reference's default.
Setter refuse_unless(const char *name, const char *what, std::function<bool(const std::string &)> is_default) {
    return [name, what, is_default](rln_refine_args &, const std::string &v) -> int {
        if (is_default(v)) return (int)RLN_OK;
        RLN_FAIL(RLN_ERR_ARG, "relion_refine option %s %s (%s) is not supported by this E-step", name, v.c_str(), what);
    };
}

bool is_none(const std::string &v) { return v == "None" || v.empty(); }

const std::map<std::string, Option> &options() {
    static const std::map<std::string, Option> table = [] {
        std::map<std::string, Option> t;
        auto V = [&](const char *n, Setter s) { t[n] = Option{true, std::move(s)}; };
        auto F = [&](const char *n, Setter s) { t[n] = Option{false, std::move(s)}; };
        // General options
        V("--i", path(&rln_refine_args::fn_data, "--i"));
        V("--o", path(&rln_refine_args::fn_out, "--o"));
        V("--iter", integer(&rln_refine_args::nr_iter, "--iter"));
        V("--tau2_fudge", [](rln_refine_args &a, const std::string &v) -> int {
            double x = 0.;
            if (to_double(v, "--tau2_fudge", x) != RLN_OK) return (int)RLN_ERR_ARG;
            a.tau2_fudge = x > 0. ? x : 1.;  // src/ml_optimiser.cpp:642-643
            return (int)RLN_OK;
        });
        V("--tau2_fudge_scheme", nullptr);
        V("--K", integer(&rln_refine_args::nr_classes, "--K"));
        V("--particle_diameter", dbl(&rln_refine_args::particle_diameter, "--particle_diameter"));
        F("--zero_mask", flag(&rln_refine_args::do_zero_mask, 1));
        F("--flatten_solvent", flag(&rln_refine_args::do_solvent, 1));
        V("--solvent_mask", path(&rln_refine_args::fn_mask, "--solvent_mask"));
        V("--solvent_mask2", refuse_unless("--solvent_mask2", "secondary solvent mask", is_none));
        V("--lowpass_mask", nullptr);
        V("--lowpass", nullptr);
        V("--tau", nullptr);
        V("--local_symmetry", refuse_unless("--local_symmetry", "local symmetry", is_none));
        F("--split_random_halves", flag(&rln_refine_args::do_split_random_halves, 1));
        V("--low_resol_join_halves", nullptr);
        F("--center_classes", nullptr);
        V("--tomograms", refuse_unless("--tomograms", "subtomogram averaging", [](const std::string &v) { return v.empty(); }));
        V("--trajectories", refuse_unless("--trajectories", "subtomogram averaging", [](const std::string &v) { return v.empty(); }));
        V("--ios", refuse_unless("--ios", "tomography optimisation sets", [](const std::string &v) { return v.empty(); }));
        // Initialisation
        V("--ref", path(&rln_refine_args::fn_ref, "--ref"));
        F("--denovo_3dref", refuse("--denovo_3dref", "de novo 3-D reference from 2-D classes"));
        V("--offset", dbl(&rln_refine_args::sigma_offset, "--offset"));
        F("--firstiter_cc", refuse("--firstiter_cc", "cross-correlation target"));
        V("--ini_high", dbl(&rln_refine_args::ini_high, "--ini_high"));
        V("--model", refuse_unless("--model", "model read from a STAR file", is_none));
        V("--sampling", refuse_unless("--sampling", "sampling read from a STAR file", is_none));
        // Orientations
        V("--oversampling", [](rln_refine_args &a, const std::string &v) -> int {
            if (to_int(v, "--oversampling", a.adaptive_oversampling) != RLN_OK) return (int)RLN_ERR_ARG;
            if (a.adaptive_oversampling != 1)
                RLN_FAIL(RLN_ERR_ARG, "relion_refine option --oversampling %s (adaptive oversampling order other than 1) "
                         "is not supported by this E-step", v.c_str());
            return (int)RLN_OK;
        });
        V("--healpix_order", integer(&rln_refine_args::healpix_order, "--healpix_order"));
        V("--psi_step", dbl(&rln_refine_args::psi_step, "--psi_step"));
        V("--limit_tilt", [](rln_refine_args &a, const std::string &v) -> int {
            if (to_double(v, "--limit_tilt", a.limit_tilt) != RLN_OK) return (int)RLN_ERR_ARG;
            // |limit_tilt| < 90 removes directions from the grid (src/healpix_sampling.cpp:1400-1460)
            if (std::fabs(a.limit_tilt) < 90.)
                RLN_FAIL(RLN_ERR_ARG, "relion_refine option --limit_tilt %s (tilt-limited grid) is not supported by this "
                         "E-step", v.c_str());
            return (int)RLN_OK;
        });
        V("--sym", [](rln_refine_args &a, const std::string &v) -> int {
            if (v != "c1" && v != "C1")
                RLN_FAIL(RLN_ERR_ARG, "relion_refine option --sym %s (point-group symmetry other than C1) is not supported "
                         "by this E-step", v.c_str());
            std::snprintf(a.sym, sizeof(a.sym), "%s", v.c_str());
            return (int)RLN_OK;
        });
        V("--relax_sym", refuse_unless("--relax_sym", "symmetry relaxation", [](const std::string &v) { return v.empty(); }));
        V("--offset_range", dbl(&rln_refine_args::offset_range_px, "--offset_range"));
        V("--offset_step", dbl(&rln_refine_args::offset_step_px, "--offset_step"));
        for (const char *n : {"--offset_range_x", "--offset_range_y", "--offset_range_z"})
            V(n, refuse_unless(n, "per-axis offset ranges of subtomograms", [](const std::string &v) {
                  return std::strtod(v.c_str(), nullptr) <= 0.;
              }));
        V("--helical_offset_step", nullptr);
        V("--perturb", [](rln_refine_args &a, const std::string &v) -> int {
            if (to_double(v, "--perturb", a.perturbation_factor) != RLN_OK) return (int)RLN_ERR_ARG;
            if (a.perturbation_factor < 0. || a.perturbation_factor > 1.)  // src/healpix_sampling.cpp:80-81
                RLN_FAIL(RLN_ERR_ARG, "--perturb: random perturbation factor should be between 0 and 1");
            return (int)RLN_OK;
        });
        F("--auto_refine", flag(&rln_refine_args::do_auto_refine, 1));
        F("--auto_sampling", nullptr);
        V("--auto_local_healpix_order", integer(&rln_refine_args::auto_local_healpix_order, "--auto_local_healpix_order"));
        V("--sigma_ang", dbl(&rln_refine_args::sigma_rot, "--sigma_ang"));  // resolved after parsing
        V("--sigma_rot", nullptr);
        V("--sigma_tilt", nullptr);
        V("--sigma_psi", nullptr);
        F("--skip_align", refuse("--skip_align", "classification without alignment"));
        F("--skip_rotate", refuse("--skip_rotate", "translation-only alignment"));
        F("--bimodal_psi", refuse("--bimodal_psi", "bimodal psi priors"));
        F("--only_sample_tilt", refuse("--only_sample_tilt", "tilt-only sampling"));
        V("--minimum_angular_sampling", nullptr);
        V("--maximum_angular_sampling", nullptr);
        // Helical reconstruction: refused when switched on; its parameters alone change nothing
        F("--helix", refuse("--helix", "helical refinement"));
        F("--ignore_helical_symmetry", nullptr);
        for (const char *n : {"--helical_nr_asu", "--helical_twist_initial", "--helical_twist_min", "--helical_twist_max",
                              "--helical_twist_inistep", "--helical_rise_initial", "--helical_rise_min",
                              "--helical_rise_max", "--helical_rise_inistep", "--helical_nstart",
                              "--helical_z_percentage", "--helical_inner_diameter", "--helical_outer_diameter",
                              "--helical_sigma_distance", "--helical_exclude_resols"})
            V(n, nullptr);
        F("--helical_symmetry_search", nullptr);
        F("--helical_keep_tilt_prior_fixed", nullptr);
        F("--skip_realspace_helical_sym", nullptr);
        V("--fourier_mask", refuse_unless("--fourier_mask", "Fourier mask of the projector", is_none));
        // Corrections
        F("--ctf", flag(&rln_refine_args::do_ctf_correction, 1));
        F("--pad_ctf", refuse("--pad_ctf", "CTF padding, disabled in the reference too"));
        F("--ctf_intact_first_peak", refuse("--ctf_intact_first_peak", "CTFs ignored until their first peak"));
        F("--ctf_uncorrected_ref", flag(&rln_refine_args::refs_are_ctf_corrected, 0));
        F("--ctf_corrected_ref", nullptr);  // removed option, warned about by the reference
        F("--ctf_phase_flipped", refuse("--ctf_phase_flipped", "phase-flipped data"));
        F("--only_flip_phases", refuse("--only_flip_phases", "phase-flip-only CTF correction"));
        F("--norm", flag(&rln_refine_args::do_norm_correction, 1));
        F("--scale", flag(&rln_refine_args::do_scale_correction, 1));
        F("--no_norm", nullptr);   // resolved after parsing (switches --norm off whatever the order)
        F("--no_scale", nullptr);
        // Stochastic gradient descent, SOM and subtomograms
        F("--grad", refuse("--grad", "gradient-based optimisation"));
        for (const char *n : {"--grad_em_iters", "--grad_ini_frac", "--grad_fin_frac", "--grad_min_resol",
                              "--grad_ini_resol", "--grad_fin_resol", "--grad_ini_subset", "--grad_fin_subset", "--mu",
                              "--grad_stepsize", "--grad_stepsize_scheme", "--grad_write_iter",
                              "--class_inactivity_threshold", "--som_ini_nodes", "--som_connectivity",
                              "--som_inactivity_threshold", "--som_neighbour_pull", "--subtomo_multi_thr"})
            V(n, nullptr);
        F("--som", refuse("--som", "self-organising maps"));
        F("--no_init_blobs", nullptr);
        F("--normalised_subtomo", nullptr);
        F("--skip_subtomo_multi", nullptr);
        F("--ctf3d_not_squared", nullptr);
        // Computation
        V("--pool", integer(&rln_refine_args::nr_pool, "--pool"));
        V("--j", integer(&rln_refine_args::nr_threads, "--j"));
        for (const char *n : {"--dont_combine_weights_via_disc", "--onthefly_shifts", "--no_parallel_disc_io",
                              "--preread_images", "--reuse_scratch", "--keep_scratch", "--fast_subsets", "--cpu"})
            F(n, nullptr);
        V("--scratch_dir", nullptr);
        V("--keep_free_scratch", nullptr);
        for (const char *n : {"--sycl", "--sycl-levelzero", "--sycl-cuda", "--sycl-hip", "--sycl-opencl", "--sycl-cpu"})
            F(n, refuse(n, "SYCL devices"));
        V("--free_gpu_memory", nullptr);
        // Expert options
        V("--pad", dbl(&rln_refine_args::padding_factor, "--pad"));
        V("--ref_angpix", nullptr);
        F("--NN", refuse("--NN", "nearest-neighbour interpolation"));
        V("--r_min_nn", nullptr);
        V("--verb", nullptr);
        V("--random_seed", integer(&rln_refine_args::random_seed, "--random_seed"));
        V("--coarse_size", integer(&rln_refine_args::max_coarse_size, "--coarse_size"));
        V("--adaptive_fraction", dbl(&rln_refine_args::adaptive_fraction, "--adaptive_fraction"));
        V("--maxsig", integer(&rln_refine_args::maximum_significants, "--maxsig"));
        V("--maskedge", integer(&rln_refine_args::width_mask_edge, "--maskedge"));
        F("--fix_sigma_noise", flag(&rln_refine_args::fix_sigma_noise, 1));
        F("--fix_sigma_offset", flag(&rln_refine_args::fix_sigma_offset, 1));
        V("--incr_size", integer(&rln_refine_args::incr_size, "--incr_size"));
        F("--print_metadata_labels", nullptr);
        F("--print_symmetry_ops", nullptr);
        V("--strict_highres_exp", dbl(&rln_refine_args::strict_highres_exp, "--strict_highres_exp"));
        V("--strict_lowres_exp", [](rln_refine_args &a, const std::string &v) -> int {
            if (to_double(v, "--strict_lowres_exp", a.strict_lowres_exp) != RLN_OK) return (int)RLN_ERR_ARG;
            if (a.strict_lowres_exp > 0.)
                RLN_FAIL(RLN_ERR_ARG, "relion_refine option --strict_lowres_exp %s (low-resolution limit of the "
                         "probabilities) is not supported by this E-step", v.c_str());
            return (int)RLN_OK;
        });
        F("--dont_check_norm", nullptr);
        F("--always_cc", refuse("--always_cc", "cross-correlation target"));
        F("--solvent_correct_fsc", nullptr);
        F("--skip_maximize", nullptr);
        V("--failsafe_threshold", nullptr);
        F("--blush", nullptr);
        F("--blush_skip_spectral_trailing", nullptr);
        F("--external_reconstruct", nullptr);
        V("--auto_iter_max", nullptr);
        F("--auto_ignore_angles", nullptr);
        F("--auto_resol_angles", nullptr);
        F("--allow_coarser_sampling", nullptr);
        F("--trust_ref_size", nullptr);
        V("--nr_parts_sigma2noise", nullptr);
        // hidden options (src/ml_optimiser.cpp:1057-1081)
        F("--no_map", flag(&rln_refine_args::do_map, 0));
        V("--minres_map", integer(&rln_refine_args::minres_map, "--minres_map"));
        V("--abort_at_resolution", nullptr);
        F("--bfactor", nullptr);
        V("--gridding_iter", nullptr);
        for (const char *n : {"--debug1", "--debug2", "--debug3", "--debug_split_random_half", "--sigma",
                              "--min_sigma2_offset"})
            V(n, nullptr);
        V("--sigma2_fudge", dbl(&rln_refine_args::sigma2_fudge, "--sigma2_fudge"));
        F("--accuracy_current_size", nullptr);
        F("--sequential_halves_recons", nullptr);
        F("--always_join_random_halves", nullptr);
        F("--use_all_data", flag(&rln_refine_args::do_use_all_data, 1));
        F("--asymmetric_padding", nullptr);
        F("--dont_skip_gridding", flag(&rln_refine_args::skip_gridding, 0));
        F("--continue", refuse("--continue", "continuation from an _optimiser.star file"));
        V("--multibody_masks", refuse("--multibody_masks", "multi-body refinement"));
        return t;
    }();
    return table;
}

void defaults(rln_refine_args &a) {
    std::memset(&a, 0, sizeof(a));
    std::snprintf(a.fn_ref, sizeof(a.fn_ref), "None");
    std::snprintf(a.fn_mask, sizeof(a.fn_mask), "None");
    std::snprintf(a.sym, sizeof(a.sym), "c1");
    std::snprintf(a.gpu_ids, sizeof(a.gpu_ids), "default");
    a.nr_iter = -1;
    a.nr_classes = 1;
    a.tau2_fudge = 1.;
    a.particle_diameter = -1.;
    a.sigma_offset = 10.;
    a.ini_high = -1.;
    a.adaptive_oversampling = 1;
    a.healpix_order = 2;
    a.psi_step = -1.;
    a.limit_tilt = -91.;
    a.offset_range_px = 6.;
    a.offset_step_px = 2.;
    a.perturbation_factor = 0.5;
    a.auto_local_healpix_order = 4;
    a.refs_are_ctf_corrected = 1;
    a.nr_pool = 1;
    a.nr_threads = 1;
    a.padding_factor = 2.;
    a.maximum_significants = -1;
    a.adaptive_fraction = 0.999;
    a.max_coarse_size = -1;
    a.width_mask_edge = 5;
    a.incr_size = 10;
    a.strict_highres_exp = -1.;
    a.strict_lowres_exp = -1.;
    a.skip_gridding = 1;
    a.random_seed = -1;
    a.do_map = 1;
    a.minres_map = 5;
    a.sigma2_fudge = 1.;
}

// A token that is a (negative) number, which relion_refine accepts as an option's value (src/args.cpp:346-357).
bool is_number(const char *s) {
    char *end = nullptr;
    std::strtod(s, &end);
    return end != s && *end == '\0';
}

}  // namespace

extern "C" int rln_refine_parse_args(int argc, const char *const *argv, rln_refine_args *a) {
    if (!a || argc < 0 || (argc > 0 && !argv)) RLN_FAIL(RLN_ERR_ARG, "rln_refine_parse_args: null argument");
    defaults(*a);
    const auto &table = options();
    double sigma_ang = -1., sigma_rot = -1., sigma_tilt = -1., sigma_psi = -1.;
    bool no_norm = false, no_scale = false;
    for (int i = 1; i < argc; ++i) {
        const std::string opt = argv[i] ? argv[i] : "";
        if (opt.compare(0, 2, "--") != 0) {
            // IOParser::checkForUnknownArguments (src/args.cpp:333-360): a token with one dash must be a number
            // following a known option; tokens without a dash are not checked
            if (opt.compare(0, 1, "-") == 0 &&
                !(is_number(opt.c_str()) && i > 1 && table.count(argv[i - 1] ? argv[i - 1] : "")))
                RLN_FAIL(RLN_ERR_ARG, "unrecognised argument '%s'", opt.c_str());
            continue;
        }
        if (opt.compare(0, 18, "--pipeline_control") == 0) {  // accepted by IOParser::checkForUnknownArguments
            if (i + 1 < argc && std::strncmp(argv[i + 1], "--", 2) != 0) ++i;
            continue;
        }
        if (opt == "--gpu") {  // a flag with an optional device list (src/ml_optimiser.cpp:966-967)
            a->do_gpu = 1;
            if (i + 1 < argc && std::strncmp(argv[i + 1], "--", 2) != 0) {
                std::snprintf(a->gpu_ids, sizeof(a->gpu_ids), "%s", argv[i + 1]);
                ++i;
            }
            continue;
        }
        const auto it = table.find(opt);
        if (it == table.end()) RLN_FAIL(RLN_ERR_ARG, "unrecognised option %s (not a relion_refine option)", opt.c_str());
        std::string value;
        if (it->second.takes_value) {
            if (i + 1 >= argc || (std::strncmp(argv[i + 1], "--", 2) == 0))
                RLN_FAIL(RLN_ERR_ARG, "option %s needs a value", opt.c_str());
            value = argv[++i];
        }
        double *sig = opt == "--sigma_ang" ? &sigma_ang : opt == "--sigma_rot" ? &sigma_rot
                      : opt == "--sigma_tilt" ? &sigma_tilt : opt == "--sigma_psi" ? &sigma_psi : nullptr;
        if (sig) {
            if (to_double(value, opt.c_str(), *sig) != RLN_OK) return RLN_ERR_ARG;
            continue;
        }
        if (opt == "--no_norm") no_norm = true;
        if (opt == "--no_scale") no_scale = true;
        if (it->second.set) {
            const int rc = it->second.set(*a, value);
            if (rc != RLN_OK) return rc;
        }
    }
    if (no_norm) a->do_norm_correction = 0;  // src/ml_optimiser.cpp:828-831
    if (no_scale) a->do_scale_correction = 0;
    // orientational prior (src/ml_optimiser.cpp:733-767); --relax_sym is refused, so its default branch is NOPRIOR
    if (sigma_ang > 0.) {
        a->sigma_rot = a->sigma_tilt = a->sigma_psi = sigma_ang;
        a->local_search = 1;
    } else if (sigma_rot > 0. || sigma_tilt > 0. || sigma_psi > 0.) {
        a->sigma_rot = sigma_rot > 0. ? sigma_rot : 0.;
        a->sigma_tilt = sigma_tilt > 0. ? sigma_tilt : 0.;
        a->sigma_psi = sigma_psi > 0. ? sigma_psi : 0.;
        a->local_search = 1;
    } else {
        a->sigma_rot = a->sigma_tilt = a->sigma_psi = 0.;
        a->local_search = 0;
    }
    if (a->nr_classes < 1) RLN_FAIL(RLN_ERR_ARG, "--K %d: at least one class", a->nr_classes);
    if (a->healpix_order < 0) RLN_FAIL(RLN_ERR_ARG, "--healpix_order %d: not a HEALPix order", a->healpix_order);
    if (!(a->padding_factor >= 1.)) RLN_FAIL(RLN_ERR_ARG, "--pad %g: padding factor below 1", a->padding_factor);
    if (!(a->offset_step_px > 0.)) RLN_FAIL(RLN_ERR_ARG, "--offset_step %g: not positive", a->offset_step_px);
    return RLN_OK;
}

extern "C" int rln_refine_image_sizes(const rln_refine_args *a, int model_ori_size, double model_pixel_size,
                                      int image_size, double image_pixel_size, int iter, int maxres_shell,
                                      double ave_pmax, int has_high_fsc_at_limit, int *current_size, int *coarse_size) {
    if (!a || !current_size || !coarse_size || model_ori_size <= 0 || !(model_pixel_size > 0.) || image_size <= 0 ||
        !(image_pixel_size > 0.) || iter < 1)
        RLN_FAIL(RLN_ERR_ARG, "rln_refine_image_sizes: invalid argument");
    // updateCurrentResolution: the resolution shell of the current resolution
    int maxres;
    if (!a->do_map) {
        maxres = model_ori_size / 2;
    } else if (a->ini_high > 0. && iter == 1) {  // iter == 0 of the reference's numbering: the first E-step
        maxres = rln::ROUND(model_ori_size * model_pixel_size / a->ini_high);
    } else {
        maxres = std::max(maxres_shell, a->minres_map);
    }
    // updateImageSizeAndResolutionPointers: getPixelFromResolution(getResolution(maxres)) is maxres again
    if (ave_pmax > 0.1 && has_high_fsc_at_limit)
        maxres += rln::ROUND(0.25 * model_ori_size / 2);
    else
        maxres += a->incr_size;
    int cur = maxres * 2;
    if (a->do_use_all_data) cur = model_ori_size;
    cur = std::min(cur, model_ori_size);
    // one optics group: remap the model's size to the group's image size, keep it even
    const double remap = (image_pixel_size * image_size) / (model_pixel_size * model_ori_size);
    int icur = 2 * rln::CEIL(0.5 * remap * cur);
    icur = std::min(image_size, icur);
    const int max_coarse = a->max_coarse_size > 0 ? (int)(remap * a->max_coarse_size) : image_size;
    int coarse;
    if (a->strict_highres_exp > 0.) {
        coarse = 2 * rln::ROUND(remap * model_ori_size * model_pixel_size / a->strict_highres_exp);
    } else if (a->adaptive_oversampling > 0) {
        const double diameter = a->particle_diameter > 0. ? a->particle_diameter
                                                          : (model_ori_size - a->width_mask_edge) * model_pixel_size;
        const double angular = 360. / (6 * rln::ROUND(std::pow(2., a->healpix_order)));
        const double rotated_distance = (angular / 360.) * rln::kPi * diameter;
        const double coarse_resolution = rotated_distance / 1.2;  // keepsafe factor of 3-D references
        coarse = 2 * rln::CEIL(remap * model_pixel_size * model_ori_size / coarse_resolution);
        coarse = std::min(max_coarse, coarse);
    } else {
        coarse = icur;
    }
    coarse = std::min(icur, coarse);
    *current_size = icur;
    *coarse_size = coarse;
    return RLN_OK;
}

extern "C" int rln_refine_estep_config(const rln_refine_args *a, int model_ori_size, double model_pixel_size,
                                       double image_pixel_size, int current_size, int coarse_size, int precision,
                                       rln_estep_config *cfg) {
    if (!a || !cfg || model_ori_size <= 0 || !(model_pixel_size > 0.) || !(image_pixel_size > 0.) || current_size <= 0 ||
        coarse_size <= 0 || coarse_size > current_size)
        RLN_FAIL(RLN_ERR_ARG, "rln_refine_estep_config: invalid argument");
    if (precision != RLN_PREC_PLAIN_FP32 && precision != RLN_PREC_SPLIT_BF16)
        RLN_FAIL(RLN_ERR_ARG, "rln_refine_estep_config: precision %d", precision);
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->ori_size = model_ori_size;
    cfg->nr_classes = a->nr_classes;
    cfg->pixel_size = image_pixel_size;
    cfg->padding_factor = a->padding_factor;
    cfg->current_size = current_size;
    cfg->coarse_size = coarse_size;
    cfg->healpix_order = a->healpix_order;
    cfg->adaptive_oversampling = a->adaptive_oversampling;
    cfg->psi_step = a->psi_step;
    // the command line's offsets are in pixels, the sampling's in Angstrom (src/ml_optimiser.cpp:2346-2355)
    cfg->offset_range = a->offset_range_px * model_pixel_size;
    cfg->offset_step = a->offset_step_px * model_pixel_size;
    cfg->local_search = a->local_search;
    cfg->precision = precision;
    cfg->sigma_rot = a->sigma_rot;
    cfg->sigma_tilt = a->sigma_tilt;
    cfg->sigma_psi = a->sigma_psi;
    cfg->sigma_cutoff = 3.;
    cfg->adaptive_fraction = a->adaptive_fraction;
    cfg->maximum_significants = a->maximum_significants;
    cfg->sigma2_fudge = a->sigma2_fudge;
    cfg->sigma2_offset = a->sigma_offset * a->sigma_offset;  // src/ml_optimiser.cpp:663-664
    cfg->do_ctf_correction = a->do_ctf_correction && a->refs_are_ctf_corrected;
    cfg->do_scale_correction = a->do_scale_correction;
    return RLN_OK;
}
