"""relion_refine's command line (include/rln_refine.h): defaults, values, refusals and the size and unit rules.

Expected values are the reference's defaults (MlOptimiser::parseInitial, src/ml_optimiser.cpp:613-1088) and its
rules worked by hand for the cases below: the current size from --ini_high in the first iteration (:5593-5596) and
from the data_vs_prior shell later, plus incr_size or a quarter of the box (:5684-5705); the coarse size from the
angular sampling and the particle diameter (:5757-5779); offsets converted from pixels to Angstrom (:2346-2355).
"""
import pytest

import relion_amd as ra
from relion_amd import refine as R

TUTORIAL = ["relion_refine", "--o", "Refine3D/job019/run", "--auto_refine", "--split_random_halves", "--i",
            "Select/job017/particles.star", "--ref", "InitialModel/job015/initial_model.mrc", "--ini_high", "40",
            "--dont_combine_weights_via_disc", "--preread_images", "--pool", "30", "--pad", "2", "--ctf",
            "--particle_diameter", "200", "--flatten_solvent", "--zero_mask", "--oversampling", "1",
            "--healpix_order", "2", "--auto_local_healpix_order", "4", "--offset_range", "5", "--offset_step", "2",
            "--sym", "C1", "--low_resol_join_halves", "40", "--norm", "--scale", "--j", "6", "--gpu", "",
            "--pipeline_control", "Refine3D/job019/"]


def test_defaults_are_the_references():
    a = R.parse_args(["relion_refine"]).as_dict()
    expect = dict(fn_ref="None", fn_mask="None", nr_iter=-1, nr_classes=1, tau2_fudge=1.0, particle_diameter=-1.0,
                  sigma_offset=10.0, ini_high=-1.0, adaptive_oversampling=1, healpix_order=2, psi_step=-1.0,
                  limit_tilt=-91.0, sym="c1", offset_range_px=6.0, offset_step_px=2.0, perturbation_factor=0.5,
                  auto_local_healpix_order=4, local_search=0, do_ctf_correction=0, refs_are_ctf_corrected=1,
                  do_norm_correction=0, do_scale_correction=0, nr_pool=1, nr_threads=1, do_gpu=0,
                  padding_factor=2.0, maximum_significants=-1, adaptive_fraction=0.999, max_coarse_size=-1,
                  width_mask_edge=5, incr_size=10, strict_highres_exp=-1.0, skip_gridding=1, random_seed=-1,
                  do_map=1, minres_map=5, sigma2_fudge=1.0, do_use_all_data=0)
    for k, v in expect.items():
        assert a[k] == v, (k, a[k], v)


def test_tutorial_command_line():
    a = R.parse_args(TUTORIAL).as_dict()
    assert a["fn_data"] == "Select/job017/particles.star" and a["fn_out"] == "Refine3D/job019/run"
    assert a["fn_ref"] == "InitialModel/job015/initial_model.mrc"
    assert (a["ini_high"], a["nr_pool"], a["nr_threads"], a["particle_diameter"]) == (40.0, 30, 6, 200.0)
    assert (a["do_ctf_correction"], a["do_norm_correction"], a["do_scale_correction"]) == (1, 1, 1)
    assert (a["do_auto_refine"], a["do_split_random_halves"], a["do_zero_mask"], a["do_solvent"]) == (1, 1, 1, 1)
    assert (a["offset_range_px"], a["offset_step_px"], a["healpix_order"], a["sym"]) == (5.0, 2.0, 2, "C1")
    assert a["do_gpu"] == 1 and a["gpu_ids"] == ""


def test_values_flags_and_overrides():
    a = R.parse_args(["x", "--limit_tilt", "-91", "--iter", "-1", "--no_norm", "--norm", "--scale", "--no_scale",
                      "--tau2_fudge", "0", "--sigma_ang", "3.5", "--maxsig", "500", "--ctf_uncorrected_ref",
                      "--gpu", "--perturb", "0"]).as_dict()
    assert a["limit_tilt"] == -91.0 and a["nr_iter"] == -1
    assert a["do_norm_correction"] == 0 and a["do_scale_correction"] == 0  # --no_* win whatever the order
    assert a["tau2_fudge"] == 1.0  # non-positive values keep 1
    assert a["local_search"] == 1 and a["sigma_rot"] == a["sigma_tilt"] == a["sigma_psi"] == 3.5
    assert a["maximum_significants"] == 500 and a["refs_are_ctf_corrected"] == 0
    assert a["do_gpu"] == 1 and a["gpu_ids"] == "default" and a["perturbation_factor"] == 0.0
    b = R.parse_args(["x", "--sigma_tilt", "2", "--sigma_psi", "4"]).as_dict()
    assert b["local_search"] == 1 and (b["sigma_rot"], b["sigma_tilt"], b["sigma_psi"]) == (0.0, 2.0, 4.0)


@pytest.mark.parametrize("argv,word", [
    (["--helix"], "--helix"), (["--sym", "D7"], "--sym"), (["--oversampling", "2"], "--oversampling"),
    (["--NN"], "--NN"), (["--firstiter_cc"], "--firstiter_cc"), (["--grad"], "--grad"), (["--som"], "--som"),
    (["--skip_align"], "--skip_align"), (["--bimodal_psi"], "--bimodal_psi"), (["--limit_tilt", "30"], "--limit_tilt"),
    (["--fourier_mask", "m.mrc"], "--fourier_mask"), (["--relax_sym", "C2"], "--relax_sym"),
    (["--ctf_intact_first_peak"], "--ctf_intact_first_peak"), (["--continue", "run_it010_optimiser.star"], "--continue"),
    (["--strict_lowres_exp", "50"], "--strict_lowres_exp"), (["--tomograms", "t.star"], "--tomograms"),
    (["--bogus_option"], "--bogus_option"), (["--healpix_order"], "--healpix_order"),
    (["--healpix_order", "three"], "--healpix_order"), (["-5"], "-5"), (["--perturb", "2"], "--perturb")])
def test_refusals_name_the_option(argv, word):
    with pytest.raises(ra.RelionError, match=word):
        R.parse_args(["relion_refine"] + argv)


def test_numbers_follow_options_and_bare_tokens_pass():
    # a single-dash number is a value only after a known option; tokens without dashes are not checked
    R.parse_args(["x", "--offset", "-3"])
    R.parse_args(["x", "--ctf", "stray"])
    R.parse_args(["x", "--ctf", "-3"])  # the previous token is a known option, as IOParser checks
    with pytest.raises(ra.RelionError):
        R.parse_args(["x", "--ctf", "stray", "-3"])


def test_first_iteration_sizes_from_ini_high():
    a = R.parse_args(["x", "--ini_high", "20", "--particle_diameter", "200", "--healpix_order", "3"])
    # maxres = ROUND(256 * 1.0 / 20) = 13, + incr_size 10 -> 46; coarse: 7.5 deg at 200 A -> 2 CEIL(256 / 10.91) = 48,
    # capped by the current size
    assert R.image_sizes(a, 256, 1.0) == (46, 46)
    a2 = R.parse_args(["x", "--ini_high", "10", "--particle_diameter", "200", "--healpix_order", "3"])
    assert R.image_sizes(a2, 256, 1.0) == (72, 48)  # maxres 26 + 10 = 36 -> 72
    # a particle box of 128 at 2 A per pixel in a model of 256 at 1 A: remap 1, sizes kept even
    assert R.image_sizes(a2, 256, 1.0, image_size=128, image_pixel_size=2.0) == (72, 48)
    # a particle box of 200 at 1 A: remap 200 / 256, current 2 CEIL(0.5 * 0.78125 * 72) = 58, coarse 2 CEIL(37.5) = 38
    assert R.image_sizes(a2, 256, 1.0, image_size=200, image_pixel_size=1.0) == (58, 38)


def test_later_iteration_sizes():
    a = R.parse_args(["x", "--ini_high", "20", "--particle_diameter", "200", "--healpix_order", "2"])
    # iteration 2: data_vs_prior shell 40 + 10 -> 100; coarse at 15 deg: 2 CEIL(256 / 21.8) = 24
    assert R.image_sizes(a, 256, 1.0, iteration=2, maxres_shell=40) == (100, 24)
    # minres_map 5 bounds the shell from below
    assert R.image_sizes(a, 256, 1.0, iteration=2, maxres_shell=2)[0] == 30
    # high FSC at the limit with Pmax > 0.1: a quarter of the half box (32) instead of incr_size
    assert R.image_sizes(a, 256, 1.0, iteration=2, maxres_shell=40, ave_pmax=0.5, has_high_fsc_at_limit=True)[0] == 144
    assert R.image_sizes(a, 256, 1.0, iteration=2, maxres_shell=125)[0] == 256  # never beyond the box
    b = R.parse_args(["x", "--use_all_data", "--healpix_order", "2", "--particle_diameter", "200"])
    assert R.image_sizes(b, 256, 1.0, iteration=3, maxres_shell=20)[0] == 256
    c = R.parse_args(["x", "--no_map", "--healpix_order", "4", "--particle_diameter", "200"])
    assert R.image_sizes(c, 256, 1.0, iteration=1) == (256, 94)  # Nyquist; 3.75 deg: 2 CEIL(256 / 5.454) = 94


def test_coarse_size_limits():
    base = ["x", "--ini_high", "5", "--particle_diameter", "200", "--healpix_order", "4"]
    # current 2 (ROUND(51.2) + 10) = 122; coarse 94 at 3.75 deg
    assert R.image_sizes(R.parse_args(base), 256, 1.0) == (122, 94)
    assert R.image_sizes(R.parse_args(base + ["--coarse_size", "64"]), 256, 1.0) == (122, 64)
    # strict_highres_exp 8 A: 2 ROUND(256 / 8) = 64
    assert R.image_sizes(R.parse_args(base + ["--strict_highres_exp", "8"]), 256, 1.0) == (122, 64)
    # no diameter: (256 - maskedge 5) * 1 A = 251 A -> 2 CEIL(256 / 6.84) = 76
    assert R.image_sizes(R.parse_args(["x", "--ini_high", "5", "--healpix_order", "4"]), 256, 1.0) == (122, 76)


def test_estep_config_units():
    a = R.parse_args(["x", "--offset_range", "4", "--offset_step", "0.5", "--offset", "3", "--ctf", "--K", "2",
                      "--healpix_order", "3", "--maxsig", "100", "--adaptive_fraction", "0.99", "--pad", "2"])
    cfg = R.estep_config(a, 128, 1.5, 64, 40, precision=ra.RLN_PREC_SPLIT_BF16)
    assert (cfg.offset_range, cfg.offset_step) == (6.0, 0.75)  # pixels x model pixel size
    assert cfg.sigma2_offset == 9.0 and cfg.do_ctf_correction == 1 and cfg.nr_classes == 2
    assert (cfg.ori_size, cfg.current_size, cfg.coarse_size, cfg.healpix_order) == (128, 64, 40, 3)
    assert cfg.maximum_significants == 100 and cfg.adaptive_fraction == 0.99
    assert cfg.precision == ra.RLN_PREC_SPLIT_BF16 and cfg.local_search == 0
    u = R.parse_args(["x", "--ctf", "--ctf_uncorrected_ref"])
    assert R.estep_config(u, 128, 1.5, 64, 40).do_ctf_correction == 0
    with pytest.raises(ra.RelionError):
        R.estep_config(a, 128, 1.5, 40, 64)  # coarse above current
