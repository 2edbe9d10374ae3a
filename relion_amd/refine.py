"""relion_refine's command line for the E-step's drop-in (include/rln_refine.h) and the E-step of a refinement's
first iteration from the files the command line names.

``parse_args`` takes relion_refine's own argument vector (``MlOptimiser::parseInitial``,
src/ml_optimiser.cpp:613-1088); ``image_sizes`` and ``estep_config`` apply the reference's size and unit rules
(src/ml_optimiser.cpp:5579-5780, 2346-2370). ``first_iteration_estep`` runs what relion_refine's first E-step does
on those files: the particles STAR file and its MRC stacks (``formats``), the reference map prepared into the
projector on the device (``Projector::computeFourierTransformMap``), the particle images prepared on the device
(``getFourierTransformsAndCtfs``) and the batched E-step, with the backprojection into a device backprojector.
Parsing, sizes and every compute step run in librln_estep.so; this module moves arguments.
"""
import ctypes as C
import os

import numpy as np

from . import (AccBackprojector, AccProjector, EstepConfig, ExpectationStep, ImagePreparation, ParticleMeta,
               RLN_PREC_PLAIN_FP32, _check, lib, sampling_counts, volume_dims)
from . import formats as F

_PATH = 1024


class RefineArgs(C.Structure):
    """rln_refine_args (include/rln_refine.h)."""
    _fields_ = [
        ("fn_data", C.c_char * _PATH), ("fn_ref", C.c_char * _PATH), ("fn_out", C.c_char * _PATH),
        ("fn_mask", C.c_char * _PATH), ("nr_iter", C.c_int), ("nr_classes", C.c_int), ("tau2_fudge", C.c_double),
        ("particle_diameter", C.c_double), ("do_zero_mask", C.c_int), ("do_solvent", C.c_int),
        ("do_split_random_halves", C.c_int), ("sigma_offset", C.c_double), ("ini_high", C.c_double),
        ("adaptive_oversampling", C.c_int), ("healpix_order", C.c_int), ("psi_step", C.c_double),
        ("limit_tilt", C.c_double), ("sym", C.c_char * 32), ("offset_range_px", C.c_double),
        ("offset_step_px", C.c_double), ("perturbation_factor", C.c_double), ("do_auto_refine", C.c_int),
        ("auto_local_healpix_order", C.c_int), ("sigma_rot", C.c_double), ("sigma_tilt", C.c_double),
        ("sigma_psi", C.c_double), ("local_search", C.c_int), ("do_ctf_correction", C.c_int),
        ("refs_are_ctf_corrected", C.c_int), ("do_norm_correction", C.c_int), ("do_scale_correction", C.c_int),
        ("nr_pool", C.c_int), ("nr_threads", C.c_int), ("do_gpu", C.c_int), ("gpu_ids", C.c_char * 256),
        ("padding_factor", C.c_double), ("maximum_significants", C.c_int), ("adaptive_fraction", C.c_double),
        ("max_coarse_size", C.c_int), ("width_mask_edge", C.c_int), ("incr_size", C.c_int),
        ("strict_highres_exp", C.c_double), ("strict_lowres_exp", C.c_double), ("fix_sigma_noise", C.c_int),
        ("fix_sigma_offset", C.c_int), ("skip_gridding", C.c_int), ("random_seed", C.c_int), ("do_map", C.c_int),
        ("minres_map", C.c_int), ("sigma2_fudge", C.c_double), ("do_use_all_data", C.c_int),
    ]

    def as_dict(self):
        return {n: (getattr(self, n).decode() if isinstance(getattr(self, n), bytes) else getattr(self, n))
                for n, _ in self._fields_}


lib.rln_refine_parse_args.restype = C.c_int
lib.rln_refine_parse_args.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(RefineArgs)]
lib.rln_refine_image_sizes.restype = C.c_int
lib.rln_refine_image_sizes.argtypes = [C.POINTER(RefineArgs), C.c_int, C.c_double, C.c_int, C.c_double, C.c_int,
                                       C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
lib.rln_refine_estep_config.restype = C.c_int
lib.rln_refine_estep_config.argtypes = [C.POINTER(RefineArgs), C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                        C.c_int, C.POINTER(EstepConfig)]


def parse_args(argv):
    """relion_refine's argument vector (argv[0] the program name) -> RefineArgs; RelionError names a refused or
    unknown option."""
    arr = (C.c_char_p * max(len(argv), 1))(*[str(a).encode() for a in argv])
    a = RefineArgs()
    _check(lib.rln_refine_parse_args(len(argv), arr, C.byref(a)))
    return a


def image_sizes(args, model_ori_size, model_pixel_size, image_size=None, image_pixel_size=None, iteration=1,
                maxres_shell=0, ave_pmax=0.0, has_high_fsc_at_limit=False):
    """(image_current_size, image_coarse_size) of an optics group for the E-step of `iteration` (1 = first)."""
    cur, coarse = C.c_int(), C.c_int()
    _check(lib.rln_refine_image_sizes(C.byref(args), model_ori_size, model_pixel_size,
                                      model_ori_size if image_size is None else image_size,
                                      model_pixel_size if image_pixel_size is None else image_pixel_size, iteration,
                                      maxres_shell, ave_pmax, int(has_high_fsc_at_limit), C.byref(cur),
                                      C.byref(coarse)))
    return cur.value, coarse.value


def estep_config(args, model_ori_size, model_pixel_size, current_size, coarse_size, image_pixel_size=None,
                 precision=RLN_PREC_PLAIN_FP32):
    cfg = EstepConfig()
    _check(lib.rln_refine_estep_config(C.byref(args), model_ori_size, model_pixel_size,
                                       model_pixel_size if image_pixel_size is None else image_pixel_size,
                                       current_size, coarse_size, precision, C.byref(cfg)))
    return cfg


def first_iteration_estep(argv, sigma2_noise, precision=RLN_PREC_PLAIN_FP32, threads=8, device=0,
                          base_dir=None, keep=False):
    """The E-step of the first iteration of ``relion_refine argv`` on the files it names, one optics group, one
    class (--K 1), 3-D reference: returns a dict with the particle set, the E-step configuration, the prepared
    inputs (fimg, fimg_nomask, fctf, metas), the per-particle results, the weighted sums and the filled
    AccBackprojector. sigma2_noise [ori_size/2 + 1] is the model's noise spectrum (relion_refine estimates it from
    the particles' power spectra before the first iteration, or reads --sigma). Priors: uniform over classes and
    directions (NOPRIOR) or the local Gaussians around each particle's angles (--sigma_ang); offsets centred on
    rlnOriginX/YPriorAngst (0 when absent, src/ml_optimiser.cpp:5941-5949)."""
    a = parse_args(argv)
    if a.nr_classes != 1:
        raise ValueError("first_iteration_estep reads one reference map (--K 1)")
    fn_data, fn_ref = a.fn_data.decode(), a.fn_ref.decode()
    if not fn_data or fn_ref in ("", "None"):
        raise ValueError("relion_refine needs --i and --ref for a 3-D refinement")
    base = os.path.dirname(os.path.abspath(fn_data)) if base_dir is None else base_dir
    ps = F.read_particles(fn_data, base_dir=base, threads=threads)
    if len(ps.optics) != 1:
        raise ValueError("first_iteration_estep handles one optics group")
    opt = next(iter(ps.optics.values()))
    N, px = opt["image_size"], opt["pixel_size"]
    vol = F.read_mrc(fn_ref)
    info = F.mrc_info(fn_ref)
    if vol.shape != (N, N, N):
        raise ValueError(f"reference {vol.shape} does not match the particles' box {N}")
    model_px = px  # --trust_ref_size aside, relion_refine requires the reference at the particles' pixel size
    ref_px = float(info.pixel_size)
    if ref_px > 0 and abs(ref_px - px) > 1e-3 * px:
        raise ValueError(f"reference pixel size {ref_px} differs from the particles' {px}")
    cur, coarse = image_sizes(a, N, model_px, N, px, iteration=1)
    cfg = estep_config(a, N, model_px, cur, coarse, px, precision)
    diameter = a.particle_diameter if a.particle_diameter > 0 else (N - a.width_mask_edge) * px
    # reference: Projector::computeFourierTransformMap at the current size
    proj = AccProjector(device)
    proj.computeFourierTransformMap(vol, cur, a.padding_factor)
    # particles: getFourierTransformsAndCtfs
    prep = ImagePreparation(N, cur, px, diameter, width_mask_edge=float(a.width_mask_edge),
                            zero_mask=bool(a.do_zero_mask), sigma2_fudge=a.sigma2_fudge,
                            random_seed=max(a.random_seed, 0), avg_norm_correction=1.0 if a.do_norm_correction else 0.0,
                            do_ctf_correction=bool(a.do_ctf_correction), ctf_premultiplied=bool(opt["ctf_premultiplied"]),
                            voltage=opt["voltage"], cs=opt["cs"], q0=opt["q0"], device=device)
    prep.set_sigma2_noise(sigma2_noise)
    metas = ps.imgprep_metas()
    out = prep.getFourierTransformsAndCtfs(ps.images, metas)
    prep_cfg = prep.config_dict()
    prep.close()
    n = len(ps)
    pm = (ParticleMeta * n)()
    c = ps.columns
    for i in range(n):
        pm[i].old_offset_x, pm[i].old_offset_y = out["old_offset"][i]
        pm[i].prior_offset_x = c["origin_x_prior_A"][i] / px
        pm[i].prior_offset_y = c["origin_y_prior_A"][i] / px
        pm[i].prior_rot, pm[i].prior_tilt, pm[i].prior_psi = c["rot"][i], c["tilt"][i], c["psi"][i]
        pm[i].scale = 1.0
        pm[i].highres_Xi2 = out["highres_Xi2"][i]
    bp = AccBackprojector(device)
    bp.setMdlDim(*volume_dims(cur, a.padding_factor), a.padding_factor)
    bp.initMdl()
    es = ExpectationStep(cfg, device)
    n_dir, _, _ = sampling_counts(cfg)
    es.set_model([proj], [bp], np.ones(1), np.asarray(sigma2_noise, np.float64),
                 pdf_direction=None if cfg.local_search else np.full((1, n_dir), 1.0 / n_dir))
    res, wsum = es.run(out["fimg"], out["fctf"], pm, fimg_nomask=out["fimg_nomask"])
    es.close()
    result = dict(args=a, particles=ps, config=cfg, imgprep_config=prep_cfg, imgprep_metas=metas, fimg=out["fimg"],
                  fimg_nomask=out["fimg_nomask"], fctf=out["fctf"], metas=pm, results=res, wsum=wsum,
                  backprojector=bp, current_size=cur, coarse_size=coarse, particle_diameter=diameter)
    if not keep:
        proj.clear()
    else:
        result["projector"] = proj
    return result


def write_data_star(path, particles, results, pixel_size):
    """The particles with this E-step's assignments in the columns relion_refine writes to _data.star
    (rlnAngleRot/Tilt/Psi, rlnOriginX/YAngst, rlnClassNumber, rlnMaxValueProbDistribution,
    rlnNrOfSignificantSamples, rlnLogLikeliContribution; src/ml_optimiser.cpp:10600-10700)."""
    st = F.StarFile.create()
    o = st.add_block("optics")
    opt = list(particles.optics.items())
    o.add_rows(len(opt))
    o.set_column("rlnOpticsGroup", [g for g, _ in opt], "{:d}")
    o.set_column("rlnImagePixelSize", [v["pixel_size"] for _, v in opt], "{:.6f}")
    o.set_column("rlnImageSize", [int(v["image_size"]) for _, v in opt], "{:d}")
    o.set_column("rlnImageDimensionality", [2] * len(opt), "{:d}")
    o.set_column("rlnVoltage", [v["voltage"] for _, v in opt], "{:.6f}")
    o.set_column("rlnSphericalAberration", [v["cs"] for _, v in opt], "{:.6f}")
    o.set_column("rlnAmplitudeContrast", [v["q0"] for _, v in opt], "{:.6f}")
    p = st.add_block("particles")
    n = len(particles)
    p.add_rows(n)
    p.set_column("rlnImageName", particles.image_names)
    c = particles.columns
    p.set_column("rlnOpticsGroup", [int(g) for g in c["optics_group"]], "{:d}")
    for lab, key in (("rlnDefocusU", "defocus_u"), ("rlnDefocusV", "defocus_v"), ("rlnDefocusAngle", "defocus_angle")):
        p.set_column(lab, [float(v) for v in c[key]], "{!r}")
    p.set_column("rlnAngleRot", [float(v) for v in results["rot"]], "{!r}")
    p.set_column("rlnAngleTilt", [float(v) for v in results["tilt"]], "{!r}")
    p.set_column("rlnAnglePsi", [float(v) for v in results["psi"]], "{!r}")
    p.set_column("rlnOriginXAngst", [float(v) * pixel_size for v in results["xoff"]], "{!r}")
    p.set_column("rlnOriginYAngst", [float(v) * pixel_size for v in results["yoff"]], "{!r}")
    p.set_column("rlnClassNumber", [int(v) for v in results["iclass"]], "{:d}")
    p.set_column("rlnMaxValueProbDistribution", [float(v) for v in results["pmax"]], "{!r}")
    p.set_column("rlnNrOfSignificantSamples", [int(v) for v in results["nr_significant"]], "{:d}")
    p.set_column("rlnLogLikeliContribution", [float(v) for v in results["dLL"]], "{!r}")
    st.write(path)
