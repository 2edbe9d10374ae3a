"""MI355X-native RELION expectation step.

Python view of ``librln_estep.so`` (C ABI in ``include/rln_estep.h``). The classes mirror the reference's
accelerator operator surface: :class:`AccProjector` (src/acc/acc_projector.h:17-104),
:class:`AccBackprojector` (src/acc/acc_backprojector.h:24-99) and :class:`ExpectationStep`, the batched
counterpart of ``MlOptimiserAccGPU::doThreadExpectationSomeParticles``
(src/acc/hip/hip_ml_optimiser.hip.cpp:250-298). All compute runs in the HIP library; importing this module
fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librln_estep.so")

RLN_OK = 0
RLN_ERR_ARG = -1
RLN_ERR_HIP = -2
RLN_ERR_STATE = -3
RLN_ERR_FILTERED_ZERO = -4
RLN_ERR_NO_SIGNIFICANT = -5
RLN_ERR_SUM_WEIGHT_ZERO = -6
RLN_ERR_MEMORY = -7
RLN_PREC_PLAIN_FP32 = 0
RLN_PREC_SPLIT_BF16 = 1


class RelionError(RuntimeError):
    """Raised for a non-zero status of the C ABI (the reference throws RelionError, src/error.h:65)."""

    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


class EstepConfig(C.Structure):
    _fields_ = [
        ("ori_size", C.c_int), ("nr_classes", C.c_int), ("pixel_size", C.c_double),
        ("padding_factor", C.c_double), ("current_size", C.c_int), ("coarse_size", C.c_int),
        ("healpix_order", C.c_int), ("adaptive_oversampling", C.c_int), ("psi_step", C.c_double),
        ("offset_range", C.c_double), ("offset_step", C.c_double), ("local_search", C.c_int),
        ("precision", C.c_int), ("sigma_rot", C.c_double), ("sigma_tilt", C.c_double), ("sigma_psi", C.c_double),
        ("sigma_cutoff", C.c_double), ("adaptive_fraction", C.c_double), ("maximum_significants", C.c_longlong),
        ("sigma2_fudge", C.c_double), ("sigma2_offset", C.c_double), ("do_ctf_correction", C.c_int),
        ("do_scale_correction", C.c_int),
    ]


class ParticleMeta(C.Structure):
    _fields_ = [
        ("old_offset_x", C.c_double), ("old_offset_y", C.c_double), ("prior_offset_x", C.c_double),
        ("prior_offset_y", C.c_double), ("prior_rot", C.c_double), ("prior_tilt", C.c_double),
        ("prior_psi", C.c_double), ("scale", C.c_double), ("highres_Xi2", C.c_double),
    ]


class ParticleResult(C.Structure):
    _fields_ = [
        ("rot", C.c_double), ("tilt", C.c_double), ("psi", C.c_double), ("xoff", C.c_double), ("yoff", C.c_double),
        ("pmax", C.c_double), ("nr_significant", C.c_double), ("dLL", C.c_double), ("sum_weight", C.c_double),
        ("min_diff2", C.c_double), ("significant_weight", C.c_double), ("coarse_sum_weight", C.c_double),
        ("coarse_min_diff2", C.c_double), ("coarse_significant_weight", C.c_double),
        ("wsum_norm_correction", C.c_double), ("signal_product", C.c_double), ("reference_power", C.c_double),
        ("best_ihidden_over", C.c_longlong), ("iclass", C.c_int), ("nr_fine_evaluated", C.c_int),
    ]


class Wsum(C.Structure):
    _fields_ = [
        ("sigma2_noise", C.POINTER(C.c_double)), ("pdf_class", C.POINTER(C.c_double)),
        ("pdf_direction", C.POINTER(C.c_double)), ("sigma2_offset", C.c_double), ("sumw_group", C.c_double),
        ("signal_product", C.c_double), ("reference_power", C.c_double), ("LL", C.c_double),
        ("ave_Pmax", C.c_double),
    ]


class SynthSpec(C.Structure):
    _fields_ = [
        ("seed", C.c_ulonglong), ("ori_size", C.c_int), ("pixel_size", C.c_double), ("nr_blobs", C.c_int),
        ("blob_radius", C.c_double), ("sigma_min", C.c_double), ("sigma_max", C.c_double),
        ("amp_min", C.c_double), ("amp_max", C.c_double), ("snr_min", C.c_double), ("snr_max", C.c_double),
        ("defocus_min", C.c_double), ("defocus_max", C.c_double), ("max_astigmatism", C.c_double),
        ("voltage", C.c_double), ("cs", C.c_double), ("q0", C.c_double), ("bfactor", C.c_double),
        ("offset_max", C.c_double), ("prior_sigma_deg", C.c_double),
    ]


class ProposalConfig(C.Structure):
    _fields_ = [("window", C.c_int), ("channels", C.c_int), ("healpix_order", C.c_int), ("n_psi_bins", C.c_int),
                ("seed", C.c_ulonglong)]


class ProposalGen(C.Structure):
    _fields_ = [("ori_size", C.c_int), ("pixel_size", C.c_double), ("voltage", C.c_double), ("cs", C.c_double),
                ("q0", C.c_double), ("defocus_min", C.c_double), ("defocus_max", C.c_double),
                ("max_astigmatism", C.c_double), ("offset_max", C.c_double),
                ("sigma2_noise", C.POINTER(C.c_double))]


class ImgPrepConfig(C.Structure):
    _fields_ = [("ori_size", C.c_int), ("current_size", C.c_int), ("pixel_size", C.c_double),
                ("particle_diameter", C.c_double), ("width_mask_edge", C.c_double), ("zero_mask", C.c_int),
                ("sigma2_fudge", C.c_double), ("random_seed", C.c_ulonglong), ("avg_norm_correction", C.c_double),
                ("do_ctf_correction", C.c_int), ("ctf_phase_flipped", C.c_int), ("only_flip_phases", C.c_int),
                ("intact_ctf_first_peak", C.c_int), ("ctf_premultiplied", C.c_int), ("voltage", C.c_double),
                ("cs", C.c_double), ("q0", C.c_double)]


class ImageMeta(C.Structure):
    _fields_ = [("norm_correction", C.c_double), ("old_offset_x", C.c_double), ("old_offset_y", C.c_double),
                ("defocus_u", C.c_double), ("defocus_v", C.c_double), ("defocus_angle", C.c_double),
                ("bfactor", C.c_double), ("kfactor", C.c_double), ("phase_shift", C.c_double),
                ("part_id", C.c_longlong)]


RESULT_DTYPE = np.dtype([(n, np.float64 if t in (C.c_double,) else (np.int64 if t is C.c_longlong else np.int32))
                         for n, t in ParticleResult._fields_])
assert RESULT_DTYPE.itemsize == C.sizeof(ParticleResult)
META_DTYPE = np.dtype([(n, np.float64) for n, _ in ParticleMeta._fields_])
assert META_DTYPE.itemsize == C.sizeof(ParticleMeta)

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p

_SIGNATURES = {
    "rln_last_error": (C.c_char_p, []),
    "rln_projector_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "rln_projector_set_mdl_dim": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _ip]),
    "rln_projector_init_mdl": (C.c_int, [_vp, _fp, _fp]),
    "rln_projector_init_mdl_complex": (C.c_int, [_vp, _dp]),
    "rln_projector_destroy": (C.c_int, [_vp]),
    "rln_projector_compute_fourier_map": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                                    _dp, _dp]),
    "rln_backprojector_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "rln_backprojector_set_mdl_dim": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                                C.POINTER(C.c_size_t)]),
    "rln_backprojector_init_mdl": (C.c_int, [_vp]),
    "rln_backprojector_get_mdl_data": (C.c_int, [_vp, _fp, _fp, _fp]),
    "rln_backprojector_get_mdl_data_noalloc": (C.c_int, [_vp, _fp, _fp, _fp]),
    "rln_backprojector_destroy": (C.c_int, [_vp]),
    "rln_backprojector_reconstruct": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_int, C.c_double, C.c_double,
                                                C.c_int, _fp, C.c_int]),
    "rln_estep_create": (C.c_int, [C.c_int, C.POINTER(EstepConfig), C.POINTER(_vp)]),
    "rln_estep_destroy": (C.c_int, [_vp]),
    "rln_estep_set_model": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), _dp, _dp, _dp, _dp]),
    "rln_estep_run": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, C.c_int, C.POINTER(ParticleMeta),
                                C.POINTER(ParticleResult), C.POINTER(Wsum)]),
    "rln_estep_keep_intermediates": (C.c_int, [_vp, C.c_int]),
    "rln_estep_set_orientation_sets": (C.c_int, [_vp, C.c_int, _ip, _ip]),
    "rln_estep_use_graphs": (C.c_int, [_vp, C.c_int]),
    "rln_estep_set_pool": (C.c_int, [_vp, C.c_int]),
    "rln_estep_set_fine_bands": (C.c_int, [_vp, C.c_int]),
    "rln_estep_set_fine_order": (C.c_int, [_vp, C.c_int, C.c_int]),
    "rln_estep_set_wsum_bands": (C.c_int, [_vp, C.c_int]),
    "rln_estep_set_fine_shape": (C.c_int, [_vp, C.c_int]),
    "rln_estep_fine_tuning": (C.c_int, [_vp, _dp, C.c_int, C.POINTER(C.c_int)]),
    "rln_estep_coarse_size": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_longlong), _ip, _ip, _ip]),
    "rln_estep_get_coarse": (C.c_int, [_vp, C.c_int, _fp, _dp, C.POINTER(C.c_ubyte)]),
    "rln_estep_fine_size": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_longlong)]),
    "rln_estep_get_fine": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_longlong), _dp, _dp]),
    "rln_estep_last_timings": (C.c_int, [_vp, _dp, C.c_int, _ip]),
    "rln_estep_last_counters": (C.c_int, [_vp, C.POINTER(C.c_longlong), C.c_int, _ip]),
    "rln_coarse_image_size": (C.c_int, [C.c_int, C.c_double, C.c_int, C.c_double, C.c_int]),
    "rln_sampling_counts": (C.c_int, [C.POINTER(EstepConfig), _ip, _ip, _ip]),
    "rln_sampling_grid": (C.c_int, [C.POINTER(EstepConfig), _dp, _dp, _dp, _dp, _dp]),
    "rln_sampling_local": (C.c_int, [C.POINTER(EstepConfig), C.c_double, C.c_double, C.c_double, _ip, _dp, C.c_int,
                                     _ip, _ip, _dp, C.c_int, _ip]),
    "rln_synth_reference": (C.c_int, [C.POINTER(SynthSpec), C.c_int, C.c_double, C.c_int, C.c_double, _dp]),
    "rln_synth_real_map": (C.c_int, [C.POINTER(SynthSpec), C.c_int, C.c_double, _fp]),
    "rln_synth_particles": (C.c_int, [C.POINTER(SynthSpec), C.c_int, C.c_double, C.c_int, C.c_longlong, C.c_int,
                                      _fp, _fp, C.POINTER(ParticleMeta), _dp, _dp]),
    "rln_ctf_value": (C.c_double, [C.c_double] * 11),
    "rln_selftest_mfma_bf16": (C.c_int, [C.c_int, _fp, _fp, _fp]),
    "rln_sampling_ang2pix": (C.c_int, [C.c_int, C.c_double, C.c_double, C.POINTER(C.c_longlong)]),
    # learned pose proposal (include/rln_proposal.h)
    "rln_proposal_create": (C.c_int, [C.c_int, C.POINTER(ProposalConfig), C.POINTER(_vp)]),
    "rln_proposal_destroy": (C.c_int, [_vp]),
    "rln_proposal_num_params": (C.c_int, [_vp, C.POINTER(C.c_longlong)]),
    "rln_proposal_num_bins": (C.c_int, [_vp, _ip]),
    "rln_proposal_get_params": (C.c_int, [_vp, _fp]),
    "rln_proposal_set_params": (C.c_int, [_vp, _fp]),
    "rln_proposal_clear_training": (C.c_int, [_vp]),
    "rln_proposal_generate": (C.c_int, [_vp, _vp, C.POINTER(ProposalGen), C.c_int, C.c_longlong, _dp]),
    "rln_proposal_add_images": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, C.c_int, _dp]),
    "rln_proposal_training_size": (C.c_int, [_vp, C.POINTER(C.c_longlong)]),
    "rln_proposal_train": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double, _dp]),
    "rln_proposal_predict": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int, _ip, _fp]),
    "rln_proposal_predict_launch": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int]),
    "rln_proposal_predict_finish": (C.c_int, [_vp, _ip, _fp]),
    "rln_proposal_pose_bin": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, _ip]),
    "rln_proposal_bin_orientations": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _ip, C.c_int, _ip]),
    "rln_proposal_orientation_sets": (C.c_int, [_vp, C.c_int, _ip, _fp, C.c_int, C.c_int, C.c_int, C.c_double, _ip,
                                                _ip, C.c_longlong]),
    "rln_proposal_images": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, C.c_int, _fp]),
    "rln_proposal_forward_backward": (C.c_int, [_vp, C.c_int, _fp, _ip, _dp, _fp, _fp]),
    # particle image preparation (include/rln_imgprep.h)
    "rln_imgprep_create": (C.c_int, [C.c_int, C.POINTER(ImgPrepConfig), C.POINTER(_vp)]),
    "rln_imgprep_destroy": (None, [_vp]),
    "rln_imgprep_set_sigma2_noise": (C.c_int, [_vp, _dp, C.c_int]),
    "rln_imgprep_run": (C.c_int, [_vp, C.c_int, _vp, C.c_int, C.POINTER(ImageMeta), _vp, _vp, _vp, C.c_int, _dp, _dp,
                                  _dp]),
    "rln_imgprep_timings": (C.c_int, [_vp, _dp]),
}

STAGES = ("prep", "coarse_diff2", "coarse_weights", "jobs", "fine_diff2", "fine_weights", "wsum_backproject",
          "host_chunk_prep", "host_first_chunk", "host_epilogue", "run_wall")


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python relion_amd/_build.py` "
                          "(or __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _float_pairs(a):
    """Complex images as interleaved float32 (re, im) pairs; float arrays as they are (already pairs)."""
    a = np.asarray(a)
    if np.iscomplexobj(a):
        return np.ascontiguousarray(a, np.complex64).view(np.float32)
    return np.ascontiguousarray(a, np.float32)


def _check(rc):
    if rc != RLN_OK:
        raise RelionError(rc, lib.rln_last_error().decode())
    return rc


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


# ------------------------------------------------------------------------------------------------
# sizes and sampling helpers
# ------------------------------------------------------------------------------------------------
def pad_size(current_size, padding_factor=2.0):
    """pad_size of Projector::initialiseData (src/projector.cpp:59-89) for r_max = current_size/2."""
    r_max = current_size // 2
    x = padding_factor * r_max
    rnd = int(x + 0.5) if x > 0 else int(x - 0.5)
    return 2 * (rnd + 1) + 1


def volume_dims(current_size, padding_factor=2.0):
    """(xdim, ydim, zdim, inity, initz, maxr) of the padded Fourier reference."""
    ps = pad_size(current_size, padding_factor)
    return ps // 2 + 1, ps, ps, -(ps // 2), -(ps // 2), current_size // 2


def coarse_image_size(ori_size, pixel_size, healpix_order, particle_diameter, current_size=None):
    """image_coarse_size rule of src/ml_optimiser.cpp:5757-5779."""
    cur = ori_size if current_size is None else current_size
    r = lib.rln_coarse_image_size(ori_size, pixel_size, healpix_order, particle_diameter, cur)
    if r < 0:
        _check(r)
    return r


def make_config(ori_size, pixel_size, healpix_order, offset_range, offset_step, coarse_size, current_size=None,
                nr_classes=1, padding_factor=2.0, local_search=False, sigma_ang=0.0, adaptive_fraction=0.999,
                maximum_significants=0, sigma2_offset=None, do_ctf_correction=True, do_scale_correction=False,
                psi_step=-1.0, sigma2_fudge=1.0, precision=RLN_PREC_PLAIN_FP32):
    cfg = EstepConfig()
    cfg.ori_size = ori_size
    cfg.nr_classes = nr_classes
    cfg.pixel_size = pixel_size
    cfg.padding_factor = padding_factor
    cfg.current_size = ori_size if current_size is None else current_size
    cfg.coarse_size = coarse_size
    cfg.healpix_order = healpix_order
    cfg.adaptive_oversampling = 1
    cfg.psi_step = psi_step
    cfg.offset_range = offset_range
    cfg.offset_step = offset_step
    cfg.local_search = 1 if local_search else 0
    cfg.precision = precision
    cfg.sigma_rot = cfg.sigma_tilt = cfg.sigma_psi = sigma_ang
    cfg.sigma_cutoff = 3.0
    cfg.adaptive_fraction = adaptive_fraction
    cfg.maximum_significants = maximum_significants
    cfg.sigma2_fudge = sigma2_fudge
    # search range = 3 sigma (src/ml_optimiser.cpp:7737-7740) unless given
    cfg.sigma2_offset = (offset_range / 3.0) ** 2 if sigma2_offset is None else sigma2_offset
    cfg.do_ctf_correction = 1 if do_ctf_correction else 0
    cfg.do_scale_correction = 1 if do_scale_correction else 0
    return cfg


def sampling_counts(cfg):
    nd, npsi, nt = C.c_int(), C.c_int(), C.c_int()
    _check(lib.rln_sampling_counts(C.byref(cfg), C.byref(nd), C.byref(npsi), C.byref(nt)))
    return nd.value, npsi.value, nt.value


def sampling_grid(cfg):
    nd, npsi, nt = sampling_counts(cfg)
    rot, tilt, psi = np.zeros(nd), np.zeros(nd), np.zeros(npsi)
    tx, ty = np.zeros(nt), np.zeros(nt)
    _check(lib.rln_sampling_grid(C.byref(cfg), _ptr(rot, C.c_double), _ptr(tilt, C.c_double), _ptr(psi, C.c_double),
                                 _ptr(tx, C.c_double), _ptr(ty, C.c_double)))
    return rot, tilt, psi, tx, ty


def sampling_local(cfg, prior_rot, prior_tilt, prior_psi):
    nd, npsi, _ = sampling_counts(cfg)
    di, dp = np.zeros(nd, np.int32), np.zeros(nd)
    pi, pp = np.zeros(npsi, np.int32), np.zeros(npsi)
    n1, n2 = C.c_int(), C.c_int()
    _check(lib.rln_sampling_local(C.byref(cfg), prior_rot, prior_tilt, prior_psi, _ptr(di, C.c_int),
                                  _ptr(dp, C.c_double), nd, C.byref(n1), _ptr(pi, C.c_int), _ptr(pp, C.c_double), npsi,
                                  C.byref(n2)))
    return di[:n1.value], dp[:n1.value], pi[:n2.value], pp[:n2.value]


def ang2pix(order, rot, tilt):
    """HEALPix NEST pixel at order containing direction (rot, tilt) in degrees."""
    out = C.c_longlong()
    _check(lib.rln_sampling_ang2pix(order, float(rot), float(tilt), C.byref(out)))
    return out.value


def selftest_mfma_bf16(A, B, device=0):
    """D = A(16x32) @ B(32x16) through one BF16 MFMA with the library's lane layout."""
    a = np.ascontiguousarray(A, np.float32)
    b = np.ascontiguousarray(B, np.float32)
    d = np.zeros((16, 16), np.float32)
    _check(lib.rln_selftest_mfma_bf16(device, _ptr(a, C.c_float), _ptr(b, C.c_float), _ptr(d, C.c_float)))
    return d


def ctf_value(defU, defV, defAngle, voltage, cs, q0, bfactor, scale, phase_shift, x, y):
    return lib.rln_ctf_value(defU, defV, defAngle, voltage, cs, q0, bfactor, scale, phase_shift, x, y)


# ------------------------------------------------------------------------------------------------
# synthetic data
# ------------------------------------------------------------------------------------------------
def synth_spec(seed, ori_size, pixel_size, nr_blobs, snr, defocus=(1.0e4, 3.0e4), offset_max=3.0,
               prior_sigma_deg=2.0, blob_radius=None, max_astigmatism=300.0):
    s = SynthSpec()
    s.seed = seed
    s.ori_size = ori_size
    s.pixel_size = pixel_size
    s.nr_blobs = nr_blobs
    s.blob_radius = 0.375 * ori_size if blob_radius is None else blob_radius
    s.sigma_min, s.sigma_max = 1.5, 3.0
    s.amp_min, s.amp_max = 0.5, 1.0
    if isinstance(snr, (tuple, list)):
        s.snr_min, s.snr_max = snr
    else:
        s.snr_min = s.snr_max = snr
    s.defocus_min, s.defocus_max = defocus
    s.max_astigmatism = max_astigmatism
    s.voltage, s.cs, s.q0, s.bfactor = 300.0, 2.7, 0.1, 0.0
    s.offset_max = offset_max
    s.prior_sigma_deg = prior_sigma_deg
    return s


def synth_reference(spec, iclass, current_size, padding_factor=2.0, share_fraction=0.0):
    xd, yd, zd, _, _, _ = volume_dims(current_size, padding_factor)
    out = np.zeros((zd, yd, xd, 2), np.float64)
    _check(lib.rln_synth_reference(C.byref(spec), iclass, share_fraction, current_size, padding_factor,
                                   _ptr(out, C.c_double)))
    return out


def synth_real_map(spec, iclass, ori_size, share_fraction=0.0):
    """The blob map of `spec` sampled in real space, float32 [ori][ori][ori] (z, y, x), origin at ori/2."""
    if ori_size != spec.ori_size:
        raise ValueError("ori_size must equal spec.ori_size")
    out = np.zeros((ori_size,) * 3, np.float32)
    _check(lib.rln_synth_real_map(C.byref(spec), iclass, share_fraction, _ptr(out, C.c_float)))
    return out


def synth_particles(spec, n, current_size, nr_classes=1, first=0, share_fraction=0.0):
    xs = current_size // 2 + 1
    fimg = np.zeros((n, current_size, xs, 2), np.float32)
    fctf = np.zeros((n, current_size, xs), np.float32)
    meta = (ParticleMeta * max(n, 1))()
    pose = np.zeros((n, 6))
    nvar = np.zeros(n)
    _check(lib.rln_synth_particles(C.byref(spec), nr_classes, share_fraction, current_size, first, n,
                                   _ptr(fimg, C.c_float), _ptr(fctf, C.c_float), meta, _ptr(pose, C.c_double),
                                   _ptr(nvar, C.c_double)))
    return fimg, fctf, meta, pose, nvar


def meta_array(meta):
    """ParticleMeta ctypes array -> numpy structured array (copy)."""
    return np.frombuffer(meta, dtype=META_DTYPE).copy()


# ------------------------------------------------------------------------------------------------
# operator surface
# ------------------------------------------------------------------------------------------------
class AccProjector:
    """Device copy of one class's padded Fourier reference (AccProjector, src/acc/acc_projector.h:17-104)."""

    def __init__(self, device=0):
        self._h = _vp()
        _check(lib.rln_projector_create(device, C.byref(self._h)))
        self.dims = None

    def setMdlDim(self, xdim, ydim, zdim, inity, initz, maxr, padding_factor):
        changed = C.c_int()
        _check(lib.rln_projector_set_mdl_dim(self._h, xdim, ydim, zdim, inity, initz, maxr, padding_factor,
                                             C.byref(changed)))
        self.dims = (xdim, ydim, zdim, inity, initz, maxr, padding_factor)
        return bool(changed.value)

    def initMdl(self, data, imag=None):
        """initMdl(Complex*) for a complex/(...,2) float64 array, or initMdl(real, imag) for two float32 arrays."""
        if imag is not None:
            re = np.ascontiguousarray(data, np.float32)
            im = np.ascontiguousarray(imag, np.float32)
            _check(lib.rln_projector_init_mdl(self._h, _ptr(re, C.c_float), _ptr(im, C.c_float)))
            return
        a = np.asarray(data)
        if np.iscomplexobj(a):
            a = np.stack([a.real, a.imag], axis=-1)
        a = np.ascontiguousarray(a, np.float64)
        _check(lib.rln_projector_init_mdl_complex(self._h, _ptr(a, C.c_double)))

    def computeFourierTransformMap(self, vol, current_size, padding_factor=2.0, do_gridding=True, min_ires=-1,
                                   return_data=False, ori_size=None, on_device=False):
        """Projector::computeFourierTransformMap on the device: a real-space map (float32 [ori]^3, or a device
        address with on_device and ori_size) becomes this projector's padded Fourier volume. Returns the power
        spectrum [ori/2+1] (and the plain data array, complex128 [S][S][S/2+1], with return_data)."""
        if on_device:
            ori, ptr, keep = int(ori_size), _vp(int(vol)), None
        else:
            keep = np.ascontiguousarray(vol, np.float32)
            ori, ptr = keep.shape[0], _vp(keep.ctypes.data)
        ps = np.zeros(ori // 2 + 1)
        data = None
        if return_data:
            P = int(np.floor(padding_factor * ori + 0.5))
            P += P % 2
            r_max = min(current_size if current_size > 0 else ori, ori) // 2
            S = 2 * (int(np.floor(P / ori * r_max + 0.5)) + 1) + 1
            data = np.zeros((S, S, S // 2 + 1, 2))
        _check(lib.rln_projector_compute_fourier_map(self._h, ptr, int(on_device), ori, current_size, padding_factor,
                                                     int(do_gridding), min_ires, _ptr(ps, C.c_double),
                                                     None if data is None else _ptr(data, C.c_double)))
        del keep
        if data is None:
            return ps
        return ps, data[..., 0] + 1j * data[..., 1]

    def clear(self):
        if self._h:
            lib.rln_projector_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.clear()
        except Exception:
            pass


class AccBackprojector:
    """Device accumulators of one class's data/weight volumes (AccBackprojector, acc_backprojector.h:24-99)."""

    def __init__(self, device=0):
        self._h = _vp()
        _check(lib.rln_backprojector_create(device, C.byref(self._h)))
        self.dims = None

    def setMdlDim(self, xdim, ydim, zdim, inity, initz, maxr, padding_factor):
        nb = C.c_size_t()
        _check(lib.rln_backprojector_set_mdl_dim(self._h, xdim, ydim, zdim, inity, initz, maxr, padding_factor,
                                                 C.byref(nb)))
        self.dims = (xdim, ydim, zdim, inity, initz, maxr, padding_factor)
        return nb.value

    def initMdl(self):
        _check(lib.rln_backprojector_init_mdl(self._h))

    def getMdlData(self):
        xd, yd, zd = self.dims[:3]
        re = np.zeros((zd, yd, xd), np.float32)
        im = np.zeros_like(re)
        w = np.zeros_like(re)
        _check(lib.rln_backprojector_get_mdl_data(self._h, _ptr(re, C.c_float), _ptr(im, C.c_float),
                                                  _ptr(w, C.c_float)))
        return re, im, w

    def getMdlDataNoAlloc(self):
        """getMdlData through the entry point that allocates no device memory (a static staging array)."""
        xd, yd, zd = self.dims[:3]
        re = np.zeros((zd, yd, xd), np.float32)
        im = np.zeros_like(re)
        w = np.zeros_like(re)
        _check(lib.rln_backprojector_get_mdl_data_noalloc(self._h, _ptr(re, C.c_float), _ptr(im, C.c_float),
                                                          _ptr(w, C.c_float)))
        return re, im, w

    def clear(self):
        if self._h:
            lib.rln_backprojector_destroy(self._h)
            self._h = _vp()

    def reconstruct(self, ori_size, max_iter_preweight=10, do_map=True, tau2=None, tau2_fudge=1.0, normalise=1.0,
                    minres_map=3, skip_gridding=True, out_address=None):
        """BackProjector::reconstruct on the device (src/backprojector.cpp:1379-1780): float32 map [ori]^3.

        skip_gridding defaults to relion_refine's setting (src/ml_optimiser.cpp:586). With out_address (a device
        pointer to ori^3 floats) the map stays on the device and None is returned."""
        t = np.ascontiguousarray(tau2 if tau2 is not None else np.zeros(0), np.float64)
        out = None if out_address is not None else np.zeros((ori_size,) * 3, np.float32)
        dst = C.cast(C.c_void_p(out_address), _fp) if out is None else _ptr(out, C.c_float)
        _check(lib.rln_backprojector_reconstruct(self._h, ori_size, int(skip_gridding), max_iter_preweight, int(do_map),
                                                 _ptr(t, C.c_double) if len(t) else None, len(t), tau2_fudge,
                                                 normalise, minres_map, dst, int(out is None)))
        return out

    def __del__(self):
        try:
            self.clear()
        except Exception:
            pass


class ExpectationStep:
    """Batched E-step engine of one device (MlDeviceBundle + MlOptimiserAccGPU)."""

    def __init__(self, cfg, device=0):
        self.cfg = cfg
        self._h = _vp()
        _check(lib.rln_estep_create(device, C.byref(cfg), C.byref(self._h)))
        self.n_dir, self.n_psi, self.n_trans = sampling_counts(cfg)
        self.nshell = cfg.ori_size // 2 + 1
        self._model = None

    def set_model(self, projectors, backprojectors, pdf_class, sigma2_noise, pdf_direction=None, data_vs_prior=None):
        K = self.cfg.nr_classes
        pa = (_vp * K)(*[p._h for p in projectors])
        ba = (_vp * K)(*[b._h for b in backprojectors])
        pc = np.ascontiguousarray(pdf_class, np.float64)
        s2 = np.ascontiguousarray(sigma2_noise, np.float64)
        pd = None if pdf_direction is None else np.ascontiguousarray(pdf_direction, np.float64)
        dv = None if data_vs_prior is None else np.ascontiguousarray(data_vs_prior, np.float64)
        _check(lib.rln_estep_set_model(self._h, pa, ba, _ptr(pc, C.c_double),
                                       None if pd is None else _ptr(pd, C.c_double), _ptr(s2, C.c_double),
                                       None if dv is None else _ptr(dv, C.c_double)))
        self._model = (projectors, backprojectors, pc, s2, pd, dv)

    def keep_intermediates(self, keep=True):
        _check(lib.rln_estep_keep_intermediates(self._h, 1 if keep else 0))

    def use_graphs(self, enable=True):
        _check(lib.rln_estep_use_graphs(self._h, 1 if enable else 0))

    def set_pool(self, max_particles):
        """Particles per chunk at most (the reference's --pool); results do not depend on it."""
        _check(lib.rln_estep_set_pool(self._h, int(max_particles)))

    def set_fine_bands(self, n_bands):
        """Radial pixel bands per fine work item (1, 2, 4, 8, 16, 32), or 0 for the engine's own tuning."""
        _check(lib.rln_estep_set_fine_bands(self._h, int(n_bands)))

    def set_fine_order(self, order=-1, wsum_order=0):
        """Order of the fine pass's work units (0: particle order, band-major; 1: orientation order, band-major;
        2: orientation order, item-major; 3: orientation order, band-major in per-XCD queues; -1: the engine's
        tuning) and of the weighted sums' units (0 band-major, 1 item-major, 2 band-major in per-XCD queues)."""
        _check(lib.rln_estep_set_fine_order(self._h, int(order), int(wsum_order)))

    def set_fine_shape(self, shape):
        """Work-item shape of the fine pass (0 .. 3: (TC, CW) = (2, 2), (1, 2), (2, 1), (1, 1)), or -1 for the
        tuning's choice."""
        _check(lib.rln_estep_set_fine_shape(self._h, int(shape)))

    def set_wsum_bands(self, n_bands):
        """Radial pixel bands per weighted-sum work unit (1, 2, 4, 8, 16, 32), or 0 for the fine pass's count."""
        _check(lib.rln_estep_set_wsum_bands(self._h, int(n_bands)))

    def fine_tuning(self):
        """Times (ms) of the fine pass's tuning candidates as an array [shape][order][band choice][grid] (-1: not
        tried), or None before any tuning."""
        n = C.c_int()
        _check(lib.rln_estep_fine_tuning(self._h, None, 0, C.byref(n)))
        if n.value == 0:
            return None
        ms = np.zeros(n.value)
        _check(lib.rln_estep_fine_tuning(self._h, _ptr(ms, C.c_double), n.value, C.byref(n)))
        return ms.reshape(4, 4, 6, 5)

    def set_orientation_lists(self, offsets, orients):
        """set_orientation_sets in the flat form: offsets[n+1], orients[offsets[n]] (int32)."""
        off = np.ascontiguousarray(offsets, np.int32)
        og = np.ascontiguousarray(orients if len(orients) else np.zeros(1), np.int32)
        _check(lib.rln_estep_set_orientation_sets(self._h, len(off) - 1, _ptr(off, C.c_int), _ptr(og, C.c_int)))

    def set_orientation_sets(self, sets):
        """Explicit coarse orientation lists for the next run: one array of global grid indices
        (idir * n_psi + ipsi) per particle; an empty array means the exhaustive grid."""
        counts = np.array([len(x) for x in sets], np.int32)
        off = np.zeros(len(sets) + 1, np.int32)
        np.cumsum(counts, out=off[1:])
        og = np.ascontiguousarray(np.concatenate([np.asarray(x, np.int32) for x in sets]) if off[-1] else
                                  np.zeros(1, np.int32), np.int32)
        _check(lib.rln_estep_set_orientation_sets(self._h, len(sets), _ptr(off, C.c_int), _ptr(og, C.c_int)))

    def run(self, fimg, fctf, meta, fimg_nomask=None, on_device=False):
        """E-step of a batch. fimg/fctf: numpy arrays (host) or integer device addresses (on_device=True)."""
        n = len(meta)
        res = (ParticleResult * max(n, 1))()
        K = self.cfg.nr_classes
        ws_s2 = np.zeros(self.nshell)
        ws_pc = np.zeros(K)
        ws_pd = np.zeros(K * self.n_dir)
        ws = Wsum()
        ws.sigma2_noise = _ptr(ws_s2, C.c_double)
        ws.pdf_class = _ptr(ws_pc, C.c_double)
        ws.pdf_direction = _ptr(ws_pd, C.c_double)
        if on_device:
            a, c = _vp(int(fimg)), _vp(int(fctf))
            an = a if fimg_nomask is None else _vp(int(fimg_nomask))
            keep = None
        else:
            fa = _float_pairs(fimg)
            fc = np.ascontiguousarray(fctf, np.float32)
            fn = fa if fimg_nomask is None else _float_pairs(fimg_nomask)
            P = self.cfg.current_size * (self.cfg.current_size // 2 + 1)
            for name, arr, per in (("fimg", fa, 2 * P), ("fimg_nomask", fn, 2 * P), ("fctf", fc, P)):
                if arr.size != n * per:
                    raise ValueError(f"{name}: {arr.size} values for {n} particles of {per}")
            a, c, an = _vp(fa.ctypes.data), _vp(fc.ctypes.data), _vp(fn.ctypes.data)
            keep = (fa, fc, fn)
        if not isinstance(meta, C.Array):
            arr = (ParticleMeta * n)()
            for i, m in enumerate(meta):
                arr[i] = m
            meta = arr
        t_call = time.perf_counter()
        _check(lib.rln_estep_run(self._h, n, a, an, c, 1 if on_device else 0, meta, res, C.byref(ws)))
        self.last_call_ms = 1000.0 * (time.perf_counter() - t_call)  # the library call alone (wall)
        del keep
        results = np.frombuffer(res, dtype=RESULT_DTYPE, count=n).copy()
        wsum = dict(sigma2_noise=ws_s2, pdf_class=ws_pc, pdf_direction=ws_pd.reshape(K, self.n_dir),
                    sigma2_offset=ws.sigma2_offset, sumw_group=ws.sumw_group, signal_product=ws.signal_product,
                    reference_power=ws.reference_power, LL=ws.LL, ave_Pmax=ws.ave_Pmax)
        return results, wsum

    def coarse(self, ip):
        n = C.c_longlong()
        nd, npsi, nt = C.c_int(), C.c_int(), C.c_int()
        _check(lib.rln_estep_coarse_size(self._h, ip, C.byref(n), C.byref(nd), C.byref(npsi), C.byref(nt)))
        d2 = np.zeros(n.value, np.float32)
        w = np.zeros(n.value)
        m = np.zeros(n.value, np.uint8)
        _check(lib.rln_estep_get_coarse(self._h, ip, _ptr(d2, C.c_float), _ptr(w, C.c_double), _ptr(m, C.c_ubyte)))
        return d2, w, m, (nd.value, npsi.value, nt.value)

    def fine(self, ip):
        n = C.c_longlong()
        _check(lib.rln_estep_fine_size(self._h, ip, C.byref(n)))
        idx = np.zeros(n.value, np.int64)
        d2 = np.zeros(n.value)
        w = np.zeros(n.value)
        _check(lib.rln_estep_get_fine(self._h, ip, _ptr(idx, C.c_longlong), _ptr(d2, C.c_double), _ptr(w, C.c_double)))
        return idx, d2, w

    def timings(self):
        ms = np.zeros(len(STAGES))
        n = C.c_int()
        _check(lib.rln_estep_last_timings(self._h, _ptr(ms, C.c_double), len(STAGES), C.byref(n)))
        return dict(zip(STAGES, ms[:n.value].tolist()))

    COUNTERS = ("particles", "coarse_poses", "coarse_pixels", "fine_poses", "fine_pixels", "fine_items",
                "wsum_orientations", "wsum_pairs", "grid_projections", "coarse_tiles", "fine_jobs", "chunks",
                "fine_graph_launches", "fine_capacity_retries", "fine_grid", "wsum_grid",
                "fine_pixels_sampled", "fine_bands", "fine_small_tc", "fine_wide_cw", "fine_order",
                "wsum_order", "wsum_bands")

    def counters(self):
        out = np.zeros(len(self.COUNTERS), np.int64)
        n = C.c_int()
        _check(lib.rln_estep_last_counters(self._h, _ptr(out, C.c_longlong), len(out), C.byref(n)))
        return dict(zip(self.COUNTERS, out.tolist()))

    def close(self):
        if self._h:
            lib.rln_estep_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ImagePreparation:
    """Particle image preparation on one device: the image part of
    ``MlOptimiser::getFourierTransformsAndCtfs`` (src/ml_optimiser.cpp:5840-6531) for 2-D single-image particles
    (norm correction, integer re-centring, soft mask, transforms, power spectrum, CTF image)."""

    META_FIELDS = tuple(n for n, _ in ImageMeta._fields_)

    def __init__(self, ori_size, current_size, pixel_size, particle_diameter, width_mask_edge=5.0, zero_mask=True,
                 sigma2_fudge=1.0, random_seed=0, avg_norm_correction=0.0, do_ctf_correction=True,
                 ctf_phase_flipped=False, only_flip_phases=False, intact_ctf_first_peak=False,
                 ctf_premultiplied=False, voltage=300.0, cs=2.7, q0=0.1, device=0):
        c = ImgPrepConfig()
        c.ori_size, c.current_size, c.pixel_size = ori_size, current_size, pixel_size
        c.particle_diameter, c.width_mask_edge, c.zero_mask = particle_diameter, width_mask_edge, int(zero_mask)
        c.sigma2_fudge, c.random_seed, c.avg_norm_correction = sigma2_fudge, random_seed, avg_norm_correction
        c.do_ctf_correction, c.ctf_phase_flipped = int(do_ctf_correction), int(ctf_phase_flipped)
        c.only_flip_phases, c.intact_ctf_first_peak = int(only_flip_phases), int(intact_ctf_first_peak)
        c.ctf_premultiplied, c.voltage, c.cs, c.q0 = int(ctf_premultiplied), voltage, cs, q0
        self.cfg = c
        self._h = _vp()
        _check(lib.rln_imgprep_create(device, C.byref(c), C.byref(self._h)))

    def config_dict(self):
        return {n: getattr(self.cfg, n) for n, _ in ImgPrepConfig._fields_}

    def set_sigma2_noise(self, sigma2):
        s2 = np.ascontiguousarray(sigma2, np.float64)
        _check(lib.rln_imgprep_set_sigma2_noise(self._h, _ptr(s2, C.c_double), len(s2)))

    @classmethod
    def meta_array(cls, metas):
        arr = (ImageMeta * max(len(metas), 1))()
        for i, m in enumerate(metas):
            for k in cls.META_FIELDS:
                setattr(arr[i], k, m[k])
        return arr

    def run_device(self, n, images, metas, fimg, fimg_nomask, fctf, images_on_device=True):
        """Device-resident variant: images/fimg/fimg_nomask/fctf are device addresses (ints); returns the host
        outputs highres_Xi2 [n], power [n][N/2+1], old_offset [n][2]."""
        N = self.cfg.ori_size
        meta = metas if isinstance(metas, C.Array) else self.meta_array(metas)
        hx = np.zeros(n)
        power = np.zeros((n, N // 2 + 1))
        off = np.zeros((n, 2))
        _check(lib.rln_imgprep_run(self._h, n, _vp(images), int(images_on_device), meta, _vp(fimg), _vp(fimg_nomask),
                                   _vp(fctf), 1, _ptr(hx, C.c_double), _ptr(power, C.c_double),
                                   _ptr(off, C.c_double)))
        return hx, power, off

    def getFourierTransformsAndCtfs(self, images, metas):
        """Host arrays in and out: images [n][N][N] float32 -> dict of fimg, fimg_nomask (complex64 [n][s][s/2+1]),
        fctf (float32), highres_Xi2, power, old_offset."""
        img = np.ascontiguousarray(images, np.float32)
        n = img.shape[0]
        N, s = self.cfg.ori_size, self.cfg.current_size
        if img.shape[1:] != (N, N):
            raise ValueError(f"images must be [n][{N}][{N}], got {img.shape}")
        xs = s // 2 + 1
        fimg = np.zeros((n, s, xs), np.complex64)
        fnm = np.zeros((n, s, xs), np.complex64)
        fctf = np.zeros((n, s, xs), np.float32)
        hx = np.zeros(n)
        power = np.zeros((n, N // 2 + 1))
        off = np.zeros((n, 2))
        meta = metas if isinstance(metas, C.Array) else self.meta_array(metas)
        _check(lib.rln_imgprep_run(self._h, n, img.ctypes.data, 0, meta, fimg.ctypes.data, fnm.ctypes.data,
                                   fctf.ctypes.data, 0, _ptr(hx, C.c_double), _ptr(power, C.c_double),
                                   _ptr(off, C.c_double)))
        return dict(fimg=fimg, fimg_nomask=fnm, fctf=fctf, highres_Xi2=hx, power=power, old_offset=off)

    def timings(self):
        ms = np.zeros(5)
        _check(lib.rln_imgprep_timings(self._h, _ptr(ms, C.c_double)))
        return dict(zip(("real", "fft", "fourier", "noise", "total"), ms.tolist()))

    def close(self):
        if self._h:
            lib.rln_imgprep_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
