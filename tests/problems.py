"""Seeded E-step problems shared by the tests: configuration, padded references and particles.

The generator (relion_amd.synth_*) is host code; the same float32 images feed the HIP path and the oracle.
"""
from dataclasses import dataclass, field

import numpy as np

import relion_amd as ra


@dataclass
class Problem:
    cfg: object
    volumes: list
    fimg: np.ndarray
    fctf: np.ndarray
    meta: object
    pose: np.ndarray
    sigma2_noise: np.ndarray
    pdf_class: np.ndarray
    pdf_direction: np.ndarray = None
    data_vs_prior: np.ndarray = None
    fimg_nomask: np.ndarray = None
    extra: dict = field(default_factory=dict)


def make_problem(box=64, px=2.0, order=2, offset_range_px=5, offset_step_px=1, diameter=100.0, n=8, snr=0.1,
                 nr_classes=1, nr_blobs=40, seed=1234, local=False, sigma_ang=0.0, prior_sigma=2.0,
                 current_size=None, scale_correction=False, perturb_meta=False, fraction=0.999, maxsig=0,
                 separate_nomask=False):
    cur = box if current_size is None else current_size
    coarse = ra.coarse_image_size(box, px, order, diameter, cur)
    cfg = ra.make_config(box, px, order, offset_range=offset_range_px * px, offset_step=offset_step_px * px,
                         coarse_size=coarse, current_size=cur, nr_classes=nr_classes, local_search=local,
                         sigma_ang=sigma_ang, adaptive_fraction=fraction, maximum_significants=maxsig,
                         do_scale_correction=scale_correction)
    spec = ra.synth_spec(seed, box, px, nr_blobs, snr=snr, offset_max=min(3.0, offset_range_px * 0.6),
                         prior_sigma_deg=prior_sigma)
    share = 0.3 if nr_classes > 1 else 0.0
    vols = [ra.synth_reference(spec, k, cur, share_fraction=share) for k in range(nr_classes)]
    fimg, fctf, meta, pose, nvar = ra.synth_particles(spec, n, cur, nr_classes=nr_classes, share_fraction=share)
    nshell = box // 2 + 1
    rng = np.random.default_rng(seed)
    # a smooth, non-constant noise spectrum around the generated noise level
    s2 = nvar.mean() * (1.0 + 0.2 * np.cos(np.arange(nshell) / 3.0))
    pdf_class = np.full(nr_classes, 1.0 / nr_classes)
    nd, _, _ = ra.sampling_counts(cfg)
    pdf_direction = np.full((nr_classes, nd), 1.0 / nd) * (1.0 + 0.1 * rng.random((nr_classes, nd)))
    dvp = None
    if scale_correction:
        dvp = np.where(np.arange(nshell) < nshell * 0.7, 10.0, 1.0)[None, :].repeat(nr_classes, 0)
    if perturb_meta:
        for i in range(n):
            meta[i].old_offset_x = float(rng.integers(-1, 2))
            meta[i].old_offset_y = float(rng.integers(-1, 2))
            meta[i].prior_offset_x = float(rng.uniform(-1, 1))
            meta[i].prior_offset_y = float(rng.uniform(-1, 1))
            meta[i].scale = float(rng.uniform(0.8, 1.2))
            meta[i].highres_Xi2 = float(rng.uniform(0, 50))
    nomask = None
    if separate_nomask:
        nomask = (fimg * np.float32(1.05)).astype(np.float32)
    return Problem(cfg, vols, fimg, fctf, meta, pose, s2, pdf_class, pdf_direction, dvp, nomask,
                   dict(spec=spec, noise_var=nvar))
