"""The E-step at the geometry of the classification, uneven-work and learned-proposal workloads.

  * classification: K = 4 classes of 128^3 blob maps (60 blobs per class, 30% shared), 1.5 A pixels, HEALPix
    order 3 global search (36,864 orientations per class), offsets 6 px at 1 px steps (113 translations),
    SNR 0.05; particle diameter 180 A (coarse size 40): 16.7 M coarse poses per particle;
  * uneven work: box 192, 1.2 A pixels, a 100-blob map, coarse order 3 with the order-4 fine pass of the
    oversampling, offsets 6 px at 2 px steps (29 translations), per-particle SNR log-uniform in [0.001, 0.3] (the
    workload's range [0.01, 0.3] widened so that 6 particles already span the spread: the oracle evaluates 128 to
    2,240 fine poses for them, where six draws from [0.01, 0.3] can all end with one significant coarse pose);
both against the FP64 oracle with the parity suite's tolerances (tests/test_gpu_parity.py: per-pose diff2 within
1e-5, identical best poses except near-ties, backprojection within 1e-4), on a few particles so that the oracle
finishes in seconds;
  * the learned proposal on box-256 data (150-blob map, 1.0 A pixels, SNR 0.03, HEALPix order 3, offsets 4 px at
    0.5 px steps): trained on projections of the reference, it seeds the split-BF16 search of 256 particles, which
    must keep the exhaustive search's best pose for at least 99% of them.
"""
import numpy as np
import pytest

import relion_amd as ra
from problems import make_problem
from test_gpu_parity import check_run

pytestmark = pytest.mark.gpu


def test_classification_box128_four_classes_order3():
    print("[configs] box 128, K = 4, order 3 global, 113 translations, 3 particles", flush=True)
    p = make_problem(box=128, px=1.5, order=3, offset_range_px=6, offset_step_px=1, diameter=180.0, n=3, snr=0.05,
                     nr_classes=4, nr_blobs=60, seed=128)
    nd, npsi, nt = ra.sampling_counts(p.cfg)
    assert (nd, npsi, nt) == (768, 48, 113) and p.cfg.coarse_size == 40
    rg, ro = check_run(p, debug=(0,))
    assert (rg["nr_fine_evaluated"] > 0).all()


def test_uneven_box192_fine_order4():
    print("[configs] box 192, log-uniform SNR 0.001-0.3, order 3 -> 4, 29 translations, 6 particles", flush=True)
    p = make_problem(box=192, px=1.2, order=3, offset_range_px=6, offset_step_px=2, diameter=180.0, n=6,
                     snr=(0.001, 0.3), nr_blobs=100, seed=193)
    nd, npsi, nt = ra.sampling_counts(p.cfg)
    assert (nd, npsi, nt) == (768, 48, 29)
    rg, ro = check_run(p, debug=(0,))
    fe = rg["nr_fine_evaluated"]
    print(f"[configs] uneven fine poses per particle: {sorted(fe.tolist())}", flush=True)
    assert fe.max() >= 10 * fe.min()


def test_proposal_box256_keeps_exhaustive_best_pose():
    from relion_amd.proposal import PoseProposal
    from test_gpu_proposal import angular_distance

    n, box = 256, 256
    print(f"[configs] proposal at box 256, order 3, {n} particles", flush=True)
    p = make_problem(box=box, px=1.0, order=3, offset_range_px=4, offset_step_px=0.5, diameter=200.0, n=n, snr=0.03,
                     nr_blobs=150, seed=2560)
    cfg = p.cfg
    dims = ra.volume_dims(box)
    pr = ra.AccProjector(0)
    pr.setMdlDim(*dims, 2.0)
    pr.initMdl(p.volumes[0])
    bp = ra.AccBackprojector(0)
    bp.setMdlDim(*dims, 2.0)
    bp.initMdl()
    cfg.precision = ra.RLN_PREC_SPLIT_BF16
    es = ra.ExpectationStep(cfg, 0)
    es.set_model([pr], [bp], p.pdf_class, p.sigma2_noise, pdf_direction=p.pdf_direction)
    res0, _ = es.run(p.fimg, p.fctf, p.meta)
    poses0 = es.counters()["coarse_poses"]
    pp = PoseProposal(window=64, channels=64, healpix_order=2, n_psi_bins=24, seed=7)
    for ep in range(5):  # fresh projections every epoch
        pp.clear_training()
        pp.generate(pr, 200000, ori_size=box, pixel_size=1.0, sigma2_noise=p.sigma2_noise, defocus=(8000.0, 25000.0),
                    max_astigmatism=300.0, offset_max=3.0, first=ep * 200000)
        loss = pp.train(epochs=1, batch=256, lr=1e-3)
        print(f"[configs] proposal epoch {ep}: loss {loss:.3f}", flush=True)
    _, n_psi, _ = ra.sampling_counts(cfg)
    bins, probs = pp.predict(p.fimg, p.fctf, n, box, topk=16)
    off, og = pp.orientation_lists(bins, probs, 3, n_psi, min_top1=0.2)
    es.set_orientation_lists(off, og)
    bp.initMdl()
    res1, _ = es.run(p.fimg, p.fctf, p.meta)
    poses1 = es.counters()["coarse_poses"]
    dist = angular_distance(res0, res1)
    kept = np.mean((dist < 1e-3) & (res0["xoff"] == res1["xoff"]) & (res0["yoff"] == res1["yoff"]))
    fallback = np.mean(np.diff(off) == 0)
    print(f"[configs] proposal box 256: loss {loss:.3f}, fallback {fallback:.1%}, coarse poses {poses1:,} vs "
          f"{poses0:,} ({poses0 / poses1:.0f}x fewer), exhaustive best pose kept for {kept:.2%}", flush=True)
    assert poses1 * 10 < poses0
    assert kept >= 0.99
    es.close()
    pp.close()
