"""Learned pose proposal on the GPU (include/rln_proposal.h) against the numpy oracle (oracle/pose_cnn.py).

  * network input images of particles within 2e-4 (absolute, unit-variance images) of the FP64 inverse DFT;
  * one forward + backward pass of the full-size network (64 x 64 input, 6 x 64 channels, 192 x 24 bins):
    logits, loss and every parameter tensor's gradient against the oracle with BF16 rounding at the same points,
    within twice the BF16 noise floor (norm-wise per tensor): the difference between two oracle runs whose
    pre-rounding values differ by 3e-7 relative, i.e. less than FP32 accumulation error (a few percent for the
    convolution gradients, ~1e-3 for the logits; a wrong index or transpose gives errors of order 1);
  * bins cover the coarse grid: every orientation of an order-3 grid belongs to one or two bins (two on psi bin
    boundaries) and a grid orientation's own bin lists it;
  * the gain: trained on projections of the reference, the proposal seeds an order-3 search that evaluates a
    fraction of the exhaustive coarse poses and keeps the exhaustive best pose for at least 99% of the particles.
"""
import numpy as np
import pytest

import oracle
import relion_amd as ra
from oracle import pose_cnn as pc
from problems import make_problem
from relion_amd import hipmem
from relion_amd.proposal import PoseProposal

pytestmark = pytest.mark.gpu


def test_proposal_input_images_match_oracle():
    p = make_problem(box=64, px=2.0, order=2, n=4, snr=0.1, seed=91)
    print("[proposal] input images, box 64, 4 particles", flush=True)
    pp = PoseProposal(window=64, channels=64, healpix_order=2, n_psi_bins=24, seed=1)
    img = pp.images(p.fimg, None, 4, 64)
    for i in range(4):
        ref = pc.input_image(p.fimg[i], 64, 64)
        err = np.abs(img[i] - ref).max()
        assert err < 2e-4, f"particle {i}: input image max abs err {err:.2e}"
    # phase flipped by the particle's CTF
    img = pp.images(p.fimg, p.fctf, 4, 64)
    for i in range(4):
        assert np.abs(img[i] - pc.input_image(p.fimg[i], 64, 64, p.fctf[i])).max() < 2e-4
    # a smaller window of a larger image
    pp32 = PoseProposal(window=32, channels=8, healpix_order=1, n_psi_bins=12, seed=1)
    img = pp32.images(p.fimg, p.fctf, 4, 64)
    for i in range(4):
        assert np.abs(img[i] - pc.input_image(p.fimg[i], 64, 32, p.fctf[i])).max() < 2e-4
    pp.close()
    pp32.close()


def rel_norm(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("W,C,n", [(64, 64, 4), (64, 64, 3), (32, 16, 3)])
def test_proposal_gradients_match_oracle(W, C, n):
    """64 x 64 input, 64 channels: the LDS-staged convolution kernels; 4 images fill the last stride-2 layer's tiles
    of 4 images (8 x 8 outputs each), 3 end on a partial one. 32 x 32 input, 16 channels: the general kernels
    (im2col GEMMs, direct data gradients) that other network sizes take."""
    order, npsi = 2, 24
    NB = 12 * 4 ** order * npsi
    print(f"[proposal] forward/backward, window {W}, {C} channels, {n} images", flush=True)
    pp = PoseProposal(window=W, channels=C, healpix_order=order, n_psi_bins=npsi, seed=2)
    assert pp.n_bins == NB and pp.n_params == pc.n_params(W, C, NB)
    rng = np.random.default_rng(8)
    w = pc.init_params(W, C, NB, 21)
    conv, Wfc, _ = pc.unpack(w, W, C, NB)
    conv = [(Wl, rng.normal(0, 0.05, C)) for Wl, _ in conv]
    w = pc.pack(conv, Wfc, rng.normal(0, 0.05, NB)).astype(np.float32)
    pp.set_params(w)
    assert np.array_equal(pp.get_params(), w)
    p = make_problem(box=64, px=2.0, order=2, n=n, snr=0.1, seed=92)
    imgs = np.stack([pc.input_image(p.fimg[i], 64, W) for i in range(n)]).astype(np.float32)
    labels = rng.integers(0, NB, n)
    loss, g, logits = pp.forward_backward(imgs, labels)
    w64, i64 = w.astype(np.float64), imgs.astype(np.float64)
    lo, go, logo = pc.loss_and_grads(w64, i64, labels, C, NB, bf=True)
    # noise floor of BF16 arithmetic: the same oracle with its pre-rounding values perturbed by 3e-7 relative
    # (below FP32 accumulation error), which moves some values across BF16 rounding boundaries
    ln, gn, logn = pc.loss_and_grads(w64, i64, labels, C, NB, bf=True, jitter=3e-7, seed=1)
    e_log, n_log = rel_norm(logits, logo), rel_norm(logn, logo)
    print(f"[proposal] logits rel err {e_log:.2e} (BF16 noise floor {n_log:.2e}), loss {loss:.6f} vs {lo:.6f}")
    assert e_log < max(2.0 * n_log, 1e-3) and abs(loss - lo) < 2e-3 * max(1.0, abs(lo))
    gc, gW, gb = pc.unpack(g, W, C, NB)
    oc, oW, ob = pc.unpack(go, W, C, NB)
    nc, nW, nb = pc.unpack(gn, W, C, NB)
    errs = []
    for l in range(len(gc)):
        errs.append((f"conv{l}.W", rel_norm(gc[l][0], oc[l][0]), rel_norm(nc[l][0], oc[l][0])))
        errs.append((f"conv{l}.b", rel_norm(gc[l][1], oc[l][1]), rel_norm(nc[l][1], oc[l][1])))
    errs += [("fc.W", rel_norm(gW, oW), rel_norm(nW, oW)), ("fc.b", rel_norm(gb, ob), rel_norm(nb, ob))]
    print("[proposal] gradient rel err (noise floor): " + ", ".join(f"{k} {v:.1e} ({f:.1e})" for k, v, f in errs))
    for k, v, f in errs:
        assert v < max(2.0 * f, 2e-3), f"{k}: gradient rel err {v:.2e} above twice the BF16 noise floor {f:.2e}"
    # the check hook leaves the weights unchanged
    assert np.array_equal(pp.get_params(), w)
    pp.close()


def test_proposal_bins_cover_the_grid():
    pp = PoseProposal(window=64, channels=8, healpix_order=2, n_psi_bins=24, seed=3)
    n_dir, n_psi = 768, 48
    cover = np.zeros(n_dir * n_psi, np.int32)
    for b in range(pp.n_bins):
        o = pp.bin_orientations(3, n_psi, b)
        assert len(o) == 4 * 3
        cover[o] += 1
    psi_idx = np.arange(n_dir * n_psi) % n_psi
    assert np.all(cover[psi_idx % 2 == 0] == 1) and np.all(cover[psi_idx % 2 == 1] == 2)
    cfg = ra.make_config(64, 2.0, 3, 10.0, 2.0, 24)
    rot, tilt, psi = ra.sampling_grid(cfg)[:3]
    rng = np.random.default_rng(4)
    for og in rng.choice(n_dir * n_psi, 200, replace=False):
        d, j = divmod(int(og), n_psi)
        b = pp.pose_bin(rot[d], tilt[d], psi[j])
        assert og in pp.bin_orientations(3, n_psi, b)
    # orientation lists: the sorted union of the top bins' orientations (empty below the confidence floor), also
    # when the grid changes between calls
    bins = rng.integers(0, pp.n_bins, size=(5, 4)).astype(np.int32)
    probs = np.array([[0.5, 0.2, 0.1, 0.1], [0.1, 0.05, 0.05, 0.05], [0.9, 0.0, 0.0, 0.0], [0.3, 0.3, 0.2, 0.1],
                      [0.25, 0.25, 0.25, 0.25]], np.float32)
    for order, npsi in ((3, n_psi), (4, 96), (3, n_psi)):
        sets = pp.orientation_sets(bins, probs, order, npsi, min_top1=0.2)
        for i in range(5):
            want = (np.unique(np.concatenate([pp.bin_orientations(order, npsi, int(b)) for b in bins[i]]))
                    if probs[i, 0] >= 0.2 else np.zeros(0))
            assert np.array_equal(sets[i], want), (order, i)
    pp.close()


def angular_distance(r0, r1):
    """Rotation angle (degrees) between the poses of two result arrays."""
    out = []
    for a, b in zip(r0, r1):
        A = oracle.euler(a["rot"], a["tilt"], a["psi"])
        B = oracle.euler(b["rot"], b["tilt"], b["psi"])
        c = (np.trace(A @ B.T) - 1.0) / 2.0
        out.append(np.degrees(np.arccos(np.clip(c, -1.0, 1.0))))
    return np.array(out)


def test_proposal_seeds_order3_search_and_keeps_best_pose():
    n = 128
    p = make_problem(box=64, px=2.0, order=3, n=n, snr=0.1, seed=93, offset_range_px=3, offset_step_px=1)
    cfg = p.cfg
    print(f"[proposal] gain test: box 64, order 3 global search, {n} particles", flush=True)
    dims = ra.volume_dims(64)
    pr = ra.AccProjector(0)
    pr.setMdlDim(*dims, 2.0)
    pr.initMdl(p.volumes[0])
    bp = ra.AccBackprojector(0)
    bp.setMdlDim(*dims, 2.0)
    bp.initMdl()
    es = ra.ExpectationStep(cfg, 0)
    es.set_model([pr], [bp], p.pdf_class, p.sigma2_noise, pdf_direction=p.pdf_direction)
    res0, _ = es.run(p.fimg, p.fctf, p.meta)
    poses0 = es.counters()["coarse_poses"]
    pp = PoseProposal(window=64, channels=64, healpix_order=2, n_psi_bins=24, seed=5)
    for ep in range(5):  # fresh projections every epoch
        pp.clear_training()
        pp.generate(pr, 100000, ori_size=64, pixel_size=2.0, sigma2_noise=p.sigma2_noise, defocus=(1.0e4, 3.0e4),
                    max_astigmatism=300.0, offset_max=3.0, first=ep * 100000)
        loss = pp.train(epochs=1, batch=256, lr=1e-3)
        print(f"[proposal] epoch {ep}: loss {loss:.3f}", flush=True)
    bins, probs = pp.predict(p.fimg, p.fctf, n, 64, topk=16)
    sets = pp.orientation_sets(bins, probs, 3, 48, min_top1=0.2)
    fallback = np.mean([len(s) == 0 for s in sets])
    es.set_orientation_sets(sets)
    res1, _ = es.run(p.fimg, p.fctf, p.meta)
    poses1 = es.counters()["coarse_poses"]
    dist = angular_distance(res0, res1)
    kept = np.mean((dist < 1e-3) & (res0["xoff"] == res1["xoff"]) & (res0["yoff"] == res1["yoff"]))
    # true-bin rank of the generated-data classifier on these particles
    tb = np.array([pp.pose_bin(*p.pose[i, :3]) for i in range(n)])
    in_top = np.mean([tb[i] in bins[i] for i in range(n)])
    print(f"[proposal] final training loss {loss:.3f}; true bin in top-16 for {in_top:.0%}; fallback {fallback:.0%}; "
          f"coarse poses {poses1:,} vs exhaustive {poses0:,} ({poses0 / poses1:.1f}x fewer); "
          f"exhaustive best pose kept for {kept:.1%}")
    assert poses1 * 3 < poses0
    assert kept >= 0.99
    es.close()
    pp.close()


def test_proposal_predict_matches_training_forward():
    """Inference runs the convolutions as implicit GEMMs (no im2col buffer; the 64-channel layers with their input
    rows staged in LDS); training runs im2col + GEMM. Both contract the same fragments in the same k order, so the
    top-k bins agree exactly. 9 particles: the last stride-2 layer's tiles of two images end on a partial one."""
    n = 9
    p = make_problem(box=64, px=2.0, order=2, n=n, snr=0.1, seed=94)
    pp = PoseProposal(window=64, channels=64, healpix_order=2, n_psi_bins=24, seed=11)
    rng = np.random.default_rng(12)
    w = pp.get_params()
    w += rng.normal(0, 0.01, w.size).astype(np.float32)  # non-zero biases
    pp.set_params(w)
    imgs = pp.images(p.fimg, p.fctf, n, 64)
    _, _, logits = pp.forward_backward(imgs, np.zeros(n, np.int32))
    bins, probs = pp.predict(p.fimg, p.fctf, n, 64, topk=16)
    z = logits.astype(np.float64)
    pr = np.exp(z - z.max(1, keepdims=True))
    pr /= pr.sum(1, keepdims=True)
    for i in range(n):
        ref = np.argsort(-z[i], kind="stable")[:16]
        assert np.array_equal(bins[i], ref), f"particle {i}: top-k bins differ"
        assert np.allclose(probs[i], pr[i][ref], rtol=1e-5, atol=1e-7)
    pp.close()


def test_proposal_predict_in_two_halves():
    """predict_launch + predict_finish return what predict returns, host and device inputs; one prediction in
    flight at a time; the launch overlaps an E-step on its own stream without changing either result."""
    import threading
    p = make_problem(box=64, px=2.0, order=2, n=24, snr=0.1, seed=95)
    print("[proposal] predict in two halves, 24 particles, box 64", flush=True)
    pp = PoseProposal(window=64, channels=64, healpix_order=2, n_psi_bins=24, seed=13)
    rng = np.random.default_rng(14)
    w = pp.get_params()
    w += rng.normal(0, 0.01, w.size).astype(np.float32)
    pp.set_params(w)
    ref_b, ref_p = pp.predict(p.fimg, p.fctf, 24, 64, topk=8)
    pp.predict_launch(p.fimg, p.fctf, 24, 64, topk=8)
    with pytest.raises(ra.RelionError):
        pp.predict_launch(p.fimg, p.fctf, 24, 64, topk=8)  # one in flight
    b, pr = pp.predict_finish()
    assert np.array_equal(b, ref_b) and np.array_equal(pr, ref_p)
    with pytest.raises(ra.RelionError):
        pp.predict_finish()  # nothing in flight
    d_img, d_ctf = hipmem.DeviceArray(p.fimg), hipmem.DeviceArray(p.fctf)
    # an E-step in a second thread while the prediction runs
    dims = ra.volume_dims(64)
    proj = ra.AccProjector(0)
    proj.setMdlDim(*dims, 2.0)
    proj.initMdl(p.volumes[0])
    bp = ra.AccBackprojector(0)
    bp.setMdlDim(*dims, 2.0)
    bp.initMdl()
    es = ra.ExpectationStep(p.cfg, 0)
    es.set_model([proj], [bp], p.pdf_class, p.sigma2_noise, pdf_direction=p.pdf_direction)
    alone, _ = es.run(p.fimg, p.fctf, p.meta)
    out = {}
    th = threading.Thread(target=lambda: out.setdefault("res", es.run(p.fimg, p.fctf, p.meta)[0]))
    pp.predict_launch(d_img.address, d_ctf.address, 24, 64, topk=8, on_device=True)
    th.start()
    b2, pr2 = pp.predict_finish()
    th.join()
    assert np.array_equal(b2, ref_b) and np.array_equal(pr2, ref_p)
    assert np.array_equal(out["res"]["best_ihidden_over"], alone["best_ihidden_over"])
    es.close()
    pp.close()


def test_proposal_trains_on_ml_assignments():
    """rln_proposal_add_images: particle images labelled with the E-step's own max-likelihood poses enter training
    with their labels. From one warm start (generated projections), fine-tuning on the particles with their ML poses
    must lower the cross entropy of those poses' bins, and by more than fine-tuning on the same images with the
    poses shuffled among the particles (the control): a mix-up of images and labels in add_images would not."""
    n = 256
    p = make_problem(box=64, px=2.0, order=3, n=n, snr=0.1, seed=96, offset_range_px=3, offset_step_px=1)
    print(f"[proposal] ML-assignment training: box 64, order 3 exhaustive E-step of {n} particles", flush=True)
    dims = ra.volume_dims(64)
    pr = ra.AccProjector(0)
    pr.setMdlDim(*dims, 2.0)
    pr.initMdl(p.volumes[0])
    bp = ra.AccBackprojector(0)
    bp.setMdlDim(*dims, 2.0)
    bp.initMdl()
    es = ra.ExpectationStep(p.cfg, 0)
    es.set_model([pr], [bp], p.pdf_class, p.sigma2_noise, pdf_direction=p.pdf_direction)
    res, _ = es.run(p.fimg, p.fctf, p.meta)
    es.close()
    ml = np.stack([res["rot"], res["tilt"], res["psi"]], 1)
    pp = PoseProposal(window=64, channels=64, healpix_order=2, n_psi_bins=24, seed=7)
    for ep in range(2):
        pp.clear_training()
        pp.generate(pr, 50000, ori_size=64, pixel_size=2.0, sigma2_noise=p.sigma2_noise, defocus=(1.0e4, 3.0e4),
                    max_astigmatism=300.0, offset_max=3.0, first=ep * 50000)
        pp.train(epochs=1, batch=256, lr=1e-3)
    w0 = pp.get_params()
    imgs = pp.images(p.fimg, p.fctf, n, 64)
    labels = np.array([pp.pose_bin(*ml[i]) for i in range(n)], np.int32)
    loss0 = pp.forward_backward(imgs, labels)[0]
    perm = np.random.default_rng(4).permutation(n)
    losses = {}
    for name, poses in (("ml", ml), ("shuffled", ml[perm])):
        pp.set_params(w0)
        pp.clear_training()
        pp.add_images(p.fimg, p.fctf, 64, poses)
        assert pp.training_size() == n
        pp.train(epochs=6, batch=64, lr=3e-4)
        losses[name] = pp.forward_backward(imgs, labels)[0]
    print(f"[proposal] cross entropy of the ML bins: warm start {loss0:.3f}, fine-tuned on ML poses "
          f"{losses['ml']:.3f}, on shuffled poses {losses['shuffled']:.3f}", flush=True)
    assert losses["ml"] < loss0 - 0.1
    assert losses["ml"] < losses["shuffled"] - 0.5
    pp.close()
