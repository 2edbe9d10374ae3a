"""Learned pose proposal on the GPU (include/rln_proposal.h) against the numpy oracle (oracle/pose_cnn.py).

  * network input images of particles within 2e-4 (absolute, unit-variance images) of the FP64 inverse DFT;
  * one forward + backward pass of the full-size network (64 x 64 input, 6 x 64 channels, 192 x 24 bins):
    logits, loss and every parameter tensor's gradient against the oracle with BF16 rounding at the same points,
    within 5e-3 relative (norm-wise per tensor; the rest is FP32 versus FP64 accumulation);
  * bins cover the coarse grid: every orientation of an order-3 grid belongs to one or two bins (two on psi bin
    boundaries) and a grid orientation's own bin lists it;
  * the gain: trained on projections of the reference, the proposal seeds an order-3 search that evaluates a
    fraction of the exhaustive coarse poses and keeps the exhaustive best pose.
"""
import numpy as np
import pytest

import relion_amd as ra
from oracle import pose_cnn as pc
from problems import make_problem
from relion_amd.proposal import PoseProposal

pytestmark = pytest.mark.gpu


def test_proposal_input_images_match_oracle():
    p = make_problem(box=64, px=2.0, order=2, n=4, snr=0.1, seed=91)
    print("[proposal] input images, box 64, 4 particles", flush=True)
    pp = PoseProposal(window=64, channels=64, healpix_order=2, n_psi_bins=24, seed=1)
    img = pp.images(p.fimg, 4, 64)
    for i in range(4):
        ref = pc.input_image(p.fimg[i], 64, 64)
        err = np.abs(img[i] - ref).max()
        assert err < 2e-4, f"particle {i}: input image max abs err {err:.2e}"
    # a smaller window of a larger image
    pp32 = PoseProposal(window=32, channels=8, healpix_order=1, n_psi_bins=12, seed=1)
    img = pp32.images(p.fimg, 4, 64)
    for i in range(4):
        assert np.abs(img[i] - pc.input_image(p.fimg[i], 64, 32)).max() < 2e-4
    pp.close()
    pp32.close()


def rel_norm(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_proposal_gradients_match_oracle():
    W, C, order, npsi = 64, 64, 2, 24
    NB = 12 * 4 ** order * npsi
    print("[proposal] forward/backward of the full network, 4 images", flush=True)
    pp = PoseProposal(window=W, channels=C, healpix_order=order, n_psi_bins=npsi, seed=2)
    assert pp.n_bins == NB and pp.n_params == pc.n_params(W, C, NB)
    rng = np.random.default_rng(8)
    w = pc.init_params(W, C, NB, 21)
    conv, Wfc, _ = pc.unpack(w, W, C, NB)
    conv = [(Wl, rng.normal(0, 0.05, C)) for Wl, _ in conv]
    w = pc.pack(conv, Wfc, rng.normal(0, 0.05, NB)).astype(np.float32)
    pp.set_params(w)
    assert np.array_equal(pp.get_params(), w)
    p = make_problem(box=64, px=2.0, order=2, n=4, snr=0.1, seed=92)
    imgs = np.stack([pc.input_image(p.fimg[i], 64, W) for i in range(4)]).astype(np.float32)
    labels = rng.integers(0, NB, 4)
    loss, g, logits = pp.forward_backward(imgs, labels)
    lo, go, logo = pc.loss_and_grads(w.astype(np.float64), imgs.astype(np.float64), labels, C, NB, bf=True)
    e_log = rel_norm(logits, logo)
    print(f"[proposal] logits rel err {e_log:.2e}, loss {loss:.6f} vs {lo:.6f}")
    assert e_log < 5e-3 and abs(loss - lo) < 5e-3 * max(1.0, abs(lo))
    gc, gW, gb = pc.unpack(g, W, C, NB)
    oc, oW, ob = pc.unpack(go, W, C, NB)
    errs = []
    for l, ((a, ab), (b, bb)) in enumerate(zip(gc, oc)):
        errs.append((f"conv{l}.W", rel_norm(a, b)))
        errs.append((f"conv{l}.b", rel_norm(ab, bb)))
    errs += [("fc.W", rel_norm(gW, oW)), ("fc.b", rel_norm(gb, ob))]
    print("[proposal] gradient rel err: " + ", ".join(f"{k} {v:.1e}" for k, v in errs))
    for k, v in errs:
        assert v < 5e-3, f"{k}: gradient rel err {v:.2e}"
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
    pp.close()
