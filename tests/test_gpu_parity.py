"""Parity of the HIP E-step (through the C ABI) against the FP64 oracle on the same seeded inputs.

Tolerances (from the E-step's accuracy contract):
  * per-pose diff2 (coarse and fine) within 1e-5 relative of the oracle;
  * max-likelihood orientation/offset (best_ihidden_over) identical, except near-ties (oracle weights of the
    two picks within 1e-4 relative);
  * backprojected data and weight volumes within 1e-4 relative (norm-wise, per class);
  * significance masks identical except entries whose oracle weight lies within 5% of the threshold (FP32
    coarse diff2 moves weights by exp(+-1e-5 * diff2));
  * sum_weight and pmax within 1e-3 relative, log(sum_weight) - min_diff2 within 1e-3 absolute: the fine diff2's
    pose-independent part is summed in FP64, so the weights see only the FP32 error of the pose-dependent part.
"""
import numpy as np
import pytest

import oracle
import relion_amd as ra
from problems import make_problem

pytestmark = pytest.mark.gpu

# the fine pass's work-item shapes (coarse translations per small item, per wave of a wide item)
FINE_SHAPES = [(2, 2), (1, 2), (2, 1), (1, 1)]


def run_gpu(p, keep=True, bands=None, shape=None):
    cfg = p.cfg
    K = cfg.nr_classes
    dims = ra.volume_dims(cfg.current_size, cfg.padding_factor)
    projs, bps = [], []
    for k in range(K):
        pr = ra.AccProjector(0)
        pr.setMdlDim(*dims, cfg.padding_factor)
        pr.initMdl(p.volumes[k])
        bp = ra.AccBackprojector(0)
        bp.setMdlDim(*dims, cfg.padding_factor)
        bp.initMdl()
        projs.append(pr)
        bps.append(bp)
    es = ra.ExpectationStep(cfg, 0)
    es.set_model(projs, bps, p.pdf_class, p.sigma2_noise, pdf_direction=p.pdf_direction,
                 data_vs_prior=p.data_vs_prior)
    es.keep_intermediates(keep)
    if bands is not None:
        es.set_fine_bands(bands)
    if shape is not None:
        es.set_fine_shape(shape)
    res, wsum = es.run(p.fimg, p.fctf, p.meta, fimg_nomask=p.fimg_nomask)
    bp_out = [b.getMdlData() for b in bps]
    return es, res, wsum, bp_out, (projs, bps)


def run_oracle(p, debug=()):
    return oracle.estep(p.cfg, p.volumes, p.pdf_class, p.sigma2_noise, p.meta, p.fimg, p.fctf,
                        fimg_nomask=p.fimg_nomask, pdf_direction=p.pdf_direction, data_vs_prior=p.data_vs_prior,
                        debug_particles=debug)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


def check_particle(es, ip, dbg, res_g, res_o):
    d2g, wg, mg, (nd, npsi, nt) = es.coarse(ip)
    d2o = dbg["coarse_diff2"]
    computed = d2o >= 0
    assert np.array_equal(d2g >= 0, computed), "coarse: computed-pose sets differ"
    r = rel(d2g[computed], d2o[computed])
    assert r.max() < 1e-5, f"coarse diff2 max rel err {r.max():.3e}"
    # significance: mismatches only near the threshold
    mo = dbg["coarse_significant"].astype(bool)
    thr = res_o["coarse_significant_weight"]
    mism = np.flatnonzero(mo != mg.astype(bool))
    if mism.size:
        wo = dbg["coarse_weight"][mism]
        assert np.all(np.abs(wo - thr) <= 0.05 * thr), f"{mism.size} coarse mask mismatches away from threshold"
    # fine pass on the poses both evaluated
    idx, fd2, fw = es.fine(ip)
    fo = dbg["fine_diff2"]
    common = fo[idx] >= 0
    assert common.mean() > 0.9, "fine pass: too few common poses"
    r = rel(fd2[common], fo[idx][common])
    assert r.max() < 1e-5, f"fine diff2 max rel err {r.max():.3e}"
    # best pose
    bo, bg = int(res_o["best_ihidden_over"]), int(res_g["best_ihidden_over"])
    if bo != bg:
        w = dbg["fine_weight"]
        assert abs(w[bo] - w[bg]) <= 1e-4 * w[bo], f"best pose differs: gpu {bg} oracle {bo}"


def check_run(p, debug=(0, 1), bp_tol=1e-4, bands=None, shape=None):
    es, rg, wg, bpg, _ = run_gpu(p, bands=bands, shape=shape)
    if bands is not None:
        assert es.counters()["fine_bands"] == bands
    if shape is not None:
        c = es.counters()
        assert (c["fine_small_tc"], c["fine_wide_cw"]) == FINE_SHAPES[shape]
    ro, wo, (bpd, bpw), dbg = run_oracle(p, debug)
    n = len(p.meta)
    for ip in debug:
        check_particle(es, ip, dbg[ip], rg[ip], ro[ip])
    same = rg["best_ihidden_over"] == ro["best_ihidden_over"]
    assert same.mean() >= 0.99 or n - same.sum() <= 1, f"best poses differ for {n - same.sum()} of {n} particles"
    for f in ("rot", "tilt", "psi", "xoff", "yoff"):
        assert np.allclose(rg[f][same], ro[f][same], atol=1e-9), f
    assert np.array_equal(rg["iclass"][same], ro["iclass"][same])
    for f in ("min_diff2", "coarse_min_diff2", "dLL"):
        assert rel(rg[f], ro[f]).max() < 1e-5, f"{f}: {rel(rg[f], ro[f]).max():.3e}"
    # the weights' normalisation on its own: log(sum_weight) - min_diff2 is dLL without the pose-independent
    # sum of log(2 pi sigma2), which dominates dLL's magnitude; FP32 diff2 differences move it by ~1e-5
    lse_g = np.log(rg["sum_weight"]) - rg["min_diff2"]
    lse_o = np.log(ro["sum_weight"]) - ro["min_diff2"]
    err_lse = np.abs(lse_g - lse_o).max()
    err_sw = max(rel(rg[f], ro[f]).max() for f in ("sum_weight", "coarse_sum_weight"))
    err_pm = rel(rg["pmax"], ro["pmax"])[same].max()
    print(f"[parity] log(sum_weight)-min_diff2 abs err {err_lse:.2e}, sum_weight rel err {err_sw:.2e}, "
          f"pmax rel err {err_pm:.2e}", flush=True)
    assert err_lse < 1e-3, f"log(sum_weight) - min_diff2: {err_lse:.3e}"
    assert err_sw < 1e-3, f"sum_weight: {err_sw:.3e}"
    assert err_pm < 1e-3, f"pmax: {err_pm:.3e}"
    assert np.array_equal(rg["nr_significant"] > 0, ro["nr_significant"] > 0)
    # weighted sums
    assert rel(wg["sigma2_noise"], wo["sigma2_noise"])[wo["sigma2_noise"] > 0].max() < 1e-3
    assert abs(wg["sumw_group"] - wo["sumw_group"]) < 1e-6 * n
    assert np.allclose(wg["pdf_class"], wo["pdf_class"], atol=1e-6 * n)
    assert np.allclose(wg["pdf_direction"], wo["pdf_direction"], atol=1e-3)
    assert abs(wg["sigma2_offset"] - wo["sigma2_offset"]) <= 1e-3 * max(1.0, abs(wo["sigma2_offset"]))
    assert abs(wg["LL"] - wo["LL"]) <= 1e-5 * abs(wo["LL"])
    if p.cfg.do_scale_correction:
        assert abs(wg["signal_product"] - wo["signal_product"]) <= 1e-3 * abs(wo["signal_product"])
        assert abs(wg["reference_power"] - wo["reference_power"]) <= 1e-3 * abs(wo["reference_power"])
    # backprojection, norm-wise per class
    for k in range(p.cfg.nr_classes):
        re, im, w = bpg[k]
        ref_d = bpd[k]
        if not ref_d.any() and not bpw[k].any():
            # no particle's significant poses in this class: nothing may have been backprojected into it
            print(f"[parity] class {k}: empty in the oracle", flush=True)
            assert not re.any() and not im.any() and not w.any(), f"class {k}: backprojected into an empty class"
            continue
        num = np.sqrt(((re - ref_d[..., 0]) ** 2 + (im - ref_d[..., 1]) ** 2).sum())
        den = np.sqrt((ref_d ** 2).sum())
        errw = np.abs(w - bpw[k]).sum() / np.abs(bpw[k]).sum()
        print(f"[parity] class {k}: BP data rel err {num / den:.2e}, weight {errw:.2e}", flush=True)
        assert num <= bp_tol * den, f"class {k}: BP data rel err {num / den:.3e}"
        assert errw <= bp_tol, f"class {k}: BP weight rel err {errw:.3e}"
    return rg, ro


def test_projector_roundtrip():
    p = make_problem(box=32, px=4.0, order=1, n=1)
    dims = ra.volume_dims(32)
    pr = ra.AccProjector(0)
    assert pr.setMdlDim(*dims, 2.0)
    assert not pr.setMdlDim(*dims, 2.0)
    pr.initMdl(p.volumes[0])
    bp = ra.AccBackprojector(0)
    nbytes = bp.setMdlDim(*dims, 2.0)
    # corner-slot accumulators: 32 floats per trilinear base voxel, allocated in whole bricks of 4 x 4 x 4
    bricks = [-(-(d - 1) // 4) for d in dims[:3]]
    assert nbytes == 4 * 32 * 64 * bricks[0] * bricks[1] * bricks[2]
    bp.initMdl()
    re, im, w = bp.getMdlData()
    assert not re.any() and not im.any() and not w.any()


@pytest.mark.parametrize("bands,shape", [(None, None), (8, None), (None, 0), (8, 1), (None, 2), (8, 3)])
def test_global_search_box64(bands, shape):
    """The reference's CPU-runnable case in shape: box 64, HEALPix order 2, offsets 5 px step 1, global search,
    noisy particles; the fine pass as tuned (one band at this size), split into 8 radial pixel bands, and in each of
    its four work-item shapes."""
    p = make_problem(box=64, px=2.0, order=2, offset_range_px=5, offset_step_px=1, n=8, snr=0.1)
    rg, ro = check_run(p, bands=bands, shape=shape)
    assert (rg["nr_fine_evaluated"] > 0).all()


def test_local_search_two_classes():
    p = make_problem(box=48, px=2.5, order=2, offset_range_px=3, offset_step_px=1, n=8, snr=0.2, nr_classes=2,
                     local=True, sigma_ang=10.0, prior_sigma=4.0, nr_blobs=30, seed=77)
    check_run(p)


def test_scale_offsets_highres_current_size():
    """current_size < ori_size, non-zero old/prior offsets, scale correction, highres_Xi2, separate unmasked image."""
    p = make_problem(box=48, px=3.0, order=1, offset_range_px=3, offset_step_px=1.5, n=6, snr=0.5,
                     current_size=40, scale_correction=True, perturb_meta=True, nr_blobs=25, seed=9,
                     separate_nomask=True)
    check_run(p)


def test_maxsig_cap():
    p = make_problem(box=32, px=4.0, order=1, offset_range_px=3, offset_step_px=1, n=6, snr=0.05, maxsig=5,
                     nr_blobs=20, seed=5)
    rg, ro = check_run(p)
    assert (ro["nr_significant"] <= 5).all()
    assert np.array_equal(rg["nr_significant"], ro["nr_significant"])


def test_four_classes_global():
    """classification-like: K = 4 classes sharing 30% of their blobs, HEALPix order 2 global search, offsets 6 px
    step 1 (113 translations)."""
    p = make_problem(box=48, px=3.0, order=2, offset_range_px=6, offset_step_px=1, n=6, snr=0.2, nr_classes=4,
                     nr_blobs=30, seed=31)
    rg, ro = check_run(p)
    assert len(np.unique(ro["iclass"])) > 1


@pytest.mark.parametrize("shape", [None, 0, 1, 2, 3])
def test_uneven_snr_fine_pass(shape):
    """uneven work: per-particle SNR log-uniform in [0.01, 0.3] spreads the significant-pose counts of the fine
    pass over orders of magnitude, so that both the small and the wide work items run, in each work-item shape."""
    p = make_problem(box=48, px=3.0, order=2, offset_range_px=3, offset_step_px=1, n=8, snr=(0.01, 0.3),
                     nr_blobs=30, seed=47)
    rg, ro = check_run(p, debug=(0, 1, 2), shape=shape)
    fe = rg["nr_fine_evaluated"]
    print(f"[uneven] fine poses per particle: min {fe.min()}, max {fe.max()}")
    assert fe.max() >= 10 * fe.min()
