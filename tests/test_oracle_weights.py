"""CPU checks that pin the oracle's weights, weighted sums, dLL and local-search priors.

The E-step's weights, spectra and likelihood are pinned here by inputs whose answers are fixed by the
mathematics, not by a second copy of the oracle's formulas:
  * a delta posterior: a particle built as the exact fine-pose projection (CTF, scale and fine translation
    applied with the already pinned projector and shift) plus small noise N, with a noise model so tight that
    every other pose is exp(-huge) less likely. Then the weighted sums collapse onto that pose and must equal
    sums over N alone: sigma2_noise per shell = sum |N|^2 over the shell, min diff2 = Xi2/2 + sum |N|^2 / (2
    sigma2), dLL = log(normalised prior of the pose) - min diff2 - sum log(2 pi sigma2), XA - scale * AA =
    sum Re(CTF P conj(N)), sigma2_offset = px^2 |prior - old - t|^2 (reference: src/ml_optimiser.cpp:8626-8698,
    9030-9058);
  * a zero reference: every pose has the same diff2, so the weights are the priors alone, mean-normalised over
    orientations and translations (src/ml_optimiser.cpp:7790-8050): their sum is the number of coarse poses,
    scaling pdf_direction changes nothing, and the offset prior is the Gaussian of old offset (pixels) plus
    translation (Angstrom) against the prior offset, times px^2 (the reference adds the two units as they are,
    :7838);
  * local-search priors: the directions kept are exactly those within sigma_cutoff * sigma of the prior
    direction (angles between view directions taken from the pinned Euler matrices), weighted by the
    normalised Gaussian of that angle; psi likewise (src/healpix_sampling.cpp:695-1076).
A dropped term, a wrong sign or index, a prior evaluated at the oversampled instead of the coarse translation,
or a missing mean normalisation fails one of them.
"""
import numpy as np
import pytest

import oracle
import relion_amd as ra


def shells(s):
    """Resolution shell ROUND(|k|) per pixel of an s x (s/2+1) half-plane; -1 beyond s/2 and on the x == 0,
    y < 0 half column that the half-plane stores twice."""
    xs = s // 2 + 1
    i = np.arange(s)
    y = np.where(i < xs, i, i - s)[:, None]
    x = np.arange(xs)[None, :]
    ires = np.floor(np.sqrt(y * y + x * x) + 0.5).astype(int)
    ires[ires >= xs] = -1
    ires[(x == 0) & (y < 0)] = -1
    return ires


def offset_prior(cfg, meta, tx, ty):
    """Gaussian offset prior of each coarse translation (Angstrom) for old/prior offsets in pixels,
    normalised to mean 1."""
    ox = meta.old_offset_x + tx - meta.prior_offset_x
    oy = meta.old_offset_y + ty - meta.prior_offset_y
    g = np.exp(-cfg.pixel_size ** 2 * (ox * ox + oy * oy) / (2.0 * cfg.sigma2_offset))
    return g / g.mean()


def _meta(old=(1.0, -1.0), prior=(0.4, -0.3), scale=1.0, xi2=0.0):
    meta = (ra.ParticleMeta * 1)()
    m = meta[0]
    m.old_offset_x, m.old_offset_y = old
    m.prior_offset_x, m.prior_offset_y = prior
    m.scale = scale
    m.highres_Xi2 = xi2
    return meta


def test_delta_posterior_pins_weights_wsum_and_dll():
    s, px = 24, 3.0
    cfg = ra.make_config(s, px, 1, 2 * px, px, s, do_scale_correction=True, sigma2_offset=20.0)
    spec = ra.synth_spec(17, s, px, 12, snr=1.0)
    vol = ra.synth_reference(spec, 0, s)
    rot, tilt, psi, tx, ty = oracle.sampling_grid(cfg)
    nd, npsi, nt = len(rot), len(psi), len(tx)
    idir, ipsi, it, r, u = 20, 5, 3, 6, 1
    orot, otilt, opsi = oracle.oversampled_orientations(cfg, idir, ipsi)
    ox, oy = oracle.oversampled_translations(cfg, it)  # pixels
    P = oracle.project(vol, s // 2, 2.0, oracle.euler(orot[r], otilt[r], opsi[r]), s)
    rng = np.random.default_rng(5)
    ctf = rng.uniform(0.3, 1.0, size=P.shape)
    scale, xi2 = 1.25, 2.5
    sig0 = np.mean(np.abs(ctf * P) ** 2) / 500.0
    nshell = s // 2 + 1
    sig2 = sig0 * (1.0 + 0.3 * np.cos(np.arange(nshell) / 2.0))
    N = np.sqrt(sig0 / 2) * (rng.normal(size=P.shape) + 1j * rng.normal(size=P.shape))
    X = oracle.shift_image(scale * ctf * P, float(s), -ox[u], -oy[u]) + N
    meta = _meta(scale=scale, xi2=xi2)
    dvp = np.where(np.arange(nshell) < 8, 10.0, 1.0)[None, :]
    res, ws, _, dbg = oracle.estep(cfg, [vol], [1.0], sig2, meta, X[None], ctf[None], pdf_direction=np.ones((1, nd)),
                                   data_vs_prior=dvp, debug_particles=[0], with_bp=False)
    res = res[0]
    # the pose
    assert res["best_ihidden_over"] == ((((idir * npsi + ipsi) * nt + it) * 8 + r) * 4 + u)
    assert (res["rot"], res["tilt"], res["psi"]) == (orot[r], otilt[r], opsi[r])
    assert res["xoff"] == pytest.approx(meta[0].old_offset_x + ox[u], abs=1e-12)
    assert res["yoff"] == pytest.approx(meta[0].old_offset_y + oy[u], abs=1e-12)
    assert res["pmax"] == pytest.approx(1.0, abs=1e-12)
    # the residual of the pose is the noise, shifted by the fine translation (|phase| = 1)
    Ns = oracle.shift_image(N, float(s), ox[u], oy[u])
    ires = shells(s)
    inv = np.zeros(P.shape)
    inv[ires > 0] = 1.0 / sig2[ires[ires > 0]]
    min_d2 = xi2 / 2 + 0.5 * np.sum(np.abs(Ns) ** 2 * inv)
    assert res["min_diff2"] == pytest.approx(min_d2, rel=1e-11)
    # weights: the pose's normalised prior (uniform directions: 1; offsets: the coarse translation's Gaussian)
    pf = offset_prior(cfg, meta[0], tx, ty)
    assert res["sum_weight"] == pytest.approx(pf[it], rel=1e-11)
    logsig = np.sum(np.log(2 * np.pi * sig2[ires[ires > 0]]))
    assert res["dLL"] == pytest.approx(np.log(pf[it]) - min_d2 - logsig, abs=1e-8)
    # weighted sums onto the one pose
    exp_s2 = np.array([np.sum(np.abs(N[ires == j]) ** 2) for j in range(nshell)])
    assert np.allclose(ws["sigma2_noise"], exp_s2, rtol=1e-10, atol=0)
    assert res["wsum_norm_correction"] == pytest.approx(exp_s2.sum(), rel=1e-10)
    sel = (ires > -1) & (dvp[0][np.maximum(ires, 0)] > 3)
    CP = ctf * P
    aa = np.sum(np.abs(CP[sel]) ** 2)
    xa = np.sum((CP[sel] * np.conj(scale * CP[sel] + Ns[sel])).real)
    assert res["reference_power"] == pytest.approx(aa, rel=1e-10)
    assert res["signal_product"] == pytest.approx(xa, rel=1e-10)
    dx = meta[0].prior_offset_x - meta[0].old_offset_x - ox[u]
    dy = meta[0].prior_offset_y - meta[0].old_offset_y - oy[u]
    assert ws["sigma2_offset"] == pytest.approx(px * px * (dx * dx + dy * dy), rel=1e-10)
    assert ws["sumw_group"] == pytest.approx(1.0, abs=1e-12)
    assert ws["pdf_class"][0] == pytest.approx(1.0, abs=1e-12)
    pd = np.zeros(nd)
    pd[idir] = 1.0
    assert np.allclose(ws["pdf_direction"][0], pd, atol=1e-12)
    assert ws["LL"] == pytest.approx(res["dLL"], abs=1e-12) and ws["ave_Pmax"] == pytest.approx(1.0, abs=1e-12)


def _zero_reference_run(pdf_direction, cfg, meta, s, K, rng):
    vol = np.zeros_like(ra.synth_reference(ra.synth_spec(3, s, cfg.pixel_size, 4, snr=1.0), 0, s))
    X = rng.normal(size=(1, s, s // 2 + 1)) + 1j * rng.normal(size=(1, s, s // 2 + 1))
    ctf = rng.uniform(0.3, 1.0, size=(1, s, s // 2 + 1))
    sig2 = np.full(s // 2 + 1, 0.7)
    return oracle.estep(cfg, [vol] * K, np.full(K, 1.0 / K), sig2, meta, X, ctf, pdf_direction=pdf_direction,
                        debug_particles=[0], with_bp=False)


def test_zero_reference_weights_are_normalised_priors():
    s, px, K = 16, 3.0, 2
    cfg = ra.make_config(s, px, 1, 2 * px, px, s, nr_classes=K, sigma2_offset=15.0)
    rot, tilt, psi, tx, ty = oracle.sampling_grid(cfg)
    nd, npsi, nt = len(rot), len(psi), len(tx)
    meta = _meta(old=(1.0, -2.0), prior=(0.5, 0.25))
    pdfd = 0.5 + np.random.default_rng(2).random((K, nd))
    res, _, _, dbg = _zero_reference_run(pdfd, cfg, meta, s, K, np.random.default_rng(9))
    d = dbg[0]
    n_coarse = K * nd * npsi * nt
    # every pose has the same diff2 ...
    assert np.ptp(d["coarse_diff2"]) <= 1e-12 * d["coarse_diff2"].max()
    # ... so the weights are the mean-normalised priors: direction prior x offset prior of the coarse translation
    po = np.repeat(pdfd / pdfd.mean(), npsi, axis=1).reshape(K, nd, npsi, 1)
    pf = offset_prior(cfg, meta[0], tx, ty).reshape(1, 1, 1, nt)
    expected = (po * pf).ravel()
    assert np.allclose(d["coarse_weight"], expected, rtol=1e-10, atol=0)
    assert res[0]["coarse_sum_weight"] == pytest.approx(n_coarse, rel=1e-10)
    # fine pass: the 32 oversampled poses of a significant coarse pose share its priors
    m = d["coarse_significant"].astype(bool)
    fw = d["fine_weight"].reshape(n_coarse, 32)
    assert np.allclose(fw[m], expected[m][:, None], rtol=1e-10, atol=0)
    assert not fw[~m].any()
    assert res[0]["sum_weight"] == pytest.approx(32 * expected[m].sum(), rel=1e-10)
    # the offset prior is not the one of the oversampled translations, nor of old offsets converted to Angstrom
    wrong = offset_prior(cfg, meta[0], tx + 0 * tx, ty)  # same translations: sanity of the helper
    assert np.allclose(wrong, pf.ravel())
    m2 = _meta(old=(1.0 * px, -2.0 * px), prior=(0.5 * px, 0.25 * px))
    assert not np.allclose(offset_prior(cfg, m2[0], tx, ty), pf.ravel(), rtol=1e-3)
    # mean normalisation: scaling pdf_direction leaves every weight unchanged
    res5, _, _, dbg5 = _zero_reference_run(5.0 * pdfd, cfg, meta, s, K, np.random.default_rng(9))
    assert np.allclose(dbg5[0]["coarse_weight"], d["coarse_weight"], rtol=1e-12, atol=0)
    assert res5[0]["sum_weight"] == pytest.approx(res[0]["sum_weight"], rel=1e-12)
    assert res5[0]["dLL"] == pytest.approx(res[0]["dLL"], abs=1e-10)


def _view_angle(rot_a, tilt_a, rot_b, tilt_b):
    """Angle (degrees) between the viewing directions of two Euler angle pairs, from the Euler matrices:
    the direction of (rot, tilt) is A^T e_z, so cos = (A_a A_b^T)[2, 2]."""
    Aa = oracle.euler(rot_a, tilt_a, 0.0)
    Ab = oracle.euler(rot_b, tilt_b, 0.0)
    return np.degrees(np.arccos(np.clip((Aa @ Ab.T)[2, 2], -1.0, 1.0)))


@pytest.mark.parametrize("sigma", [7.5, 12.0, 20.0])
def test_local_priors_are_truncated_gaussians(sigma):
    cfg = ra.make_config(64, 2.0, 2, 10.0, 2.0, 24, local_search=True, sigma_ang=sigma)
    rot, tilt, psi, _, _ = oracle.sampling_grid(cfg)
    rng = np.random.default_rng(int(sigma * 10))
    for _ in range(6):
        pr, pt, pp = rng.uniform(-180, 180), rng.uniform(0, 180), rng.uniform(-180, 360)
        di, dp, pi, ppr = oracle.sampling_local(cfg, pr, pt, pp)
        ang = np.array([_view_angle(rot[d], tilt[d], pr, pt) for d in range(len(rot))])
        keep = np.flatnonzero(ang < 3.0 * sigma)
        assert np.array_equal(di, keep)
        g = np.exp(-0.5 * (ang[keep] / sigma) ** 2)
        assert np.allclose(dp, g / g.sum(), rtol=1e-9, atol=0)
        dpsi = np.mod(psi - pp, 360.0)
        dpsi = np.minimum(dpsi, 360.0 - dpsi)
        keep = np.flatnonzero(dpsi < 3.0 * sigma)
        assert np.array_equal(pi, keep)
        g = np.exp(-0.5 * (dpsi[keep] / sigma) ** 2)
        assert np.allclose(ppr, g / g.sum(), rtol=1e-9, atol=0)


def test_local_priors_fall_back_to_the_nearest_sample():
    sigma = 0.01
    cfg = ra.make_config(64, 2.0, 2, 10.0, 2.0, 24, local_search=True, sigma_ang=sigma)
    rot, tilt, psi, _, _ = oracle.sampling_grid(cfg)
    pr, pt, pp = 33.3, 71.7, 101.1  # no sample within 3 sigma
    di, dp, pi, ppr = oracle.sampling_local(cfg, pr, pt, pp)
    ang = np.array([_view_angle(rot[d], tilt[d], pr, pt) for d in range(len(rot))])
    assert ang.min() > 3 * sigma
    assert list(di) == [int(np.argmin(ang))] and list(dp) == [1.0]
    dpsi = np.abs(psi - pp)
    assert list(pi) == [int(np.argmin(dpsi))] and list(ppr) == [1.0]
