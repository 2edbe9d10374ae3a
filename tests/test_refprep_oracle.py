"""Pins of the reference-preparation oracle (oracle/refprep.py): closed forms of deltas, the central-slice theorem
with the E-step's image normalisation, the product's analytic Fourier reference, and the power spectrum."""
import numpy as np
import pytest

from oracle import refprep as rp


def test_sizes_follow_the_reference_rules():
    # padoridim = ROUND(pad * ori) made even; pad_size = 2 (ROUND(pad r_max) + 1) + 1
    assert rp.sizes(64, 64, 2.0)[:4] == (128, 2.0, 32, 131)
    assert rp.sizes(64, 40, 2.0)[3] == 2 * (40 + 1) + 1
    pd, pad, _, _, _ = rp.sizes(33, 33, 2.0)
    assert pd == 66 and pad == 2.0
    pd, pad, _, _, _ = rp.sizes(30, 30, 1.5)
    assert pd == 46 and pad == pytest.approx(46 / 30)


def test_centred_delta_is_flat():
    ori = 16
    vol = np.zeros((ori,) * 3)
    vol[ori // 2, ori // 2, ori // 2] = 1.0
    data, _ = rp.compute_fourier_map(vol, ori, 2.0, do_gridding=True)
    _, _, _, pad_size, max_r2 = rp.sizes(ori, ori, 2.0)
    h = pad_size // 2
    z = np.arange(pad_size) - h
    Z, Y, X = np.meshgrid(z, z, np.arange(h + 1), indexing="ij")
    inside = (Z * Z + Y * Y + X * X <= max_r2) & (Z > -ori) & (Y > -ori)
    assert np.allclose(data[inside], 1.0 / ori ** 2)
    assert np.all(data[~inside] == 0)


def test_displaced_delta_is_ramp_over_gridding_factor():
    ori, a = 16, (2, -3, 1)  # x, y, z displacement
    vol = np.zeros((ori,) * 3)
    vol[ori // 2 + a[2], ori // 2 + a[1], ori // 2 + a[0]] = 1.0
    data, _ = rp.compute_fourier_map(vol, ori, 2.0, do_gridding=True)
    pd = 2 * ori
    r = np.sqrt(sum(v * v for v in a)) / pd
    g = (np.sin(np.pi * r) / (np.pi * r)) ** 2
    h = data.shape[0] // 2
    for (kz, ky, kx) in ((0, 0, 1), (3, -2, 5), (-7, 4, 2)):
        expect = np.exp(-2j * np.pi * (kx * a[0] + ky * a[1] + kz * a[2]) / pd) / ori ** 2 / g
        assert data[kz + h, ky + h, kx] == pytest.approx(expect, rel=1e-12)
    # without gridding the ramp is unscaled
    data2, _ = rp.compute_fourier_map(vol, ori, 2.0, do_gridding=False)
    assert data2[h, h, 1] == pytest.approx(np.exp(-2j * np.pi * a[0] / pd) / ori ** 2, rel=1e-12)


def test_central_section_is_the_transform_of_the_projection():
    # identity rotation: data(pad kx, pad ky, 0) = FT2(sum_z vol) / ori^2 with the image's centring sign
    rng = np.random.default_rng(3)
    ori = 24
    vol = rng.normal(size=(ori,) * 3)
    data, _ = rp.compute_fourier_map(vol, ori, 2.0, do_gridding=False)
    proj = vol.sum(axis=0)
    F2 = np.fft.rfft2(proj) / ori ** 2
    s = np.where(((np.arange(ori)[:, None] + np.arange(ori // 2 + 1)[None, :]) & 1) == 1, -1.0, 1.0)
    F2 = F2 * s
    h = data.shape[0] // 2
    for ky in range(-ori // 2 + 1, ori // 2):
        for kx in range(0, ori // 2):
            if 4 * (kx * kx + ky * ky) > rp.sizes(ori, ori, 2.0)[4]:
                continue
            assert data[h, h + 2 * ky, 2 * kx] == pytest.approx(F2[ky % ori, kx], rel=1e-9, abs=1e-12)


def test_blob_map_matches_the_analytic_reference():
    # a map of Gaussian blobs sampled in real space, without gridding correction, against the product's analytic
    # Fourier-space reference of the same blobs (an independent formula)
    import relion_amd as ra
    ori = 32
    spec = ra.synth_spec(5, ori, 2.0, 12, snr=1.0, blob_radius=5.0)
    vol = ra.synth_real_map(spec, 0, ori)
    data, _ = rp.compute_fourier_map(vol, ori, 2.0, do_gridding=False)
    ana = ra.synth_reference(spec, 0, ori, 2.0)
    ana = ana[..., 0] + 1j * ana[..., 1]
    assert ana.shape == data.shape
    err = np.abs(data - ana).max() / np.abs(ana).max()
    assert err < 1e-3, err


def test_power_spectrum_of_a_cosine():
    ori = 32
    z, y, x = np.mgrid[:ori, :ori, :ori]
    # padding factor 1 keeps the wave periodic in the transformed box: one frequency, |k| = 6.4 -> shell 6
    vol = np.cos(2 * np.pi * (5 * x + 4 * y) / ori)
    data, ps = rp.compute_fourier_map(vol, ori, 1.0, do_gridding=False)
    big = np.argsort(ps)[-1]
    assert big == 6
    assert ps[6] > 1e12 * np.delete(ps, 6).max()
    # the wave's frequencies are (5, 4, 0) and (-5, -4, 0); the half transform keeps (5, 4, 0), where the unscaled
    # transform is ori^3 / 2 and data = transform / ori^2
    h = data.shape[0] // 2
    assert abs(data[h, h + 4, 5]) == pytest.approx(ori / 2)
