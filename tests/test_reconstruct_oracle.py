"""Pins of the reconstruction oracle (oracle/reconstruct.py): the blob transform against scipy's Bessel functions,
a round trip through reference preparation, scale equivariance of the gridding iterations, and the MAP limits."""
import numpy as np
import pytest
from scipy import special

from oracle import reconstruct as rc
from oracle import refprep as rp


def test_blob_transform_matches_scipy_bessel_functions():
    a, alpha = rc.BLOB_RADIUS, rc.BLOB_ALPHA
    w = np.linspace(0.0, 0.5, 2001)[1:]
    t = 2 * np.pi * a * w
    s = np.sqrt(np.abs(alpha ** 2 - t ** 2))
    lo = t <= alpha
    ref = np.empty_like(w)
    ref[lo] = special.iv(1.5, s[lo]) / s[lo] ** 1.5
    ref[~lo] = special.jv(1.5, s[~lo]) / s[~lo] ** 1.5
    ref *= (2 * np.pi) ** 1.5 * a ** 3 / special.i0(alpha)
    ok = np.abs(s) > 1e-6
    assert np.allclose(rc.kaiser_ft(w)[ok], ref[ok], rtol=1e-9, atol=1e-12 * np.abs(ref).max())
    # value at 0: the blob's volume integral (2 pi / alpha)^{3/2} a^3 I_{3/2}(alpha) / I_0(alpha)
    v0 = (2 * np.pi / alpha) ** 1.5 * a ** 3 * special.iv(1.5, alpha) / special.i0(alpha)
    assert rc.kaiser_ft(0.0) == pytest.approx(v0, rel=1e-12)
    # the table holds the value at the floor of the requested position
    tab = rc.blob_table()
    assert rc.blob_lookup(tab, 0.123456) == tab[int(0.123456 / (0.5 / len(tab)))]
    assert rc.blob_lookup(tab, 0.5) == 0.0


def _smooth_map(ori, seed):
    rng = np.random.default_rng(seed)
    c = np.arange(ori) - ori // 2
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    vol = np.zeros((ori,) * 3)
    for _ in range(6):
        p = rng.uniform(-ori / 6, ori / 6, size=3)
        s = rng.uniform(2.0, 3.0)
        vol += rng.uniform(0.5, 1.0) * np.exp(-((X - p[0]) ** 2 + (Y - p[1]) ** 2 + (Z - p[2]) ** 2) / (2 * s * s))
    return vol


def test_round_trip_through_reference_preparation():
    # data = transform of a band-limited map (no gridding correction), uniform weight c: the reconstruction is the map
    # divided by c and by the final sinc^2 gridding correction
    ori, c = 24, 3.0
    vol = _smooth_map(ori, 1)
    data, _ = rp.compute_fourier_map(vol, ori, 2.0, do_gridding=False)
    weight = np.full(data.shape, c)
    out = rc.reconstruct(data, weight, ori, 2.0, do_map=False, max_iter_preweight=10)
    cc = np.arange(ori) - ori // 2
    r = np.sqrt(cc[:, None, None] ** 2 + cc[None, :, None] ** 2 + cc[None, None, :] ** 2)
    s = np.where(r > 0, np.sin(np.pi * r / (2 * ori)) / (np.pi * r / (2 * ori) + (r == 0)), 1.0)
    expect = vol / c / (s * s)
    inner = r < ori / 2 - 4
    err = np.abs(out - expect)[inner].max() / np.abs(expect).max()
    assert err < 2e-2, err


def test_gridding_iterations_are_scale_equivariant():
    ori = 16
    rng = np.random.default_rng(2)
    data, _ = rp.compute_fourier_map(_smooth_map(ori, 3), ori, 2.0)
    weight = rng.uniform(0.5, 2.0, size=data.shape)
    base = rc.reconstruct(data, weight, ori, 2.0, do_map=False, max_iter_preweight=4)
    both = rc.reconstruct(3.0 * data, 3.0 * weight, ori, 2.0, do_map=False, max_iter_preweight=4)
    assert np.allclose(both, base, rtol=1e-8, atol=1e-12 * np.abs(base).max())
    wonly = rc.reconstruct(data, 2.0 * weight, ori, 2.0, do_map=False, max_iter_preweight=4)
    assert np.allclose(wonly, base / 2.0, rtol=1e-8, atol=1e-12 * np.abs(base).max())
    norm = rc.reconstruct(data, weight, ori, 2.0, do_map=False, max_iter_preweight=4, normalise=5.0)
    assert np.allclose(norm, base, rtol=1e-8, atol=1e-12 * np.abs(base).max())


def test_map_limits():
    ori = 16
    data, _ = rp.compute_fourier_map(_smooth_map(ori, 4), ori, 2.0)
    weight = np.full(data.shape, 2.0)
    plain = rc.reconstruct(data, weight, ori, 2.0, do_map=False, max_iter_preweight=4)
    huge = rc.reconstruct(data, weight, ori, 2.0, tau2=np.full(ori // 2 + 1, 1e30), do_map=True, minres_map=0,
                          max_iter_preweight=4)
    assert np.allclose(huge, plain, rtol=1e-9, atol=1e-12 * np.abs(plain).max())
    # constant tau2 adds r = 1 / (pad^3 * tau2) to the uniform weight: the map scales by c / (c + r)
    tau = 0.05
    r = 1.0 / (8.0 * tau)
    reg = rc.reconstruct(data, weight, ori, 2.0, tau2=np.full(ori // 2 + 1, tau), do_map=True, minres_map=0,
                         max_iter_preweight=4)
    assert np.allclose(reg, plain * 2.0 / (2.0 + r), rtol=1e-6, atol=1e-9 * np.abs(plain).max())


def test_skip_gridding_round_trip_and_floor():
    # without gridding iterations a uniform weight c divides the data exactly: the map over c, sinc^2-corrected
    ori, c = 24, 3.0
    vol = _smooth_map(ori, 5)
    data, _ = rp.compute_fourier_map(vol, ori, 2.0, do_gridding=False)
    out = rc.reconstruct(data, np.full(data.shape, c), ori, 2.0, do_map=False, skip_gridding=True, r_max=ori // 2)
    cc = np.arange(ori) - ori // 2
    r = np.sqrt(cc[:, None, None] ** 2 + cc[None, :, None] ** 2 + cc[None, None, :] ** 2)
    s = np.where(r > 0, np.sin(np.pi * r / (2 * ori)) / (np.pi * r / (2 * ori) + (r == 0)), 1.0)
    inner = r < ori / 2 - 4
    expect = vol / c / (s * s)
    assert np.abs(out - expect)[inner].max() / np.abs(expect).max() < 2e-3
    # max_iter_preweight 0 takes the same path
    same = rc.reconstruct(data, np.full(data.shape, c), ori, 2.0, do_map=False, max_iter_preweight=0)
    assert np.array_equal(same, out)
    # zero weight everywhere: no division at all
    ones = rc.reconstruct(data, np.ones(data.shape), ori, 2.0, do_map=False, skip_gridding=True)
    zero = rc.reconstruct(data, np.zeros(data.shape), ori, 2.0, do_map=False, skip_gridding=True)
    assert np.allclose(zero, ones, rtol=1e-12, atol=1e-14 * np.abs(ones).max())
    # one element with a negligible weight among weights of 1 takes 1/1000 of its shell's average instead, so the
    # map of that element alone is 1000 / (1 - 1/n) times the map with a weight of 1 (n elements in the shell)
    single = np.zeros_like(data)
    S = data.shape[0]
    kz, ky, kx = S // 2 + 3, S // 2 - 5, 4  # logical (3, -5, 4), r = 7.07, shell FLOOR(r / 2) = 3
    single[kz, ky, kx] = 1.0 + 0.5j
    w1 = np.ones(data.shape)
    wt = w1.copy()
    wt[kz, ky, kx] = 1e-12
    a = rc.reconstruct(single, w1, ori, 2.0, do_map=False, skip_gridding=True)
    b = rc.reconstruct(single, wt, ori, 2.0, do_map=False, skip_gridding=True)
    big = np.abs(a) > 1e-3 * np.abs(a).max()
    ratio = b[big] / a[big]
    assert np.allclose(ratio, ratio[0], rtol=1e-9)
    n = sum(1 for z in range(-8, 9) for y in range(-8, 9) for x in range(0, 9)
            if int(np.floor(np.sqrt(z * z + y * y + x * x) / 2.0)) == 3)
    assert ratio[0] == pytest.approx(1000.0 / (1.0 - 1.0 / n), rel=1e-9)
