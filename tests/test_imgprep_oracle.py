"""Pins of the image-preparation oracle (oracle/imgprep.py) against the reference's known answer, closed forms
and invariants, so that a dropped term, a wrong sign or a transposed axis fails here."""
import numpy as np
import pytest

import oracle
from oracle import imgprep as ip


# ------------------------------------------------------------------------------------------------- CTF image
def test_ctf_image_matches_pinned_c_restatement():
    # the C restatement holds the reference's known answer (tests/ctf.cpp:5-10); compare at every pixel
    N, s, px = 64, 48, 1.3
    args = dict(defU=14000.0, defV=11000.0, defAngle=37.0, bfactor=25.0, kfactor=0.9, phase_shift=12.0)
    img = ip.ctf_image(s, N, px, 300.0, 2.7, 0.1, **args)
    pi, pj = ip.half_indices(s, s // 2 + 1)
    for i in range(0, s, 5):
        for j in range(0, s // 2 + 1, 3):
            v = oracle.ctf_value(args["defU"], args["defV"], args["defAngle"], 300.0, 2.7, 0.1, args["bfactor"],
                                 args["kfactor"], args["phase_shift"], pj[i, j] / (N * px), pi[i, j] / (N * px))
            assert img[i, j] == pytest.approx(v, rel=1e-12, abs=1e-15)


def test_ctf_known_answer_through_image():
    # CTF::setValues(10000, 12000, 90, 300, 2.7, 0.1, 0, 1, 0); getCTF(10, 10) == 0.59154: the pixel (10, 10) of an
    # image with N * pixel_size = 1 carries frequency (10, 10)
    img = ip.ctf_image(32, 32, 1.0 / 32, 300.0, 2.7, 0.1, 10000.0, 12000.0, 90.0, 0.0, 1.0, 0.0)
    assert img[10, 10] == pytest.approx(0.59154, rel=1.2e-5)


def test_ctf_flags():
    N, s, px = 64, 64, 1.0
    base = dict(defU=9000.0, defV=9500.0, defAngle=10.0, bfactor=40.0, kfactor=1.1, phase_shift=0.0)
    plain = ip.ctf_image(s, N, px, 300.0, 2.7, 0.1, **base)
    assert np.allclose(ip.ctf_image(s, N, px, 300.0, 2.7, 0.1, **base, do_abs=True), np.abs(plain))
    assert np.array_equal(ip.ctf_image(s, N, px, 300.0, 2.7, 0.1, **base, flip=True) / 1.1, np.sign(plain))
    assert np.allclose(ip.ctf_image(s, N, px, 300.0, 2.7, 0.1, **base, premultiplied=True), plain ** 2)
    # intact first peak: 1 * damping * scale at low frequency, where |gamma| < pi/2
    intact = ip.ctf_image(s, N, px, 300.0, 2.7, 0.1, **base, intact=True)
    u2 = (np.arange(4) / (N * px)) ** 2
    assert np.allclose(intact[0, :4], 1.1 * np.exp(-40.0 / 4 * u2))
    assert np.allclose(intact[:, 20:], plain[:, 20:])
    # no CTF value is smaller than 1e-8 in magnitude
    zero_q0 = ip.ctf_image(8, 8, 1.0, 300.0, 2.7, 0.0, 1e4, 1e4, 0.0, 0.0, 1.0, 0.0)
    assert zero_q0[0, 0] == 1e-8  # -sin(0) = 0 -> +1e-8


# ------------------------------------------------------------------------------------------------- translation
def test_translation_equals_circular_shift_of_compact_image():
    rng = np.random.default_rng(1)
    img = np.zeros((32, 32))
    img[10:22, 9:20] = rng.normal(size=(12, 11))
    for dx, dy in ((3, -2), (-5, 4), (0, 0), (7, 7)):
        assert np.array_equal(ip.translate_int(img, dx, dy), np.roll(np.roll(img, dy, 0), dx, 1))


def test_translation_zero_fills_edges():
    img = np.ones((16, 16))
    out = ip.translate_int(img, 3, -2)
    assert out[:, :3].sum() == 0 and out[-2:, :].sum() == 0
    assert out[:-2, 3:].min() == 1.0
    assert ip.translate_int(img, 16, 0).sum() == 0


# ------------------------------------------------------------------------------------------------- soft mask
def test_mask_passes_constants_and_backgrounds():
    N = 48
    c = np.full((N, N), 2.5)
    assert np.allclose(ip.soft_mask(c, 15.0, 5.0), c)
    rng = np.random.default_rng(2)
    img = np.full((N, N), -1.25)
    y, x = np.mgrid[:N, :N] - N // 2
    inside = np.sqrt(x * x + y * y) < 15.0
    img[inside] = rng.normal(size=inside.sum())
    assert np.allclose(ip.soft_mask(img, 15.0, 5.0), img)  # background already at its average


def test_mask_affine_equivariance_and_identity_beyond_corners():
    rng = np.random.default_rng(3)
    img = rng.normal(size=(40, 40))
    a, b = 1.7, -0.3
    assert np.allclose(ip.soft_mask(a * img + b, 12.0, 4.0), a * ip.soft_mask(img, 12.0, 4.0) + b)
    assert np.array_equal(ip.soft_mask(img, 40.0, 5.0), img)


def test_mask_regions():
    rng = np.random.default_rng(4)
    N, R, W = 40, 10.0, 5.0
    img = rng.normal(size=(N, N))
    noise = rng.normal(size=(N, N))
    out = ip.soft_mask(img, R, W, noise)
    y, x = np.mgrid[:N, :N] - N // 2
    r = np.sqrt(x * x + y * y)
    assert np.array_equal(out[r < R], img[r < R])
    assert np.array_equal(out[r > R + W], noise[r > R + W])
    band = (r >= R) & (r <= R + W)
    lo = np.minimum(img, noise)[band] - 1e-12
    hi = np.maximum(img, noise)[band] + 1e-12
    assert np.all((out[band] >= lo) & (out[band] <= hi))  # convex blend inside the edge


# ------------------------------------------------------------------------------------------------- transforms
def test_centred_delta_is_flat():
    N = 32
    img = np.zeros((N, N))
    img[N // 2, N // 2] = 1.0
    F = ip.window_center(ip.forward(img), N)
    assert np.allclose(F, 1.0 / N ** 2)


def test_shifted_delta_is_phase_ramp_with_rows_as_y():
    N, s = 32, 24
    a, b = 3, -5  # x, y offsets from the centre
    img = np.zeros((N, N))
    img[N // 2 + b, N // 2 + a] = 1.0
    F = ip.window_center(ip.forward(img), s)
    kyy, kxx = ip.half_indices(s, s // 2 + 1)
    expect = np.exp(-2j * np.pi * (kxx * a + kyy * b) / N) / N ** 2
    assert np.allclose(F, expect, atol=1e-15)


def test_cosine_wave_spectrum():
    N = 64
    a, b = 9, -7  # |k| = 11.4 -> shell 11
    y, x = np.mgrid[:N, :N]
    img = np.cos(2 * np.pi * (a * x + b * y) / N)
    spec, hx = ip.spectrum(ip.forward(img), 16)
    expect = np.zeros(N // 2 + 1)
    expect[11] = 0.25
    assert np.allclose(spec, expect, atol=1e-14)
    assert hx == pytest.approx(0.25)  # shell 11 >= 16/2 + 1
    assert ip.spectrum(ip.forward(img), 32)[1] < 1e-25  # shell 11 is below 32/2 + 1


def test_window_is_restriction_of_full_transform():
    rng = np.random.default_rng(5)
    N, s = 32, 20
    F = ip.forward(rng.normal(size=(N, N)))
    full = ip.window_center(F, N)
    win = ip.window_center(F, s)
    for i in range(s):
        iy = i if i <= s // 2 else i - s
        assert np.allclose(win[i], full[iy % N, : s // 2 + 1])


# ------------------------------------------------------------------------------------------------- noise
def test_noise_hermitian_columns_and_spectrum():
    N = 64
    sigma2 = np.linspace(4.0, 0.5, N // 2 + 1)
    F = ip.noise_fourier(17, N, sigma2, 1.0)
    # x = 0 and x = N/2 columns Hermitian: irfft2 then rfft2 returns them unchanged
    back = np.fft.rfft2(np.fft.irfft2(F, s=(N, N), norm="forward"), norm="forward")
    assert np.allclose(back, F, atol=1e-12)
    # per-component variance: sum over shells of |F|^2 ~ 2 sigma2 per component
    pyy, pxx = ip.half_indices(N, N // 2 + 1)
    ires = np.rint(np.sqrt(pyy ** 2 + pxx ** 2)).astype(int)
    for sh in (5, 12, 25):
        sel = (ires == sh) & (pxx > 0) & (pxx < N // 2)
        ratio = np.mean(np.abs(F[sel]) ** 2) / (2 * sigma2[sh])
        assert 0.6 < ratio < 1.4, (sh, ratio)
    assert np.all(F[ires > N // 2] == 0)


def test_noise_generator_is_seeded_and_standard_normal():
    z1 = np.concatenate(ip.noise_pair(5, np.arange(200000)))
    z2 = np.concatenate(ip.noise_pair(5, np.arange(200000)))
    z3 = np.concatenate(ip.noise_pair(6, np.arange(200000)))
    assert np.array_equal(z1, z2) and not np.array_equal(z1, z3)
    assert abs(z1.mean()) < 0.01 and abs(z1.std() - 1.0) < 0.01
    assert abs(np.mean(z1 ** 4) - 3.0) < 0.05


# ------------------------------------------------------------------------------------------------- whole stage
def _cfg(**kw):
    c = dict(ori_size=32, current_size=24, pixel_size=2.0, particle_diameter=40.0, width_mask_edge=4.0, zero_mask=1,
             sigma2_fudge=1.0, random_seed=7, avg_norm_correction=0.0, do_ctf_correction=1, ctf_phase_flipped=0,
             only_flip_phases=0, intact_ctf_first_peak=0, ctf_premultiplied=0, voltage=300.0, cs=2.7, q0=0.1)
    c.update(kw)
    return c


def _meta(**kw):
    m = dict(norm_correction=1.0, old_offset_x=0.0, old_offset_y=0.0, defocus_u=12000.0, defocus_v=11000.0,
             defocus_angle=20.0, bfactor=0.0, kfactor=1.0, phase_shift=0.0, part_id=0)
    m.update(kw)
    return m


def test_prepare_rounds_offsets_and_applies_norm():
    rng = np.random.default_rng(6)
    img = rng.normal(size=(2, 32, 32))
    cfg = _cfg(avg_norm_correction=2.0, particle_diameter=1000.0)
    out = ip.prepare(img, [_meta(old_offset_x=2.5, old_offset_y=-1.5, norm_correction=4.0),
                           _meta(old_offset_x=-0.49, old_offset_y=0.51)], cfg)
    assert np.array_equal(out["old_offset"], [[3, -2], [0, 1]])
    ref = ip.window_center(ip.forward(0.5 * ip.translate_int(img[0], 3, -2)), 24)
    assert np.allclose(out["fimg_nomask"][0], ref)
    assert np.allclose(out["fimg"], out["fimg_nomask"])  # mask radius beyond the corners


def test_prepare_power_sums_to_parseval():
    # Parseval over the half-plane: columns 1..N/2-1 count twice, columns 0 and N/2 once
    rng = np.random.default_rng(8)
    N = 32
    img = rng.normal(size=(1, N, N))
    out = ip.prepare(img, [_meta()], _cfg(current_size=N, particle_diameter=1000.0))
    F = np.fft.rfft2(img[0]) / N ** 2
    total = (np.abs(F[:, 1:N // 2]) ** 2).sum() * 2 + (np.abs(F[:, [0, N // 2]]) ** 2).sum()
    assert total == pytest.approx((img[0] ** 2).sum() / N ** 2)
    pyy, pxx = ip.half_indices(N, N // 2 + 1)
    ires = np.rint(np.sqrt(pyy ** 2 + pxx ** 2))
    kept = (ires > 0) & (ires <= N // 2) & ~((pxx == 0) & (pyy < 0))
    assert out["power"][0].sum() == pytest.approx((np.abs(F[kept]) ** 2).sum())
    assert out["highres_Xi2"][0] == 0.0


def test_prepare_returns_the_estep_input_convention():
    # A noise-free particle of the E-step's synthetic generator (a centred half-complex slice times its CTF),
    # taken to real space and moved by -d pixels, prepared with old offset +d, gives back the generator's image:
    # rows are y, the centring sign and 1/N^2 scaling match, and the old offset re-centres the particle.
    import relion_amd as ra
    # blobs within 4 px of the centre (sigma <= 3 px) and a near-focus CTF, whose point spread stays within a few
    # pixels: nothing reaches an edge after the shifts below, so zero fill and circular shifts agree
    N = 64
    spec = ra.synth_spec(11, N, 2.0, 20, snr=1e12, offset_max=0.0, blob_radius=4.0, defocus=(100.0, 200.0),
                         max_astigmatism=0.0)
    F, fctf, _, _, _ = ra.synth_particles(spec, 2, N)
    F = F[..., 0].astype(np.float64) + 1j * F[..., 1]
    i = np.arange(N)[:, None]
    j = np.arange(N // 2 + 1)[None, :]
    sign = np.where(((i + j) & 1) == 1, -1.0, 1.0)
    metas, imgs = [], []
    for b, (dx, dy) in enumerate(((3, -2), (-4, 5))):
        img = np.fft.irfft2(F[b] * sign, s=(N, N)) * N * N
        imgs.append(np.roll(np.roll(img, -dy, 0), -dx, 1))
        metas.append(_meta(old_offset_x=dx, old_offset_y=dy, part_id=b))
    out = ip.prepare(np.array(imgs), metas, _cfg(ori_size=N, current_size=N, pixel_size=2.0,
                                                  particle_diameter=1000.0, do_ctf_correction=0))
    for b in range(2):
        err = np.abs(out["fimg"][b] - F[b])
        assert err.max() <= 1e-5 * np.abs(F[b]).max()  # FP32 generator output and edge content of ~1e-7
        assert np.allclose(out["fimg_nomask"][b], out["fimg"][b])
