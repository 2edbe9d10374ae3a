"""CPU restatement (FP64 numpy) of the image part of MlOptimiser::getFourierTransformsAndCtfs.

Test infrastructure only: tests, the smoke check and the CPU leg of the benchmark use it to check
``rln_imgprep_run``; nothing in the product imports it.

Follows src/ml_optimiser.cpp:5840-6531 for 2-D particles with one image each:
norm correction (:6225-6237), rounding of the old offset (:6085) and integer translation with zero fill
(selfTranslate DONT_WRAP, :6247; src/transformations.h:254-460), the unmasked transform (:6278-6289), the soft
mask (src/mask.cpp:43-96) with an average or noise background (:6291-6370), the masked transform with its power
spectrum and high-resolution power (:6379-6420), windowFourierTransform (src/fftw.h:809-860), CenterFFTbySign
(src/fftw.h:390-402), FourierTransformer's 1/N^2 forward scaling (src/fftw.cpp:335-358) and the CTF image
(CTF::getFftwImage, src/ctf.cpp:314-453; CTF::getCTF, src/ctf.h:184-240).

The noise background uses the library's counter-based generator (restated here), not the reference's random
sequence, whose instance cannot be reproduced; its distribution is checked separately.

Pins (tests/test_imgprep_oracle.py): the CTF against the FP64 C restatement, which holds the reference's known
answer (tests/ctf.cpp:5-10); translation against circular shifts of compact images and zero fill at the edges;
the mask against invariants (constants and backgrounds pass unchanged, affine equivariance, identity for radii
beyond the corners); the transform against a centred delta (flat 1/N^2), shifted deltas (phase ramps) and a
cosine wave (one spectrum bin of 1/4); the window against the full-size transform.
"""
import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def mix64(z):
    """splitmix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def noise_pair(seed, e):
    """Two N(0,1) deviates per element index e (Box-Muller on two 53-bit uniforms)."""
    with np.errstate(over="ignore"):
        h1 = mix64(mix64(np.uint64(seed)) + np.asarray(e, dtype=np.uint64))
    h2 = mix64(h1)
    u1 = ((h1 >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    u2 = ((h2 >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u1))
    return r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)


def half_indices(n_rows, n_cols):
    """(ip, jp) of a half-complex array with n_rows rows: ip = i for i < n_cols, else i - n_rows."""
    i = np.arange(n_rows)[:, None]
    j = np.arange(n_cols)[None, :]
    ip = np.where(i < n_cols, i, i - n_rows)
    return np.broadcast_to(ip, (n_rows, n_cols)), np.broadcast_to(j, (n_rows, n_cols))


def noise_fourier(seed, N, sigma2, fudge):
    """Fourier noise of one image: N(0, sigma) real and imaginary parts, sigma = sqrt(fudge*sigma2[ires]) for
    ires = ROUND(|k|) < N/2+1 (src/ml_optimiser.cpp:6318-6335); x = 0 and x = N/2 columns Hermitian."""
    xh = N // 2 + 1
    ip, jp = half_indices(N, xh)
    gi = np.arange(N)[:, None] * np.ones((1, xh), np.int64)
    edge = (jp == 0) | (jp == N // 2)
    conj = edge & (ip < 0)
    real_only = edge & ((ip == 0) | (ip == N // 2))
    gi = np.where(conj, -ip, gi)
    ires = np.rint(np.sqrt(ip * ip + jp * jp)).astype(np.int64)
    ok = ires < xh
    sg = np.sqrt(fudge * np.asarray(sigma2, np.float64)[np.minimum(ires, xh - 1)])
    zr, zi = noise_pair(seed, gi.astype(np.uint64) * np.uint64(xh) + jp.astype(np.uint64))
    re = np.where(ok, sg * zr, 0.0)
    im = np.where(ok & ~real_only, np.where(conj, -sg, sg) * zi, 0.0)
    return re + 1j * im


def noise_image(seed, N, sigma2, fudge):
    """Real-space noise: unnormalised inverse transform (FFTW backward, as FourierTransformer)."""
    return np.fft.irfft2(noise_fourier(seed, N, sigma2, fudge), s=(N, N), norm="forward")


def ctf_image(s, N, pixel_size, voltage, cs, q0, defU, defV, defAngle, bfactor, kfactor, phase_shift,
              do_abs=False, flip=False, intact=False, premultiplied=False):
    """CTF::getFftwImage at s x (s/2+1), frequencies (jp, ip) / (N * pixel_size) (src/ctf.cpp:441-451), getCTF with
    damping (src/ctf.h:184-240), then the square for CTF-premultiplied data (src/ml_optimiser.cpp:6487-6493)."""
    kV = voltage * 1e3
    lam = 12.2643247 / np.sqrt(kV * (1.0 + kV * 0.978466e-6))
    K1 = np.pi / 2 * 2 * lam
    K2 = np.pi / 2 * (cs * 1e7) * lam ** 3
    K3 = np.arctan(q0 / np.sqrt(1 - q0 * q0))
    K4 = -bfactor / 4.0
    K5 = np.deg2rad(phase_shift)
    az = np.deg2rad(defAngle)
    c, sn = np.cos(az), np.sin(az)
    Axx = c * c * (-defU) + sn * sn * (-defV)
    Axy = c * sn * (-defU) - sn * c * (-defV)
    Ayy = sn * sn * (-defU) + c * c * (-defV)
    ip, jp = half_indices(s, s // 2 + 1)
    X = jp / (N * pixel_size)
    Y = ip / (N * pixel_size)
    u2 = X * X + Y * Y
    gamma = K1 * (Axx * X * X + 2.0 * Axy * X * Y + Ayy * Y * Y) + K2 * u2 * u2 - K5 - K3
    r = -np.sin(gamma)
    if intact:
        r = np.where(np.abs(gamma) < np.pi / 2, 1.0, r)
    r = r * np.exp(K4 * u2)
    if do_abs:
        r = np.abs(r)
    elif flip:
        r = np.where(r < 0, -1.0, 1.0)
    r = r * kfactor
    r = np.where(np.abs(r) < 1e-8, np.where(r >= 0, 1.0, -1.0) * 1e-8, r)
    if premultiplied:
        r = r * r
    return r


def translate_int(img, dx, dy):
    """selfTranslate by an integer (dx, dy) without wrapping: out(y, x) = in(y - dy, x - dx), zero outside."""
    N0, N1 = img.shape
    out = np.zeros_like(img)
    ys, yd = (slice(0, N0 - dy), slice(dy, N0)) if dy >= 0 else (slice(-dy, N0), slice(0, N0 + dy))
    xs, xd = (slice(0, N1 - dx), slice(dx, N1)) if dx >= 0 else (slice(-dx, N1), slice(0, N1 + dx))
    if abs(dy) < N0 and abs(dx) < N1:
        out[yd, xd] = img[ys, xs]
    return out


def soft_mask(img, radius, width, noise=None):
    """softMaskOutsideMap (src/mask.cpp:43-96) on an image with its origin at (N/2, N/2)."""
    N0, N1 = img.shape
    if radius < 0:
        radius = N1 / 2.0
    radius_p = radius + width
    y = np.arange(N0)[:, None] - N0 // 2
    x = np.arange(N1)[None, :] - N1 // 2
    r = np.sqrt((x * x + y * y).astype(np.float64))
    inside = r < radius
    outer = r > radius_p
    rc = 0.5 + 0.5 * np.cos(np.pi * (radius_p - r) / width)
    if noise is None:
        w = np.where(inside, 0.0, np.where(outer, 1.0, rc))
        sw = w.sum()
        bg = (w * img).sum() / sw if sw > 0 else 0.0
        add = np.full_like(img, bg)
    else:
        add = noise
    out = np.where(outer, add, (1 - rc) * img + rc * add)
    return np.where(inside, img, out)


def forward(img):
    """FourierTransformer::FourierTransform: half-complex transform scaled by 1/N^2."""
    return np.fft.rfft2(img) / img.size


def window_center(F, s):
    """windowFourierTransform to s x (s/2+1) followed by CenterFFTbySign."""
    N = F.shape[0]
    ip, jp = half_indices(s, s // 2 + 1)
    out = F[np.where(ip >= 0, ip, ip + N), jp]
    i = np.arange(s)[:, None]
    j = np.arange(s // 2 + 1)[None, :]
    return np.where(((i + j) & 1) == 1, -out, out)


def spectrum(F, s):
    """Power per shell of the full transform (x = 0 column: y >= 0 only) and its sum from shell s/2+1 on."""
    N = F.shape[0]
    xh = N // 2 + 1
    ip, jp = half_indices(N, xh)
    ires = np.rint(np.sqrt(ip * ip + jp * jp)).astype(np.int64)
    ok = (ires > 0) & (ires < xh) & ~((jp == 0) & (ip < 0))
    p = np.abs(F) ** 2
    spec = np.bincount(ires[ok], weights=p[ok], minlength=xh)[:xh]
    hx = float(p[ok & (ires >= s // 2 + 1)].sum()) if s < N else 0.0
    return spec, hx


def prepare(images, metas, cfg, sigma2=None):
    """All outputs of rln_imgprep_run for a stack of real-space images [n][N][N].

    cfg: dict with ori_size, current_size, pixel_size, particle_diameter, width_mask_edge, zero_mask, sigma2_fudge,
    random_seed, avg_norm_correction, do_ctf_correction, ctf_phase_flipped, only_flip_phases,
    intact_ctf_first_peak, ctf_premultiplied, voltage, cs, q0. metas: sequence of dicts with norm_correction,
    old_offset_x, old_offset_y, defocus_u, defocus_v, defocus_angle, bfactor, kfactor, phase_shift, part_id.
    """
    N, s = cfg["ori_size"], cfg["current_size"]
    n = len(metas)
    xs = s // 2 + 1
    fimg = np.zeros((n, s, xs), np.complex128)
    fnm = np.zeros_like(fimg)
    fctf = np.ones((n, s, xs))
    power = np.zeros((n, N // 2 + 1))
    hx = np.zeros(n)
    off = np.zeros((n, 2))
    radius = cfg["particle_diameter"] / (2.0 * cfg["pixel_size"])
    for b, m in enumerate(metas):
        img = np.asarray(images[b], np.float64)
        if cfg["avg_norm_correction"] > 0:
            img = img * (cfg["avg_norm_correction"] / m["norm_correction"])
        dx = int(np.floor(m["old_offset_x"] + 0.5)) if m["old_offset_x"] > 0 else int(np.ceil(m["old_offset_x"] - 0.5))
        dy = int(np.floor(m["old_offset_y"] + 0.5)) if m["old_offset_y"] > 0 else int(np.ceil(m["old_offset_y"] - 0.5))
        off[b] = dx, dy
        img = translate_int(img, dx, dy)
        fnm[b] = window_center(forward(img), s)
        noise = None
        if not cfg["zero_mask"]:
            noise = noise_image(cfg["random_seed"] + m["part_id"], N, sigma2, cfg["sigma2_fudge"])
        F = forward(soft_mask(img, radius, cfg["width_mask_edge"], noise))
        power[b], hx[b] = spectrum(F, s)
        fimg[b] = window_center(F, s)
        if cfg["do_ctf_correction"]:
            fctf[b] = ctf_image(s, N, cfg["pixel_size"], cfg["voltage"], cfg["cs"], cfg["q0"], m["defocus_u"],
                                m["defocus_v"], m["defocus_angle"], m["bfactor"], m["kfactor"], m["phase_shift"],
                                bool(cfg["ctf_phase_flipped"]), bool(cfg["only_flip_phases"]),
                                bool(cfg["intact_ctf_first_peak"]), bool(cfg["ctf_premultiplied"]))
    return dict(fimg=fimg, fimg_nomask=fnm, fctf=fctf, power=power, highres_Xi2=hx, old_offset=off)
