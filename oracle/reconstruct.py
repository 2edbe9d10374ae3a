"""CPU restatement (FP64 numpy) of BackProjector::reconstruct for a 3-D reference, 2-D data, trilinear
interpolation, the default (windowing) path.

Test infrastructure only: tests and the CPU leg of the benchmark use it to check
``rln_backprojector_reconstruct``; nothing in the product imports it.

Follows src/backprojector.cpp:1379-1780: decenter of weight and data (Projector::decenter, src/projector.h:249-259);
MAP regularisation of the weight per shell ROUND(r / pad) from minres_map on (:1440-1503). With skip_gridding (the
relion_refine default, src/ml_optimiser.cpp:586) or no iterations: Fconv = data / max(weight, 1/1000 of the radial
average of the weight) (:1507-1572). Otherwise division of both by normalise; max_iter_preweight iterations of
Fconv = Fnewweight * Fweight, blob convolution in real space (convoluteBlobRealSpace, :2483-2528, with the tabulated
Kaiser-Bessel transform TabFtBlob, src/tabfuncs.cpp:95-121, radius 2 * 1.9, alpha 15, order 0, 10,000 steps over
[0, 0.5)) and Fnewweight /= max(1e-6, |Fconv|) inside the sphere (:1580-1640); Fconv = data * Fnewweight
(:1650-1660; the 1e20 cap there belongs to single-precision builds only). Then windowToOridimRealSpace (:2530-2640:
window to the padded box, CenterFFTbySign, inverse transform, real-space window to the original box, division by
pad^3 * ori, softMaskOutsideMap with radius ori/2 and width 3) and griddingCorrect (src/projector.cpp:595-628).

The blob's transform is Lewitt's closed form for order 0, (2 pi)^{3/2} a^3 I_{3/2}(s) / (I_0(alpha) s^{3/2}) with
s = sqrt(alpha^2 - (2 pi a w)^2) (J_{3/2} and |...| beyond), written with the closed forms of the half-integer
Bessel functions.
"""
import numpy as np

BLOB_RADIUS = 2 * 1.9
BLOB_ALPHA = 15.0
BLOB_TABLE = 10000


def _i32_over_pow(s):
    """I_{3/2}(s) / s^{3/2} for s >= 0 (series below 1e-3)."""
    s = np.asarray(s, np.float64)
    small = s < 1e-3
    ss = np.where(small, 1.0, s)
    big = np.sqrt(2.0 / (np.pi * ss)) * (np.cosh(ss) - np.sinh(ss) / ss) / ss ** 1.5
    ser = np.sqrt(2.0 / np.pi) / 3.0 * (1.0 + s * s / 10.0)
    return np.where(small, ser, big)


def _j32_over_pow(s):
    """J_{3/2}(s) / s^{3/2} for s > 0 (series below 1e-3)."""
    s = np.asarray(s, np.float64)
    small = s < 1e-3
    ss = np.where(small, 1.0, s)
    big = np.sqrt(2.0 / (np.pi * ss)) * (np.sin(ss) / ss - np.cos(ss)) / ss ** 1.5
    ser = np.sqrt(2.0 / np.pi) / 3.0 * (1.0 - s * s / 10.0)
    return np.where(small, ser, big)


def kaiser_ft(w, a=BLOB_RADIUS, alpha=BLOB_ALPHA):
    """Fourier transform of the order-0 Kaiser-Bessel blob of radius a and taper alpha at frequency w."""
    w = np.asarray(w, np.float64)
    t = 2.0 * np.pi * a * w
    s = np.sqrt(np.abs(alpha * alpha - t * t))
    c = (2.0 * np.pi) ** 1.5 * a ** 3 / np.i0(alpha)
    return c * np.where(t > alpha, _j32_over_pow(s), _i32_over_pow(s))


def blob_table(n=BLOB_TABLE):
    return kaiser_ft(np.arange(n) * (0.5 / n))


def blob_lookup(table, val):
    """TabFtBlob::operator(): the entry at floor(|val| / sampling), 0 beyond the table."""
    idx = (np.abs(val) / (0.5 / len(table))).astype(np.int64)
    return np.where(idx < len(table), table[np.minimum(idx, len(table) - 1)], 0.0)


def fftw_freq(n_rows, n_cols):
    """Logical (kz, ky, kx) of a half-complex [n][n][n_cols] array: index < n_cols maps to itself, else - n."""
    k = np.arange(n_rows)
    kp = np.where(k < n_cols, k, k - n_rows)
    return np.meshgrid(kp, kp, np.arange(n_cols), indexing="ij")


def decenter(centred, n, max_r2):
    """Projector::decenter: the centred [S][S][S/2+1] array (z, y from -(S/2)) into the half-complex layout of an
    n^3 box, entries with r^2 <= max_r2."""
    S = centred.shape[0]
    h = S // 2
    nc = n // 2 + 1
    kz, ky, kx = fftw_freq(n, nc)
    r2 = kz * kz + ky * ky + kx * kx
    ok = (r2 <= max_r2) & (np.abs(kz) <= h) & (np.abs(ky) <= h) & (kx <= h)
    out = np.zeros((n, n, nc), centred.dtype)
    out[ok] = centred[kz[ok] + h, ky[ok] + h, kx[ok]]
    return out


def soft_mask_3d(vol, radius=-1.0, width=3.0):
    """softMaskOutsideMap (src/mask.cpp:43-96) with the average background."""
    n = vol.shape[0]
    if radius < 0:
        radius = n / 2.0
    radius_p = radius + width
    c = np.arange(n) - n // 2
    r = np.sqrt(c[:, None, None] ** 2 + c[None, :, None] ** 2 + c[None, None, :] ** 2)
    inside = r < radius
    outer = r > radius_p
    rc = 0.5 + 0.5 * np.cos(np.pi * (radius_p - r) / width)
    w = np.where(inside, 0.0, np.where(outer, 1.0, rc))
    bg = (w * vol).sum() / w.sum() if w.sum() > 0 else 0.0
    out = np.where(outer, bg, (1 - rc) * vol + rc * bg)
    return np.where(inside, vol, out)


def reconstruct(data, weight, ori_size, padding_factor=2.0, tau2=None, tau2_fudge=1.0, normalise=1.0, minres_map=3,
                max_iter_preweight=10, do_map=True, skip_gridding=False, r_max=None):
    """vol_out [ori]^3 from the backprojector's centred data (complex) and weight arrays [S][S][S/2+1].

    skip_gridding or max_iter_preweight <= 0 takes the path without gridding iterations (:1507-1572): data divided by
    max(weight, 1/1000 of the weight's average over shell FLOOR(r / pad) below r_max), no normalise."""
    S = data.shape[0]
    h = S // 2
    pad = padding_factor
    rr = h - 1  # pad_size = 2 (ROUND(pad r_max) + 1) + 1, max_r2 = ROUND(pad r_max)^2
    max_r2 = rr * rr
    if r_max is None:
        r_max = int(np.floor(rr / pad + 0.5))
    assert int(np.floor(pad * r_max + 0.5)) == rr, "r_max does not match the backprojector's size"
    ovs = pad ** 3
    nc = S // 2 + 1
    Fweight = decenter(np.asarray(weight, np.float64), S, max_r2)
    kz, ky, kx = fftw_freq(S, nc)
    r2 = kz * kz + ky * ky + kx * kx
    sph = r2 < max_r2
    if do_map:
        ires = np.floor(np.sqrt(r2) / pad + 0.5).astype(np.int64)
        t = np.asarray(tau2, np.float64)[np.minimum(ires, len(tau2) - 1)]
        invw = Fweight
        with np.errstate(divide="ignore"):
            inv = np.where(t > 0, 1.0 / (ovs * tau2_fudge * np.where(t > 0, t, 1.0)),
                           np.where(invw > 1e-20, 1.0 / (0.001 * np.where(invw > 1e-20, invw, 1.0)), 0.0))
        Fweight = np.where(sph & (ires >= minres_map), Fweight + inv, Fweight)
    if skip_gridding or max_iter_preweight <= 0:
        Fconv = decenter(np.asarray(data, np.complex128), S, max_r2)
        round_max_r2 = int(np.floor(r_max * pad * r_max * pad + 0.5))
        rad = np.sqrt(r2)
        ires = np.floor(rad / pad).astype(np.int64)
        inner = r2 < round_max_r2
        assert ires[inner].max() < r_max
        tot = np.bincount(ires[inner], weights=Fweight[inner], minlength=r_max)[:r_max]
        cnt = np.bincount(ires[inner], minlength=r_max)[:r_max].astype(np.float64)
        assert np.all((cnt > 0) | (tot > 0))
        avg = tot / (1000.0 * cnt)
        wt = np.maximum(Fweight, avg[np.minimum(ires, r_max - 1)])
        Fconv = np.where(wt != 0, Fconv / np.where(wt != 0, wt, 1.0), Fconv)
        return window_to_oridim(Fconv, ori_size, pad)
    Fweight = Fweight / normalise
    Fdata = decenter(np.asarray(data, np.complex128), S, max_r2) / normalise
    Fnew = np.where(sph, 1.0, 0.0)
    table = blob_table()
    norm0 = table[0]
    c = np.arange(S)
    cw = np.where(c < S // 2, c, c - S)
    rv = np.sqrt(cw[:, None, None] ** 2 + cw[None, :, None] ** 2 + cw[None, None, :] ** 2) / (ori_size * pad)
    blob = blob_lookup(table, rv) / norm0
    for _ in range(max_iter_preweight):
        Fconv = Fnew * Fweight
        M = np.fft.irfftn(Fconv, s=(S, S, S), axes=(0, 1, 2)) * S ** 3  # unnormalised inverse
        M *= blob
        Fconv = np.fft.rfftn(M) / S ** 3                # normalised forward
        w = np.maximum(1e-6, np.abs(Fconv))
        Fnew = np.where(sph, Fnew / w, Fnew)
    Fconv = Fdata * Fnew  # no 1e20 cap: that is RELION_SINGLE_PRECISION only, the default CPU build is FP64
    return window_to_oridim(Fconv, ori_size, pad)


def window_to_oridim(Fconv, ori_size, pad):
    """windowToOridimRealSpace (src/backprojector.cpp:2530-2640) followed by griddingCorrect."""
    S = Fconv.shape[0]
    P = int(np.floor(pad * ori_size + 0.5))
    P += P % 2
    nP = P // 2 + 1
    kz, ky, kx = fftw_freq(P, nP)
    Fw = np.zeros((P, P, nP), np.complex128)
    if nP <= Fconv.shape[2]:  # shrink: every element of the smaller transform from the larger one
        Fw = Fconv[kz % S, ky % S, kx]
    else:                     # grow: the larger one's corners stay zero beyond (S/2)^2
        sz, sy, sx = fftw_freq(S, S // 2 + 1)
        lim = (S // 2 + 1 - 1) ** 2
        ok = sz * sz + sy * sy + sx * sx <= lim
        Fw[sz[ok] % P, sy[ok] % P, sx[ok]] = Fconv[ok]
    i = np.arange(P)
    sgn = np.where(((i[:, None, None] + i[None, :, None] + np.arange(nP)[None, None, :]) & 1) == 1, -1.0, 1.0)
    M = np.fft.irfftn(Fw * sgn, s=(P, P, P), axes=(0, 1, 2)) * P ** 3
    o = P // 2 - ori_size // 2
    vol = M[o:o + ori_size, o:o + ori_size, o:o + ori_size] / (pad ** 3 * ori_size)
    vol = soft_mask_3d(vol)
    c = np.arange(ori_size) - ori_size // 2
    r = np.sqrt(c[:, None, None] ** 2 + c[None, :, None] ** 2 + c[None, None, :] ** 2) / (ori_size * pad)
    with np.errstate(invalid="ignore", divide="ignore"):
        s = np.where(r > 0, np.sin(np.pi * r) / (np.pi * r), 1.0)
    return vol / (s * s)
