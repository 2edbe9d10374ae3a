"""CPU restatement (FP64 numpy) of Projector::computeFourierTransformMap for a 3-D reference and 2-D data.

Test infrastructure only: tests and the CPU leg of the benchmark use it to check
``rln_projector_compute_fourier_map``; nothing in the product imports it.

Follows src/projector.cpp:116-593 (the CPU branch) and Projector::griddingCorrect (:595-628), trilinear
interpolator: padded box padoridim = ROUND(pad * ori), made even, and pad = padoridim / ori (:131-134); gridding
correction of the real-space map by sinc^2(r / padoridim) (r = 0 untouched); zero padding around the centre; forward
transform scaled by 1 / padoridim^3 (src/fftw.cpp:335-358); CenterFFTbySign; data(kp, ip, jp) = Faux * pad^3 * ori
for r^2 <= ROUND(r_max * pad)^2 (:506-516), r_max = min(current_size, ori) / 2; power spectrum: |data|^2 / 2 summed
per shell ROUND(sqrt(r^2) / pad) and divided by the count (:517-567); r^2 <= ROUND(min_ires * pad)^2 zeroed after the
spectrum.

Pins (tests/test_refprep_oracle.py): a delta at the centre gives a flat 1/ori^2; a displaced delta gives the phase
ramp divided by its gridding factor; the central section at identity equals the 2-D transform of the map's
projection along z (the central-slice theorem, with the E-step's 1/ori^2 image normalisation); a sampled Gaussian-blob
map matches the product's analytic Fourier reference; the spectrum of a single cosine sits in one shell.
"""
import numpy as np


def sizes(ori_size, current_size, padding_factor):
    """(padoridim, pad, r_max, pad_size, max_r2) as the reference derives them."""
    padoridim = int(np.floor(padding_factor * ori_size + 0.5))
    padoridim += padoridim % 2
    pad = padoridim / ori_size
    r_max = min(current_size if current_size > 0 else ori_size, ori_size) // 2
    rr = int(np.floor(pad * r_max + 0.5))
    pad_size = 2 * (rr + 1) + 1
    return padoridim, pad, r_max, pad_size, rr * rr


def gridding_correct(vol, padoridim):
    """Divide by sinc^2(r / padoridim) (trilinear interpolation), r from the centre (Xmipp origin)."""
    n = vol.shape[0]
    c = np.arange(n) - n // 2
    r = np.sqrt(c[:, None, None] ** 2 + c[None, :, None] ** 2 + c[None, None, :] ** 2)
    rv = r / padoridim
    with np.errstate(invalid="ignore", divide="ignore"):
        s = np.where(r > 0, np.sin(np.pi * rv) / (np.pi * rv), 1.0)
    return vol / (s * s)


def compute_fourier_map(vol, current_size, padding_factor=2.0, do_gridding=True, min_ires=-1):
    """Projector data [pad_size][pad_size][pad_size/2+1] (z, y from -(pad_size/2), x from 0) and the power
    spectrum [ori/2+1] of a real-space map [ori][ori][ori] (z, y, x; origin at ori/2)."""
    vol = np.asarray(vol, np.float64)
    ori = vol.shape[0]
    padoridim, pad, r_max, pad_size, max_r2 = sizes(ori, current_size, padding_factor)
    if do_gridding:
        vol = gridding_correct(vol, padoridim)
    mpad = np.zeros((padoridim,) * 3)
    o = padoridim // 2 - ori // 2
    mpad[o:o + ori, o:o + ori, o:o + ori] = vol
    F = np.fft.rfftn(mpad) / padoridim ** 3
    kk = np.arange(padoridim)
    sgn = np.where(((kk[:, None, None] + kk[None, :, None] + np.arange(padoridim // 2 + 1)[None, None, :]) & 1) == 1,
                   -1.0, 1.0)
    F = F * sgn
    normfft = pad ** 3 * ori
    h = pad_size // 2
    data = np.zeros((pad_size, pad_size, h + 1), np.complex128)
    zp = np.arange(pad_size) - h
    xp = np.arange(h + 1)
    Z, Y, X = np.meshgrid(zp, zp, xp, indexing="ij")
    r2 = Z * Z + Y * Y + X * X
    # the transform holds frequencies -(padoridim/2 - 1) .. padoridim/2 in z and y (FOR_ALL_ELEMENTS_IN_FFTW_TRANSFORM):
    # -padoridim/2 has no element of its own and stays zero
    inside = (r2 <= max_r2) & (Z > -padoridim // 2) & (Y > -padoridim // 2)
    src = F[Z % padoridim, Y % padoridim, np.minimum(X, padoridim // 2)]  # X beyond it lies outside the sphere
    data[inside] = src[inside] * normfft
    ires = np.floor(np.sqrt(r2[inside]) / pad + 0.5).astype(np.int64)
    nb = ori // 2 + 1
    ps = np.bincount(ires, weights=np.abs(data[inside]) ** 2 / 2.0, minlength=nb)[:nb]
    cnt = np.bincount(ires, minlength=nb)[:nb].astype(np.float64)
    ps = np.where(cnt < 1.0, 0.0, ps / np.maximum(cnt, 1.0))
    if min_ires > 0:
        mr = int(np.floor(min_ires * pad + 0.5))
        data[inside & (r2 <= mr * mr)] = 0.0
    return data, ps
