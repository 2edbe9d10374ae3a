"""TEST INFRASTRUCTURE ONLY -- numpy (FP64) restatement of the learned pose proposal of include/rln_proposal.h.

The proposal has no counterpart in the reference (it restricts the coarse enumeration of
src/acc/acc_projector_plan_impl.h:99-410), so this restatement follows the header's own definition: the network
input image, the six 3x3 convolutions with ReLU, the fully connected layer, the softmax cross entropy and their
gradients, in the header's parameter layout. With bf16=True it rounds to BF16 where the HIP kernels store BF16
(inputs, weights, post-ReLU activations, output and hidden gradients), so that the remaining difference is FP32
versus FP64 accumulation.

Pinned in tests by: central finite differences of the loss (gradients), a direct loop convolution (im2col
layout), and numpy's inverse real FFT (input image). Only tests/ import this module.
"""
import numpy as np

STRIDES = (1, 2, 1, 2, 1, 2)


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layers(W, C):
    out, H, cin = [], W, 1
    for s in STRIDES:
        Ho = (H - 1) // s + 1
        out.append(dict(cin=cin, H=H, Ho=Ho, s=s, Kp=16 if cin == 1 else 9 * cin))
        H, cin = Ho, C
    return out


def n_params(W, C, NB):
    L = layers(W, C)
    F = C * (W // 8) ** 2
    return sum(C * l["Kp"] + C for l in L) + NB * F + NB


def unpack(w, W, C, NB):
    w = np.asarray(w, np.float64)
    L = layers(W, C)
    conv, off = [], 0
    for l in L:
        Wl = w[off:off + C * l["Kp"]].reshape(C, l["Kp"])
        off += C * l["Kp"]
        bl = w[off:off + C]
        off += C
        conv.append((Wl, bl))
    F = C * (W // 8) ** 2
    Wfc = w[off:off + NB * F].reshape(NB, F)
    off += NB * F
    bfc = w[off:off + NB]
    off += NB
    assert off == w.size
    return conv, Wfc, bfc


def pack(conv, Wfc, bfc):
    parts = []
    for Wl, bl in conv:
        parts += [Wl.ravel(), bl.ravel()]
    parts += [Wfc.ravel(), bfc.ravel()]
    return np.concatenate(parts)


def im2col(X, s, Ho, Kp):
    """X [n][H][H][cin] -> col [n*Ho*Ho][Kp], k = (3*ky + kx) * cin + c, zero padding 1."""
    n, H, _, cin = X.shape
    Xp = np.zeros((n, H + 2, H + 2, cin))
    Xp[:, 1:H + 1, 1:H + 1, :] = X
    col = np.zeros((n, Ho, Ho, Kp))
    for ky in range(3):
        for kx in range(3):
            tap = 3 * ky + kx
            col[..., tap * cin:(tap + 1) * cin] = Xp[:, ky:ky + s * (Ho - 1) + 1:s, kx:kx + s * (Ho - 1) + 1:s, :]
    return col.reshape(n * Ho * Ho, Kp)


def col2im(dcol, n, H, cin, s, Ho):
    """Adjoint of im2col: dX [n][H][H][cin]."""
    d = dcol.reshape(n, Ho, Ho, -1)
    dXp = np.zeros((n, H + 2, H + 2, cin))
    for ky in range(3):
        for kx in range(3):
            tap = 3 * ky + kx
            dXp[:, ky:ky + s * (Ho - 1) + 1:s, kx:kx + s * (Ho - 1) + 1:s, :] += d[..., tap * cin:(tap + 1) * cin]
    return dXp[:, 1:H + 1, 1:H + 1, :]


def _rounder(bf, jitter, rng):
    """BF16 rounding (or identity); with jitter > 0 the value is first perturbed by a relative N(0, jitter) to
    model FP32 accumulation error, which moves some values across BF16 rounding boundaries."""
    if not bf:
        return lambda a: np.asarray(a, np.float64)
    if jitter <= 0.0:
        return bf16
    return lambda a: bf16(np.asarray(a, np.float64) * (1.0 + jitter * rng.standard_normal(np.shape(a))))


def forward(w, images, C, NB, bf=False, jitter=0.0, rng=None):
    """Logits [n][NB] and the activations needed by backward."""
    r = _rounder(bf, jitter, rng)
    n, W, _ = images.shape
    conv, Wfc, bfc = unpack(w, W, C, NB)
    q = bf16 if bf else (lambda a: np.asarray(a, np.float64))  # exact roundings of stored inputs and weights
    X = q(images)[..., None]
    cache = []
    for (Wl, bl), l in zip(conv, layers(W, C)):
        col = im2col(X, l["s"], l["Ho"], l["Kp"])
        Z = col @ q(Wl).T + bl
        Xn = r(np.maximum(Z, 0.0)).reshape(n, l["Ho"], l["Ho"], C)
        cache.append((X, col))
        X = Xn
    feats = X.reshape(n, -1)
    logits = feats @ q(Wfc).T + bfc
    return logits, (cache, X, feats)


def loss_and_grads(w, images, labels, C, NB, bf=False, jitter=0.0, seed=0):
    """Mean cross entropy over the images, its gradient in the parameter layout, and the logits."""
    rng = np.random.default_rng(seed)
    r = _rounder(bf, jitter, rng)
    images = np.asarray(images, np.float64)
    n, W, _ = images.shape
    conv, Wfc, bfc = unpack(w, W, C, NB)
    logits, (cache, X6, feats) = forward(w, images, C, NB, bf, jitter, rng)
    m = logits.max(1, keepdims=True)
    e = np.exp(logits - m)
    s = e.sum(1, keepdims=True)
    lab = np.asarray(labels)
    loss = float(np.mean(np.log(s[:, 0]) + m[:, 0] - logits[np.arange(n), lab]))
    p = e / s
    p[np.arange(n), lab] -= 1.0
    dlog = r(p / n)
    gWfc = dlog.T @ feats
    gbfc = dlog.sum(0)
    q = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    dX = (dlog @ q(Wfc)).reshape(X6.shape)
    dZ = r(dX * (X6 > 0))
    gconv = [None] * len(conv)
    L = layers(W, C)
    for i in range(len(conv) - 1, -1, -1):
        Wl, _ = conv[i]
        Xin, col = cache[i]
        l = L[i]
        dZm = dZ.reshape(-1, C)
        gconv[i] = (dZm.T @ col, dZm.sum(0))
        if i == 0:
            break
        dcol = dZm @ q(Wl)
        dXin = col2im(dcol, n, l["H"], l["cin"], l["s"], l["Ho"])
        dZ = r(dXin * (Xin > 0))
    return loss, pack(gconv, gWfc, gbfc), logits


def init_params(W, C, NB, seed):
    """He-normal convolutions and N(0, 1/F) fully connected weights, zero biases (any init is valid input)."""
    rng = np.random.default_rng(seed)
    conv = []
    for l in layers(W, C):
        Wl = np.zeros((C, l["Kp"]))
        Wl[:, :9 * l["cin"]] = rng.normal(0, np.sqrt(2.0 / (9 * l["cin"])), (C, 9 * l["cin"]))
        conv.append((Wl, np.zeros(C)))
    F = C * (W // 8) ** 2
    return pack(conv, rng.normal(0, np.sqrt(1.0 / F), (NB, F)), np.zeros(NB))


def window(fimg, s, W, ctf=None):
    """The W x (W/2+1) low-frequency window of an s x (s/2+1) FFTW half-complex image (complex), multiplied by
    the sign of the CTF (same layout) when given."""
    xs = s // 2 + 1
    F = np.asarray(fimg, np.float64).reshape(s, xs, 2)
    F = F[..., 0] + 1j * F[..., 1]
    if ctf is not None:
        F = np.where(np.asarray(ctf).reshape(s, xs) < 0, -F, F)
    rows = [ky if ky >= 0 else ky + s for ky in list(range(W // 2)) + list(range(-W // 2, 0))]
    return F[rows, :W // 2 + 1]


def input_image(fimg, s, W, ctf=None):
    """Network input: normalised sum_{ky, kx} h(kx) Re[F e^{2 pi i (kx u + ky v)/W}] (v-major) of the
    (phase-flipped) window."""
    img = np.fft.irfft2(window(fimg, s, W, ctf), s=(W, W)) * (W * W)
    img = img - img.mean()
    return img / np.sqrt((img ** 2).mean())


def input_image_direct(fimg, s, W):
    """input_image by the explicit double sum (used to pin the irfft2 form)."""
    Fw = window(fimg, s, W)
    ky = np.array(list(range(W // 2)) + list(range(-W // 2, 0)))
    kx = np.arange(W // 2 + 1)
    h = np.where((kx == 0) | (kx == W // 2), 1.0, 2.0)
    u = np.arange(W)
    ey = np.exp(2j * np.pi * np.outer(ky, u) / W)   # [ky][v]
    ex = np.exp(2j * np.pi * np.outer(kx, u) / W)   # [kx][u]
    G = ey.T @ Fw                                   # [v][kx]
    img = np.real((G * h[None, :]) @ ex)            # [v][u]
    img = img - img.mean()
    return img / np.sqrt((img ** 2).mean())
