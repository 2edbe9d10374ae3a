"""CPU pins of the pose-proposal oracle (oracle/pose_cnn.py) and of the product's HEALPix ang2pix.

  * gradients of the FP64 restatement against central finite differences of its loss;
  * the im2col convolution against a direct loop over taps and channels;
  * the network input image against an explicit double-sum inverse DFT (and numpy's irfft2 form);
  * ang2pix of the product against the reference's own HEALPix (golden fixture from src/Healpix_2.15a), the
    pixel-centre round trip and the NEST hierarchy (the order-N pixel of a direction is its order-(N+1) pixel
    divided by 4).
"""
import os

import numpy as np

import relion_amd as ra
from oracle import pose_cnn as pc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "healpix_ref.npz")


def test_oracle_gradients_match_finite_differences():
    W, C, NB = 16, 8, 48
    rng = np.random.default_rng(3)
    w = pc.init_params(W, C, NB, 11)
    # non-zero biases so that every term of the gradient is exercised
    conv, Wfc, bfc = pc.unpack(w, W, C, NB)
    conv = [(Wl, rng.normal(0, 0.05, C)) for Wl, _ in conv]
    w = pc.pack(conv, Wfc, rng.normal(0, 0.05, NB))
    imgs = rng.normal(size=(3, W, W))
    labels = np.array([5, 17, 40])
    loss, g, _ = pc.loss_and_grads(w, imgs, labels, C, NB)
    assert np.isfinite(loss) and g.shape == w.shape
    # padding columns of the first layer carry no gradient
    assert np.all(g[:C * 16].reshape(C, 16)[:, 9:] == 0)
    nz = np.flatnonzero(np.abs(g) > 1e-6)
    idx = rng.choice(nz, 60, replace=False)
    eps = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        fd = (pc.loss_and_grads(wp, imgs, labels, C, NB)[0] - pc.loss_and_grads(wm, imgs, labels, C, NB)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])) + 1e-7, (i, fd, g[i])


def test_oracle_convolution_matches_direct_loop():
    rng = np.random.default_rng(4)
    n, H, cin, C = 2, 9, 3, 5
    X = rng.normal(size=(n, H, H, cin))
    for s in (1, 2):
        Ho = (H - 1) // s + 1
        Kp = 9 * cin
        Wl = rng.normal(size=(C, Kp))
        out = (pc.im2col(X, s, Ho, Kp) @ Wl.T).reshape(n, Ho, Ho, C)
        ref = np.zeros_like(out)
        for b in range(n):
            for oy in range(Ho):
                for ox in range(Ho):
                    for ky in range(3):
                        for kx in range(3):
                            iy, ix = oy * s + ky - 1, ox * s + kx - 1
                            if 0 <= iy < H and 0 <= ix < H:
                                for c in range(cin):
                                    ref[b, oy, ox] += Wl[:, (3 * ky + kx) * cin + c] * X[b, iy, ix, c]
        assert np.allclose(out, ref, atol=1e-12)
        # col2im is the adjoint of im2col: <im2col(X), D> == <X, col2im(D)>
        D = rng.normal(size=(n * Ho * Ho, Kp))
        lhs = np.sum(pc.im2col(X, s, Ho, Kp) * D)
        rhs = np.sum(X * pc.col2im(D, n, H, cin, s, Ho))
        assert abs(lhs - rhs) < 1e-9 * abs(lhs)


def test_input_image_is_the_inverse_dft_of_the_window():
    rng = np.random.default_rng(5)
    s, W = 40, 16
    fimg = rng.normal(size=(s * (s // 2 + 1) * 2))
    a = pc.input_image(fimg, s, W)
    b = pc.input_image_direct(fimg, s, W)
    assert np.allclose(a, b, atol=1e-10)
    assert abs(a.mean()) < 1e-12 and abs((a ** 2).mean() - 1.0) < 1e-12
    # a single real coefficient at (kx, ky) = (2, 3) gives cos(2 pi (2u + 3v)/W) up to normalisation
    f = np.zeros((s, s // 2 + 1, 2))
    f[3, 2, 0] = 1.0
    img = pc.input_image(f.ravel(), s, W)
    u = np.arange(W)
    ref = np.cos(2 * np.pi * (2 * u[None, :] + 3 * u[:, None]) / W)
    assert np.allclose(img, ref / np.sqrt((ref ** 2).mean()), atol=1e-12)


def test_ang2pix_matches_reference_healpix():
    g = np.load(GOLDEN)
    z, phi = g["a_z"], g["a_phi"]
    # the 4000 random directions; the trailing special cases (poles, ring boundary, equator corner) lie exactly on
    # pixel boundaries, which the degree round trip of the (rot, tilt) interface does not preserve
    ok = np.arange(len(z)) < 4000
    tilt = np.degrees(np.arccos(z))
    rot = np.degrees(phi)
    for order in range(5):
        mine = np.array([ra.ang2pix(order, r, t) for r, t, k in zip(rot, tilt, ok) if k])
        assert np.array_equal(mine, g[f"a_pix{order}"][ok]), f"order {order}"


def test_ang2pix_inverts_the_grid_and_nests():
    for order in range(1, 5):
        cfg = ra.make_config(64, 2.0, order, 10.0, 2.0, 24)
        rot, tilt = ra.sampling_grid(cfg)[:2]
        pix = np.array([ra.ang2pix(order, r, t) for r, t in zip(rot, tilt)])
        assert np.array_equal(pix, np.arange(len(rot)))
        parent = np.array([ra.ang2pix(order - 1, r, t) for r, t in zip(rot, tilt)])
        assert np.array_equal(parent, np.arange(len(rot)) >> 2)
