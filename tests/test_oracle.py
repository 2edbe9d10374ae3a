"""CPU checks that pin the oracle (and the product's host-side sampling / CTF / generator code).

Each check is chosen so that a plausible mistake (transposed rotation, wrong sign of a shift, off-by-one in the
Xmipp origin, wrong corner order) fails: reference golden vectors (HEALPix built from the reference's own
sources; the reference's CTF known-answer test), closed forms (trilinear interpolation is exact on linear data,
the integer-shift theorem of the DFT), exact invariants (backprojection is the adjoint of projection) and
brute force on tiny inputs (the exact-match pose has the minimum squared difference).
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
import relion_amd as ra
from problems import make_problem

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "healpix_ref.npz")


# ------------------------------------------------------------------------------------------------
# HEALPix against the reference's own implementation (golden vectors from oracle/_ref)
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("order", range(5))
def test_healpix_oracle_matches_reference_golden(order):
    g = np.load(GOLDEN)
    n = 12 * 4 ** order
    for p in range(n):
        z, phi = C.c_double(), C.c_double()
        oracle.lib.orc_hp_pix2ang_z_phi(order, p, C.byref(z), C.byref(phi))
        assert z.value == g[f"z{order}"][p] and phi.value == g[f"phi{order}"][p]
        ix, iy, f = C.c_int(), C.c_int(), C.c_int()
        oracle.lib.orc_hp_nest2xyf(order, p, C.byref(ix), C.byref(iy), C.byref(f))
        assert (ix.value, iy.value, f.value) == (g[f"ix{order}"][p], g[f"iy{order}"][p], g[f"face{order}"][p])
        assert oracle.lib.orc_hp_xyf2nest(order, ix.value, iy.value, f.value) == p


@pytest.mark.parametrize("order", range(5))
def test_product_directions_match_reference_golden(order):
    g = np.load(GOLDEN)
    cfg = ra.make_config(64, 2.0, order, 10.0, 2.0, 24)
    rot, tilt, _, _, _ = ra.sampling_grid(cfg)
    # HealpixSampling::getDirectionFromHealPix: rot = phi (wrapped to (-180,180]), tilt = acos(z)
    exp_tilt = np.degrees(np.arccos(g[f"z{order}"]))
    exp_rot = np.degrees(g[f"phi{order}"])
    exp_rot = np.where(exp_rot > 180.0, exp_rot - 360.0, exp_rot)
    assert np.allclose(tilt, exp_tilt, atol=1e-12, rtol=0)
    assert np.allclose(rot, exp_rot, atol=1e-12, rtol=0)


def test_healpix_live_reference_when_built():
    ref = oracle.ref_healpix()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    for order in (5, 6):
        for p in range(0, 12 * 4 ** order, 97):
            a, b, c, d = C.c_double(), C.c_double(), C.c_double(), C.c_double()
            oracle.lib.orc_hp_pix2ang_z_phi(order, p, C.byref(a), C.byref(b))
            ref.ref_hp_pix2ang_z_phi(order, p, C.byref(c), C.byref(d))
            assert (a.value, b.value) == (c.value, d.value)


# ------------------------------------------------------------------------------------------------
# CTF: the reference's only unit test (tests/ctf.cpp:5-10) plus closed forms
# ------------------------------------------------------------------------------------------------
def test_ctf_reference_known_answer():
    # CTF::setValues(10000, 12000, 90, 300, 2.7, 0.1, 0, 1, 0); getCTF(10, 10) == Approx(0.59154)
    args = (10000.0, 12000.0, 90.0, 300.0, 2.7, 0.1, 0.0, 1.0, 0.0, 10.0, 10.0)
    for v in (oracle.ctf_value(*args), ra.ctf_value(*args)):
        assert v == pytest.approx(0.59154, rel=1.2e-5)


def test_ctf_closed_forms():
    # at zero frequency the CTF is the amplitude contrast Q0: -sin(-atan(Q0/sqrt(1-Q0^2))) = Q0
    for q0 in (0.07, 0.1, 0.2):
        assert oracle.ctf_value(15000, 14000, 30, 300, 2.7, q0, 0, 1, 0, 0.0, 0.0) == pytest.approx(q0, abs=1e-15)
    # without astigmatism the CTF is rotationally symmetric
    rng = np.random.default_rng(0)
    for _ in range(20):
        r, a1, a2 = rng.uniform(0.01, 0.3), rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi)
        v1 = oracle.ctf_value(20000, 20000, 17, 300, 2.7, 0.1, 0, 1, 0, r * np.cos(a1), r * np.sin(a1))
        v2 = oracle.ctf_value(20000, 20000, 17, 300, 2.7, 0.1, 0, 1, 0, r * np.cos(a2), r * np.sin(a2))
        assert v1 == pytest.approx(v2, abs=1e-12)
    # astigmatism: swapping defocus U/V equals rotating the azimuth by 90 degrees
    for _ in range(20):
        x, y = rng.uniform(-0.2, 0.2, 2)
        a = oracle.ctf_value(18000, 15000, 30, 300, 2.7, 0.1, 0, 1, 0, x, y)
        b = oracle.ctf_value(15000, 18000, 120, 300, 2.7, 0.1, 0, 1, 0, x, y)
        assert a == pytest.approx(b, abs=1e-12)
        assert a == pytest.approx(ra.ctf_value(18000, 15000, 30, 300, 2.7, 0.1, 0, 1, 0, x, y), abs=1e-14)


# ------------------------------------------------------------------------------------------------
# Euler angles
# ------------------------------------------------------------------------------------------------
def _rz(t):
    c, s = np.cos(t), np.sin(t)
    return np.array([[c, s, 0], [-s, c, 0], [0, 0, 1.0]])


def _ry(t):
    c, s = np.cos(t), np.sin(t)
    return np.array([[c, 0, -s], [0, 1, 0], [s, 0, c]])


def test_euler_is_zyz_composition():
    rng = np.random.default_rng(1)
    for _ in range(50):
        a, b, g = rng.uniform(-180, 180), rng.uniform(0, 180), rng.uniform(-180, 180)
        A = oracle.euler(a, b, g)
        expected = _rz(np.radians(g)) @ _ry(np.radians(b)) @ _rz(np.radians(a))
        assert np.allclose(A, expected, atol=1e-14)
        assert np.allclose(A @ A.T, np.eye(3), atol=1e-14)
        assert np.linalg.det(A) == pytest.approx(1.0, abs=1e-13)


# ------------------------------------------------------------------------------------------------
# Sampling grids: closed counts and the order-1 oversampling rules
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("order,rng_px,step_px,ntrans", [(2, 5, 1, 81), (3, 6, 1, 113), (3, 4, 0.5, 197), (4, 6, 2, 29)])
def test_sampling_counts(order, rng_px, step_px, ntrans):
    cfg = ra.make_config(128, 1.0, order, float(rng_px), float(step_px), 32)
    nd, npsi, nt = oracle.sampling_counts(cfg)
    assert nd == 12 * 4 ** order and npsi == 6 * 2 ** order * 1 and nt == ntrans
    assert ra.sampling_counts(cfg) == (nd, npsi, nt)
    og, pg = oracle.sampling_grid(cfg), ra.sampling_grid(cfg)
    for a, b in zip(og, pg):
        assert np.array_equal(a, b)


def test_oversampled_orientations_are_nest_children():
    cfg = ra.make_config(64, 2.0, 2, 10.0, 2.0, 24)
    rot, tilt, psi, _, _ = oracle.sampling_grid(cfg)
    step = psi[1] - psi[0]
    for idir in (0, 17, 100, 191):
        r8, t8, p8 = oracle.oversampled_orientations(cfg, idir, 5)
        children = []
        for p in range(4 * idir, 4 * idir + 4):  # NEST children of pixel idir at order+1
            z, phi = C.c_double(), C.c_double()
            oracle.lib.orc_hp_pix2ang_z_phi(3, p, C.byref(z), C.byref(phi))
            rr = np.degrees(phi.value)
            children.append((rr - 360.0 if rr > 180.0 else rr, np.degrees(np.arccos(z.value))))
        got = sorted(set(zip(np.round(r8, 9), np.round(t8, 9))))
        assert got == sorted(set((round(a, 9), round(b, 9)) for a, b in children))
        assert np.allclose(sorted(set(np.round(p8, 9))), [psi[5] - step / 4, psi[5] + step / 4])


def test_oversampled_translations():
    cfg = ra.make_config(64, 2.0, 2, 10.0, 2.0, 24)
    _, _, _, tx, ty = oracle.sampling_grid(cfg)
    for it in (0, 40, 80):
        x, y = oracle.oversampled_translations(cfg, it)
        assert sorted(set(np.round(x * 2.0, 12))) == [tx[it] - 0.5, tx[it] + 0.5]
        assert sorted(set(np.round(y * 2.0, 12))) == [ty[it] - 0.5, ty[it] + 0.5]


def test_local_priors_product_matches_oracle():
    cfg = ra.make_config(128, 1.5, 3, 6.0, 1.5, 40, local_search=True, sigma_ang=7.5)
    rng = np.random.default_rng(4)
    for _ in range(30):
        pr = (rng.uniform(-180, 180), rng.uniform(0, 180), rng.uniform(0, 360))
        o = oracle.sampling_local(cfg, *pr)
        p = ra.sampling_local(cfg, *pr)
        for a, b in zip(o, p):
            assert np.array_equal(a, b)
        assert o[1].sum() == pytest.approx(1.0) and o[3].sum() == pytest.approx(1.0)
        assert 0 < len(o[0]) < 100 and 0 < len(o[2]) <= 7


def test_coarse_image_size_rule():
    # image_coarse_size = 2*CEIL(ori*px / (angle/360*pi*D/1.2)) (src/ml_optimiser.cpp:5763-5769)
    assert ra.coarse_image_size(64, 2.0, 2, 100.0) == 24
    assert ra.coarse_image_size(256, 1.0, 3, 200.0) == 48
    assert ra.coarse_image_size(64, 2.0, 0, 100.0) == 6


# ------------------------------------------------------------------------------------------------
# Projection, backprojection and shifts
# ------------------------------------------------------------------------------------------------
def _linear_hermitian_volume(current_size, pad, coef, const):
    xd, yd, zd, iy, iz, _ = ra.volume_dims(current_size, pad)
    z, y, x = np.meshgrid(np.arange(zd) + iz, np.arange(yd) + iy, np.arange(xd), indexing="ij")
    # V(k) = const + i (a . k): V(-k) = conj(V(k)), linear in k
    v = const + 1j * (coef[0] * x + coef[1] * y + coef[2] * z)
    return np.stack([v.real, v.imag], -1), (xd, yd, zd, iy, iz)


def test_projection_exact_on_linear_volume():
    """Trilinear interpolation reproduces linear data exactly, so every projected pixel must equal
    V(pad * A^T (x, y, 0)); a transposed matrix, a wrong origin or a missing Hermitian conjugate fails."""
    s, pad = 24, 2.0
    coef, const = (0.3, -0.7, 1.1), 2.5
    vol, _ = _linear_hermitian_volume(s, pad, coef, const)
    rng = np.random.default_rng(2)
    for _ in range(10):
        A = oracle.euler(*rng.uniform(-180, 180, 3))
        img = oracle.project(vol, s // 2, pad, A, s)
        for i in range(s):
            y = i if i <= s // 2 else i - s
            for x in range(s // 2 + 1):
                if x * x + y * y > (s // 2) ** 2:
                    assert img[i, x] == 0
                    continue
                p3 = pad * (A.T @ np.array([x, y, 0.0]))
                expected = const + 1j * np.dot(coef, p3)
                assert abs(img[i, x] - expected) < 1e-9


def test_projection_identity_is_central_section():
    s, pad = 16, 2.0
    rng = np.random.default_rng(3)
    xd, yd, zd, iy, iz, _ = ra.volume_dims(s, pad)
    vol = rng.normal(size=(zd, yd, xd, 2))
    img = oracle.project(vol, s // 2, pad, np.eye(3), s)
    for i in range(s):
        y = i if i <= s // 2 else i - s
        for x in range(s // 2 + 1):
            if x * x + y * y <= (s // 2) ** 2:
                v = vol[0 - iz, int(pad * y) - iy, int(pad * x)]
                assert img[i, x] == v[0] + 1j * v[1]


def test_backprojection_is_adjoint_of_projection():
    s, pad = 20, 2.0
    rng = np.random.default_rng(5)
    xd, yd, zd, iy, iz, _ = ra.volume_dims(s, pad)
    V = rng.normal(size=(zd, yd, xd, 2))
    for _ in range(5):
        A = oracle.euler(*rng.uniform(-180, 180, 3))
        I = rng.normal(size=(s, s // 2 + 1)) + 1j * rng.normal(size=(s, s // 2 + 1))
        I[s // 2 + 1:, 0] = 0  # the duplicated x == 0 half column is not backprojected
        PV = oracle.project(V, s // 2, pad, A, s)
        BI = np.zeros((zd, yd, xd, 2))
        BW = np.zeros((zd, yd, xd))
        oracle.backproject(BI, BW, s // 2, pad, A, I, np.ones(I.shape))
        lhs = np.sum((np.conj(I) * PV).real)
        rhs = np.sum(V[..., 0] * BI[..., 0] + V[..., 1] * BI[..., 1])
        assert lhs == pytest.approx(rhs, rel=1e-10)
        # trilinear weights of every backprojected pixel sum to one
        inside = sum(1 for i in range(s) for x in range(s // 2 + 1)
                     if x * x + (i if i <= s // 2 else i - s) ** 2 <= (s // 2) ** 2 and not (i > s // 2 and x == 0))
        assert BW.sum() == pytest.approx(inside, rel=1e-12)


def test_shift_equals_circular_shift_for_integer_offsets():
    rng = np.random.default_rng(6)
    s = 16
    img = rng.normal(size=(s, s))
    F = np.fft.rfft2(img)
    for dx, dy in ((1, 0), (0, 2), (-3, 1), (2, -2)):
        got = oracle.shift_image(F, float(s), float(dx), float(dy))
        expected = np.fft.rfft2(np.roll(img, shift=(dy, dx), axis=(0, 1)))
        assert np.allclose(got, expected, atol=1e-10)


# ------------------------------------------------------------------------------------------------
# E-step: brute force on tiny inputs and invariants
# ------------------------------------------------------------------------------------------------
def test_exact_pose_has_minimum_diff2():
    """A particle that is exactly the projection at grid pose (o*, t*) (CTF 1, no noise) has diff2 == Xi2/2
    there, and larger values everywhere else; pins the pose indexing and the shift direction."""
    box, px = 24, 4.0
    cfg = ra.make_config(box, px, 1, 8.0, 4.0, box, do_ctf_correction=False)
    spec = ra.synth_spec(9, box, px, 10, snr=1.0)
    vol = ra.synth_reference(spec, 0, box)
    rot, tilt, psi, tx, ty = oracle.sampling_grid(cfg)
    nd, npsi, nt = len(rot), len(psi), len(tx)
    idir, ipsi, it = 13, 7, 6
    A = oracle.euler(rot[idir], tilt[idir], psi[ipsi])
    P = oracle.project(vol, box // 2, 2.0, A, box)
    X = oracle.shift_image(P, float(box), -tx[it] / px, -ty[it] / px)  # particle shifted by -t
    meta = (ra.ParticleMeta * 1)()
    meta[0].scale = 1.0
    meta[0].highres_Xi2 = 3.0
    s2 = np.ones(box // 2 + 1)
    res, _, _, dbg = oracle.estep(cfg, [vol], [1.0], s2, meta, X[None], np.ones((1, box, box // 2 + 1)),
                                  debug_particles=[0], with_bp=False)
    d2 = dbg[0]["coarse_diff2"]
    best = (idir * npsi + ipsi) * nt + it
    assert d2[best] == pytest.approx(1.5, abs=1e-9)
    assert np.argmin(d2) == best and np.sort(d2)[1] > 1.5 + 1e-3


def test_significance_set_is_minimal():
    p = make_problem(box=32, px=4.0, order=1, offset_range_px=3, offset_step_px=1, n=3, snr=0.02, nr_blobs=20, seed=8)
    res, _, _, dbg = oracle.estep(p.cfg, p.volumes, p.pdf_class, p.sigma2_noise, p.meta, p.fimg, p.fctf,
                                  pdf_direction=p.pdf_direction, debug_particles=[0, 1, 2], with_bp=False)
    f = p.cfg.adaptive_fraction
    for ip in range(3):
        w = dbg[ip]["coarse_weight"]
        m = dbg[ip]["coarse_significant"].astype(bool)
        tot = w.sum()
        assert tot == pytest.approx(res[ip]["coarse_sum_weight"], rel=1e-12)
        sig = np.sort(w[m])[::-1]
        assert sig.sum() > f * tot                      # the significant set holds the requested fraction
        assert sig[:-1].sum() <= f * tot                # and no smaller descending prefix does
        assert np.all(w[~m] <= sig[-1])                 # it is a top set
        assert res[ip]["nr_significant"] == m.sum()


def test_estep_invariants_and_recovery():
    p = make_problem(box=48, px=3.0, order=2, offset_range_px=3, offset_step_px=1, n=6, snr=2.0, nr_blobs=25, seed=10)
    res, ws, (bpd, bpw), _ = oracle.estep(p.cfg, p.volumes, p.pdf_class, p.sigma2_noise, p.meta, p.fimg, p.fctf,
                                          pdf_direction=p.pdf_direction)
    n = len(res)
    assert np.all((res["pmax"] > 0) & (res["pmax"] <= 1.0))
    assert p.cfg.adaptive_fraction * n <= ws["sumw_group"] <= n * (1 + 1e-12)
    assert ws["pdf_class"].sum() == pytest.approx(ws["sumw_group"])
    assert ws["pdf_direction"].sum() == pytest.approx(ws["sumw_group"])
    assert np.all(bpw >= 0)
    # high-SNR particles are assigned within the fine sampling of their true orientation
    for i in range(n):
        A = oracle.euler(*p.pose[i, :3])
        B = oracle.euler(res[i]["rot"], res[i]["tilt"], res[i]["psi"])
        ang = np.degrees(np.arccos(np.clip((np.trace(A @ B.T) - 1) / 2, -1, 1)))
        assert ang < 10.0
        assert abs(res[i]["xoff"] - p.pose[i, 3]) <= 0.75 and abs(res[i]["yoff"] - p.pose[i, 4]) <= 0.75
