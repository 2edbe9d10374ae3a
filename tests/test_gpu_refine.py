"""The relion_refine drop-in end to end: a command line naming a particles STAR file, its MRC stacks and a
reference map runs the first iteration's E-step on the device (relion_amd.refine.first_iteration_estep: parse,
sizes, reference preparation, image preparation, E-step, backprojection), against the chained FP64 oracles on
the same data (oracle/refprep.py, oracle/imgprep.py, the E-step oracle).

Data: a 64^3 blob map at 2 A per pixel; 8 particles made on the host by rotating the map (trilinear), projecting
along z, multiplying by the CTF of the defocus written to the STAR file, shifting by whole pixels and adding white
noise. The device and the oracles read the same numbers (MRC float32 is exact). The prepared inputs differ by FP32
transform rounding (within 5e-6 of the image norm, tests/test_gpu_imgprep.py), which the E-step carries on, hence
tolerances of 1e-4 relative on diff2 and dLL and 1e-3 on the backprojected volumes; the best poses are the same.
"""
import numpy as np
import pytest
from scipy import ndimage

import oracle
import relion_amd as ra
from oracle import imgprep as orc_prep
from oracle import refprep as orc_ref
from relion_amd import formats as F
from relion_amd import refine as R

pytestmark = pytest.mark.gpu


def euler(rot, tilt, psi):
    a, b, g = np.radians([rot, tilt, psi])
    ca, sa, cb, sb, cg, sg = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(g), np.sin(g)
    return np.array([[cg * cb * ca - sg * sa, cg * cb * sa + sg * ca, -cg * sb],
                     [-sg * cb * ca - cg * sa, -sg * cb * sa + cg * ca, sg * sb],
                     [sb * ca, sb * sa, cb]])


def make_files(tmp, n=8, N=64, px=2.0, seed=5):
    rng = np.random.default_rng(seed)
    spec = ra.synth_spec(seed, N, px, 40, snr=1.0)
    vol = ra.synth_real_map(spec, 0, N)
    c = np.array([N // 2] * 3, np.float64)
    imgs, metas = np.zeros((n, N, N), np.float32), []
    for i in range(n):
        rot, tilt, psi = rng.uniform(0, 360), np.degrees(np.arccos(rng.uniform(-1, 1))), rng.uniform(0, 360)
        A = euler(rot, tilt, psi)
        # rotated map v'(r) = v(A^T r) in (x, y, z); arrays are (z, y, x)
        P = np.diag([1, 1, 1])[::-1]
        M = P @ A.T @ P
        rv = ndimage.affine_transform(vol, M, offset=c - M @ c, order=1)
        img = rv.sum(0)
        du = rng.uniform(10000, 20000)
        m = dict(defocus_u=du, defocus_v=du - 200.0, defocus_angle=rng.uniform(0, 180), bfactor=0.0, kfactor=1.0,
                 phase_shift=0.0, norm_correction=1.0, old_offset_x=0.0, old_offset_y=0.0, part_id=i)
        ctf = orc_prep.ctf_image(N, N, px, 300.0, 2.7, 0.1, m["defocus_u"], m["defocus_v"], m["defocus_angle"], 0.0,
                                 1.0, 0.0)
        Fi = np.fft.rfft2(np.fft.ifftshift(img)) * ctf
        img = np.fft.fftshift(np.fft.irfft2(Fi, s=(N, N)))
        img = np.roll(img, (rng.integers(-2, 3), rng.integers(-2, 3)), (0, 1))
        img += rng.normal(scale=0.5 * img.std(), size=img.shape)
        imgs[i] = img
        metas.append(m)
    F.write_mrc(tmp / "ref.mrc", vol, pixel_size=px)
    F.write_mrc(tmp / "stack.mrcs", imgs, pixel_size=px)
    F.write_particles(tmp / "particles.star", [f"{i + 1:06d}@stack.mrcs" for i in range(n)], metas,
                      [dict(group=1, pixel_size=px, image_size=N, voltage=300.0, cs=2.7, q0=0.1)])
    noise_var = float(np.mean([(0.5 * im.std()) ** 2 for im in imgs]))
    return vol, imgs, metas, noise_var


def test_first_iteration_estep_from_files_matches_oracles(tmp_path):
    N, px = 64, 2.0
    vol, imgs, metas, nv = make_files(tmp_path, N=N, px=px)
    argv = ["relion_refine", "--i", str(tmp_path / "particles.star"), "--ref", str(tmp_path / "ref.mrc"), "--o",
            str(tmp_path / "run"), "--ctf", "--particle_diameter", "100", "--healpix_order", "2", "--offset_range", "3",
            "--offset_step", "1", "--ini_high", "8", "--zero_mask", "--perturb", "0", "--pad", "2", "--gpu", "0",
            "--pool", "8", "--j", "4"]
    sigma2 = np.full(N // 2 + 1, nv / (2.0 * N * N))  # white noise of variance nv after the 1/N^2 transform
    print(f"[refine] {len(metas)} particles, box {N}, files in {tmp_path}", flush=True)
    out = R.first_iteration_estep(argv, sigma2)
    cur, coarse = out["current_size"], out["coarse_size"]
    # ini_high 8 A: ROUND(128 / 8) + 10 = 26 shells -> 52; 15 deg at 100 A -> 2 CEIL(128 / 10.91) = 24
    assert (cur, coarse) == (52, 24)
    cfg = out["config"]
    assert (cfg.offset_range, cfg.offset_step, cfg.healpix_order, cfg.do_ctf_correction) == (6.0, 2.0, 2, 1)
    pc = out["imgprep_config"]
    assert (pc["current_size"], pc["particle_diameter"], pc["zero_mask"], pc["do_ctf_correction"]) == (52, 100.0, 1, 1)
    # oracles on the same numbers
    data, _ = orc_ref.compute_fourier_map(vol, cur, 2.0)
    ref_vol = np.stack([data.real, data.imag], -1)
    prep = orc_prep.prepare(imgs, metas, pc, sigma2=sigma2)
    meta = (ra.ParticleMeta * len(metas))()
    for i in range(len(metas)):
        meta[i].old_offset_x, meta[i].old_offset_y = prep["old_offset"][i]
        meta[i].scale = 1.0
        meta[i].highres_Xi2 = prep["highres_Xi2"][i]
    nd, _, _ = ra.sampling_counts(cfg)
    ro, wo, (bpd, bpw), _ = oracle.estep(cfg, [ref_vol], np.ones(1), sigma2, meta, prep["fimg"], prep["fctf"],
                                         fimg_nomask=prep["fimg_nomask"], pdf_direction=np.full((1, nd), 1.0 / nd))
    rg = out["results"]
    same = rg["best_ihidden_over"] == ro["best_ihidden_over"]
    print(f"[refine] best poses equal for {same.sum()} of {len(same)}; pmax {np.round(rg['pmax'], 3)}", flush=True)
    assert same.all()
    for f in ("min_diff2", "coarse_min_diff2", "dLL"):
        err = np.max(np.abs(rg[f] - ro[f]) / np.abs(ro[f]))
        assert err < 1e-4, f"{f}: {err:.2e}"
    assert np.allclose(rg["pmax"], ro["pmax"], rtol=1e-3, atol=1e-6)
    re, im, w = out["backprojector"].getMdlData()
    num = np.sqrt(((re - bpd[0, ..., 0]) ** 2 + (im - bpd[0, ..., 1]) ** 2).sum())
    den = np.sqrt((bpd[0] ** 2).sum())
    errw = np.abs(w - bpw[0]).sum() / np.abs(bpw[0]).sum()
    print(f"[refine] backprojection rel err {num / den:.2e}, weight {errw:.2e}", flush=True)
    assert num <= 1e-3 * den and errw <= 1e-3
    # the assignments written as relion_refine's _data.star columns read back as computed
    R.write_data_star(tmp_path / "run_it001_data.star", out["particles"], rg, px)
    st = F.StarFile.read(tmp_path / "run_it001_data.star")
    pb = st.block("particles")
    assert np.allclose(pb.floats("rlnAngleRot"), rg["rot"]) and np.allclose(pb.floats("rlnAnglePsi"), rg["psi"])
    assert np.allclose(pb.floats("rlnOriginXAngst"), rg["xoff"] * px)
    assert np.array_equal(pb.ints("rlnClassNumber"), rg["iclass"])
    out["backprojector"].clear()
