"""GPU image preparation (rln_imgprep_run through the C ABI) against the FP64 oracle (oracle/imgprep.py).

Tolerances: transforms from FP32 FFTs, per image ||gpu - oracle|| <= 5e-6 ||oracle|| and max error <= 2e-5 of
the largest component; CTF images to FP32 rounding (rel 1e-6, abs 1e-7, FP64 evaluation on both sides); power
spectra and high-resolution power to rel 2e-5 per shell with significant power; rounded offsets exact.
"""
import numpy as np
import pytest

import relion_amd as ra
from relion_amd import hipmem
from oracle import imgprep as orc

pytestmark = pytest.mark.gpu


def _images(n, N, seed, offset=3.0):
    """Gaussian blobs near the centre plus white noise, and per-particle metadata dicts."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[:N, :N] - N // 2
    imgs = np.zeros((n, N, N), np.float32)
    metas = []
    for b in range(n):
        im = rng.normal(scale=0.3, size=(N, N))
        for _ in range(6):
            cx, cy = rng.uniform(-N / 5, N / 5, size=2)
            im += rng.uniform(0.5, 2.0) * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * rng.uniform(1.5, 3.0) ** 2))
        imgs[b] = im
        metas.append(dict(norm_correction=rng.uniform(0.5, 2.0), old_offset_x=rng.uniform(-offset, offset),
                          old_offset_y=rng.uniform(-offset, offset), defocus_u=rng.uniform(8000, 25000),
                          defocus_v=0.0, defocus_angle=rng.uniform(0, 180), bfactor=rng.uniform(0, 60),
                          kfactor=rng.uniform(0.8, 1.2), phase_shift=rng.uniform(0, 30), part_id=1000 + b))
        metas[-1]["defocus_v"] = metas[-1]["defocus_u"] - rng.uniform(0, 400)
    return imgs, metas


def _compare(out, ref, label):
    for b in range(len(ref["fimg"])):
        for key in ("fimg", "fimg_nomask"):
            g, r = out[key][b].astype(np.complex128), ref[key][b]
            err = np.linalg.norm(g - r)
            assert err <= 5e-6 * np.linalg.norm(r), (label, key, b, err / np.linalg.norm(r))
            assert np.abs(g - r).max() <= 2e-5 * np.abs(r).max(), (label, key, b)
        assert np.allclose(out["fctf"][b], ref["fctf"][b], rtol=1e-6, atol=1e-7), (label, "fctf", b)
        big = ref["power"][b] > 1e-6 * ref["power"][b].max()
        assert np.allclose(out["power"][b][big], ref["power"][b][big], rtol=2e-5), (label, "power", b)
        assert out["highres_Xi2"][b] == pytest.approx(ref["highres_Xi2"][b], rel=2e-5, abs=1e-12), (label, b)
    assert np.array_equal(out["old_offset"], ref["old_offset"]), label


@pytest.mark.parametrize("N,s,zero_mask", [(64, 48, True), (64, 64, True), (96, 64, False), (128, 100, False)])
def test_imgprep_matches_oracle(N, s, zero_mask):
    imgs, metas = _images(10, N, seed=N + s)
    print(f"[imgprep] box {N} window {s} zero_mask {zero_mask}", flush=True)
    prep = ra.ImagePreparation(N, s, pixel_size=1.5, particle_diameter=0.55 * N * 1.5, width_mask_edge=5.0,
                               zero_mask=zero_mask, sigma2_fudge=1.3, random_seed=99, avg_norm_correction=1.1)
    sigma2 = 0.05 + np.exp(-np.arange(N // 2 + 1) / 10.0)
    if not zero_mask:
        prep.set_sigma2_noise(sigma2)
    out = prep.getFourierTransformsAndCtfs(imgs, metas)
    ref = orc.prepare(imgs, metas, prep.config_dict(), sigma2=sigma2)
    _compare(out, ref, f"N{N} s{s}")
    if s == N:
        assert np.all(out["highres_Xi2"] == 0.0)
    else:
        assert np.all(out["highres_Xi2"] > 0.0)
    prep.close()


@pytest.mark.parametrize("flags", [dict(ctf_premultiplied=True), dict(ctf_phase_flipped=True),
                                   dict(only_flip_phases=True), dict(intact_ctf_first_peak=True),
                                   dict(do_ctf_correction=False)])
def test_imgprep_ctf_flags(flags):
    N, s = 64, 64
    imgs, metas = _images(4, N, seed=7)
    print(f"[imgprep] CTF flags {flags}", flush=True)
    prep = ra.ImagePreparation(N, s, pixel_size=1.1, particle_diameter=60.0, **flags)
    out = prep.getFourierTransformsAndCtfs(imgs, metas)
    ref = orc.prepare(imgs, metas, prep.config_dict())
    _compare(out, ref, str(flags))
    if not flags.get("do_ctf_correction", True):
        assert np.all(out["fctf"] == 1.0)
    prep.close()


def test_imgprep_offsets_beyond_edges_and_empty_batch():
    N = 64
    imgs, metas = _images(6, N, seed=11)
    for m, (ox, oy) in zip(metas, ((63.4, 0.0), (-64.0, 2.0), (31.5, -31.5), (-0.5, 0.5), (100.0, -100.0),
                                   (0.49, -0.49))):
        m["old_offset_x"], m["old_offset_y"] = ox, oy
    print("[imgprep] offsets beyond the box edges", flush=True)
    prep = ra.ImagePreparation(N, 48, pixel_size=2.0, particle_diameter=80.0)
    out = prep.getFourierTransformsAndCtfs(imgs, metas)
    ref = orc.prepare(imgs, metas, prep.config_dict())
    _compare(out, ref, "edges")
    empty = prep.getFourierTransformsAndCtfs(np.zeros((0, N, N), np.float32), [])
    assert empty["fimg"].shape == (0, 48, 25)
    with pytest.raises(ra.RelionError):
        bad = ra.ImagePreparation(N, 48, pixel_size=2.0, particle_diameter=80.0, avg_norm_correction=1.0)
        metas[0]["norm_correction"] = 0.0
        bad.getFourierTransformsAndCtfs(imgs, metas)
    prep.close()


def test_imgprep_chunks_and_device_buffers():
    # more particles than one chunk holds at box 32 (4096), device-resident input and outputs
    N, s, n = 32, 32, 4100
    imgs, metas = _images(n, N, seed=5, offset=2.0)
    print(f"[imgprep] {n} particles at box {N}: chunk boundary, device buffers", flush=True)
    prep = ra.ImagePreparation(N, s, pixel_size=3.0, particle_diameter=70.0, zero_mask=False, random_seed=3)
    sigma2 = np.full(N // 2 + 1, 0.02)
    prep.set_sigma2_noise(sigma2)
    host = prep.getFourierTransformsAndCtfs(imgs, metas)
    xs = s // 2 + 1
    d_img = hipmem.DeviceArray(imgs)
    d_f = hipmem.DeviceArray(np.zeros((n, s, xs), np.complex64))
    d_u = hipmem.DeviceArray(np.zeros((n, s, xs), np.complex64))
    d_c = hipmem.DeviceArray(np.zeros((n, s, xs), np.float32))
    hx, power, off = prep.run_device(n, d_img.address, metas, d_f.address, d_u.address, d_c.address)
    hipmem.synchronize()
    assert np.array_equal(d_f.to_host(), host["fimg"]) and np.array_equal(d_u.to_host(), host["fimg_nomask"])
    assert np.array_equal(d_c.to_host(), host["fctf"]) and np.array_equal(off, host["old_offset"])
    # spectra are summed with FP64 shared-memory atomics, whose order differs between runs
    assert np.allclose(power, host["power"], rtol=1e-12, atol=0)
    pick = [0, 1, 4095, 4096, 4099]
    ref = orc.prepare(imgs[pick], [metas[i] for i in pick], prep.config_dict(), sigma2=sigma2)
    _compare({k: v[pick] for k, v in host.items()}, ref, "chunks")
    t = prep.timings()
    print(f"[imgprep] timings {t}")
    assert t["total"] > 0.0
    prep.close()
