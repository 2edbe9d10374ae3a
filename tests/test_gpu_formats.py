"""Particles read from MRC stacks and a particles STAR file (relion_amd.formats) give the image preparation on the
GPU (rln_imgprep_run) exactly the inputs, and so the outputs, of the same particles held in memory (bit for bit,
power spectra to 1e-12: they are sums of FP64 atomics)."""
import numpy as np
import pytest

import relion_amd as ra
from relion_amd import formats as F

pytestmark = pytest.mark.gpu


def test_star_mrc_particles_feed_image_preparation(tmp_path):
    rng = np.random.default_rng(12)
    N, n, px = 64, 40, 1.5
    y, x = np.mgrid[:N, :N] - N // 2
    imgs = rng.normal(scale=0.3, size=(n, N, N)).astype(np.float32)
    imgs += np.exp(-(x ** 2 + y ** 2) / 50.0).astype(np.float32)
    metas = [dict(norm_correction=rng.uniform(0.8, 1.2), old_offset_x=float(rng.integers(-3, 4)),
                  old_offset_y=float(rng.integers(-3, 4)), defocus_u=rng.uniform(9000, 20000), defocus_v=0.0,
                  defocus_angle=rng.uniform(0, 180), bfactor=0.0, kfactor=1.0, phase_shift=0.0, part_id=i)
             for i in range(n)]
    for m in metas:
        m["defocus_v"] = m["defocus_u"] - 300.0
    # two stacks, particles listed out of stack order
    F.write_mrc(tmp_path / "a.mrcs", imgs[:25], pixel_size=px)
    F.write_mrc(tmp_path / "b.mrcs", imgs[25:], pixel_size=px)
    order = rng.permutation(n)
    names = [f"{i + 1:06d}@a.mrcs" if i < 25 else f"{i - 24:06d}@b.mrcs" for i in order]
    F.write_particles(tmp_path / "particles.star", names, [metas[i] for i in order],
                      [dict(group=1, pixel_size=px, image_size=N, voltage=300.0, cs=2.7, q0=0.1)])
    print(f"[formats] {n} particles from 2 stacks into the image preparation, box {N}", flush=True)
    ps = F.read_particles(tmp_path / "particles.star", base_dir=tmp_path, threads=4)
    assert np.array_equal(ps.images, imgs[order])
    loaded = ps.imgprep_metas()
    for r, i in zip(loaded, order):
        r["part_id"] = int(i)
    opt = ps.optics[1]
    prep = ra.ImagePreparation(N, N, pixel_size=opt["pixel_size"], particle_diameter=0.6 * N * px,
                               voltage=opt["voltage"], cs=opt["cs"], q0=opt["q0"])
    out_file = prep.getFourierTransformsAndCtfs(ps.images, loaded)
    out_mem = prep.getFourierTransformsAndCtfs(imgs[order], [metas[i] for i in order])
    for key in ("fimg", "fimg_nomask", "fctf", "old_offset"):
        assert np.array_equal(out_file[key], out_mem[key]), key
    # spectra are sums of FP64 atomics, whose order varies from run to run
    for key in ("power", "highres_Xi2"):
        assert np.allclose(out_file[key], out_mem[key], rtol=1e-12, atol=0.0), key
    prep.close()
