"""GPU reference preparation (rln_projector_compute_fourier_map through the C ABI) against the FP64 oracle
(oracle/refprep.py), and the prepared projector in an E-step.

Tolerances: the data array from an FP32 3-D FFT, max |gpu - oracle| <= 2e-5 of the largest component; power
spectrum within 1e-4 relative on shells holding more than 1e-6 of the peak power.
"""
import numpy as np
import pytest

import relion_amd as ra
from relion_amd import hipmem
from oracle import refprep as orc

pytestmark = pytest.mark.gpu


def _map(ori, seed):
    rng = np.random.default_rng(seed)
    spec = ra.synth_spec(seed, ori, 2.0, 20, snr=1.0, blob_radius=0.3 * ori)
    vol = ra.synth_real_map(spec, 0, ori) + rng.normal(scale=0.05, size=(ori,) * 3).astype(np.float32)
    return spec, vol.astype(np.float32)


@pytest.mark.parametrize("ori,current,grid,min_ires", [(32, 32, True, -1), (32, 20, False, -1), (48, 48, True, 3),
                                                       (48, 30, True, -1)])
def test_compute_fourier_map_matches_oracle(ori, current, grid, min_ires):
    _, vol = _map(ori, ori + current)
    print(f"[refprep] box {ori} current {current} gridding {grid} min_ires {min_ires}", flush=True)
    pr = ra.AccProjector(0)
    ps, data = pr.computeFourierTransformMap(vol, current, 2.0, do_gridding=grid, min_ires=min_ires,
                                             return_data=True)
    ref, ref_ps = orc.compute_fourier_map(vol, current, 2.0, do_gridding=grid, min_ires=min_ires)
    assert data.shape == ref.shape
    err = np.abs(data - ref).max() / np.abs(ref).max()
    print(f"[refprep] max rel err {err:.2e}")
    assert err <= 2e-5
    assert np.array_equal(data == 0, ref == 0)
    big = ref_ps > 1e-6 * ref_ps.max()
    assert np.allclose(ps[big], ref_ps[big], rtol=1e-4)
    assert np.all(ps[~big] <= 1e-5 * ref_ps.max())
    pr.clear()


def test_compute_fourier_map_from_device_memory():
    ori = 32
    _, vol = _map(ori, 9)
    print("[refprep] device-resident map", flush=True)
    d_vol = hipmem.DeviceArray(vol)
    pr = ra.AccProjector(0)
    ps_d, data_d = pr.computeFourierTransformMap(d_vol.address, ori, 2.0, return_data=True, ori_size=ori,
                                                 on_device=True)
    ps_h, data_h = pr.computeFourierTransformMap(vol, ori, 2.0, return_data=True)
    assert np.array_equal(data_d, data_h)
    assert np.allclose(ps_d, ps_h, rtol=1e-12)
    pr.clear()


def test_prepared_projector_drives_the_estep_like_the_analytic_one():
    # Particles from the analytic Fourier reference of a blob map; the E-step with a projector prepared on the GPU
    # from the same map sampled in real space (no gridding correction, so both describe the same transform) finds the
    # same best poses as with the analytic projector.
    import sys
    import os
    sys.path.insert(0, os.path.dirname(__file__))
    from problems import make_problem
    p = make_problem(box=32, px=4.0, order=1, offset_range_px=2, offset_step_px=1, n=12, snr=0.5, nr_blobs=20, seed=21)
    print("[refprep] E-step with a GPU-prepared projector, box 32, 12 particles", flush=True)
    vol = ra.synth_real_map(p.extra["spec"], 0, 32)
    cfg = p.cfg
    dims = ra.volume_dims(cfg.current_size)
    results = []
    for prepared in (False, True):
        pr = ra.AccProjector(0)
        if prepared:
            pr.computeFourierTransformMap(vol, cfg.current_size, 2.0, do_gridding=False)
        else:
            pr.setMdlDim(*dims, 2.0)
            pr.initMdl(p.volumes[0])
        bp = ra.AccBackprojector(0)
        bp.setMdlDim(*dims, 2.0)
        bp.initMdl()
        es = ra.ExpectationStep(cfg, 0)
        es.set_model([pr], [bp], p.pdf_class, p.sigma2_noise, pdf_direction=p.pdf_direction)
        res, _ = es.run(p.fimg, p.fctf, p.meta)
        results.append(res)
        es.close()
    same = np.mean(results[0]["best_ihidden_over"] == results[1]["best_ihidden_over"])
    print(f"[refprep] same best pose for {same:.0%} of particles")
    assert same >= 0.9
