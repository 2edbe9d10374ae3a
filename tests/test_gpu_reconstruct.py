"""GPU reconstruction (rln_backprojector_reconstruct through the C ABI) against the FP64 oracle
(oracle/reconstruct.py), on backprojectors filled by an E-step.

Tolerance: both sides compute in FP64 from the FP32 accumulators and the GPU map is stored as FP32, so
max |gpu - oracle| <= 1e-6 of the oracle's largest absolute value, on both the path without gridding iterations and the
gridding path (with shells of zero tau2, whose MAP term 1 / (0.001 weight) spans many decades).
"""
import os
import sys

import numpy as np
import pytest

import relion_amd as ra
from relion_amd import hipmem
from oracle import reconstruct as orc

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.dirname(__file__))
from problems import make_problem  # noqa: E402

BOX = 32
_filled = {}


def _filled_backprojector(current):
    """One E-step over 16 particles (box 32, HEALPix order 1) accumulated into a fresh backprojector."""
    if current not in _filled:
        p = make_problem(box=BOX, px=4.0, order=1, offset_range_px=2, offset_step_px=1, n=16, snr=0.5, nr_blobs=20,
                         seed=31 + current, current_size=current)
        print(f"[reconstruct] E-step fill, box {BOX}, current size {current}, 16 particles", flush=True)
        dims = ra.volume_dims(p.cfg.current_size)
        pr = ra.AccProjector(0)
        pr.setMdlDim(*dims, 2.0)
        pr.initMdl(p.volumes[0])
        bp = ra.AccBackprojector(0)
        bp.setMdlDim(*dims, 2.0)
        bp.initMdl()
        es = ra.ExpectationStep(p.cfg, 0)
        es.set_model([pr], [bp], p.pdf_class, p.sigma2_noise, pdf_direction=p.pdf_direction)
        es.run(p.fimg, p.fctf, p.meta)
        es.close()
        re, im, w = bp.getMdlData()
        _filled[current] = (bp, pr, re.astype(np.float64) + 1j * im.astype(np.float64), w.astype(np.float64),
                            dims[5])
    return _filled[current]


def _tau2():
    # a decaying signal spectrum whose last shells are 0 (the 1 / (0.001 weight) branch)
    t = 50.0 * np.exp(-np.arange(BOX // 2 + 1) / 4.0)
    t[-3:] = 0.0
    return t


@pytest.mark.parametrize("current,skip,iters,do_map", [(32, True, 10, True), (20, True, 0, False),
                                                       (32, False, 10, True), (20, False, 5, True),
                                                       (32, False, 3, False)])
def test_reconstruct_matches_oracle(current, skip, iters, do_map):
    bp, _, data, weight, r_max = _filled_backprojector(current)
    print(f"[reconstruct] current {current} skip_gridding {skip} iterations {iters} MAP {do_map}", flush=True)
    kw = dict(max_iter_preweight=iters, do_map=do_map, tau2=_tau2(), tau2_fudge=2.0, normalise=7.5, minres_map=3)
    gpu = bp.reconstruct(BOX, skip_gridding=skip, **kw)
    ref = orc.reconstruct(data, weight, BOX, 2.0, skip_gridding=skip, r_max=r_max, **kw)
    err = np.abs(gpu - ref).max() / np.abs(ref).max()
    print(f"[reconstruct] max rel err {err:.2e}")
    assert np.abs(ref).max() > 0
    assert err <= 1e-6


def test_reconstruct_to_device_memory_and_repeat():
    bp, _, _, _, _ = _filled_backprojector(32)
    print("[reconstruct] device output, repeated call", flush=True)
    host = bp.reconstruct(BOX, max_iter_preweight=4, skip_gridding=False, tau2=_tau2())
    d_out = hipmem.DeviceArray(np.zeros((BOX,) * 3, np.float32))
    assert bp.reconstruct(BOX, max_iter_preweight=4, skip_gridding=False, tau2=_tau2(),
                          out_address=d_out.address) is None
    assert np.array_equal(d_out.to_host(), host)
    # the accumulated data and weight are left as they were
    assert np.array_equal(bp.reconstruct(BOX, max_iter_preweight=4, skip_gridding=False, tau2=_tau2()), host)


def test_reconstruct_empty_backprojector_and_argument_errors():
    print("[reconstruct] empty backprojector", flush=True)
    dims = ra.volume_dims(BOX)
    bp = ra.AccBackprojector(0)
    bp.setMdlDim(*dims, 2.0)
    bp.initMdl()
    assert not np.any(bp.reconstruct(BOX, do_map=False))
    with pytest.raises(ra.RelionError):
        bp.reconstruct(BOX + 1, do_map=False)  # odd box
    with pytest.raises(ra.RelionError):
        bp.reconstruct(BOX, do_map=True, tau2=None)  # MAP without a spectrum
    with pytest.raises(ra.RelionError):
        bp.reconstruct(BOX, do_map=False, normalise=0.0)
    bp.clear()
