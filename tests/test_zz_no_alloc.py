"""rln_backprojector_get_mdl_data_noalloc against rln_backprojector_get_mdl_data.

The entry point without device allocation folds the backprojector's slots a few brick layers at a time through a
static staging array; its volumes must equal the allocating entry point's bit for bit (same kernel, same summation
order per voxel). Two calls in a row: first a box-128 backprojector (65 brick layers, several passes through the
staging and a partial last one), then a smaller box-32 one, so that stale staging contents from the first call would
show. Both backprojectors are filled by an E-step. A volume whose 4-plane slab exceeds the staging is refused.
"""
import numpy as np
import pytest

import relion_amd as ra
from problems import make_problem

pytestmark = pytest.mark.gpu


def filled_backprojector(box, n, seed):
    p = make_problem(box=box, px=256.0 / box, order=1, offset_range_px=2, offset_step_px=1, n=n, snr=0.5,
                     nr_blobs=20, seed=seed)
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
    return bp, pr


@pytest.mark.parametrize("box,n,seed", [(128, 3, 71), (32, 4, 72)])
def test_noalloc_readback_equals_allocating_readback(box, n, seed):
    print(f"[no-alloc] box {box}, {n} particles, seed {seed}", flush=True)
    bp, _ = filled_backprojector(box, n, seed)
    re0, im0, w0 = bp.getMdlData()
    re1, im1, w1 = bp.getMdlDataNoAlloc()
    assert np.abs(w0).max() > 0 and np.abs(re0).max() > 0
    assert np.array_equal(re0, re1) and np.array_equal(im0, im1) and np.array_equal(w0, w1)
    # a second call on the same data gives the same volumes again
    re2, _, w2 = bp.getMdlDataNoAlloc()
    assert np.array_equal(re0, re2) and np.array_equal(w0, w2)


def test_noalloc_refuses_volumes_beyond_its_staging():
    print("[no-alloc] 700 x 1000 x 4 volume (4 planes need 33.6 MB of staging)", flush=True)
    bp = ra.AccBackprojector(0)
    bp.setMdlDim(700, 1000, 4, -500, -2, 10, 2.0)
    with pytest.raises(ra.RelionError, match="staging"):
        bp.getMdlDataNoAlloc()
