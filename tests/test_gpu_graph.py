"""The fine pass as captured graphs: replays agree with direct launches of the same kernels.

Three E-step runs replay the captured graphs (with host work and synchronisation between the runs); each is
compared with a run that launches the same kernels directly: identical max-likelihood poses, diff2 and fine
weights; sums of weights within 1e-12 relative (FP64 atomics add in a different order); backprojected volumes
within 1e-5 relative (FP32 atomics); the noise power per shell within 1e-6 relative (each item sums its pixels
into FP32 shared-memory bins with atomics, whose order differs between runs).
"""
import numpy as np
import pytest

import relion_amd as ra
from problems import make_problem

pytestmark = pytest.mark.gpu


def test_graph_replays_match_direct_launches():
    p = make_problem(box=48, px=3.0, order=2, offset_range_px=3, offset_step_px=1, n=12, snr=(0.01, 0.3),
                     nr_blobs=30, seed=53)
    print("[graph] box 48 order 2 global, 12 particles of uneven SNR", flush=True)
    cfg = p.cfg
    dims = ra.volume_dims(cfg.current_size)
    pr = ra.AccProjector(0)
    pr.setMdlDim(*dims, cfg.padding_factor)
    pr.initMdl(p.volumes[0])
    bp = ra.AccBackprojector(0)
    bp.setMdlDim(*dims, cfg.padding_factor)
    es = ra.ExpectationStep(cfg, 0)
    es.set_model([pr], [bp], p.pdf_class, p.sigma2_noise, pdf_direction=p.pdf_direction)

    def run():
        bp.initMdl()
        res, wsum = es.run(p.fimg, p.fctf, p.meta)
        return res, wsum, bp.getMdlData(), es.counters()

    es.use_graphs(False)
    r0, w0, v0, c0 = run()
    assert c0["fine_graph_launches"] == 0
    es.use_graphs(True)
    for rep in range(3):
        r1, w1, v1, c1 = run()
        print(f"[graph] replay {rep}: {c1['fine_graph_launches']} graph launches")
        assert c1["fine_graph_launches"] > 0
        for f in ("best_ihidden_over", "min_diff2", "coarse_min_diff2", "rot", "tilt", "psi", "xoff", "yoff",
                  "iclass", "nr_fine_evaluated"):
            assert np.array_equal(r0[f], r1[f]), f
        for f in ("sum_weight", "coarse_sum_weight", "pmax", "dLL"):
            assert np.allclose(r0[f], r1[f], rtol=1e-12, atol=0), f
        for a, b in zip(v0, v1):
            assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(a)
        assert np.allclose(w0["sigma2_noise"], w1["sigma2_noise"], rtol=1e-6, atol=0)
    es.close()
