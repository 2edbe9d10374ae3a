"""Numerical guards of the E-step through the C ABI, driven by inputs (the reference's error classes,
src/error.h:260-280; GPU guards src/acc/acc_ml_optimiser_impl.h:2242-2256, 2282-2299, 2505-2520; CPU guards
src/ml_optimiser.cpp:8054-8104):
  * every orientation prior zero: no pose is evaluated (zero-prior orientations are skipped, as at
    src/ml_optimiser.cpp:7228), the sum of weights is zero -> RLN_ERR_SUM_WEIGHT_ZERO, on the plain and the
    split-BF16 coarse pass, and the oracle returns the same status;
  * a NaN pixel inside the coarse window: every coarse diff2 of the particle is NaN -> zero / NaN sum of the
    coarse weights -> RLN_ERR_SUM_WEIGHT_ZERO;
  * a NaN pixel beyond the coarse window: the coarse pass succeeds and the fine pass's sum is NaN ->
    RLN_ERR_SUM_WEIGHT_ZERO, and the failed particle adds nothing to the backprojector (no NaN reaches it).
After an error the engine serves the next batch, which matches the oracle.

RLN_ERR_NO_SIGNIFICANT and RLN_ERR_FILTERED_ZERO are kept for the reference's error classes but no input reaches
them on this path: the significance rule (the descending cumulative sum crossing adaptive_fraction * sum,
src/ml_optimiser.cpp:8145-8168) always keeps the largest positive weight once the sum is positive, and a positive
sum has a positive weight.
"""
import numpy as np
import pytest

import relion_amd as ra
from problems import make_problem
from test_gpu_parity import run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def _problem(precision):
    p = make_problem(box=32, px=4.0, order=1, offset_range_px=3, offset_step_px=1, n=4, snr=0.5, nr_blobs=20, seed=12)
    p.cfg.precision = precision
    return p


@pytest.mark.parametrize("precision", [ra.RLN_PREC_PLAIN_FP32, ra.RLN_PREC_SPLIT_BF16])
def test_zero_orientation_priors_give_sum_weight_zero(precision):
    print(f"[errors] zero pdf_direction, precision {precision}", flush=True)
    p = _problem(precision)
    p.pdf_direction = np.zeros_like(p.pdf_direction)
    with pytest.raises(ra.RelionError) as ei:
        run_gpu(p)
    assert ei.value.code == ra.RLN_ERR_SUM_WEIGHT_ZERO
    with pytest.raises(RuntimeError, match="status -6"):
        run_oracle(p)


@pytest.mark.parametrize("where", ["coarse", "fine"])
def test_nan_pixel_gives_sum_weight_zero_and_engine_recovers(where):
    print(f"[errors] NaN pixel in the {where} window", flush=True)
    p = _problem(ra.RLN_PREC_PLAIN_FP32)
    s, sc = p.cfg.current_size, p.cfg.coarse_size
    good = p.fimg.copy()
    bad = p.fimg.copy()
    if where == "coarse":
        bad[2, 1, 1, :] = np.nan          # x = 1, y = 1: inside the coarse window
    else:
        x = s // 2 - 1                    # y = 0, x beyond the coarse window, inside the current size
        assert x > sc // 2
        bad[2, 0, x, :] = np.nan
    dims = ra.volume_dims(s, p.cfg.padding_factor)
    pr = ra.AccProjector(0)
    pr.setMdlDim(*dims, p.cfg.padding_factor)
    pr.initMdl(p.volumes[0])
    bp = ra.AccBackprojector(0)
    bp.setMdlDim(*dims, p.cfg.padding_factor)
    bp.initMdl()
    es = ra.ExpectationStep(p.cfg, 0)
    es.set_model([pr], [bp], p.pdf_class, p.sigma2_noise, pdf_direction=p.pdf_direction)
    with pytest.raises(ra.RelionError) as ei:
        es.run(bad, p.fctf, p.meta)
    assert ei.value.code == ra.RLN_ERR_SUM_WEIGHT_ZERO
    assert "particle 2" in str(ei.value)
    re, im, w = bp.getMdlData()
    assert np.isfinite(re).all() and np.isfinite(im).all() and np.isfinite(w).all()
    p.fimg = bad
    with pytest.raises(RuntimeError, match="status -6"):
        run_oracle(p)
    # the engine serves the next batch: clean backprojector, good images, same as the oracle
    bp.initMdl()
    rg, _ = es.run(good, p.fctf, p.meta)
    p.fimg = good
    ro, _, (bpd, _), _ = run_oracle(p)
    assert np.array_equal(rg["best_ihidden_over"], ro["best_ihidden_over"])
    re, im, w = bp.getMdlData()
    err = np.sqrt(((re - bpd[0, ..., 0]) ** 2 + (im - bpd[0, ..., 1]) ** 2).sum() / (bpd[0] ** 2).sum())
    assert err < 1e-4, f"BP after recovery: rel err {err:.2e}"
