"""Split-BF16 MFMA coarse pass (RLN_PREC_SPLIT_BF16) against the FP64 oracle.

Tolerances for this path (it trades the coarse cross term's precision for MFMA throughput):
  * coarse diff2 within 1e-5 relative of the oracle, as the plain-FP32 path (measured 1.5e-6 to 3.7e-6 on these
    cases; the box-256 exhaustive case of tests/test_gpu_fullsize.py holds the same 1e-5); the maximum is printed;
  * the fine pass, weights and backprojection run in the same FP32/FP64 kernels as the plain path, so the
    max-likelihood poses must agree with the oracle for >= 99% of particles (all but one in these small runs)
    and the backprojected data within 1e-4 norm-wise.
"""
import numpy as np
import pytest

import relion_amd as ra
from problems import make_problem
from test_gpu_parity import rel, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def test_mfma_lane_layout_exact():
    rng = np.random.default_rng(0)
    for _ in range(4):
        A = rng.integers(-8, 9, size=(16, 32)).astype(np.float32)
        B = rng.integers(-8, 9, size=(32, 16)).astype(np.float32)
        D = ra.selftest_mfma_bf16(A, B)
        assert np.array_equal(D, A @ B)


def _split(p):
    p.cfg.precision = ra.RLN_PREC_SPLIT_BF16
    return p


@pytest.mark.parametrize("kind", ["global64", "local2class", "scale_offsets"])
def test_split_bf16_estep(kind):
    if kind == "global64":
        p = make_problem(box=64, px=2.0, order=2, offset_range_px=5, offset_step_px=1, n=8, snr=0.1)
    elif kind == "local2class":
        p = make_problem(box=48, px=2.5, order=2, offset_range_px=3, offset_step_px=1, n=8, snr=0.2, nr_classes=2,
                         local=True, sigma_ang=10.0, prior_sigma=4.0, nr_blobs=30, seed=77)
    else:
        p = make_problem(box=48, px=3.0, order=1, offset_range_px=3, offset_step_px=1.5, n=6, snr=0.5,
                         current_size=40, scale_correction=True, perturb_meta=True, nr_blobs=25, seed=9,
                         separate_nomask=True)
    _split(p)
    es, rg, wg, bpg, _ = run_gpu(p)
    ro, wo, (bpd, bpw), dbg = run_oracle(p, (0, 1))
    worst = 0.0
    for ip in (0, 1):
        d2g, _, _, _ = es.coarse(ip)
        d2o = dbg[ip]["coarse_diff2"]
        ok = d2o >= 0
        assert np.array_equal(d2g >= 0, ok)
        worst = max(worst, rel(d2g[ok], d2o[ok]).max())
    print(f"split-BF16 coarse diff2 max rel err vs FP64 oracle ({kind}): {worst:.2e}")
    assert worst < 1e-5
    same = rg["best_ihidden_over"] == ro["best_ihidden_over"]
    assert len(same) - same.sum() <= max(1, int(0.01 * len(same)))
    assert rel(rg["min_diff2"], ro["min_diff2"]).max() < 1e-5  # fine pass is FP32 direct
    for k in range(p.cfg.nr_classes):
        re, im, w = bpg[k]
        err = np.sqrt(((re - bpd[k][..., 0]) ** 2 + (im - bpd[k][..., 1]) ** 2).sum() / (bpd[k] ** 2).sum())
        print(f"split-BF16 class {k}: BP data rel err {err:.2e}", flush=True)
        assert err <= 1e-4
