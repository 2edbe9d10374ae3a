"""The E-step at the headline geometry (box 256, 1.0 A pixels, current size 256, HEALPix order 3, offsets 4 px
at 0.5 px steps: 197 translations, particle diameter 200 A: coarse size 48) against the FP64 oracle, with the
tolerances of the parity suite (tests/test_gpu_parity.py): per-pose coarse and fine diff2 within 1e-5, identical
best poses except near-ties, dLL within 1e-5, backprojected volumes (68.4 M voxels) within 1e-4 norm-wise.
At this size the pose-independent part of a fine diff2 (the noise power sum(|q X|^2), about 1.6e4) is far larger
than the differences between poses that the weights exponentiate; the fine kernels therefore sum only the
pose-dependent part in FP32 and add the rest in FP64, so that the weights, and the backprojection they scale,
keep the oracle's relative values.
A local search in plain FP32 and an exhaustive global search in split BF16 (7.26 M coarse poses per particle), on
a few particles each so that the oracle finishes in seconds."""
import pytest

from problems import make_problem
from test_gpu_parity import check_run

pytestmark = pytest.mark.gpu


def _headline(n, **kw):
    return make_problem(box=256, px=1.0, order=3, offset_range_px=4, offset_step_px=0.5, diameter=200.0, n=n,
                        snr=0.03, nr_blobs=60, seed=256, **kw)


def test_full_size_local_search_plain():
    print("[fullsize] box 256 local search (sigma 7.5 deg), plain FP32, 6 particles, fine pass in 16 bands", flush=True)
    p = _headline(6, local=True, sigma_ang=7.5)
    check_run(p, debug=(0,), bp_tol=1e-4, bands=16)


def test_full_size_global_search_split_bf16():
    import relion_amd as ra
    print("[fullsize] box 256 exhaustive global search, split BF16, 3 particles", flush=True)
    p = _headline(3)
    p.cfg.precision = ra.RLN_PREC_SPLIT_BF16
    check_run(p, debug=(0,), bp_tol=1e-4)
