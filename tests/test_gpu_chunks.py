"""A batch split into many chunks gives the results of the same batch in one chunk.

The E-step processes particles in chunks (at most `set_pool` particles each). From the third chunk on, a chunk's
lists are built and packed into page-locked memory on a second host thread while the device runs the chunk before,
alternating between two staging buffers; these tests run 13 particles as one chunk and as five chunks (1, 3, 3, 3,
3: the first chunk of a run holds at most a quarter of the pool) and require the same per-particle results. Per-particle values do not depend on chunking except through the
order of floating-point atomic sums (FP64 sums of weights over segments), hence the tolerances; the best poses and
the minimum diff2 are equal. Sums over particles (weighted sums, backprojection) are compared at the tolerance of
their FP32 / FP64 atomic accumulation order, which varies from run to run as well. The split case mixes exhaustive
fallbacks (empty lists) with explicit lists, so that the coarse tiles of both kinds are placed in every chunk.
"""
import numpy as np
import pytest

import relion_amd as ra
from problems import make_problem

pytestmark = pytest.mark.gpu


def engine(p, precision):
    cfg = p.cfg
    cfg.precision = precision
    dims = ra.volume_dims(cfg.current_size, cfg.padding_factor)
    pr = ra.AccProjector(0)
    pr.setMdlDim(*dims, cfg.padding_factor)
    pr.initMdl(p.volumes[0])
    bp = ra.AccBackprojector(0)
    bp.setMdlDim(*dims, cfg.padding_factor)
    es = ra.ExpectationStep(cfg, 0)
    return es, pr, bp


def run(es, pr, bp, p, sets):
    bp.initMdl()
    es.set_model([pr], [bp], p.pdf_class, p.sigma2_noise, pdf_direction=p.pdf_direction)
    if sets is not None:
        es.set_orientation_sets(sets)
    res, wsum = es.run(p.fimg, p.fctf, p.meta)
    return res, wsum, bp.getMdlData(), es.counters()["chunks"]


@pytest.mark.parametrize("precision", [ra.RLN_PREC_PLAIN_FP32, ra.RLN_PREC_SPLIT_BF16], ids=["plain", "split"])
def test_chunks_match_one_batch(precision):
    n = 13
    p = make_problem(box=64, px=2.0, order=2, n=n, snr=0.1, seed=31)
    print(f"[chunks] box 64 order 2 global, {n} particles, precision {precision}", flush=True)
    es, pr, bp = engine(p, precision)
    sets = None
    if precision == ra.RLN_PREC_SPLIT_BF16:
        n_or = es.n_dir * es.n_psi
        rng = np.random.default_rng(3)
        sets = [np.zeros(0, np.int32) if i % 4 == 1 else np.sort(rng.choice(n_or, 150, replace=False)).astype(np.int32)
                for i in range(n)]
    es.set_pool(4096)
    r1, w1, v1, c1 = run(es, pr, bp, p, sets)
    es.set_pool(3)
    r5, w5, v5, c5 = run(es, pr, bp, p, sets)
    print(f"[chunks] {c1} chunk(s) vs {c5} chunks", flush=True)
    assert c1 == 1 and c5 == 5
    for f in ("best_ihidden_over", "iclass", "rot", "tilt", "psi", "xoff", "yoff", "min_diff2", "coarse_min_diff2",
              "nr_fine_evaluated"):
        assert np.array_equal(r1[f], r5[f]), f
    for f in ("sum_weight", "coarse_sum_weight", "pmax", "dLL", "significant_weight", "coarse_significant_weight",
              "signal_product", "reference_power", "wsum_norm_correction", "nr_significant"):
        assert np.allclose(r1[f], r5[f], rtol=1e-9, atol=0), f
    # sigma2_noise: per work item, shell sums are FP32 atomics in LDS (their order varies from run to run)
    for f, rtol in (("sigma2_noise", 1e-5), ("pdf_class", 1e-9), ("pdf_direction", 1e-9)):
        a, b = np.asarray(w1[f]), np.asarray(w5[f])
        rel = np.max(np.abs(a - b) / np.maximum(np.abs(a), 1e-300))
        print(f"[chunks] {f}: max relative difference {rel:.3g}", flush=True)
        assert np.allclose(a, b, rtol=rtol, atol=1e-300), f
    for f in ("sigma2_offset", "sumw_group", "signal_product", "reference_power", "LL", "ave_Pmax"):
        assert np.isclose(w1[f], w5[f], rtol=1e-9, atol=0), f
    # backprojection: FP32 atomics in another order
    for a, b, name in zip(v1, v5, ("data re", "data im", "weight")):
        err = np.linalg.norm(a - b) / np.linalg.norm(a)
        assert err <= 1e-5, f"{name}: relative norm difference {err:.3g}"
    with pytest.raises(ra.RelionError):
        es.set_pool(0)
    es.close()


def test_fine_orders_match_particle_order():
    """The fine pass's unit orders (jobs in particle order or sorted by class and coarse orientation; units band-major
    or item-major) and the weighted sums' unit order give the same results up to the order of atomic sums. Two
    classes at order 2 (9,216 keys, so the job keys are shifted to fit the histogram) and 40 particles, so that many
    jobs share a coarse direction; four bands, so that the unit orders differ; the weighted sums also over 16 bands."""
    n = 40
    p = make_problem(box=64, px=2.0, order=2, n=n, snr=0.1, nr_classes=2, seed=47)
    print(f"[orders] box 64 order 2 global, 2 classes, {n} particles", flush=True)
    cfg = p.cfg
    dims = ra.volume_dims(cfg.current_size, cfg.padding_factor)
    prs, bps = [], []
    for k in range(2):
        pr = ra.AccProjector(0)
        pr.setMdlDim(*dims, cfg.padding_factor)
        pr.initMdl(p.volumes[k])
        bp = ra.AccBackprojector(0)
        bp.setMdlDim(*dims, cfg.padding_factor)
        prs.append(pr)
        bps.append(bp)
    es = ra.ExpectationStep(cfg, 0)
    es.set_fine_bands(4)
    out = {}
    for order, wim, wb in ((0, 0, 0), (1, 0, 0), (2, 0, 0), (3, 0, 0), (2, 1, 0), (1, 1, 0), (3, 2, 0), (1, 0, 16)):
        es.set_fine_order(order, wim)
        es.set_wsum_bands(wb)
        for bp in bps:
            bp.initMdl()
        es.set_model(prs, bps, p.pdf_class, p.sigma2_noise, pdf_direction=p.pdf_direction)
        res, wsum = es.run(p.fimg, p.fctf, p.meta)
        c = es.counters()
        assert (c["fine_order"], c["wsum_order"], c["fine_bands"], c["wsum_bands"]) == (order, wim, 4, wb or 4)
        out[(order, wim, wb)] = (res, wsum, [bp.getMdlData() for bp in bps])
    r0, w0, v0 = out[(0, 0, 0)]
    for key, (r1, w1, v1) in out.items():
        print(f"[orders] order {key[0]}, weighted-sum item-major {key[1]}, weighted-sum bands {key[2]}", flush=True)
        for f in ("best_ihidden_over", "iclass", "nr_fine_evaluated", "coarse_min_diff2"):
            assert np.array_equal(r0[f], r1[f]), f
        for f in ("min_diff2", "sum_weight", "pmax", "dLL", "significant_weight", "signal_product",
                  "reference_power", "wsum_norm_correction"):
            assert np.allclose(r0[f], r1[f], rtol=1e-9, atol=0), f
        for f, rtol in (("sigma2_noise", 1e-5), ("pdf_class", 1e-9), ("pdf_direction", 1e-9)):
            assert np.allclose(np.asarray(w0[f]), np.asarray(w1[f]), rtol=rtol, atol=1e-300), f
        for k in range(2):
            for a, b, name in zip(v0[k], v1[k], ("data re", "data im", "weight")):
                err = np.linalg.norm(a - b) / np.linalg.norm(a)
                assert err <= 1e-5, f"class {k} {name}: relative norm difference {err:.3g}"
    with pytest.raises(ra.RelionError):
        es.set_fine_order(4, 0)
    with pytest.raises(ra.RelionError):
        es.set_fine_order(1, 3)
    with pytest.raises(ra.RelionError):
        es.set_wsum_bands(3)
    es.close()


def test_large_chunks_decode_on_host_threads():
    """Chunks of at least 1,024 particles decode their per-particle results on several host threads and sum the
    sigma2_noise shells on the device: 1,100 particles at the default pool (a first chunk of 1,024, then 76) give
    the results of a pool of 1,000 (chunks of 250 and 850, each decoded on one thread), up to the order of
    floating-point sums."""
    n = 1100
    p = make_problem(box=32, px=4.0, order=1, n=n, snr=0.2, seed=61, offset_range_px=3, offset_step_px=1)
    print(f"[chunks] box 32 order 1 global, {n} particles, default pool vs pool 1,000", flush=True)
    es, pr, bp = engine(p, ra.RLN_PREC_PLAIN_FP32)
    r1, w1, v1, c1 = run(es, pr, bp, p, None)
    es.set_pool(1000)
    r2, w2, v2, c2 = run(es, pr, bp, p, None)
    print(f"[chunks] {c1} chunks vs {c2} chunks", flush=True)
    assert c1 == 2 and c2 == 2
    for f in ("best_ihidden_over", "iclass", "nr_fine_evaluated", "rot", "tilt", "psi", "xoff", "yoff"):
        assert np.array_equal(r1[f], r2[f]), f
    for f in ("min_diff2", "sum_weight", "pmax", "dLL", "coarse_min_diff2"):
        assert np.allclose(r1[f], r2[f], rtol=1e-9, atol=0), f
    for f, rtol in (("sigma2_noise", 1e-5), ("pdf_class", 1e-9), ("LL", 1e-9)):
        assert np.allclose(np.asarray(w1[f]), np.asarray(w2[f]), rtol=rtol, atol=1e-300), f
    for a, b, name in zip(v1, v2, ("data re", "data im", "weight")):
        err = np.linalg.norm(a - b) / np.linalg.norm(a)
        assert err <= 1e-5, f"{name}: relative norm difference {err:.3g}"
    es.close()
