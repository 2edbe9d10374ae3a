"""Explicit per-particle coarse orientation lists (the entry point the learned pose proposal seeds).

Pinned by identities with the exhaustive global search on the same inputs:
  * a list holding the whole grid reproduces the exhaustive run bit for bit;
  * an empty list means the exhaustive grid for that particle;
  * the coarse diff2 of a listed orientation equals the exhaustive run's diff2 of that orientation exactly
    (the per-pose sums do not depend on which other orientations are evaluated);
  * a list that holds the exhaustive best orientation and its grid neighbours keeps the exhaustive best pose.
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
    bp.initMdl()
    es = ra.ExpectationStep(cfg, 0)
    es.set_model([pr], [bp], p.pdf_class, p.sigma2_noise, pdf_direction=p.pdf_direction)
    es.keep_intermediates(True)
    return es, (pr, bp)


@pytest.mark.parametrize("precision", [ra.RLN_PREC_PLAIN_FP32, ra.RLN_PREC_SPLIT_BF16], ids=["plain", "split"])
def test_orientation_sets_match_exhaustive(precision):
    p = make_problem(box=64, px=2.0, order=2, n=6, snr=0.1, seed=77)
    print(f"[sets] box 64 order 2 global, 6 particles, precision {precision}", flush=True)
    es, keep = engine(p, precision)
    n_or = es.n_dir * es.n_psi
    nt = es.n_trans
    res0, _ = es.run(p.fimg, p.fctf, p.meta)
    full = [es.coarse(i)[0].reshape(n_or, nt) for i in range(6)]

    # the whole grid as an explicit list, except particle 0 (empty list = exhaustive)
    sets = [np.zeros(0, np.int32)] + [np.arange(n_or, dtype=np.int32) for _ in range(5)]
    es.set_orientation_sets(sets)
    res1, _ = es.run(p.fimg, p.fctf, p.meta)
    for f in ("best_ihidden_over", "min_diff2", "sum_weight", "pmax", "rot", "tilt", "psi", "xoff", "yoff"):
        assert np.array_equal(res0[f], res1[f]), f
    # lists around the exhaustive best: its direction's grid neighbourhood (by angle) x all psi near its psi
    rng = np.random.default_rng(5)
    rot, tilt, psi = ra.sampling_grid(p.cfg)[:3]
    dirs = np.stack([np.sin(np.radians(tilt)) * np.cos(np.radians(rot)),
                     np.sin(np.radians(tilt)) * np.sin(np.radians(rot)), np.cos(np.radians(tilt))], 1)
    sets = []
    for i in range(6):
        ob = int(res0["best_ihidden_over"][i]) // 32 // nt  # coarse orientation of the best pose
        d0, p0 = divmod(ob, es.n_psi)
        near_d = np.flatnonzero(dirs @ dirs[d0] > np.cos(np.radians(35.0)))
        dpsi = np.abs((psi - psi[p0] + 180.0) % 360.0 - 180.0)
        near_p = np.flatnonzero(dpsi <= 35.0)
        lst = (near_d[:, None] * es.n_psi + near_p[None, :]).ravel()
        extra = rng.choice(n_or, 40, replace=False)
        sets.append(np.unique(np.concatenate([lst, extra])).astype(np.int32))
    es.set_orientation_sets(sets)
    res2, _ = es.run(p.fimg, p.fctf, p.meta)
    for i in range(6):
        d2, _, _, (nd, npsi, nt2) = es.coarse(i)
        assert (nd, npsi, nt2) == (len(sets[i]), 1, nt)
        d2 = d2.reshape(len(sets[i]), nt)
        assert np.array_equal(d2, full[i][sets[i]]), f"particle {i}: listed diff2 differ from the exhaustive run"
    # with a list, the orientation part of best_ihidden_over indexes the list
    same = np.mean([sets[i][int(res2["best_ihidden_over"][i]) // (32 * nt)] ==
                    int(res0["best_ihidden_over"][i]) // (32 * nt) for i in range(6)])
    print(f"[sets] restricted lists keep the exhaustive best coarse orientation for {same:.0%} of particles")
    assert same == 1.0
    assert np.allclose(res2["rot"], res0["rot"]) and np.allclose(res2["psi"], res0["psi"])
    es.close()
