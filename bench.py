#!/usr/bin/env python3
"""E-step throughput on one MI355X: particles/s of the expectation step at box 256, HEALPix order 3.

Data (stated choices marked *): one 256^3 map of 150 Gaussian blobs at 1.0 A/pixel, 10,000 particles (256x256,
SNR 0.03, defocus U(0.8, 2.5) um, 300 kV, Cs 2.7 mm, Q0 0.1, offsets U(+-3 px)), offsets 4 A range, 0.5 A
step, oversampling 1, padding 2, adaptive_fraction 0.999, *particle diameter 200 A (coarse image 48 px),
*current_size = box. Synthetic and seeded.

Headline (the learned proposal on this data): HEALPix order 3 global search.
  * plain: the exhaustive coarse search in plain FP32 (36,864 orientations x 197 translations per particle),
    timed on the first --plain-particles particles (particles are independent, so particles/s carries over);
  * proposal: the pose-proposal network (64x64 input, 6 conv x 64 channels, 192 x 24 bins) trained on
    --train-samples projections of the reference with CTF and the model's noise, its top-16 bins seed the coarse
    search (exhaustive fallback below top-1 probability 0.2), coarse cross term in split-BF16 on MFMA. Timed
    on all particles; the step includes the network's inference.
  Reported beside it: the miss rate (particles whose max-likelihood pose differs from the exhaustive search's)
  and the FSC between the two paths' backprojections, both over the plain subset; training time (untimed).
Secondary (precision variant): HEALPix order 3 local search (*sigma_ang 7.5 deg, *priors = true pose + 3 deg
noise), plain FP32 vs split-BF16 coarse pass on all particles.

Also reported (extra block): image preparation (real-space particles -> masked/unmasked transforms, power
spectra and CTF images) on --prep-particles device-resident particles.

A step is one E-step over the particles: projections of the coarse grid, coarse pass, weights and
significance, fine pass, weighted sums and backprojection; particle images are resident in HBM before the
timed region. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 peak
PEAK_BF16_TFLOPS = 2500.0  # MI355X published dense BF16 MFMA peak
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=int, default=10000)
    ap.add_argument("--plain-particles", type=int, default=2000,
                    help="particles of the exhaustive plain-FP32 baseline (and of the miss-rate / FSC check)")
    ap.add_argument("--box", type=int, default=256)
    ap.add_argument("--train-samples", type=int, default=200000)
    ap.add_argument("--train-epochs", type=int, default=5)
    ap.add_argument("--finetune-samples", type=int, default=25000,
                    help="fresh projections of the warm-started training of the next iteration, which also sees "
                         "the particles labelled with this iteration's max-likelihood poses (0: skip); 25,000, "
                         "50,000 and 100,000 kept the exhaustive best pose alike on the 2,000 particles checked")
    ap.add_argument("--topk", type=int, default=16)
    ap.add_argument("--min-top1", type=float, default=0.2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="CPU oracle budget: batches of one particle per thread until this much time (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--format-particles", type=int, default=4096,
                    help="particles written to and read from MRC stacks and a STAR file in the formats leg")
    ap.add_argument("--prep-particles", type=int, default=8192)
    ap.add_argument("--lists-out", default=None,
                    help="save the proposal's orientation lists of the last timed step (npz), for profiling runs")
    ap.add_argument("--lists-in", default=None,
                    help="replay saved orientation lists instead of training and running the proposal network "
                         "(the E-step workload of the run that saved them; for counter-collection runs)")
    ap.add_argument("--fine-order", type=int, default=-1,
                    help="order of the fine pass's work units (0 particle order, 1 orientation order band-major, "
                         "2 orientation order item-major, 3 orientation order band-major in per-XCD queues; -1 tuned)")
    ap.add_argument("--fine-shape", type=int, default=-1,
                    help="work-item shape of the fine pass (0-3: (TC, CW) = (2, 2), (1, 2), (2, 1), (1, 1); -1 tuned)")
    ap.add_argument("--wsum-order", type=int, default=0,
                    help="weighted sums' unit order: 0 band-major, 1 item-major, 2 band-major in per-XCD queues")
    ap.add_argument("--pool", type=int, default=0,
                    help="particles per E-step chunk (the reference's --pool; 0: the engine's default)")
    ap.add_argument("--wsum-bands", type=int, default=0,
                    help="radial pixel bands per weighted-sum work unit (0: the fine pass's tuned count)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the fine stage directly instead of through captured graphs (same kernels)")
    ap.add_argument("--classify-particles", type=int, default=1000,
                    help="particles of the 4-class box-128 classification leg (one step per precision)")
    ap.add_argument("--uneven-particles", type=int, default=4096,
                    help="particles of the box-192 uneven-SNR leg (one step)")
    ap.add_argument("--paths", default="proposal,plain,local,classify,uneven,prep,refprep,recon,formats",
                    help="comma list: proposal (headline), plain (exhaustive baseline), local (precision variant), "
                         "prep (image preparation before the E-step), refprep (reference preparation), recon "
                         "(reconstruction from the filled backprojector), formats (particles from STAR and MRC files)")
    ap.add_argument("--workload", default="refine256", choices=["refine256", "uneven192"],
                    help="refine256: the headline (box 256); uneven192: box 192, 20,000 particles of log-uniform "
                         "SNR in [0.01, 0.3], global order-3 coarse search, order-4 fine pass (batched fine pass)")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run one device-to-device copy of this size (reference for traffic counters)")
    ap.add_argument("--traffic-kib", type=float, default=None,
                    help="measured HBM traffic per launch of the dominant kernel (KiB, counter-calibrated)")
    return ap.parse_args()


def log(msg, args):
    if not args.quiet:
        print(f"[bench] {msg}", file=sys.stderr, flush=True)


def engine(cfg, args):
    """An E-step engine on device 0 with the run's launch options (graphs, unit orders)."""
    import relion_amd as ra
    es = ra.ExpectationStep(cfg, 0)
    if args.no_graphs:
        es.use_graphs(False)
    es.set_fine_order(args.fine_order, args.wsum_order)
    es.set_fine_shape(args.fine_shape)
    es.set_wsum_bands(args.wsum_bands)
    if args.pool > 0:
        es.set_pool(args.pool)
    return es


# work-item shapes of the fine pass: (coarse translations per small item, per wave of a wide item)
FINE_SHAPES = ((2, 2), (1, 2), (2, 1), (1, 1))


def tuning_summary(es):
    """The fine pass's tuning on the first chunk: the fastest time (ms) per unit order and per work-item shape, and
    the choice kept."""
    t = es.fine_tuning()
    if t is None:
        return None
    names = ("particle order, band-major", "orientation order, band-major", "orientation order, item-major",
             "orientation order, band-major, per-XCD queues")
    best = {names[o]: (round(float(t[:, o][t[:, o] >= 0].min()), 3) if (t[:, o] >= 0).any() else None)
            for o in range(t.shape[1])}
    by_shape = {f"tc{tc}_cw{cw}": (round(float(t[s][t[s] >= 0].min()), 3) if (t[s] >= 0).any() else None)
                for s, (tc, cw) in enumerate(FINE_SHAPES)}
    c = es.counters()
    return {"fastest_ms_by_order": best, "fastest_ms_by_shape": by_shape, "kept": {"order": names[c["fine_order"]], "bands": c["fine_bands"],
                                                   "grid": c["fine_grid"], "small_tc": c["fine_small_tc"],
                                                   "wide_cw": c["fine_wide_cw"]}}


def rotm(a, b, c):
    a, b, c = np.radians([a, b, c])
    ca, sa, cb, sb, cg, sg = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    return np.array([[cg * cb * ca - sg * sa, cg * cb * sa + sg * ca, -cg * sb],
                     [-sg * cb * ca - cg * sa, -sg * cb * sa + cg * ca, sg * sb],
                     [sb * ca, sb * sa, cb]])


def angle_between(r0, r1):
    A, B = rotm(*r0), rotm(*r1)
    return float(np.degrees(np.arccos(np.clip((np.trace(A @ B.T) - 1) / 2, -1, 1))))


def fsc_maps(m0, m1, r_max):
    """Fourier shell correlation of two real-space maps per shell ROUND(|k|) (original-box Fourier pixels),
    shells 0 .. r_max."""
    F0, F1 = np.fft.rfftn(m0.astype(np.float64)), np.fft.rfftn(m1.astype(np.float64))
    n = m0.shape[0]
    k = np.fft.fftfreq(n) * n
    kx = np.arange(n // 2 + 1)
    r = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + kx[None, None, :] ** 2)
    shell = np.rint(r).astype(np.int64).ravel()
    keep = shell <= r_max
    sh = shell[keep]
    a, b = F0.ravel()[keep], F1.ravel()[keep]
    num = np.bincount(sh, np.real(a * np.conj(b)), r_max + 1)
    d0 = np.bincount(sh, np.abs(a) ** 2, r_max + 1)
    d1 = np.bincount(sh, np.abs(b) ** 2, r_max + 1)
    return num / np.sqrt(np.maximum(d0 * d1, 1e-300))


def uneven(args):
    """--workload uneven192 on its own: the uneven leg with 20,000 particles (or --particles), one JSON line."""
    from relion_amd import hipmem
    hipmem.set_device(0)
    n = args.particles if args.particles != 10000 else 20000
    out = uneven_leg(args, n, args.steps, args.warmup)
    print(json.dumps(out), flush=True)
    return 0


def uneven_leg(args, n, steps, warmup):
    """Box 192, 1.2 A, 100 blobs, n particles with per-particle SNR log-uniform in [0.01, 0.3]: the
    significant-pose counts of the fine pass spread over orders of magnitude. Global HEALPix order-3 coarse
    search (split-BF16), order-4 fine pass (oversampling 1), offsets 6 px range 2 px step; the fine pass runs
    from device-side plans as persistent kernels in captured graphs with autotuned work units and grids."""
    import relion_amd as ra
    from relion_amd import hipmem
    box, px, order = 192, 1.2, 3
    diameter = 0.8 * box * px
    coarse = ra.coarse_image_size(box, px, order, diameter)
    cfg = ra.make_config(box, px, order, offset_range=6 * px, offset_step=2 * px, coarse_size=coarse,
                         precision=ra.RLN_PREC_SPLIT_BF16)
    n_dir, _, _ = ra.sampling_counts(cfg)
    spec = ra.synth_spec(args.seed, box, px, 100, snr=(0.01, 0.3), defocus=(8000.0, 25000.0), offset_max=3.0)
    t0 = time.time()
    vol = ra.synth_reference(spec, 0, box)
    fimg, fctf, meta, pose, nvar = ra.synth_particles(spec, n, box)
    log(f"generated reference + {n} particles in {time.time() - t0:.1f} s (coarse {coarse}px)", args)
    sigma2 = np.full(box // 2 + 1, float(np.median(nvar)))
    d_img, d_ctf = hipmem.DeviceArray(fimg), hipmem.DeviceArray(fctf)
    del fimg, fctf
    dims = ra.volume_dims(box)
    proj = ra.AccProjector(0)
    proj.setMdlDim(*dims, 2.0)
    proj.initMdl(vol)
    bp = ra.AccBackprojector(0)
    bp.setMdlDim(*dims, 2.0)
    es = engine(cfg, args)
    pdf_dir = np.full(n_dir, 1.0 / n_dir)

    def step():
        bp.initMdl()
        es.set_model([proj], [bp], [1.0], sigma2, pdf_direction=pdf_dir)
        return es.run(d_img.address, d_ctf.address, meta, on_device=True)[0]

    for _ in range(warmup):
        step()
    hipmem.synchronize()
    t = time.time()
    for _ in range(steps):
        res = step()
    hipmem.synchronize()
    el = time.time() - t
    tm, cn = es.timings(), es.counters()
    fe = res["nr_fine_evaluated"] // 32
    out = {"metric": "E-step particles/s, box 192 uneven-SNR batched fine pass", "value": round(n * steps / el, 2),
           "unit": "particles/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": round(1000 * el / steps, 2), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "bf16x3+f32", "data": "synthetic",
           "config": {"workload": f"box{box} order3 global, order4 fine, {n} particles, SNR log-U[0.01,0.3]",
                      "coarse_size": coarse, "global_batch": n, "parallelism": "1 GPU"},
           "significant_coarse_poses_per_particle": {"min": int(fe.min()), "median": int(np.median(fe)),
                                                     "p99": int(np.percentile(fe, 99)), "max": int(fe.max())},
           "stages_ms": {k: round(v, 2) for k, v in tm.items()}, "counters": cn}
    log(f"[uneven192] {n} particles: {out['value']:.0f} particles/s, stages {json.dumps(out['stages_ms'])}", args)
    es.close()
    return out


def classify_leg(args, n):
    """Classification: K = 4 classes of 128^3 blob maps (60 blobs each, 30% shared), 1.5 A pixels, n particles at
    SNR 0.05 with uniform class labels, HEALPix order 3 global search over all classes (4 x 36,864 orientations),
    offsets 6 px at 1 px steps (113 translations), particle diameter 180 A (coarse 40 px): 16.7 M coarse poses per
    particle. One timed step each of the plain-FP32 and the split-BF16 coarse pass after one warm-up."""
    import relion_amd as ra
    from relion_amd import hipmem
    box, px, order, K = 128, 1.5, 3, 4
    coarse = ra.coarse_image_size(box, px, order, 180.0)
    spec = ra.synth_spec(args.seed + 1, box, px, 60, snr=0.05, offset_max=3.0)
    t0 = time.time()
    vols = [ra.synth_reference(spec, k, box, share_fraction=0.3) for k in range(K)]
    fimg, fctf, meta, pose, nvar = ra.synth_particles(spec, n, box, nr_classes=K, share_fraction=0.3)
    log(f"[classify] generated {K} references + {n} particles in {time.time() - t0:.1f} s (coarse {coarse}px)", args)
    d_img, d_ctf = hipmem.DeviceArray(fimg), hipmem.DeviceArray(fctf)
    del fimg, fctf
    sigma2 = np.full(box // 2 + 1, float(nvar.mean()))
    dims = ra.volume_dims(box)
    projs, bps = [], []
    for k in range(K):
        pr = ra.AccProjector(0)
        pr.setMdlDim(*dims, 2.0)
        pr.initMdl(vols[k])
        bp = ra.AccBackprojector(0)
        bp.setMdlDim(*dims, 2.0)
        projs.append(pr)
        bps.append(bp)
    out = {"box": box, "classes": K, "particles": n, "coarse_size": coarse,
           "coarse_poses_per_particle": K * 36864 * 113}
    for name, prec in (("plain", ra.RLN_PREC_PLAIN_FP32), ("split", ra.RLN_PREC_SPLIT_BF16)):
        cfg = ra.make_config(box, px, order, offset_range=6 * px, offset_step=px, coarse_size=coarse, nr_classes=K,
                             precision=prec)
        n_dir, _, _ = ra.sampling_counts(cfg)
        es = engine(cfg, args)
        pdf_dir = np.full((K, n_dir), 1.0 / (K * n_dir))

        def step():
            for b in bps:
                b.initMdl()
            es.set_model(projs, bps, np.full(K, 1.0 / K), sigma2, pdf_direction=pdf_dir)
            return es.run(d_img.address, d_ctf.address, meta, on_device=True)[0]

        step()
        hipmem.synchronize()
        t = time.time()
        res = step()
        hipmem.synchronize()
        el = time.time() - t
        tm = es.timings()
        acc = round(float(np.mean(res["iclass"] == pose[:, 5].astype(int))), 4)  # both 1-based
        out[name] = {"particles_per_s": round(n / el, 1), "ms_per_step": round(1000 * el, 1),
                     "stages_ms": {k: round(v, 1) for k, v in tm.items()}, "class_accuracy": acc}
        log(f"[classify] {name}: {n / el:.0f} particles/s, stages {json.dumps(out[name]['stages_ms'])}", args)
        es.close()
    out["split_vs_plain"] = round(out["split"]["particles_per_s"] / out["plain"]["particles_per_s"], 2)
    d_img.free()
    d_ctf.free()
    return out


def committed_traffic(name, kernel):
    """HBM bytes per launch of `kernel` (its name without spaces, e.g. k_fine_diff2<0,1,1>) from a committed counter
    summary (two rocprofv3 --pmc runs of the same workload, FETCH_SIZE and WRITE_SIZE scaled per kernel by the
    request-size calibrations), or (None, None)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                return int(float(parts[2])), f"profiles/{name}"
    return None, None


def refprep_leg(args, spec, box):
    """Reference preparation (Projector::computeFourierTransformMap) of the box^3 blob map at padding 2 with
    gridding correction, timed over --steps calls after one warm-up; the FP64 oracle on the host beside it when
    --cpu-seconds > 0. HBM bytes counted per call: the map read (4 box^3), the padded map written and read by the
    FFT (8 P^3), the transform written and read (16 P^2 (P/2+1)), the plain window written and read (16 S^2 (S/2+1))
    and the corner-packed volume written (64 per base voxel); the FFT's internal passes are not counted."""
    import relion_amd as ra
    from relion_amd import hipmem

    vol = ra.synth_real_map(spec, 0, box)
    pr = ra.AccProjector(0)
    pr.computeFourierTransformMap(vol, box, 2.0)
    hipmem.synchronize()
    d_vol = hipmem.DeviceArray(vol)
    steps = max(1, args.steps)
    t0 = time.time()
    for _ in range(steps):
        pr.computeFourierTransformMap(d_vol.address, box, 2.0, ori_size=box, on_device=True)
    hipmem.synchronize()
    ms = 1000.0 * (time.time() - t0) / steps
    P, S = 2 * box, 2 * (box + 1) + 1
    nbytes = 4 * box ** 3 + 8 * P ** 3 + 16 * P * P * (P // 2 + 1) + 16 * S * S * (S // 2 + 1) + 64 * (S // 2) * (S - 1) ** 2
    out = {"box": box, "padded_box": P, "ms_per_call": round(ms, 2), "calls": steps,
           "roofline": {"bound": "hbm", "bytes_per_call": nbytes, "achieved": round(nbytes / (ms * 1e-3) / 1e9, 1),
                        "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}}
    log(f"[refprep] box {box}: {ms:.1f} ms per call", args)
    if args.cpu_seconds > 0:
        from oracle import refprep as orc
        t = time.time()
        orc.compute_fourier_map(vol, box, 2.0)
        out["cpu_oracle_s"] = round(time.time() - t, 2)
        out["cpu_oracle_kind"] = "port (numpy FP64, one process)"
        log(f"[refprep] oracle {out['cpu_oracle_s']} s", args)
    pr.clear()
    return out


def recon_leg(args, bp, box, tau2):
    """Reconstruction (BackProjector::reconstruct) of the box^3 map from the backprojector the last E-step filled,
    output left on the device: relion_refine's default path without gridding iterations, and the path with 10
    gridding iterations, each timed over --steps calls after one warm-up; the FP64 oracle of the default path on the
    host beside it when --cpu-seconds > 0. HBM bytes counted per call of the default path: the slots folded into the
    plain arrays (24 per plain voxel), the decentred weight, gridding weight and data written and read by the
    radial average, the division and the window (112 per half-complex S^3 element), the padded transform written and
    read by the inverse FFT (32 P^2 (P/2+1)), its real output written and read (8 P^3 + 8 box^3) and the map written
    (4 box^3); the FFT's internal passes are not counted."""
    import relion_amd as ra  # noqa: F401
    from relion_amd import hipmem

    d_out = hipmem.DeviceArray(np.zeros((box,) * 3, np.float32))
    steps = max(1, args.steps)
    out = {"box": box}
    for name, skip, iters in (("skip_gridding", True, 10), ("gridding_10_iterations", False, 10)):
        kw = dict(max_iter_preweight=iters, do_map=True, tau2=tau2, tau2_fudge=1.0, normalise=1.0, minres_map=3,
                  skip_gridding=skip, out_address=d_out.address)
        bp.reconstruct(box, **kw)
        hipmem.synchronize()
        t0 = time.time()
        for _ in range(steps):
            bp.reconstruct(box, **kw)
        hipmem.synchronize()
        ms = 1000.0 * (time.time() - t0) / steps
        out[name] = {"ms_per_call": round(ms, 2), "calls": steps}
        log(f"[recon] box {box} {name}: {ms:.1f} ms per call", args)
    P, S = 2 * box, 2 * (box + 1) + 1
    nv, nS = S * S * (S // 2 + 1), S * S * (S // 2 + 1)
    nbytes = 24 * nv + 112 * nS + 32 * P * P * (P // 2 + 1) + 8 * P ** 3 + 12 * box ** 3
    ms = out["skip_gridding"]["ms_per_call"]
    out["skip_gridding"]["roofline"] = {"bound": "hbm", "bytes_per_call": nbytes,
                                        "achieved": round(nbytes / (ms * 1e-3) / 1e9, 1), "peak": PEAK_HBM_GBS,
                                        "unit": "GB/s", "frac": round(nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
    if args.cpu_seconds > 0:
        from oracle import reconstruct as orc
        re, im, w = bp.getMdlData()
        t = time.time()
        orc.reconstruct(re.astype(np.float64) + 1j * im, w.astype(np.float64), box, 2.0, tau2=tau2, minres_map=3,
                        do_map=True, skip_gridding=True)
        out["cpu_oracle_skip_gridding_s"] = round(time.time() - t, 2)
        out["cpu_oracle_kind"] = "port (numpy FP64, one process)"
        log(f"[recon] oracle {out['cpu_oracle_skip_gridding_s']} s", args)
    d_out.free()
    return out


def formats_leg(args, box, px):
    """Particle input from disk: --format-particles images of box^2 float32 in MRC stacks of 1,024 with a particles
    STAR file (optics + particles tables) written to a temporary directory, then read back by read_particles (STAR
    parse, per-particle metadata, images by 16 threads, stacks in the page cache as just written). CPU leg beside
    it: one process reading every image with numpy (np.fromfile at the image's offset), the reference's
    per-particle Image::read pattern."""
    import shutil
    import tempfile

    from relion_amd import formats as F

    n = args.format_particles
    tmp = tempfile.mkdtemp(prefix="rln_formats_")
    try:
        rng = np.random.default_rng(args.seed)
        block = rng.normal(size=(64, box, box)).astype(np.float32)
        names, metas = [], []
        for k0 in range(0, n, 1024):
            k = min(1024, n - k0)
            stack = np.tile(block, (-(-k // 64), 1, 1))[:k]
            fn = f"stack{k0 // 1024:03d}.mrcs"
            F.write_mrc(os.path.join(tmp, fn), stack, pixel_size=px)
            names += [f"{i + 1:06d}@{fn}" for i in range(k)]
        for i in range(n):
            metas.append(dict(norm_correction=1.0, old_offset_x=float(i % 5 - 2), old_offset_y=float(i % 3 - 1),
                              defocus_u=15000.0 + i, defocus_v=14800.0 + i, defocus_angle=30.0, bfactor=0.0,
                              kfactor=1.0, phase_shift=0.0))
        star = os.path.join(tmp, "particles.star")
        F.write_particles(star, names, metas, [dict(group=1, pixel_size=px, image_size=box, voltage=300.0, cs=2.7,
                                                    q0=0.1)])
        t = time.time()
        F.read_particles(star, base_dir=tmp, load_images=False)
        t_star = time.time() - t
        t = time.time()
        ps = F.read_particles(star, base_dir=tmp, threads=16)
        t_all = time.time() - t
        gb = ps.images.nbytes / 1e9
        del ps
        t = time.time()
        m = 0
        while m < n and time.time() - t < max(1.0, args.cpu_seconds):
            idx, fn = names[m].split("@")
            np.fromfile(os.path.join(tmp, fn), np.float32, count=box * box, offset=1024 + (int(idx) - 1) * box * box * 4)
            m += 1
        t_np = time.time() - t
        out = {"particles": n, "box": box, "stacks": -(-n // 1024), "star_parse_ms": round(1000 * t_star, 1),
               "read_particles_ms": round(1000 * t_all, 1), "particles_per_s": round(n / t_all, 1),
               "image_GB_per_s": round(gb / t_all, 2),
               "cpu_per_image_numpy": {"particles_per_s": round(m / t_np, 1), "sample": f"first {m} images",
                                       "cores": 1}}
        log(f"[formats] {n} particles: STAR {1000 * t_star:.1f} ms, with images {1000 * t_all:.1f} ms "
            f"({gb / t_all:.2f} GB/s); numpy per image {m / t_np:.0f} particles/s", args)
        return out
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def prep_leg(args, box, px, diameter, d_img, nvar):
    """Image preparation on device-resident real-space particles: norm correction, re-centring, soft mask with a
    noise background, forward transforms, power spectra, CTF images. The real-space stack is the inverse
    transform of the first 512 synthetic particles, tiled to --prep-particles. HBM roofline of the three
    streaming kernels and the library FFT: 4 N^2 + 8 N^2 (real) + 8 N^2 + 16 N (N/2+1) (FFT) + 16 N (N/2+1) +
    20 s (s/2+1) (Fourier) bytes per particle, plus the noise images (8 N (N/2+1) + 4 N^2 + 4 N^2 read)."""
    import relion_amd as ra
    from relion_amd import hipmem

    N = s = box
    xs = s // 2 + 1
    k = 512
    sub = hipmem.DeviceArray(np.zeros((k, N, xs, 2), np.float32))
    hipmem.copy_device(sub.address, d_img.address, sub.nbytes)
    F = sub.to_host()
    sub.free()
    F = F[..., 0].astype(np.float64) + 1j * F[..., 1]
    sign = np.where(((np.arange(N)[:, None] + np.arange(xs)[None, :]) & 1) == 1, -1.0, 1.0)
    real = (np.fft.irfft2(F * sign, s=(N, N)) * (N * N)).astype(np.float32)
    n = args.prep_particles
    reps = -(-n // k)
    d_real = hipmem.DeviceArray(np.tile(real, (reps, 1, 1))[:n])
    d_f = hipmem.DeviceArray(np.zeros((n, s, xs, 2), np.float32))
    d_u = hipmem.DeviceArray(np.zeros((n, s, xs, 2), np.float32))
    d_c = hipmem.DeviceArray(np.zeros((n, s, xs), np.float32))
    rng = np.random.default_rng(args.seed)
    metas = ra.ImagePreparation.meta_array([dict(norm_correction=rng.uniform(0.8, 1.2), old_offset_x=rng.uniform(-3, 3),
                                                 old_offset_y=rng.uniform(-3, 3), defocus_u=rng.uniform(8e3, 2.5e4),
                                                 defocus_v=rng.uniform(8e3, 2.5e4), defocus_angle=rng.uniform(0, 180),
                                                 bfactor=0.0, kfactor=1.0, phase_shift=0.0, part_id=i)
                                            for i in range(n)])
    prep = ra.ImagePreparation(N, s, px, diameter, zero_mask=False, random_seed=args.seed, avg_norm_correction=1.0)
    prep.set_sigma2_noise(np.full(N // 2 + 1, float(np.mean(nvar))))
    prep.run_device(n, d_real.address, metas, d_f.address, d_u.address, d_c.address)  # warm-up (FFT plans)
    hipmem.synchronize()
    steps = max(1, args.steps)
    t0 = time.time()
    ms = {"real": 0.0, "fft": 0.0, "fourier": 0.0, "noise": 0.0, "total": 0.0}
    for _ in range(steps):
        prep.run_device(n, d_real.address, metas, d_f.address, d_u.address, d_c.address)
        for kk, v in prep.timings().items():
            ms[kk] += v / steps
    hipmem.synchronize()
    wall = (time.time() - t0) / steps
    hw = N * (N // 2 + 1)
    per = (4 * N * N + 8 * N * N) + (8 * N * N + 16 * hw) + (16 * hw + 20 * s * xs) + (8 * hw + 8 * N * N)
    kernel_ms = ms["real"] + ms["fft"] + ms["fourier"] + ms["noise"]
    gbs = per * n / (kernel_ms * 1e-3) / 1e9
    log(f"[prep] {n} particles: {json.dumps({kk: round(v, 2) for kk, v in ms.items()})} ms, wall {wall * 1e3:.1f} ms",
        args)
    prep.close()
    return {"particles": n, "box": N, "window": s, "background": "noise", "particles_per_s": round(n / (ms["total"] * 1e-3), 1),
            "stages_ms": {kk: round(v, 3) for kk, v in ms.items()}, "wall_ms": round(wall * 1e3, 2),
            "roofline": {"bound": "hbm", "bytes_per_particle": per, "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS,
                         "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 4)}}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU: run with --gpus 1 "
                                   "(sharding over several GPUs is designed, not run)"}))
        return 2
    if args.workload == "uneven192":
        return uneven(args)
    import relion_amd as ra
    from relion_amd import hipmem
    from relion_amd.proposal import PoseProposal

    if hipmem.device_count() < 1:
        print(json.dumps({"error": "no HIP device visible"}))
        return 2
    hipmem.set_device(0)
    paths = args.paths.split(",")
    box, px, order = args.box, 256.0 / args.box if args.box != 256 else 1.0, 3
    diameter = 200.0 * box / 256.0
    coarse = ra.coarse_image_size(box, px, order, diameter)
    cfg_g = ra.make_config(box, px, order, offset_range=4.0, offset_step=0.5, coarse_size=coarse)
    cfg_l = ra.make_config(box, px, order, offset_range=4.0, offset_step=0.5, coarse_size=coarse, local_search=True,
                           sigma_ang=7.5)
    n_dir, n_psi, n_tr = ra.sampling_counts(cfg_g)
    spec = ra.synth_spec(args.seed, box, px, 150, snr=0.03, defocus=(8000.0, 25000.0), offset_max=3.0,
                         prior_sigma_deg=3.0)
    t0 = time.time()
    vol = ra.synth_reference(spec, 0, box)
    n = args.particles
    n_plain = min(args.plain_particles, n)
    fimg, fctf, meta, pose, nvar = ra.synth_particles(spec, n, box)
    log(f"generated reference + {n} particles in {time.time() - t0:.1f} s (coarse {coarse}px)", args)
    sigma2 = np.full(box // 2 + 1, float(nvar.mean()))
    pdf_dir = np.full(n_dir, 1.0 / n_dir)
    d_img = hipmem.DeviceArray(fimg)
    d_ctf = hipmem.DeviceArray(fctf)
    del fimg, fctf
    meta_plain = (ra.ParticleMeta * n_plain)(*[meta[i] for i in range(n_plain)])
    dims = ra.volume_dims(box)
    proj = ra.AccProjector(0)
    proj.setMdlDim(*dims, 2.0)
    proj.initMdl(vol)
    bp = ra.AccBackprojector(0)
    bp.setMdlDim(*dims, 2.0)

    if args.copy_calibration_mb > 0:
        nb = args.copy_calibration_mb << 20
        src = hipmem.DeviceArray(np.zeros(nb, np.uint8))
        dst = hipmem.DeviceArray(np.zeros(nb, np.uint8))
        hipmem.copy_device(dst.address, src.address, nb)
        hipmem.synchronize()
        log(f"copy calibration: {nb} bytes device-to-device", args)
        del src, dst

    def timed(name, step, n_part, steps, warmup, es):
        for w in range(warmup):
            t = time.time()
            step()
            log(f"[{name}] warmup {w}: {time.time() - t:.2f} s", args)
        hipmem.synchronize()
        t_start = time.time()
        for _ in range(steps):
            out = step()
        hipmem.synchronize()
        elapsed = time.time() - t_start
        timings, counters = es.timings(), es.counters()
        log(f"[{name}] timed {steps} steps: {elapsed:.2f} s; stages "
            f"{json.dumps({k: round(v, 1) for k, v in timings.items()})}", args)
        log(f"[{name}] counters {json.dumps(counters)}", args)
        return dict(out=out, elapsed=elapsed, timings=timings, counters=counters, value=n_part * steps / elapsed,
                    ms=1000.0 * elapsed / steps, steps=steps, warmup=warmup, particles=n_part)

    runs, extra = {}, {}
    pp = None
    # ---------------------------------------------------------------- learned proposal (training is untimed)
    replay = None
    if "proposal" in paths and args.lists_in:
        z = np.load(args.lists_in)
        replay = (z["off"], z["og"])
        if len(replay[0]) != n + 1:
            raise SystemExit(f"{args.lists_in} holds lists for {len(replay[0]) - 1} particles, not {n}")
    if "proposal" in paths and replay is not None:
        cfg_g.precision = ra.RLN_PREC_SPLIT_BF16
        es_p = engine(cfg_g, args)

        def proposal_step(count=n, mt=meta):
            bp.initMdl()
            es_p.set_model([proj], [bp], [1.0], sigma2, pdf_direction=pdf_dir)
            off, og = replay
            es_p.set_orientation_lists(off[:count + 1], og[:off[count]])
            res, _ = es_p.run(d_img.address, d_ctf.address, mt, on_device=True)
            return res, off[:count + 1]

        runs["proposal"] = timed("proposal (replayed lists)", proposal_step, n, args.steps, args.warmup, es_p)
        extra["fine_tuning"] = tuning_summary(es_p)
        extra["proposal_lists_replayed_from"] = args.lists_in
    elif "proposal" in paths:
        pp = PoseProposal(window=64, channels=64, healpix_order=2, n_psi_bins=24, seed=args.seed)
        # each epoch sees fresh projections (new poses, CTFs, shifts and noise)
        t_gen = t_train = 0.0
        for ep in range(args.train_epochs):
            t = time.time()
            pp.clear_training()
            pp.generate(proj, args.train_samples, ori_size=box, pixel_size=px, sigma2_noise=sigma2,
                        defocus=(8000.0, 25000.0), max_astigmatism=300.0, offset_max=3.0,
                        first=ep * args.train_samples)
            hipmem.synchronize()
            t_gen += time.time() - t
            t = time.time()
            loss = pp.train(epochs=1, batch=256, lr=1e-3)
            t_train += time.time() - t
            log(f"proposal epoch {ep}: {args.train_samples} fresh projections, loss {loss:.3f}", args)
        log(f"proposal: generation {t_gen:.1f} s, training {t_train:.1f} s", args)
        extra["proposal_training"] = {"samples_per_epoch": args.train_samples, "epochs": args.train_epochs,
                                      "batch": 256, "generate_s": round(t_gen, 2), "train_s": round(t_train, 2),
                                      "last_epoch_loss": round(loss, 4)}
        cfg_g.precision = ra.RLN_PREC_SPLIT_BF16
        es_p = engine(cfg_g, args)
        pred_ms = []
        pred_split = []  # (network, orientation lists, hand-over to the engine) per step, ms

        def proposal_step(count=n, mt=meta):
            bp.initMdl()
            es_p.set_model([proj], [bp], [1.0], sigma2, pdf_direction=pdf_dir)
            t1 = time.time()
            bins, probs = pp.predict(d_img.address, d_ctf.address, count, box, topk=args.topk, on_device=True)
            t2 = time.time()
            off, og = pp.orientation_lists(bins, probs, order, n_psi, min_top1=args.min_top1)
            t3 = time.time()
            es_p.set_orientation_lists(off, og)
            t4 = time.time()
            pred_ms.append(1000.0 * (t4 - t1))
            pred_split.append((1000.0 * (t2 - t1), 1000.0 * (t3 - t2), 1000.0 * (t4 - t3)))
            res, _ = es_p.run(d_img.address, d_ctf.address, mt, on_device=True)
            return res, off

        runs["proposal"] = timed("proposal", proposal_step, n, args.steps, args.warmup, es_p)
        extra["proposal_predict_ms"] = round(float(np.mean(pred_ms[-args.steps:])), 2)
        sp = np.mean(np.array(pred_split[-args.steps:]), axis=0)
        extra["proposal_predict_split_ms"] = {"network": round(float(sp[0]), 2), "orientation_lists": round(float(sp[1]), 2),
                                              "set_orientation_lists": round(float(sp[2]), 2)}
        extra["fine_tuning"] = tuning_summary(es_p)

        def pipelined_steps(steps):
            """`steps` proposal steps over the batch, the inference and orientation lists of step k + 1 running
            (GPU: the proposal's stream; host: this thread) while step k's E-step runs (its own stream, called from
            a second thread). Every step does the same work as proposal_step; only the order overlaps."""
            import threading

            def lists_of():
                pp.predict_launch(d_img.address, d_ctf.address, n, box, topk=args.topk, on_device=True)
                bins, probs = pp.predict_finish()
                return pp.orientation_lists(bins, probs, order, n_psi, min_top1=args.min_top1)

            lists = lists_of()
            res = None
            for k in range(steps):
                t_a = time.time()
                bp.initMdl()
                es_p.set_model([proj], [bp], [1.0], sigma2, pdf_direction=pdf_dir)
                es_p.set_orientation_lists(*lists)
                phase_ms["setup"] += 1000.0 * (time.time() - t_a)
                box_ = {}

                def run():
                    try:
                        box_["t0"] = time.time()
                        box_["res"] = es_p.run(d_img.address, d_ctf.address, meta, on_device=True)[0]
                        box_["t1"] = time.time()
                    except BaseException as exc:  # re-raised in the caller
                        box_["exc"] = exc

                t_a = time.time()
                th = threading.Thread(target=run)
                th.start()
                nxt = lists_of() if k + 1 < steps else None
                phase_ms["next_lists"] += 1000.0 * (time.time() - t_a)
                th.join()
                t_b = time.time()
                phase_ms["estep_run"] += 1000.0 * (t_b - t_a)
                if "exc" in box_:
                    raise box_["exc"]
                phase_ms["estep_call"] += 1000.0 * (box_["t1"] - box_["t0"])
                phase_ms["library_call"] += es_p.last_call_ms
                phase_ms["thread_start"] += 1000.0 * (box_["t0"] - t_a)
                phase_ms["thread_end"] += 1000.0 * (t_b - box_["t1"])
                res, cur = box_["res"], lists
                if nxt is not None:
                    lists = nxt
            return res, cur[0]

        # host wall time per step: set-up, the next batch's inference and lists (this thread), the E-step from
        # thread start to join, of it the E-step wrapper call and the library call, and the thread's start / end
        phase_ms = {"setup": 0.0, "next_lists": 0.0, "estep_run": 0.0, "estep_call": 0.0, "library_call": 0.0,
                    "thread_start": 0.0, "thread_end": 0.0}
        hipmem.synchronize()
        t0 = time.time()
        pl_out = pipelined_steps(args.steps)
        hipmem.synchronize()
        el = time.time() - t0
        runs["proposal_pipelined"] = dict(out=pl_out, elapsed=el, timings=es_p.timings(), counters=es_p.counters(),
                                          value=n * args.steps / el, ms=1000.0 * el / args.steps, steps=args.steps,
                                          warmup=args.warmup, particles=n)
        extra["proposal_pipelined_host_ms_per_step"] = {k: round(v / args.steps, 2) for k, v in phase_ms.items()}
        log(f"[proposal_pipelined] timed {args.steps} steps: {el:.2f} s; per step "
            f"{json.dumps(extra['proposal_pipelined_host_ms_per_step'])}", args)
        off = runs["proposal"]["out"][1]
        if args.lists_out:
            bins, probs = pp.predict(d_img.address, d_ctf.address, n, box, topk=args.topk, on_device=True)
            o2, g2 = pp.orientation_lists(bins, probs, order, n_psi, min_top1=args.min_top1)
            np.savez(args.lists_out, off=o2, og=g2)
            log(f"orientation lists saved to {args.lists_out}", args)
        extra["proposal_fallback_fraction"] = round(float(np.mean(np.diff(off) == 0)), 4)
        extra["proposal_orientations_per_particle"] = round(float(np.mean(np.diff(off)[np.diff(off) > 0])), 1)
    # ---------------------------------------------------------------- exhaustive plain FP32 baseline
    if "plain" in paths:
        cfg_g.precision = ra.RLN_PREC_PLAIN_FP32
        es_x = engine(cfg_g, args)

        def plain_step():
            bp.initMdl()
            es_x.set_model([proj], [bp], [1.0], sigma2, pdf_direction=pdf_dir)
            res, _ = es_x.run(d_img.address, d_ctf.address, meta_plain, on_device=True)
            return res

        runs["plain"] = timed("plain", plain_step, n_plain, 1, 0, es_x)
        map_plain = bp.reconstruct(box, do_map=False, skip_gridding=True)
        es_x.close()
        res_x = runs["plain"]["out"]
        lim = int(0.8 * (box // 2))

        def versus_exhaustive():
            """The proposal path on the plain subset (untimed): best poses against the exhaustive search's, and the
            FSC between the two paths' reconstructions (BackProjector::reconstruct on the device, no MAP term)."""
            res_p, off_p = proposal_step(n_plain, meta_plain)
            map_prop = bp.reconstruct(box, do_map=False, skip_gridding=True)
            ang = np.array([angle_between((a["rot"], a["tilt"], a["psi"]), (b["rot"], b["tilt"], b["psi"]))
                            for a, b in zip(res_x, res_p)])
            same = (ang < 1e-3) & (np.abs(res_x["xoff"] - res_p["xoff"]) < 1e-6) & \
                   (np.abs(res_x["yoff"] - res_p["yoff"]) < 1e-6)
            f = fsc_maps(map_plain, map_prop, box // 2)
            return {"particles": n_plain, "kept_best_pose": round(float(same.mean()), 4),
                    "miss_rate": round(float(1.0 - same.mean()), 4),
                    "misses_within_fine_step_deg": round(float(np.mean(ang[~same] < 3.75)) if (~same).any() else 1.0, 4),
                    "fallback_fraction": round(float(np.mean(np.diff(off_p) == 0)), 4),
                    "fsc_min_to_0.8_nyquist": round(float(f[1:lim + 1].min()), 5),
                    "fsc_mean_to_0.8_nyquist": round(float(f[1:lim + 1].mean()), 5),
                    "fsc_of": "reconstructions (skip-gridding path, no MAP term) of the two paths' backprojectors"}

        if pp is not None:
            extra["proposal_vs_exhaustive"] = versus_exhaustive()
            log(f"proposal vs exhaustive: {json.dumps(extra['proposal_vs_exhaustive'])}", args)
            if args.finetune_samples > 0:
                # the next iteration: warm start from this network, fresh projections of the reference plus the
                # particles labelled with this iteration's max-likelihood poses (the last timed step's results)
                res_ml = runs["proposal"]["out"][0]
                poses_ml = np.stack([res_ml["rot"], res_ml["tilt"], res_ml["psi"]], 1)
                hipmem.synchronize()
                t = time.time()
                pp.clear_training()
                pp.generate(proj, args.finetune_samples, ori_size=box, pixel_size=px, sigma2_noise=sigma2,
                            defocus=(8000.0, 25000.0), max_astigmatism=300.0, offset_max=3.0,
                            first=args.train_epochs * args.train_samples)
                pp.add_images(d_img.address, d_ctf.address, box, poses_ml, on_device=True)
                hipmem.synchronize()
                t_g = time.time() - t
                t = time.time()
                loss2 = pp.train(epochs=1, batch=256, lr=3e-4)
                hipmem.synchronize()
                t_t = time.time() - t
                log(f"proposal fine-tune: {args.finetune_samples} projections + {n} ML-labelled particles, "
                    f"generation {t_g:.2f} s, training {t_t:.2f} s, loss {loss2:.3f}", args)
                runs["proposal_finetuned"] = timed("proposal_finetuned", proposal_step, n, args.steps, args.warmup,
                                                   es_p)
                vs = versus_exhaustive()
                log(f"fine-tuned proposal vs exhaustive: {json.dumps(vs)}", args)
                it_s = runs["proposal_finetuned"]["ms"] / 1000.0 + t_g + t_t
                v_plain, v_ft = runs["plain"]["value"], runs["proposal_finetuned"]["value"]

                def break_even(mult):
                    # particles per iteration at which n / (n / v_ft + t_ft) reaches mult * v_plain
                    d = 1.0 / (mult * v_plain) - 1.0 / v_ft
                    return int(np.ceil((t_g + t_t) / d)) if d > 0 else None

                extra["proposal_finetune"] = {
                    "samples": args.finetune_samples, "ml_labelled_particles": n, "epochs": 1, "lr": 3e-4,
                    "generate_s": round(t_g, 3), "train_s": round(t_t, 3), "loss": round(loss2, 4),
                    "vs_exhaustive": vs}
                extra["amortised_per_iteration"] = {
                    "particles": n, "estep_s": round(runs["proposal_finetuned"]["ms"] / 1000.0, 4),
                    "training_s": round(t_g + t_t, 3), "particles_per_s": round(n / it_s, 1),
                    "vs_exhaustive_plain": round(n / it_s / v_plain, 2),
                    "break_even_particles_1x_plain": break_even(1.0),
                    "break_even_particles_5x_plain": break_even(5.0),
                    "note": "an iteration after the first: E-step with the warm-started network plus its training "
                            "(fresh projections and this iteration's max-likelihood assignments); the first "
                            "iteration's cold training is proposal_training"}
    if "proposal" in runs:
        es_p.close()
    # ---------------------------------------------------------------- precision variant: local search
    if "local" in paths:
        for name, prec in (("local_plain", ra.RLN_PREC_PLAIN_FP32), ("local_split", ra.RLN_PREC_SPLIT_BF16)):
            cfg_l.precision = prec
            es_l = engine(cfg_l, args)

            def local_step():
                bp.initMdl()
                es_l.set_model([proj], [bp], [1.0], sigma2)
                res, _ = es_l.run(d_img.address, d_ctf.address, meta, on_device=True)
                return res

            runs[name] = timed(name, local_step, n, args.steps, args.warmup, es_l)
            es_l.close()

    # ---------------------------------------------------------------- classification and uneven-work workloads
    if "classify" in paths:
        extra["classification_box128_k4"] = classify_leg(args, args.classify_particles)
    if "uneven" in paths:
        u = uneven_leg(args, args.uneven_particles, 1, 1)
        extra["uneven_box192"] = {k: u[k] for k in ("value", "ms_per_step", "config", "stages_ms",
                                                    "significant_coarse_poses_per_particle")}
        extra["uneven_box192"]["counters"] = {k: u["counters"][k] for k in ("fine_poses", "fine_grid", "fine_bands")}
    # ---------------------------------------------------------------- reconstruction (the consumer of the backprojector)
    if "recon" in paths and runs:
        extra["reconstruction"] = recon_leg(args, bp, box, 1e3 * np.exp(-np.arange(box // 2 + 1) / (box / 16.0)))
    # ---------------------------------------------------------------- reference preparation (the producer of the projector)
    if "refprep" in paths:
        extra["reference_preparation"] = refprep_leg(args, spec, box)
    # ---------------------------------------------------------------- particle input from STAR / MRC files
    if "formats" in paths:
        extra["particle_input"] = formats_leg(args, box, px)
    # ---------------------------------------------------------------- image preparation (the stage before the E-step)
    if "prep" in paths:
        extra["image_preparation"] = prep_leg(args, box, px, diameter, d_img, nvar)
    if not runs:  # preparation stages alone
        print(json.dumps({"metric": "preparation stages, box %d" % box, "n_gpus": 1, "data": "synthetic", **extra}),
              flush=True)
        return 0
    head = ("proposal_pipelined" if "proposal_pipelined" in runs else "proposal" if "proposal" in runs else
            ("local_split" if "local_split" in runs else next(iter(runs))))
    R = runs[head]
    timings, counters = R["timings"], R["counters"]
    is_prop = head in ("proposal", "proposal_pipelined")
    res = R["out"][0] if is_prop else R["out"]
    errs = np.array([angle_between(pose[i, :3], (res[i]["rot"], res[i]["tilt"], res[i]["psi"]))
                     for i in range(0, len(res), max(1, len(res) // 500))])

    # roofline of the dominant kernel of the reported path, on the algorithmic work of the E-step's cost model:
    #   coarse diff2: F = 4 P_c N_pose FLOP (the cross term Re sum conj(X_t) CTF P, 2 FMA per complex pixel)
    #     against FP32 peak (plain) or BF16 MFMA peak / 3 (split-BF16: hi*hi + hi*lo + lo*hi);
    #   fine diff2: B_proj = 8 P N_or bytes, the complex-FP32 projections of the N_or fine orientations (8 per fine
    #     job) at the fine image size P = s (s/2 + 1), against HBM peak;
    #   weighted sums / backprojection: B_proj of the backprojected orientations plus the read-modify-write of
    #     the three FP32 accumulators at the backprojected pixels (24 B per pixel and orientation).
    # The fine pass actually gathers one 64-byte corner block of the packed reference per sampled pixel and
    # orientation; that request rate is reported beside the roofline (gather_requests).
    stage = max(("fine_diff2", "coarse_diff2", "wsum_backproject"), key=lambda k: timings[k])
    split = is_prop or head == "local_split"
    P_f = counters["fine_pixels"]
    gather = None
    if stage == "coarse_diff2":
        work = 4.0 * counters["coarse_poses"] * counters["coarse_pixels"]
        bound, unit = "mfma", "TFLOP/s"
        peak = PEAK_BF16_TFLOPS / 3.0 if split else PEAK_FP32_TFLOPS
        achieved = work / (timings[stage] * 1e-3) / 1e12
        alg = {"flop_per_step": work, "formula": "4 * P_c * N_pose"}
    else:
        if stage == "fine_diff2":
            n_or = 8 * counters["fine_jobs"]
            work = 8.0 * P_f * n_or
            alg = {"bytes_per_step": work, "formula": "8 * P * N_or", "P": P_f, "N_or": n_or}
            req = 64.0 * n_or * counters["fine_pixels_sampled"]
            gather = {"bytes_per_step": req, "GB_per_s": round(req / (timings[stage] * 1e-3) / 1e9, 1),
                      "requests_per_s": round(req / 64.0 / (timings[stage] * 1e-3), 1),
                      "note": "one 64-byte corner block per sampled (pixel, fine orientation)"}
        else:
            n_or = counters["wsum_orientations"]
            work = (8.0 + 24.0) * P_f * n_or
            alg = {"bytes_per_step": work, "formula": "(8 + 24) * P * N_or_backprojected", "P": P_f, "N_or": n_or}
        bound, unit, peak = "hbm", "GB/s", PEAK_HBM_GBS
        achieved = work / (timings[stage] * 1e-3) / 1e9
    traffic, traffic_src = None, None
    if args.traffic_kib is not None:
        traffic, traffic_src = round(args.traffic_kib * 1024.0), "--traffic-kib"
    elif is_prop and box == 256 and n == 10000 and stage == "fine_diff2":
        shape = f"k_fine_diff2<0,{counters['fine_small_tc']},{counters['fine_wide_cw']}>"
        traffic, traffic_src = committed_traffic("estep_proposal_box256_order3_10k_traffic.md", shape)
    roofline = {"kernel": stage, "bound": bound, "achieved": round(achieved, 3), "peak": round(peak, 1),
                "unit": unit, "frac": round(achieved / peak, 4), "traffic": traffic, "traffic_source": traffic_src,
                "algorithmic": alg, "stage_ms": round(timings[stage], 2), "launches": counters["chunks"],
                "avg_launch_ms": round(timings[stage] / max(1, counters["chunks"]), 3)}
    if gather is not None:
        roofline["gather_requests"] = gather

    cpu = None
    if args.cpu_seconds > 0:
        import oracle
        threads = args.cpu_threads or len(os.sched_getaffinity(0))
        ccfg = cfg_g if is_prop else cfg_l
        m, dt = 0, 0.0
        while dt < args.cpu_seconds and m < n:
            k = min(threads, n - m)
            sub_img, sub_ctf, sub_meta, _, _ = ra.synth_particles(spec, k, box, first=m)
            t = time.time()
            oracle.estep(ccfg, [vol], [1.0], sigma2, sub_meta, sub_img, sub_ctf,
                         pdf_direction=None if ccfg.local_search else pdf_dir[None, :], threads=threads)
            dt += time.time() - t
            m += k
        search = "exhaustive global" if is_prop else "local"
        cpu = {"value": round(m / dt, 4), "unit": "particles/s", "cores": threads, "kind": "port",
               "sample": f"first {m} particles of the same data, {search} search ({dt:.1f} s), FP64 oracle, one "
                         "particle per OpenMP thread"}
        log(f"cpu oracle: {m} particles in {dt:.1f} s on {threads} threads", args)

    def summary(v):
        return {"particles_per_s": round(v["value"], 2), "ms_per_step": round(v["ms"], 2),
                "particles": v["particles"], "steps": v["steps"], "warmup": v["warmup"],
                "stages_ms": {s_: round(t_, 2) for s_, t_ in v["timings"].items()},
                "coarse_poses": v["counters"]["coarse_poses"]}

    speed = {}
    if "proposal" in runs and "plain" in runs:
        speed["proposal_vs_exhaustive_plain"] = round(runs["proposal"]["value"] / runs["plain"]["value"], 2)
    if "proposal_pipelined" in runs and "plain" in runs:
        speed["proposal_pipelined_vs_exhaustive_plain"] = round(runs["proposal_pipelined"]["value"] /
                                                                runs["plain"]["value"], 2)
    if "proposal_pipelined" in runs:
        speed["pipelined_vs_sequential_proposal"] = round(runs["proposal_pipelined"]["value"] /
                                                          runs["proposal"]["value"], 3)
    if "local_split" in runs and "local_plain" in runs:
        speed["local_split_vs_local_plain"] = round(runs["local_split"]["value"] / runs["local_plain"]["value"], 3)
    out = {
        "metric": "E-step particles/s, box 256 refine: plain FP32 vs proposal+split-BF16",
        "value": round(R["value"], 3), "unit": "particles/s", "n_gpus": 1, "steps": R["steps"],
        "warmup": R["warmup"], "ms_per_step": round(R["ms"], 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16x3+f32" if split else "f32", "data": "synthetic",
        "config": {"workload": (f"refine box{box} order{order} "
                                f"{'global search seeded by the pose proposal' if is_prop else 'local search'}"
                                f" {n} particles"),
                   "path": {"proposal": "proposal top-16 + split-bf16 coarse + fp32 fine",
                            "proposal_pipelined": "proposal top-16 + split-bf16 coarse + fp32 fine; the next batch's "
                                                  "inference and orientation lists overlap this batch's E-step",
                            "local_split": "split-bf16 coarse + fp32 fine"}.get(head, head),
                   "global_batch": n, "coarse_size": coarse, "particle_diameter_A": diameter, "parallelism": "1 GPU"},
        "roofline": roofline, "cpu_baseline": cpu,
        "stages_ms": {k: round(v, 2) for k, v in timings.items()}, "counters": counters,
        "paths": {k: summary(v) for k, v in runs.items()},
        "speedup": speed,
        "accuracy_vs_true_pose": {"median_angular_error_deg": round(float(np.median(errs)), 3),
                                  "frac_within_3deg": round(float((errs < 3.0).mean()), 4)},
    }
    out.update(extra)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
