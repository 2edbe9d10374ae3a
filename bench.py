#!/usr/bin/env python3
"""E-step throughput on one MI355X: particles/s of the expectation step at box 256, HEALPix order 3.

Workload (stated choices marked *): one 256^3 map of 150 Gaussian blobs at 1.0 A/pixel, 10,000 particles
(256x256, SNR 0.03, defocus U(0.8, 2.5) um, 300 kV, Cs 2.7 mm, Q0 0.1, offsets U(+-3 px)), HEALPix order 3
local search (*sigma_ang 7.5 deg = twice the fine angular step, cut at 3 sigma; *priors = true pose + 3 deg
Gaussian noise), offsets 4 A range, 0.5 A step, oversampling 1, padding 2, adaptive_fraction 0.999,
*particle diameter 200 A (coarse image 48 px), *current_size = box. Data are synthetic (seeded).

A step is one E-step over all particles: projections of the coarse grid, coarse pass, weights and
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
    ap.add_argument("--box", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="CPU oracle budget: batches of one particle per thread until this much time (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--paths", default="plain,split",
                    help="comma list of E-step paths to time: plain (FP32 fallback), split (split-BF16 MFMA coarse)")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run one device-to-device copy of this size (reference for traffic counters)")
    ap.add_argument("--traffic-kib", type=float, default=None,
                    help="measured HBM traffic per launch of the dominant kernel (KiB, counter-calibrated)")
    return ap.parse_args()


def log(msg, args):
    if not args.quiet:
        print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs the E-step on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md, not run)"}))
        return 2
    import relion_amd as ra
    from relion_amd import hipmem

    if hipmem.device_count() < 1:
        print(json.dumps({"error": "no HIP device visible"}))
        return 2
    hipmem.set_device(0)
    box, px, order = args.box, 256.0 / args.box if args.box != 256 else 1.0, 3
    diameter = 200.0 * box / 256.0
    coarse = ra.coarse_image_size(box, px, order, diameter)
    cfg = ra.make_config(box, px, order, offset_range=4.0, offset_step=0.5, coarse_size=coarse, local_search=True,
                         sigma_ang=7.5)
    spec = ra.synth_spec(args.seed, box, px, 150, snr=0.03, defocus=(8000.0, 25000.0), offset_max=3.0,
                         prior_sigma_deg=3.0)
    t0 = time.time()
    vol = ra.synth_reference(spec, 0, box)
    n = args.particles
    fimg, fctf, meta, pose, nvar = ra.synth_particles(spec, n, box)
    log(f"generated reference + {n} particles in {time.time() - t0:.1f} s (coarse {coarse}px)", args)
    sigma2 = np.full(box // 2 + 1, float(nvar.mean()))
    d_img = hipmem.DeviceArray(fimg)
    d_ctf = hipmem.DeviceArray(fctf)
    del fimg, fctf
    dims = ra.volume_dims(box)
    proj = ra.AccProjector(0)
    proj.setMdlDim(*dims, 2.0)
    proj.initMdl(vol)
    bp = ra.AccBackprojector(0)
    bp.setMdlDim(*dims, 2.0)
    precisions = {"plain": ra.RLN_PREC_PLAIN_FP32, "split": ra.RLN_PREC_SPLIT_BF16}

    def time_path(name):
        cfg.precision = precisions[name]
        es = ra.ExpectationStep(cfg, 0)

        def step():
            bp.initMdl()
            es.set_model([proj], [bp], [1.0], sigma2)   # per-iteration model: invalidates the grid projections
            return es.run(d_img.address, d_ctf.address, meta, on_device=True)

        for w in range(args.warmup):
            t = time.time()
            step()
            log(f"[{name}] warmup {w}: {time.time() - t:.2f} s", args)
        hipmem.synchronize()
        t_start = time.time()
        for _ in range(args.steps):
            res, wsum = step()
        hipmem.synchronize()
        elapsed = time.time() - t_start
        timings, counters = es.timings(), es.counters()
        log(f"[{name}] timed {args.steps} steps: {elapsed:.2f} s; stages "
            f"{json.dumps({k: round(v, 1) for k, v in timings.items()})}", args)
        log(f"[{name}] counters {json.dumps(counters)}", args)
        es.close()
        return dict(res=res, elapsed=elapsed, timings=timings, counters=counters,
                    value=n * args.steps / elapsed, ms=1000.0 * elapsed / args.steps)

    if args.copy_calibration_mb > 0:
        nb = args.copy_calibration_mb << 20
        src = hipmem.DeviceArray(np.zeros(nb, np.uint8))
        dst = hipmem.DeviceArray(np.zeros(nb, np.uint8))
        hipmem.copy_device(dst.address, src.address, nb)
        hipmem.synchronize()
        log(f"copy calibration: {nb} bytes device-to-device", args)
        del src, dst
    runs = {name: time_path(name) for name in args.paths.split(",")}
    head = "split" if "split" in runs else "plain"
    R = runs[head]
    res, timings, counters = R["res"], R["timings"], R["counters"]
    value, ms_per_step = R["value"], R["ms"]

    # accuracy against the generating poses (orientation error of the max-likelihood pose)
    def rotm(a, b, c):
        a, b, c = np.radians([a, b, c])
        ca, sa, cb, sb, cg, sg = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
        return np.array([[cg * cb * ca - sg * sa, cg * cb * sa + sg * ca, -cg * sb],
                         [-sg * cb * ca - cg * sa, -sg * cb * sa + cg * ca, sg * sb],
                         [sb * ca, sb * sa, cb]])
    errs = []
    for i in range(0, n, max(1, n // 500)):
        A = rotm(*pose[i, :3])
        B = rotm(res[i]["rot"], res[i]["tilt"], res[i]["psi"])
        errs.append(np.degrees(np.arccos(np.clip((np.trace(A @ B.T) - 1) / 2, -1, 1))))
    errs = np.array(errs)

    # roofline of the dominant kernel of the reported path
    stage = max(("fine_diff2", "coarse_diff2", "wsum_backproject"), key=lambda k: timings[k])
    if stage == "coarse_diff2":
        work = 4.0 * counters["coarse_poses"] * counters["coarse_pixels"]      # FLOP, cross-term count
        bound, unit = "mfma", "TFLOP/s"
        peak = PEAK_BF16_TFLOPS / 3.0 if head == "split" else PEAK_FP32_TFLOPS
        achieved = work / (timings[stage] * 1e-3) / 1e12
    else:
        # gather-bound: 64 bytes of the corner-packed reference per (fine orientation, pixel) sample, plus the
        # 128-byte slot line read and written per backprojected sample in the weighted sums
        if stage == "fine_diff2":
            work = 64.0 * 8 * counters["fine_jobs"] * counters["fine_pixels"]
        else:
            work = (64.0 + 256.0) * counters["wsum_orientations"] * counters["fine_pixels"]
        bound, unit, peak = "hbm", "GB/s", PEAK_HBM_GBS
        achieved = work / (timings[stage] * 1e-3) / 1e9
    roofline = {"kernel": stage, "bound": bound, "achieved": round(achieved, 3), "peak": round(peak, 1),
                "unit": unit, "frac": round(achieved / peak, 4),
                "traffic": None if args.traffic_kib is None else round(args.traffic_kib * 1024.0),
                "stage_ms": round(timings[stage], 2), "launches": counters["chunks"],
                "avg_launch_ms": round(timings[stage] / max(1, counters["chunks"]), 3)}

    cpu = None
    if args.cpu_seconds > 0:
        import oracle
        threads = args.cpu_threads or len(os.sched_getaffinity(0))
        m, dt = 0, 0.0
        while dt < args.cpu_seconds and m < n:
            k = min(threads, n - m)
            sub_img, sub_ctf, sub_meta, _, _ = ra.synth_particles(spec, k, box, first=m)
            t = time.time()
            oracle.estep(cfg, [vol], [1.0], sigma2, sub_meta, sub_img, sub_ctf, threads=threads)
            dt += time.time() - t
            m += k
        cpu = {"value": round(m / dt, 4), "unit": "particles/s", "cores": threads, "kind": "port",
               "sample": f"first {m} particles of the same workload ({dt:.1f} s), FP64 oracle, one particle per "
                         "OpenMP thread"}
        log(f"cpu oracle: {m} particles in {dt:.1f} s on {threads} threads", args)

    out = {
        "metric": "E-step particles/s, box 256 refine: plain FP32 vs proposal+split-BF16",
        "value": round(value, 3), "unit": "particles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3+f32" if head == "split" else "f32", "data": "synthetic",
        "config": {"workload": f"refine box{box} order{order} local-search {n} particles",
                   "path": "split-bf16 coarse + fp32 fine" if head == "split" else "plain-fp32",
                   "global_batch": n, "coarse_size": coarse, "sigma_ang_deg": 7.5, "particle_diameter_A": diameter,
                   "parallelism": "1 GPU"},
        "roofline": roofline, "cpu_baseline": cpu,
        "stages_ms": {k: round(v, 2) for k, v in timings.items()}, "counters": counters,
        "paths": {k: {"particles_per_s": round(v["value"], 1), "ms_per_step": round(v["ms"], 2),
                      "stages_ms": {s_: round(t_, 2) for s_, t_ in v["timings"].items()}} for k, v in runs.items()},
        "speedup_vs_plain": (round(runs["split"]["value"] / runs["plain"]["value"], 3)
                             if "split" in runs and "plain" in runs else None),
        "accuracy": {"median_angular_error_deg": round(float(np.median(errs)), 3),
                     "frac_within_3deg": round(float((errs < 3.0).mean()), 4)},
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
