"""Benchmark of the MI355X SIFT front end + guided matcher on the realistic-shape workload.

Workload (one "step"): 8 synthetic nadir photos of 2640x1978 uint8 cropped from one seeded
field mosaic along a 2x4 flight grid (75 % front / 65 % side overlap, +-3 deg rotation, +-2 %
scale, noise sigma 2).  SIFT per photo: OpenCV-default thresholds (contrast 0.04, edge 10),
detect int(8000*2.2) keypoints, grid-spread to 8000, describe, RootSIFT uint8.  Guided
matching: every overlapping pair, prior = true mapping + N(0, 3 px), 60 px window, top-2,
ratio 0.8, one query per target.

Prints ONE JSON line.  ``value`` = SIFT images/s with the photos resident in HBM, consecutive
steps rotating over --inflight feature batches (workspaces) on as many streams, so that one
step's latency-bound stages overlap the next step's pyramid (each step does the whole pipeline);
``matching`` holds matched pairs/s of the fast and the plain FP32 GPU forms and whether the
match set equals the FP64 oracle; ``precision`` holds the float-descriptor brute-force section
(200 pairs, BF16 and FP8 screening, distance error against float64, match-set equality);
``roofline`` is for the dominant kernel (the pyramid's LDS strip blur); ``flight400`` holds the
400-photo mixed-resolution flight (SIFT in graph-captured chunks against the HBM roofline, guided
matching of its ~3400 pairs in both forms, a sample against the FP64 oracle, and match_pair's
whole chain through match_pairs); ``cpu_baseline`` times the reference's CPU pool shape
(os.cpu_count() - 1 single-threaded workers) on a bounded sample: SIFT with the C restatement of
gsift's CPU path and guided matching with the float32 numpy matcher.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12        # MI355X HBM3E, bytes/s
INT8_PEAK = 5.0e15       # MI355X dense INT8 MFMA, ops/s
NFEAT = 8000
CONTRAST, EDGE = 0.04, 10.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_flight(seed=2, workers=8):
    from fieldstitch_amd import synth
    return synth.flight(seed=seed, workers=max(1, min(workers, os.cpu_count() or 1)))


def pyramid_bytes(H, W):
    """Algorithmic HBM bytes of the pyramid stage of one image: u8 in + base layer out +
    per octave pixel 5 x (read G_i, write G_i+1, write D_i) x 4 B + 2 B downsample."""
    from fieldstitch_amd import gsift
    pl = gsift.plan(H, W)
    sumP = sum(r * c for r, c in pl)
    P0 = pl[0][0] * pl[0][1]
    return H * W + 4 * P0 + 62 * sumP


def sift_bytes(H, W):
    """Algorithmic bytes of the whole SIFT front end of one image: 118 * sumP + 4 * P0 + H*W
    (pyramid + downsample + gradient of 3 layers + extrema over 5 DoG)."""
    from fieldstitch_amd import gsift
    pl = gsift.plan(H, W)
    sumP = sum(r * c for r, c in pl)
    P0 = pl[0][0] * pl[0][1]
    return H * W + 4 * P0 + 118 * sumP


# ---- CPU baseline: the reference's CPU worker pool shape -------------------------------------
# A spawn pool of os.cpu_count() - 1 single-threaded workers (sysinfo.py:28-35, pipeline.py:14-18);
# photos and features reach the workers through files they memory-map, so the timed maps carry
# only indices.

_W = {}


def _cpu_pool_init(photo_file, feat_file):
    os.environ["OMP_NUM_THREADS"] = os.environ["OPENBLAS_NUM_THREADS"] = os.environ["MKL_NUM_THREADS"] = "1"
    _W["photos"] = np.load(photo_file, mmap_mode="r")
    with np.load(feat_file) as f:
        _W["feat"] = {k: f[k] for k in f.files}


def _cpu_noop(k):
    return k


def cpu_sift_photo(k):
    """One photo through the CPU SIFT front end: the C restatement of gsift's CPU path (detect
    int(2.2 nfeat), spread to nfeat, response order, describe, RootSIFT)."""
    from oracle import sift_oracle as so
    from fieldstitch_amd import features
    img = np.ascontiguousarray(_W["photos"][k % len(_W["photos"])])
    t = time.perf_counter()
    kp, st = so.detect(img, int(NFEAT * 2.2), CONTRAST, EDGE)
    sel = features.spread_select(kp["x"], kp["y"], kp["response"], img.shape[1], img.shape[0], NFEAT)
    sel = sel[np.argsort(-kp["response"][sel], kind="stable")]
    so.rootsift_u8(so.describe(st, kp, sel))
    return time.perf_counter() - t


def cpu_match_pair(p):
    """One guided pair with the reference's float32 numpy matcher (vectorised restatement of
    matching.py:113-152, every candidate in the 60 px window)."""
    from oracle import match_oracle as mo
    f = _W["feat"]
    i, j = int(f["pairs"][p, 0]), int(f["pairs"][p, 1])
    t = time.perf_counter()
    mo.guided_f32_fast(f["desc%d" % i], f["xy%d" % i], f["desc%d" % j], f["xy%d" % j], f["Hm"][p],
                       tuple(f["wh%d" % j]), RADIUS, RATIO)
    return time.perf_counter() - t


def cpu_baseline(images, host, pairs, Hm, shapes, match_reps=16):
    """SIFT images/s and guided-matching pairs/s of the CPU pool on the host cores: 2 photos
    per worker (the benchmark's photos cycled), then the benchmark's pairs match_reps times in
    chunks of 8 (pipeline.py:536)."""
    import multiprocessing as mp
    import tempfile
    n_w = max(1, (os.cpu_count() or 1) - 1)
    try:
        import cv2  # noqa: F401
        have_cv2 = True
    except Exception:  # noqa: BLE001 - OpenCV is optional on the box
        have_cv2 = False
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        photo_file = os.path.join(tmp, "photos.npy")
        np.save(photo_file, np.stack(images))
        feat = {"pairs": np.asarray(pairs, np.int64), "Hm": np.stack([np.asarray(h, np.float64) for h in Hm])}
        for i, (u8, xy) in enumerate(host):
            feat["desc%d" % i], feat["xy%d" % i] = u8, xy
            feat["wh%d" % i] = np.array([shapes[i][1], shapes[i][0]])
        feat_file = os.path.join(tmp, "feat.npz")
        np.savez(feat_file, **feat)
        ctx = mp.get_context("spawn")
        with ctx.Pool(n_w, initializer=_cpu_pool_init, initargs=(photo_file, feat_file)) as pool:
            pool.map(_cpu_noop, range(4 * n_w))                 # every worker started and importing done
            pool.map(cpu_match_pair, range(min(n_w, len(pairs))))   # imports of the matching leg
            n_photo = 2 * n_w
            t0 = time.perf_counter()
            per_photo = pool.map(cpu_sift_photo, range(n_photo), chunksize=1)
            t_sift = time.perf_counter() - t0
            tasks = list(range(len(pairs))) * match_reps
            t0 = time.perf_counter()
            per_pair = pool.map(cpu_match_pair, tasks, chunksize=8)
            t_match = time.perf_counter() - t0
    return {"value": n_photo / t_sift, "unit": "images/s", "cores": n_w, "kind": "port",
            "pool": "spawn pool of os.cpu_count() - 1 = %d single-threaded workers" % n_w,
            "sample": "%d photos (the %d benchmark photos cycled, 2 per worker) through the C restatement of "
                      "gsift's CPU path (detect, spread, describe, RootSIFT): wall %.1f s, %.1f s CPU per photo; "
                      "OpenCV %s" % (n_photo, len(images), t_sift, float(np.mean(per_photo)),
                                      "importable but not timed" if have_cv2 else "not installed"),
            "matching": {"value": len(tasks) / t_match, "unit": "pairs/s", "cores": n_w, "kind": "port",
                         "sample": "the %d guided pairs x %d through the float32 numpy guided matcher "
                                   "(oracle.match_oracle.guided_f32_fast) in the same pool, chunksize 8: wall %.2f s, "
                                   "%.3f s CPU per pair" % (len(pairs), match_reps, t_match, float(np.mean(per_pair)))},
            "host_cpu_count": os.cpu_count(), "host_cpus_usable": len(os.sched_getaffinity(0)),
            "cpu_model": cpu_model()}


def cpu_model():
    """The host CPU's model name (first 'model name' of /proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


RADIUS, RATIO = 60.0, 0.8


def bench_matching(fb, fl, shapes, args):
    """Guided matching of every overlapping pair on the device-resident features: the MFMA
    form and the plain FP32 form, timed separately; both checked against the FP64 oracle."""
    import torch
    from fieldstitch_amd.matching import PairMatcher
    from oracle import match_oracle as mo
    n_det, n_sel = fb.counts()
    images, host = [], []
    for i, (H, W) in enumerate(shapes):
        xy, u8, resp, _ = fb.results(i)
        xy64 = xy.astype(np.float64)     # reduce 1, zero distortion: und = xy
        host.append((u8, xy64))
        images.append((fb.u8[i, :len(xy)].contiguous(), torch.from_numpy(xy64).cuda(), (W, H)))
    pairs = list(fl.pairs)
    Hm = [fl.prior[p] for p in pairs]
    pm = PairMatcher(images, pairs, Hm, radius=RADIUS, kc=0, ratio=RATIO)
    from fieldstitch_amd import autotune
    out = {"pairs": len(pairs), "radius_px": RADIUS, "ratio": RATIO, "kc": "all in window",
           "autotune": autotune.tune_matcher(pm, "mfma", reps=3)}
    results = {}
    for form in ("mfma", "plain"):
        for _ in range(max(1, args.warmup)):
            pm.run(form)
        torch.cuda.synchronize()
        # eager launches with stage events: setup / contraction / compaction breakdown
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
        t0 = time.perf_counter()
        for k in range(args.steps):
            pm.run(form, events=evs[k])
        torch.cuda.synchronize()
        t_eager = (time.perf_counter() - t0) / args.steps
        st = [float(np.mean([evs[k][i].elapsed_time(evs[k][i + 1]) for k in range(args.steps)])) for i in range(3)]
        # the timed figure: the launch sequence replayed as a HIP graph
        pm.capture(form)
        for _ in range(max(1, args.warmup)):
            pm.run(form, graph=True)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record()
        for k in range(args.steps):
            pm.run(form, graph=True)
        e1.record()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / args.steps
        results[form] = pm.results()
        out[form] = {"pairs_per_s": len(pairs) / t, "ms_per_step": t * 1e3,
                     "gpu_ms_per_step": e0.elapsed_time(e1) / args.steps, "eager_ms_per_step": t_eager * 1e3,
                     "setup_ms": st[0], "contraction_ms": st[1], "compact_ms": st[2]}
    # exactness: both forms against the FP64 oracle, pair by pair
    equal_forms = all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
                      for a, b in zip(results["mfma"], results["plain"]))
    n_equal, n_matches, cand = 0, 0, 0
    t_cpu = time.perf_counter()
    for p, (i, j) in enumerate(pairs):
        q, t = mo.guided_fp64_fast(host[i][0], host[i][1], host[j][0], host[j][1], Hm[p], (shapes[j][1], shapes[j][0]),
                                   RADIUS, RATIO)
        got = results["mfma"][p]
        n_equal += int(np.array_equal(got[0], q) and np.array_equal(got[1], t))
        n_matches += len(q)
    t_cpu = time.perf_counter() - t_cpu
    # algorithmic contraction work: every query against the targets inside its prior window
    # (128 multiply-adds each); the MFMA form computes dense 32 x 32 blocks over whole grid
    # cells, so the MFMA work it executes is larger than this
    from scipy.spatial import cKDTree
    n_win = 0
    for p, (i, j) in enumerate(pairs):
        xa = host[i][1]
        pr = np.c_[xa, np.ones(len(xa))] @ np.asarray(Hm[p], np.float64).T
        pr = pr[:, :2] / pr[:, 2:3]
        n_win += int(cKDTree(host[j][1]).query_ball_point(pr, RADIUS, return_length=True).sum())
    ops = 2.0 * 128 * n_win
    t_m = out["mfma"]["gpu_ms_per_step"] / 1e3
    out["window_candidates"] = n_win
    out["algorithmic_int8_tops"] = ops / t_m / 1e12
    out["mfma_roofline_frac"] = ops / t_m / INT8_PEAK
    out["mfma_roofline_note"] = ("algorithmic ops = 2 * 128 * (targets inside each query's window) over the "
                                 "MFMA form's graph-replay time, against the dense INT8 peak")
    # compulsory bytes: every descriptor of the batch read once (u8 rows)
    comp = 128.0 * sum(len(h[0]) for h in host)
    out["compulsory_bytes"] = comp
    out["compulsory_hbm_frac"] = comp / t_m / HBM_PEAK
    # the CPU oracle's FP64 brute-force guided matching of the same pairs, one process (it also
    # decides the exactness flag above)
    out["cpu_baseline"] = {"value": len(pairs) / t_cpu, "unit": "pairs/s", "cores": 1, "kind": "port",
                           "sample": "the %d pairs, FP64 numpy guided matching (oracle.match_oracle."
                                     "guided_fp64_fast), %.2f s" % (len(pairs), t_cpu)}
    out["match_set_equals_fp64_oracle"] = bool(n_equal == len(pairs) and equal_forms)
    out["pairs_equal_fp64"] = n_equal
    out["matches_total"] = n_matches
    out["speedup_vs_plain_fp32"] = out["plain"]["ms_per_step"] / out["mfma"]["ms_per_step"]
    out["value"] = out["mfma"]["pairs_per_s"]
    out["unit"] = "matched pairs/s"
    return out, host, pairs, Hm


CHAIN_CTX = {"prior_tol": 0.18, "ransac_thr": 1.0, "min_inliers": 20, "ratio": 0.8, "use_prior": True,
             "guided": True, "n_coarse": 1000, "guided_radius": 0.03, "guided_k": 16}


def bench_chain(fl, desc_dev, xy_dev, counts, shapes, pairs, reps=2):
    """match_pair's whole chain (coarse, prior, homography RANSAC, guided, F verification,
    fallback, similarity) for every pair in one match_pairs call, on device-resident features,
    with nadir cameras that see the synthetic flight (ground-plane prior on; the pipeline's
    thresholds, pipeline.py:478-482).  Timed end to end (host glue and the count reads between
    stages included)."""
    import torch
    from fieldstitch_amd import match_flow, synth
    ctx = dict(CHAIN_CTX, **synth.flight_cameras(fl, seed=7))
    cfg = match_flow.FlowConfig(ctx)
    images = {i: match_flow.FlowImage(desc_dev[i][:counts[i]], xy_dev[i][:counts[i]], shapes[i][1], shapes[i][0])
              for i in range(len(shapes))}
    stats = {}
    res = match_flow.run_pairs(pairs, images, cfg, stats=stats)          # warm-up (allocations, kernels)
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        res = match_flow.run_pairs(pairs, images, cfg, stats=stats)     # stage_ms: the last rep's
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    t = sum(ts) / reps
    matched = [r for r in res if r[2] is not None]
    return {"pairs": len(pairs), "ms": t * 1e3, "ms_reps": [round(x * 1e3, 2) for x in ts],
            "pairs_per_s": len(pairs) / t, "stage_pairs": stats,
            "matched_pairs": len(matched), "verified_matches_mean": float(np.mean([r[5] for r in matched])) if matched
            else 0.0, "errors": int(sum(1 for r in res if isinstance(r[4], tuple) and r[4] and r[4][0] == "error"))}


def bench_precision(args):
    """Float-descriptor brute-force matching of 200 pairs, n ~ U[2000, 12000] per set (SIFT-like
    descriptors, half of the smaller set planted as perturbed copies): BF16 and FP8 screening
    with the float64 certificate.  Reference: the exact float64 GPU pass (K = 0) on every pair,
    which the tests pin to the FP64 oracle, plus the CPU FP64 oracle on the two smallest pairs."""
    import torch
    from fieldstitch_amd import synth
    from fieldstitch_amd.matching import FloatMatcher
    from oracle import match_oracle as mo
    rng = np.random.default_rng(3)
    n_pair = args.precision_pairs
    sizes = rng.integers(2000, 12001, (n_pair, 2))
    t0 = time.time()
    sets = []
    for na, nb in sizes:
        da, db, _ = synth.descriptor_pair(rng, int(na), int(nb))
        sets += [da, db]
    log("precision: %d pairs generated in %.1f s" % (n_pair, time.time() - t0))
    pairs = [(2 * p, 2 * p + 1) for p in range(n_pair)]
    flops = 2.0 * 128 * float(sum(int(a) * int(b) for a, b in sizes))
    out = {"pairs": n_pair, "descriptors": int(sizes.sum()), "algorithmic_tflop": flops / 1e12}
    ref = FloatMatcher(sets, pairs, ratio=RATIO, precision="bf16", K=0)
    ref.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ref.run()
    ref_res = ref.results()
    out["fp64_exact_ms"] = (time.perf_counter() - t0) * 1e3
    peaks = {"bf16": 2.5e15, "fp8": 5.0e15}
    for prec in ("bf16", "fp8"):
        fm = FloatMatcher(sets, pairs, ratio=RATIO, precision=prec, K=8, device="cuda")
        fm.desc = ref.desc                       # same device descriptors
        for _ in range(max(1, args.warmup)):
            fm.run()
        torch.cuda.synchronize()
        steps = max(1, args.steps // 2)
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(steps)]
        t0 = time.perf_counter()
        for k in range(steps):
            fm.run(events=evs[k])
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / steps
        st = [float(np.mean([evs[k][i].elapsed_time(evs[k][i + 1]) for k in range(steps)])) for i in range(4)]
        res = fm.results()
        stats = fm.stats()
        same = all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) for a, b in zip(res, ref_res))
        out[prec] = {"ms_per_batch": t * 1e3, "pairs_per_s": n_pair / t,
                     "stage_ms": {"quantize": st[0], "screen": st[1], "rescore_fallback": st[2], "unique": st[3]},
                     "screen_tflops": flops / (st[1] / 1e3) / 1e12,
                     "mfma_roofline_frac": flops / (st[1] / 1e3) / peaks[prec],
                     "max_abs_distance_error": stats["max_abs_distance_error"],
                     "mean_abs_distance_error": stats["mean_abs_distance_error"],
                     "fallback_queries": stats["fallback_queries"],
                     "match_set_equals_fp64": bool(same)}
    # CPU FP64 oracle on the two smallest pairs
    small = np.argsort(sizes[:, 0] * sizes[:, 1])[:2]
    ok = True
    for p in small:
        q, t, _ = mo.match_fp64(sets[2 * p], sets[2 * p + 1], RATIO)
        ok &= bool(np.array_equal(q, ref_res[p][0]) and np.array_equal(t, ref_res[p][1]))
    out["cpu_fp64_oracle_pairs_checked"] = [int(p) for p in small]
    out["cpu_fp64_oracle_equal"] = ok
    out["matches_total"] = int(sum(len(r[0]) for r in ref_res))
    return out


def bench_flight400(args, tracks=True):
    """The 400-photo flight workload: a 400-photo flight (20 x 20 grid, 75 % / 65 % overlap) with resolutions mixed
    50/50 between 2640x1978 and 1320x989.  Photos resident in HBM; SIFT in chunks of one fixed
    shape composition through graph-captured FeatureBatches, --inflight chunks in flight on as
    many streams (launch shapes autotuned first);
    candidate pairs from the GPS grid on the GPU (16 nearest + temporal); guided matching of all
    pairs in one batch (MFMA and plain forms, graph replays), a sample checked against the FP64
    oracle."""
    import torch
    from fieldstitch_amd import synth, features, autotune
    from fieldstitch_amd import pairs as fpairs
    from fieldstitch_amd.matching import PairMatcher
    from oracle import match_oracle as mo
    t0 = time.time()
    fl = synth.flight(seed=4, rows=20, cols=20, W=2640, H=1978, lowres_frac=0.5,
                      workers=max(1, min(args.gen_workers, os.cpu_count() or 1)),
                      progress=lambda d, t: log("flight400: rendered %d / %d photos (%.0f s)" % (d, t, time.time() - t0)))
    n = len(fl.images)
    shapes = [im.shape for im in fl.images]
    log("flight400: %d photos generated in %.1f s" % (n, time.time() - t0))
    dev = [torch.from_numpy(im).cuda() for im in fl.images]
    big = [i for i in range(n) if shapes[i] == (1978, 2640)]
    small = [i for i in range(n) if shapes[i] != (1978, 2640)]
    CB = CS = args.flight_chunk
    chunks, bi, si = [], 0, 0
    while bi < len(big) or si < len(small):
        chunks.append(big[bi:bi + CB] + small[si:si + CS])
        bi += CB
        si += CS
    batches = {}
    # batch sets in flight: chunk c runs on set c % args.inflight and its stream (inputs copied,
    # run and results gathered there, so a set's buffers are reused in stream order)
    streams = [torch.cuda.Stream() for _ in range(args.inflight)]
    def batch_for(idx, slot=0):
        key = (tuple(shapes[i] for i in idx), slot)
        if key not in batches:
            fb = features.FeatureBatch(list(key[0]), nfeat=NFEAT, contrast=CONTRAST, edge=EDGE)
            bufs = [torch.empty(sh, dtype=torch.uint8, device="cuda") for sh in key[0]]
            sels = [fb.selection_device(s_)[:NFEAT] for s_ in range(len(key[0]))]   # views, fixed addresses
            batches[key] = (fb, bufs, sels)
        return batches[key]
    # autotune the launch shapes on the first chunk
    fb0, bufs0, _ = batch_for(chunks[0])
    for b, i in zip(bufs0, chunks[0]):
        b.copy_(dev[i])
    tuned = autotune.tune_features(fb0, bufs0, reps=2)
    u8_all = torch.empty((n, NFEAT, 128), dtype=torch.uint8, device="cuda")
    xy_all = torch.empty((n, NFEAT, 2), dtype=torch.float64, device="cuda")
    cnt_all = torch.empty(n, dtype=torch.int32, device="cuda")

    chunk_idx = [torch.tensor(idx, device="cuda") for idx in chunks]
    main = torch.cuda.current_stream()

    def sift_pass():
        for s_ in streams:
            s_.wait_stream(main)
        for c, idx in enumerate(chunks):
            slot = c % args.inflight
            fb, bufs, sels = batch_for(idx, slot)
            with torch.cuda.stream(streams[slot]):
                for b, i in zip(bufs, idx):
                    b.copy_(dev[i])
                fb.run(bufs, graph=True, stream=streams[slot])
                it = chunk_idx[c]
                u8_all.index_copy_(0, it, fb.u8[:, :NFEAT])
                cnt_all.index_copy_(0, it, torch.clamp(fb.n_out, max=NFEAT))
                # selected (x, y) of every photo of the chunk in one gather (entries past a
                # photo's count are stale and clamped, as the matcher only reads its count)
                sel = torch.stack(sels).clamp_(0, fb.max_kp - 1)
                g = torch.gather(fb.kp[:, :2, :], 2, sel[:, None, :].expand(-1, 2, -1))
                xy_all.index_copy_(0, it, g.transpose(1, 2).double())
        for s_ in streams:
            main.wait_stream(s_)

    sift_pass()                                   # warm-up (captures the graphs)
    for fb, _, _ in batches.values():
        fb.check_capacity()
    torch.cuda.synchronize()
    steps = max(1, args.steps // 5)
    t1 = time.perf_counter()
    for _ in range(steps):
        sift_pass()
    torch.cuda.synchronize()
    t_sift = (time.perf_counter() - t1) / steps
    alg = sum(sift_bytes(*sh) for sh in shapes)
    counts = cnt_all.cpu().numpy()
    # candidate pairs from the GPS grid (photo centres in mosaic units, ground half-extent of
    # the long side = 1320 for both resolutions)
    centres = np.array([(A @ np.array([(sh[1] - 1) / 2.0, (sh[0] - 1) / 2.0, 1.0]))[:2]
                        for A, sh in zip(fl.A, shapes)])
    t2 = time.perf_counter()
    pl = fpairs.select_pairs(centres, np.full(n, 1320.0), 16, True)
    t_pairs = time.perf_counter() - t2
    rng = np.random.default_rng(5)
    Hm = []
    for i, j in pl:
        P = np.linalg.inv(fl.A[j]) @ fl.A[i]
        d = rng.normal(0.0, 3.0, 2)
        P[0, 2] += d[0]
        P[1, 2] += d[1]
        Hm.append(P)
    images = [(u8_all[i, :counts[i]], xy_all[i, :counts[i]], (shapes[i][1], shapes[i][0])) for i in range(n)]
    pm = PairMatcher(images, pl, Hm, radius=RADIUS, kc=0, ratio=RATIO)
    out_m = {"autotune": autotune.tune_matcher(pm, "mfma", reps=3)}
    res = {}
    for form in ("mfma", "plain"):
        pm.run(form, graph=True)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        for _ in range(steps):
            pm.run(form, graph=True)
        torch.cuda.synchronize()
        tm = (time.perf_counter() - t3) / steps
        res[form] = pm.results()
        out_m[form] = {"ms_per_flight": tm * 1e3, "pairs_per_s": len(pl) / tm}
    same_forms = all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
                     for a, b in zip(res["mfma"], res["plain"]))
    sample = rng.choice(len(pl), size=min(24, len(pl)), replace=False)
    ok = 0
    for p in sample:
        i, j = pl[p]
        q, t = mo.guided_fp64_fast(u8_all[i, :counts[i]].cpu().numpy(), xy_all[i, :counts[i]].cpu().numpy(),
                                   u8_all[j, :counts[j]].cpu().numpy(), xy_all[j, :counts[j]].cpu().numpy(),
                                   Hm[p], (shapes[j][1], shapes[j][0]), RADIUS, RATIO)
        ok += int(np.array_equal(q, res["mfma"][p][0]) and np.array_equal(t, res["mfma"][p][1]))
    log("flight400: matching timed (%d pairs), chain next" % len(pl))
    chain = bench_chain(fl, u8_all, xy_all, counts, shapes, pl, reps=3)
    log("flight400: chain %.1f ms" % chain["ms"])
    if not tracks:
        return _flight400_record(args, n, t_sift, alg, counts, tuned, t_pairs, pl, out_m, same_forms, sample, ok,
                                 None, big, small, CB, CS, chain)
    # tracks (recon.py:14-56) from the flight's guided matches (the pipeline feeds RANSAC-
    # verified ones; the guided matches stand in for them here), device pass timed
    from fieldstitch_amd import tracks as trk
    from oracle import tracks_oracle as to
    tpairs = [(i, j, q, t) for (i, j), (q, t) in zip(pl, res["mfma"])]
    toff, tij, tstart, tq, tt = trk.pack_pairs(tpairs, counts)
    dtr = [torch.from_numpy(np.ascontiguousarray(a_)).cuda() for a_ in (toff, tij, tstart, tq, tt)]
    tb = trk.TrackBuilder("cuda")
    tb.run(*dtr)
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    for _ in range(steps):
        g_tid, g_img, g_feat, g_ntr = tb.run(*dtr)
    t_tr = (time.perf_counter() - t4) / steps
    ref_tr = to.build_tracks(tpairs, counts, n)
    tracks_equal = (np.array_equal(g_tid.cpu().numpy(), ref_tr[0]) and np.array_equal(g_img.cpu().numpy(), ref_tr[1])
                    and np.array_equal(g_feat.cpu().numpy(), ref_tr[2]) and g_ntr == ref_tr[3])
    out_tracks = {"matches": int(len(tq)), "observations": int(len(g_tid)), "tracks": int(g_ntr),
                  "ms_per_flight_incl_count_readback": t_tr * 1e3, "equal_to_oracle": bool(tracks_equal)}
    return _flight400_record(args, n, t_sift, alg, counts, tuned, t_pairs, pl, out_m, same_forms, sample, ok,
                             out_tracks, big, small, CB, CS, chain)


def _flight400_record(args, n, t_sift, alg, counts, tuned, t_pairs, pl, out_m, same_forms, sample, ok, out_tracks,
                      big, small, CB, CS, chain):
    steps = max(1, args.steps // 5)
    rec = {
        "metric": "images/s (SIFT) + matched pairs/s vs plain FP32 GPU form, match set = FP64 oracle",
        "value": n / t_sift, "unit": "images/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": t_sift * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (f64 blur accumulation)", "data": "synthetic (seeded field mosaic, 400 photos)",
        "config": {"workload": "400 photos, 20x20 grid 75/65 overlap, 2640x1978 and 1320x989 mixed 50/50; "
                               "SIFT contrast 0.04 edge 10 nfeat 8000 in graph-captured chunks of %d+%d, %d chunks "
                               "in flight; guided matching of GPS-selected pairs" % (CB, CS, args.inflight),
                   "photos": n, "large": len(big), "small": len(small), "pairs": len(pl),
                   "batches_in_flight": args.inflight},
        "sift": {"ms_per_flight": t_sift * 1e3, "algorithmic_bytes": alg,
                 "sift_roofline_frac": alg / t_sift / HBM_PEAK, "keypoints_selected_mean": float(counts.mean()),
                 "autotune": tuned["chosen"]},
        "pair_selection_ms": t_pairs * 1e3,
        "matching": {"pairs": len(pl), **out_m, "speedup_vs_plain_fp32":
                     out_m["plain"]["ms_per_flight"] / out_m["mfma"]["ms_per_flight"],
                     "forms_equal": bool(same_forms), "fp64_oracle_sample": int(len(sample)),
                     "fp64_oracle_equal": int(ok),
                     "compulsory_bytes": 128.0 * float(counts.sum()),
                     "compulsory_hbm_frac": 128.0 * float(counts.sum()) / (out_m["mfma"]["ms_per_flight"] / 1e3)
                                            / HBM_PEAK},
        "match_pairs_chain": chain,
    }
    if out_tracks is not None:
        rec["matching"]["tracks_from_these_matches"] = out_tracks
    return rec


def bench_filter(args):
    """The learned match filter on 400-photo flights (20 x 20 grid, 2640x1978 and 1320x989
    mixed 50/50, nfeat 8000, every overlapping pair guided-matched): trained with Adam (lr 1e-3,
    batch 65536) on the plain pipeline's RANSAC-verified inliers of the first n - 2 flights,
    threshold chosen on the next for 99.7 % inlier recall, evaluated against the 99.5 % target
    on the last.  Reports the pruned fraction and the inlier fraction of what reaches
    verification with and without the network."""
    import torch
    from fieldstitch_amd import filter_mlp as fm
    t0 = time.time()
    n_fl = max(3, args.filter_flights)
    workers = max(1, min(args.gen_workers, os.cpu_count() or 1))
    data = []
    for k, seed in enumerate(range(40, 40 + n_fl)):
        data.append(fm.flight_candidates(seed, workers=workers, rows=20, cols=20, lowres_frac=0.5,
                                         keep_matcher=k == n_fl - 1))
        log("filter: flight %d: %d pairs, %d candidates, %d plain, %d inliers (%.0f s)" %
            (k, data[-1]["n_pairs"], data[-1]["X"].shape[0], data[-1]["n_plain"], data[-1]["n_inliers"],
             time.time() - t0))
    n_tr = n_fl - 2
    Xtr = torch.cat([d["X"] for d in data[:n_tr]])
    ytr = torch.cat([d["y"] for d in data[:n_tr]])
    for d in data[:n_tr]:
        d["X"] = d["y"] = None            # the training copy is all that is needed
    net = fm.MatchFilterMLP(seed=1, lr=1e-3)
    batch = args.filter_batch
    epochs = args.filter_epochs
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    losses = fm.train(net, Xtr, ytr, epochs=epochs, batch=batch)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    steps = epochs * (Xtr.shape[0] // batch)
    pv = net.predict(data[n_tr]["X"]).cpu().numpy()
    # the threshold keeps 99.7 % of the validation flight's inliers: margin for the 99.5 % target
    tau = fm.threshold_for_recall(pv, data[n_tr]["label"], 0.997)
    Xte = data[n_tr + 1]["X"]
    net.predict(Xte[:65536])
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    pt = net.predict(Xte)
    torch.cuda.synchronize()
    t_inf = time.perf_counter() - t2
    ev = fm.evaluate(pt.cpu().numpy(), tau, data[n_tr + 1])
    stage = filter_stage_times(data[n_tr + 1], net, tau)
    flops_fwd = 2.0 * (288 * 256 + 256 * 256 + 256)
    return {
        "metric": "learned match filter: verified-inlier recall vs the plain set (target >= 99.5 %)",
        "value": ev["inlier_recall"], "unit": "recall", "n_gpus": 1, "higher_is_better": True,
        "data": "synthetic (%d seeded 400-photo flights, 2640x1978 / 1320x989, RANSAC labels from the plain "
                "pipeline)" % n_fl,
        "config": {"workload": "MLP 264-256-256-1 BF16, Adam lr 1e-3, batch %d, %d epochs; train %d flights, "
                               "threshold (99.7 %% recall) on 1, test on 1" % (batch, epochs, n_tr),
                   "pairs_per_flight": [d["n_pairs"] for d in data],
                   "train_candidates": int(Xtr.shape[0]), "test_candidates": int(Xte.shape[0])},
        "train": {"seconds": t_train, "steps": steps, "samples_per_s": steps * batch / t_train,
                  "tflops": 3 * flops_fwd * steps * batch / t_train / 1e12, "epoch_loss": losses},
        "threshold": tau,
        "inference": {"candidates_per_s": Xte.shape[0] / t_inf, "tflops": flops_fwd * Xte.shape[0] / t_inf / 1e12},
        "test": ev,
        "match_and_verify": stage,
    }


def filter_stage_times(d, net, tau, reps=5):
    """Matching + verification of the test flight's pairs with and without the pruning stage:
    guided matching (PairMatcher.run, or run_filtered with the network and tau) through the host
    match lists, then homography RANSAC (a seed per pair, positions gathered on the device, as
    match_pairs verifies) over every pair's list with the labelling threshold and hypotheses.
    Median of `reps` runs each; also the matches reaching verification and the inliers found."""
    import torch
    from fieldstitch_amd import match_flow
    pm, xy = d["matcher"], d["xy"]
    n_hyp, thr, seed = d["ransac"]
    sz = [torch.as_tensor(np.asarray(v, np.float32)).cuda() for v in d["sizes"]]
    an = [torch.as_tensor(np.asarray(v, np.float32)).cuda() for v in d["angles"]]
    pos = match_flow.Positions({i: x for i, x in enumerate(xy)})     # every photo's positions on the device
    seeds = [match_flow.pair_seed(i, j, 0x1) for i, j in pm.pairs]

    def once(filtered):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if filtered:
            pm.run_filtered("mfma", net, tau, sz, an)
        else:
            pm.run("mfma")
        res = pm.results()
        t1 = time.perf_counter()
        _, _, cnt = match_flow.ransac_rows("homography", pos, [pos.rows(i, q) for (i, _), (q, _) in zip(pm.pairs, res)],
                                           [pos.rows(j, t) for (_, j), (_, t) in zip(pm.pairs, res)],
                                           [thr] * len(res), seeds, n_hyp)
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1, sum(len(q) for q, _ in res), int(np.sum(cnt))

    out = {}
    for name, filtered in (("plain", False), ("filtered", True)):
        once(filtered)
        runs = [once(filtered) for _ in range(reps)]
        m = float(np.median([r[0] for r in runs]))
        v = float(np.median([r[1] for r in runs]))
        out[name] = {"match_ms": 1e3 * m, "verify_ms": 1e3 * v, "total_ms": 1e3 * (m + v),
                     "matches_to_verify": runs[-1][2], "inliers": runs[-1][3]}
    out["pairs"] = len(pm.pairs)
    out["queries_scored"], out["queries_kept"] = pm.filter_counts()
    out["speedup"] = out["plain"]["total_ms"] / out["filtered"]["total_ms"]
    return out


def pcie_inclusive_steps(fbs, streams, images, steps):
    """SIFT steps whose photos are uploaded from pinned host memory every step: a copy stream
    fills one of len(fbs) + 1 device image sets while earlier steps compute (events order each
    set's reuse after the step that read it and each step after its upload).  Two compute
    streams, the copy stream and the default stream fit the 4 hardware queues a process gets
    here (GPU_MAX_HW_QUEUES), so the copies do not queue behind a compute stream.  Returns
    images/s and ms per step including the uploads, and the upload bandwidth alone."""
    import torch
    nb = len(fbs) + 1
    hosts = [torch.from_numpy(im).pin_memory() for im in images]
    sets = [[torch.empty(h.shape, dtype=torch.uint8, device="cuda") for h in hosts] for _ in range(nb)]
    cs = torch.cuda.Stream()
    done_copy = [torch.cuda.Event() for _ in range(nb)]
    done_use = [None] * nb
    main = torch.cuda.current_stream()

    def run(k_steps, upload=True):
        cs.wait_stream(main)
        for s_ in streams:
            s_.wait_stream(main)
        for k in range(k_steps):
            b = k % nb
            with torch.cuda.stream(cs):
                if done_use[b] is not None:
                    cs.wait_event(done_use[b])
                if upload:
                    for d, h in zip(sets[b], hosts):
                        d.copy_(h, non_blocking=True)
                done_copy[b].record(cs)
            st = streams[k % len(fbs)]
            st.wait_event(done_copy[b])
            fbs[k % len(fbs)].run(sets[b], stream=st)
            ev = torch.cuda.Event()
            ev.record(st)
            done_use[b] = ev
        for s_ in streams + [cs]:
            main.wait_stream(s_)

    run(nb)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(steps)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    # the same schedule without the uploads (photos already in the device sets)
    t0 = time.perf_counter()
    run(steps, upload=False)
    torch.cuda.synchronize()
    t_res = (time.perf_counter() - t0) / steps
    # upload alone: the same copies back to back on the copy stream
    nbytes = sum(h.numel() for h in hosts)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    with torch.cuda.stream(cs):
        for k in range(steps):
            for d, h in zip(sets[k % nb], hosts):
                d.copy_(h, non_blocking=True)
    torch.cuda.synchronize()
    t_up = (time.perf_counter() - t1) / steps
    return {"images_per_s": len(images) / t, "ms_per_step": t * 1e3, "resident_ms_per_step": t_res * 1e3,
            "upload_ms_per_step": t_up * 1e3,
            "upload_gb_per_s": nbytes / t_up / 1e9, "bytes_per_step": nbytes,
            "batches_in_flight": len(fbs),
            "note": "photos copied from pinned host memory on a copy stream each step, overlapped with the "
                    "SIFT of earlier steps; value above has them resident in HBM"}


def copy_probe(mib):
    """Three device-to-device copies of `mib` MiB: a kernel with known HBM traffic (read mib,
    write mib per dispatch) whose FETCH_SIZE / WRITE_SIZE readings scale the counters of the
    SIFT kernels to bytes."""
    import torch
    a = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": a.numel() * 4, "dispatches": 3}), flush=True)


def pyramid_launches(shapes):
    """Blur launches of one step: per group of equal shapes, 1 base + 5 per octave."""
    from fieldstitch_amd import gsift
    return sum(1 + 5 * len(gsift.plan(H, W)) for H, W in set(shapes))


def bench_tracks(args):
    """Track building (recon.py:14-56) from the verified match lists of a 400-photo survey
    (20 x 20 raster, 8000 features per photo, every two photos within one raster step a pair:
    1482 pairs, 3.6 M matches of which 2 % wrong; seeded, oracle.tracks_oracle
    .synthetic_flight_matches).  Match lists resident on the GPU; the timed region is the
    libfsx pass (union-find, photo-clash check, numbering, sorted output) without the count
    readback.  Checked equal to the oracle; the oracle (numpy + scipy connected_components, one
    core) is the CPU baseline on the same input."""
    import ctypes as C
    import torch
    from fieldstitch_amd import _lib, tracks
    from oracle import tracks_oracle as to
    t0 = time.time()
    pairs, nfeat = to.synthetic_flight_matches()
    offsets, ij, start, q, t = tracks.pack_pairs(pairs, nfeat)
    n_edges, n_nodes = len(q), int(offsets[-1])
    log("tracks: %d pairs, %d matches, generated in %.1f s" % (len(pairs), n_edges, time.time() - t0))
    d = [torch.from_numpy(np.ascontiguousarray(a_)).cuda() for a_ in (offsets, ij, start, q, t)]
    tb = tracks.TrackBuilder("cuda")
    tid, img, feat, n_tr = tb.run(*d)
    res = torch.empty(4, dtype=torch.int64, device="cuda")
    vp = lambda x: C.c_void_p(x.data_ptr())
    o = tb._out

    def launch():
        _lib.call("fsx_build_tracks", len(nfeat), vp(d[0]), n_nodes, len(pairs), vp(d[1]), vp(d[2]), vp(d[3]),
                  vp(d[4]), n_edges, vp(tb._ws), tb._ws_bytes, o.shape[1], vp(o[0]), vp(o[1]), vp(o[2]), vp(res),
                  C.c_void_p(_lib.stream_handle(None)))
    for _ in range(args.warmup):
        launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        launch()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    r = res.cpu().tolist()
    g = (tid.cpu().numpy(), img.cpu().numpy(), feat.cpu().numpy(), n_tr)
    same_again = int(r[0]) == len(g[0]) and int(r[1]) == n_tr and bool(torch.equal(o[0, :r[0]], tid))
    t1 = time.time()
    ref = to.build_tracks(pairs, nfeat, len(nfeat))
    t_cpu = time.time() - t1
    equal = all(np.array_equal(x, y) for x, y in zip(g[:3], ref[:3])) and g[3] == ref[3]
    # drop-in form: host match lists in, numpy arrays out (packing, PCIe both ways)
    t2 = time.time()
    tracks.build_tracks(pairs, nfeat, len(nfeat))
    t_host = time.time() - t2
    n_obs = len(g[0])
    compulsory = 8.0 * n_edges + 24.0 * n_obs + 8.0 * n_nodes
    return {
        "metric": "matches/s into tracks (build_tracks), equal to the oracle",
        "value": n_edges / (ms / 1e3), "unit": "matches/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic (seeded 400-photo survey match lists)",
        "config": {"workload": "build_tracks, 400 photos x 8000 features, %d pairs, %d matches (2 %% wrong)"
                               % (len(pairs), n_edges)},
        "n_observations": n_obs, "n_tracks": n_tr, "equal_to_oracle": bool(equal and same_again),
        "roofline": {"bound": "hbm", "achieved": compulsory / (ms / 1e3) / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": compulsory / (ms / 1e3) / HBM_PEAK, "traffic": None,
                     "note": "compulsory bytes: matches read once (8 B), observations written once (3 x int64), "
                             "one write + read of the per-feature parent (8 B/feature); the union-find's "
                             "random parent accesses are served mostly from L2"},
        "drop_in_ms": t_host * 1e3,
        "cpu_baseline": {"value": n_edges / t_cpu, "unit": "matches/s", "cores": 1, "kind": "port",
                         "sample": "the same match lists, oracle.tracks_oracle.build_tracks (numpy + scipy "
                                   "connected_components), %.2f s" % t_cpu},
    }


def bench_clahe(args):
    """CLAHE (clip 2.0, 8 x 8 tiles; features.py:60-71) of the 8-photo 2640 x 1978 batch,
    photos resident on the GPU: one LUT launch and one apply launch per batch.  Algorithmic
    bytes 3 per pixel (read for the histograms, read + write for the mapping).  Checked equal to
    the CPU restatement of OpenCV's algorithm on one photo, which is also the CPU baseline."""
    import torch
    from fieldstitch_amd import clahe
    from oracle import clahe_oracle as co
    rng = np.random.default_rng(0)
    n, H, W = 8, 1978, 2640
    base = (rng.normal(110, 35, (H, W)) + 40 * np.sin(np.arange(W) / 53.0)[None, :]).clip(0, 255).astype(np.uint8)
    imgs = np.stack([np.roll(base, 97 * k, axis=1) for k in range(n)])
    d = torch.from_numpy(imgs).cuda()
    out = torch.empty_like(d)
    c = clahe.createCLAHE(2.0, (8, 8))
    for _ in range(args.warmup):
        c.apply(d, out=out)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        c.apply(d, out=out)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    t0 = time.time()
    ref = co.apply(imgs[0])
    t_cpu = time.time() - t0
    equal = bool(np.array_equal(out[0].cpu().numpy(), ref))
    alg = 3.0 * n * H * W
    return {
        "metric": "photos/s through CLAHE (clip 2.0, 8 x 8 tiles)", "value": n / (ms / 1e3), "unit": "photos/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8 / f32", "data": "synthetic",
        "config": {"workload": "CLAHE, 8 photos 2640x1978 uint8"},
        "equal_to_oracle_photo0": equal,
        "roofline": {"bound": "hbm", "achieved": alg / (ms / 1e3) / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": alg / (ms / 1e3) / HBM_PEAK, "traffic": None,
                     "note": "3 B/px: histogram read, mapping read + write; both launches in the time"},
        "cpu_baseline": {"value": 1.0 / t_cpu, "unit": "photos/s", "cores": 1, "kind": "port",
                         "sample": "one photo, oracle.clahe_oracle.apply (numpy), %.2f s" % t_cpu},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-flight400", action="store_true", help="skip the 400-photo flight section")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic_k_blur_8x2640x1978.json"))
    ap.add_argument("--gen-workers", type=int, default=16,
                    help="processes rendering the synthetic photos (1 under a profiler)")
    ap.add_argument("--no-precision", action="store_true", help="skip the float-descriptor precision section")
    ap.add_argument("--filter", action="store_true",
                    help="only train and evaluate the learned match filter (400-photo synthetic flights)")
    ap.add_argument("--filter-epochs", type=int, default=8)
    ap.add_argument("--filter-flights", type=int, default=7, help="400-photo flights: n-2 train, 1 threshold, 1 test")
    ap.add_argument("--filter-batch", type=int, default=65536)
    ap.add_argument("--clahe", action="store_true", help="CLAHE of the 8-photo batch instead of the headline")
    ap.add_argument("--tracks", action="store_true",
                    help="track building from a 400-photo survey's match lists instead of the headline")
    ap.add_argument("--flight400", action="store_true",
                    help="only run the 400-photo mixed-resolution flight (graph-captured SIFT chunks + guided matching)")
    ap.add_argument("--precision-pairs", type=int, default=200)
    ap.add_argument("--flight-chunk", type=int, default=25,
                    help="--flight400: photos of each resolution per feature batch (chunk)")
    ap.add_argument("--inflight", type=int, default=3,
                    help="feature batches in flight (workspaces and streams) in the timed SIFT steps")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 3 plain device copies of MIB MiB (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.inflight < 1 or args.flight_chunk < 1:
        ap.error("--inflight and --flight-chunk must be >= 1")
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.flight400:
        print(json.dumps(bench_flight400(args)), flush=True)
        return
    if args.filter:
        print(json.dumps(bench_filter(args)), flush=True)
        return
    if args.tracks:
        print(json.dumps(bench_tracks(args)), flush=True)
        return
    if args.clahe:
        print(json.dumps(bench_clahe(args)), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this library has no multi-GPU path; photos would shard by flight-grid block",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from fieldstitch_amd import features

    t0 = time.time()
    fl = make_flight(workers=args.gen_workers)
    log("flight: %d photos %s, %d pairs, generated in %.1f s" % (len(fl.images), fl.images[0].shape,
                                                                  len(fl.pairs), time.time() - t0))
    dev = [torch.from_numpy(im).cuda() for im in fl.images]
    shapes = [im.shape for im in fl.images]
    # batches in flight (--inflight, 3 by default): consecutive steps rotate over that many
    # workspaces on as many (default-priority) streams, so one step's latency-bound stages
    # (small octaves, sort, spread, descriptors) overlap the next step's memory-bound pyramid
    fbs = [features.FeatureBatch(shapes, nfeat=NFEAT, contrast=CONTRAST, edge=EDGE) for _ in range(args.inflight)]
    fb = fbs[0]
    streams = [torch.cuda.Stream() for _ in range(args.inflight)]
    main = torch.cuda.current_stream()

    def sift_steps(k_steps):
        for s_ in streams:
            s_.wait_stream(main)
        for k in range(k_steps):
            fbs[k % args.inflight].run(dev, stream=streams[k % args.inflight])
        for s_ in streams:
            main.wait_stream(s_)

    torch.cuda.synchronize()
    sift_steps(max(args.warmup, args.inflight))
    for f_ in fbs:
        f_.check_capacity()
    torch.cuda.synchronize()

    # timed region: K steps of the SIFT front end on all photos, as a caller runs it (no stage
    # events)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    sift_steps(args.steps)
    torch.cuda.synchronize()
    t_sift = (time.perf_counter() - t_start) / args.steps
    for f_ in fbs:
        f_.check_capacity()
    # every workspace holds the same features (identical inputs): the matcher reads fbs[0]
    for i in range(len(shapes)):
        a_, b_ = fbs[0].results(i), fbs[-1].results(i)
        if not (np.array_equal(a_[0], b_[0]) and np.array_equal(a_[1], b_[1])):
            raise RuntimeError("the in-flight batches disagree on photo %d" % i)
    # the same steps with the photos arriving from pinned host memory over PCIe each step
    # (reported beside value, which has them resident in HBM)
    pcie = pcie_inclusive_steps(fbs[:2], streams[:2], fl.images, args.steps)
    # stage breakdown (and the pyramid time of the roofline) from K more steps with per-stage
    # events
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(8)] for _ in range(args.steps)]
    for k in range(args.steps):
        fb.run(dev, events=evs[k])
    torch.cuda.synchronize()
    stages = {}
    for k in range(args.steps):
        for name, v in features.FeatureBatch.stage_times(evs[k]).items():
            stages[name] = stages.get(name, 0.0) + v / args.steps
    n_det, n_sel = fb.counts()
    n_img = len(fl.images)
    pyr_b = sum(pyramid_bytes(*s) for s in shapes)
    sift_b = sum(sift_bytes(*s) for s in shapes)
    pyr_s = stages["pyramid"] / 1e3
    n_launch = pyramid_launches(shapes)
    achieved = pyr_b / pyr_s
    traffic = None
    if os.path.exists(args.traffic):
        with open(args.traffic) as f:
            traffic = json.load(f).get("bytes_per_launch")
    res = {
        "metric": "images/s (SIFT) + matched pairs/s vs plain FP32 GPU form, match set = FP64 oracle",
        "value": n_img / t_sift,
        "unit": "images/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_sift * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 blur accumulation)",
        "data": "synthetic (seeded field mosaic, 8 photos)",
        "config": {"workload": "8 photos 2640x1978 u8, 2x4 flight grid 75/65 overlap; SIFT contrast 0.04 edge 10, "
                               "nfeat 8000 (detect 17600), RootSIFT u8",
                   "photos": n_img, "nfeat": NFEAT, "pairs": len(fl.pairs), "batches_in_flight": args.inflight},
        "sift": {"stage_ms": {k: round(v, 4) for k, v in stages.items()},
                 "stage_ms_note": "separate runs with per-stage events, one batch in flight (their sum is "
                                  "serial_ms_per_step; ms_per_step has batches_in_flight of them)",
                 "serial_ms_per_step": round(sum(stages.values()), 4),
                 "keypoints_detected_mean": float(n_det.mean()), "keypoints_selected_mean": float(n_sel.mean()),
                 "algorithmic_bytes_per_step": sift_b,
                 "sift_roofline_frac": sift_b / t_sift / HBM_PEAK},
        "pcie_inclusive": pcie,
        "roofline": {"kernel": "k_blur (pyramid: base + 5 layers x octaves, all photos)", "bound": "hbm",
                     "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "launches_per_step": n_launch, "algorithmic_bytes_per_launch": pyr_b / n_launch,
                     "avg_launch_us": stages["pyramid"] * 1e3 / n_launch,
                     "algorithmic_bytes_per_step": pyr_b, "time_ms_per_step": stages["pyramid"],
                     "note": "k_blur is VALU-issue-bound: float64 taps in the reference's order, ~86 % VALU busy by "
                             "counters, ~2.3 ms of VALU issue per step at 100 % (profiles/pmc_k_blur_8x2640x1978.md)"},
    }
    log("sift timed: %.3f ms per step; matching next" % (t_sift * 1e3))
    res["matching"], host, pairs, Hm = bench_matching(fb, fl, shapes, args)
    # match_pair's whole chain on the same photos' device-resident features
    xy_dev = [torch.from_numpy(h[1]).cuda() for h in host]
    res["matching"]["match_pairs_chain"] = bench_chain(fl, [fb.u8[i, :len(host[i][0])] for i in range(n_img)],
                                                        xy_dev, [len(h[0]) for h in host], shapes, pairs)
    if not args.no_precision:
        log("precision section")
        res["precision"] = bench_precision(args)
    if not args.no_flight400:
        log("flight400 section")
        del fbs, fb
        torch.cuda.empty_cache()
        res["flight400"] = bench_flight400(args, tracks=False)
        res["config"]["flight400_workload"] = res["flight400"]["config"]["workload"]
    if not args.no_cpu_baseline:
        log("cpu baseline: pool of %d workers" % max(1, (os.cpu_count() or 1) - 1))
        res["cpu_baseline"] = cpu_baseline(fl.images, host, pairs, Hm, shapes)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
