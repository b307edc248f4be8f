#!/usr/bin/env python3
"""Benchmark of the LaueMatching indexing hot path on one MI355X.

Workload (default): the plain GPU drop-in at full size -- 100M uniform random
orientations (float64 3x3, seeded), FCC a = 0.36 nm with every FCC-allowed
reflection |h|,|k|,|l| <= 8 (1240 rows), 5-30 keV, 2048 x 2048 detector at
0.2 mm pitch 500 mm from the sample centred at 2theta = 90 deg,
MaxNrLaueSpots 64, one uint8 image with 5 grains, MinNrSpots 5,
MinIntensity 1000 (the reference defaults).

A step scores one image against the resident table: the compare kernel, hit
compaction, the device-to-host copy of the compact list and its ascending
sort.  The spot table (forward simulation, untimed setup) and the image are
resident in HBM before the timed region.  Prints ONE JSON line on stdout;
progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(db, u8, im_scale, min_sp, min_int, sample, budget_s=12.0):
    """The CPU oracle (C + OpenMP restatement of the reference DoFwd=0 loop)
    timed on all host cores over a prefix of the same table and image."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob

    table = db.export_table(0, sample)
    cores = len(os.sched_getaffinity(0))
    passes, t_cpu = 0, 0.0
    while t_cpu < budget_s and passes < 2000:
        t0 = time.perf_counter()
        ob.compare(table, db.max_spots, u8, db.geom.nrPxX, im_scale, min_sp, min_int,
                   threads=cores)
        t_cpu += time.perf_counter() - t0
        passes += 1
    rate = sample * passes / t_cpu
    return {"value": rate, "unit": "orient×image pairs/s", "cores": cores, "kind": "port",
            "sample": f"first {sample} orientations of the same table and image, "
                      f"{passes} passes of the oracle compare (OpenMP, {cores} threads), "
                      f"{t_cpu:.1f} s"}


def cpu_forward_baseline(O, hk, p, rows):
    """The CPU oracle's forward simulation (C + OpenMP restatement of the
    reference DoFwd=1 loop, every pair projected) on all host cores over the
    first `rows` orientations of the same DB."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob

    cores = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    ob.forward(O[:rows], hk, ob.geom(p), p.MaxNrLaueSpots, threads=cores)
    dt = time.perf_counter() - t0
    out = {"pairs_per_s": rows * len(hk) / dt, "cores": cores, "kind": "port",
           "sample": f"first {rows} orientations x {len(hk)} hkls, {dt:.1f} s"}
    log(f"cpu forward simulation {out['pairs_per_s']:.3e} pairs/s ({cores} cores, {rows} rows)")
    return out


def oracle_parity(db, cases, chunk=10_000_000):
    """Every GPU result of the run against the CPU oracle over the WHOLE
    table: the table is exported in chunks of `chunk` rows (ForwardFile
    records), each chunk scored by the oracle for every case, and each case's
    GPU hits in that row range compared bit for bit (ids, nspots, float64
    totInt and score, nrResults).  cases: (name, u8, imScale, MinNrSpots,
    MinIntensity, gpu Hits)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob

    cores = len(os.sched_getaffinity(0))
    n = db.n
    ok = {c[0]: True for c in cases}
    nres = {c[0]: 0 for c in cases}
    t0 = time.perf_counter()
    for first in range(0, n, chunk):
        cnt = min(chunk, n - first)
        table = db.export_table(first, cnt)
        for name, u8, s, ms, mi, h in cases:
            matched, ns, tot, nr = ob.compare(table, db.max_spots, u8, db.geom.nrPxX, s, ms, mi,
                                              threads=cores)
            rows, mA = ob.collect(matched)
            ri = rows.astype(np.int64)
            sel = (h.ids >= first) & (h.ids < first + cnt)
            ok[name] &= bool(np.array_equal(h.ids[sel], rows + np.uint64(first))
                             and np.array_equal(h.score[sel], mA)
                             and np.array_equal(h.nspots[sel], ns[ri])
                             and np.array_equal(h.totint[sel], tot[ri]))
            nres[name] += int(nr)
        del table
    for name, u8, s, ms, mi, h in cases:
        ok[name] &= nres[name] == int(h.nr_results)
    out = {"rows": int(n), "cases": len(cases), "all_equal": all(ok.values()),
           "per_case": ok, "seconds": time.perf_counter() - t0}
    log(f"oracle parity over all {n} rows: {sum(ok.values())}/{len(ok)} cases bit-exact "
        f"({out['seconds']:.1f} s)")
    return out


def _batch_image(args):
    """One synthetic batch image: seeded grain count U{1..20}, grains drawn
    from the DB rows (regenerated from the DB seed)."""
    seed, n, db_seed, i = args
    from lauematching_amd import synth

    rng = np.random.default_rng([seed, i])
    k = int(rng.integers(1, 21))
    rows = np.sort(rng.choice(n, k, replace=False))
    p = synth.baseline_params()
    img = synth.synth_image(p, synth.fcc_hkls(8, "box"), synth.orientation_rows(db_seed, rows),
                            seed=int(rng.integers(1 << 30)))
    return img, rows


def make_batch_images(count, n, db_seed, seed=1234, workers=16):
    """Generate batch images in worker processes (before any GPU use)."""
    from concurrent.futures import ProcessPoolExecutor

    jobs = [(seed, n, db_seed, i) for i in range(count)]
    with ProcessPoolExecutor(min(workers, count)) as ex:
        return list(ex.map(_batch_image, jobs))


def run_batches(db, batch_imgs, widths, p, steps, warmup, parity_cases, spots):
    """Scores batches of B images in one pass over the table; returns per-B
    rates, checks each batched image against scoring it alone and adds three
    images per width to the oracle parity cases."""
    import torch
    from lauematching_amd import api

    u8s, scales = [], []
    for img, _ in batch_imgs:
        u, s = api.quantize_u8(img)
        u8s.append(u)
        scales.append(s)
    out = []
    for B in widths:
        dev = torch.from_numpy(np.stack(u8s[:B])).to("cuda:0")
        torch.cuda.synchronize()
        for _ in range(warmup):
            hs = db.score_batch_dev(dev.data_ptr(), B, scales[:B], p.MinNrSpots, p.MinIntensity)
        torch.cuda.synchronize()
        kms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            hs = db.score_batch_dev(dev.data_ptr(), B, scales[:B], p.MinNrSpots, p.MinIntensity)
            kms.append(hs[0].kernel_ms)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / steps * 1e3
        ok = True
        for i in sorted({0, B // 2, B - 1}):
            one = db.score_dev(dev[i].data_ptr(), scales[i], p.MinNrSpots, p.MinIntensity)
            ok &= (np.array_equal(one.ids, hs[i].ids) and np.array_equal(one.score, hs[i].score)
                   and one.nr_results == hs[i].nr_results)
        found = sum(set(rows.tolist()) <= set(hs[i].ids.tolist())
                    for i, (_, rows) in enumerate(batch_imgs[:B]))
        for i in sorted({0, B // 2, B - 1}):
            parity_cases.append((f"batch{B}_img{i}", u8s[i], scales[i], p.MinNrSpots,
                                 p.MinIntensity, hs[i]))
        kms_m = float(np.mean(kms))
        # algorithmic bytes: the table once, the B images once, 16 B per hit;
        # the walks over the orientation-major table fetch one 128-byte line
        # of the interleaved images per spot (the line-fetch bound)
        alg = 2.0 * db.n + 4.0 * spots + B * float(u8s[0].size) + 16.0 * sum(
            int(h.nr_results) for h in hs)
        lines = 128.0 * spots
        out.append({"images": B, "ms_per_batch": ms, "kernel_ms": kms_m,
                    "pairs_per_s": db.n * B / (ms / 1e3),
                    "algorithmic_GBps": alg / (kms_m / 1e3) / 1e9,
                    "algorithmic_frac_of_hbm": alg / (kms_m / 1e3) / HBM_PEAK,
                    "line_fetch_GBps": lines / (kms_m / 1e3) / 1e9,
                    "line_fetch_frac_of_hbm": lines / (kms_m / 1e3) / HBM_PEAK,
                    "line_fetch_floor_ms": lines / HBM_PEAK * 1e3,
                    "bit_exact_vs_single": bool(ok),
                    "images_with_all_grains_recovered": int(found),
                    "hits_per_image": [int(len(h.ids)) for h in hs]})
        log(f"batch B={B}: {ms:.2f} ms/batch, {out[-1]['pairs_per_s']:.3e} pairs/s, "
            f"{out[-1]['algorithmic_GBps']:.0f} GB/s algorithmic, line fetches "
            f"{out[-1]['line_fetch_GBps']:.0f} GB/s, exact={ok}, grains recovered in {found}/{B}")
        del dev
    return out


HEX_LAT = (0.295, 0.295, 0.468165, 90.0, 90.0, 120.0)


def serve_params(crystal, det):
    from lauematching_amd import synth

    over = {}
    if det == 1024:
        over.update(NrPxX=1024, NrPxY=1024, PxX=0.0004, PxY=0.0004)
    if crystal == "hex":
        over.update(LatticeParameter=HEX_LAT, SpaceGroup=194)
    return synth.baseline_params(**over)


def _serve_image(args):
    crystal, det, n, db_seed, i = args
    from lauematching_amd import synth

    p = serve_params(crystal, det)
    hk = synth.fcc_hkls(8, "box") if crystal == "fcc" else synth.hcp_hkls(6)
    rng = np.random.default_rng([11, i, det, crystal == "hex"])
    rows = rng.choice(n, int(rng.integers(1, 11)), replace=False)
    ops = None if crystal == "fcc" else synth.hexagonal_ops()
    img = synth.synth_image(p, hk, synth.orientation_rows(db_seed, rows, ops),
                            seed=int(rng.integers(1 << 30)))
    return crystal, det, img


def make_serve_pool(combos, per, workers=16):
    """Images per (crystal, detector, DB size, DB seed), made in worker
    processes before anything touches the GPU."""
    from concurrent.futures import ProcessPoolExecutor

    jobs = [(c, d, n, s, i) for (c, d, n, s) in combos for i in range(per)]
    with ProcessPoolExecutor(min(workers, len(jobs))) as ex:
        return list(ex.map(_serve_image, jobs))


def poisson_run(srv, reqs, rate, duration, seed):
    """Submit requests at Poisson arrival times (rate per second) from one
    thread while another collects results.  A request's latency runs from its
    scheduled arrival to its result: the server's submit -> result time plus
    any delay of the submit call itself (admission slots full, submitter
    behind schedule)."""
    import queue
    import threading

    srv.reset_stats()
    rng = np.random.default_rng(seed)
    nreq = max(12 if rate < 5 else 20, int(rate * duration))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    picks = rng.integers(0, len(reqs), nreq)
    tickets = queue.Queue()
    lat = np.zeros(nreq)

    def collect():
        for _ in range(nreq):
            k, lag, t = tickets.get()
            _, ms = srv.wait(t)
            lat[k] = lag * 1e3 + ms

    th = threading.Thread(target=collect)
    th.start()
    t0 = time.perf_counter()
    for k in range(nreq):
        dt = t0 + arrivals[k] - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        lag = max(0.0, time.perf_counter() - (t0 + arrivals[k]))
        db_id, u8, s = reqs[picks[k]]
        tickets.put((k, lag, srv.submit(db_id, u8, s, 5, 1000.0)))
    th.join()
    wall = time.perf_counter() - t0
    st = srv.stats()
    return {"rate": rate, "requests": nreq, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()),
            "images_per_s": nreq / wall, "mean_batch": st["mean_batch"]}


def closed_loop(srv, reqs, count, window=128, clients=4):
    """Sustained throughput: `clients` submitting threads keep `window`
    requests outstanding (a new one as soon as one completes), `count`
    requests in all; one thread collects results in completion order (polling
    the outstanding tickets), so a slow request does not hold back the slots
    of those that finish after it.  Reports the rate over the whole loop and
    over the middle half of the completions (the sustained rate, without the
    fill and drain)."""
    import collections
    import threading

    srv.reset_stats()
    tickets = collections.deque()
    lock = threading.Condition()
    slots = threading.Semaphore(window)
    nxt = [0]

    def submit():
        while True:
            with lock:
                k = nxt[0]
                if k >= count:
                    return
                nxt[0] += 1
            slots.acquire()
            db_id, u8, s = reqs[k % len(reqs)]
            t = srv.submit(db_id, u8, s, 5, 1000.0)
            with lock:
                tickets.append(t)
                lock.notify_all()

    finish = []

    def collect():
        pending = []
        while len(finish) < count:
            with lock:
                pending.extend(tickets)
                tickets.clear()
            still = []
            for t in pending:
                if srv.poll(t) is None:
                    still.append(t)
                else:
                    finish.append(time.perf_counter())
                    slots.release()
            if len(still) == len(pending):
                time.sleep(0.0005)
            pending = still

    ths = [threading.Thread(target=submit) for _ in range(clients)]
    ths += [threading.Thread(target=collect)]
    t0 = time.perf_counter()
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    dt = time.perf_counter() - t0
    st = srv.stats()
    q1, q3 = finish[count // 4], finish[(3 * count) // 4 - 1]
    return {"images": count, "window": window, "clients": clients, "images_per_s": count / dt,
            "sustained_images_per_s": (3 * count // 4 - count // 4 - 1) / max(q3 - q1, 1e-9),
            "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_batch": st["mean_batch"]}


def run_serving(ctx, dbs, reqs, rates, duration, nmax):
    """Continuous batching versus a plain server (one image per launch, one
    stream: the reference daemon's serialised GPU gate)."""
    from lauematching_amd.serve import Server

    res = {}
    for name, kw in [("batched", dict(max_batch=256, workers=2, use_graphs=True)),
                     ("plain", dict(max_batch=1, workers=1, use_graphs=False))]:
        srv = Server(ctx, **kw)
        ids = [srv.add_db(db) for db in dbs]
        rq = [(ids[i], u8, s) for i, u8, s in reqs]
        # warm-up (buffers, captured launch sequences of each width), not measured
        closed_loop(srv, rq, 4 * kw["max_batch"] + 16, window=2 * kw["max_batch"])
        for t in [srv.submit(*rq[k % len(rq)], 5, 1000.0) for k in range(3)]:
            srv.wait(t)
        runs = [poisson_run(srv, rq, r, duration, 100 + j) for j, r in enumerate(rates)]
        # closed loop: two full batches outstanding for one DB, one per DB
        # when there are more (requests split over the DBs; the plain server
        # takes one image per launch whatever the window)
        mb = kw["max_batch"]
        cl = closed_loop(srv, rq, nmax, window=max(2 * mb, mb * len(dbs)) if mb > 1 else 128)
        st = srv.stats()
        srv.close()
        res[name] = {"poisson": runs, "closed_loop": cl, "graph_launches": st["graph_launches"]}
        log(f"serving {name}: " + ", ".join(
            f"{r['rate']}/s p50 {r['p50_ms']:.1f} p99 {r['p99_ms']:.1f} ms" for r in runs) +
            f"; closed loop {cl['images_per_s']:.0f} img/s, sustained "
            f"{cl['sustained_images_per_s']:.0f} img/s (batch {cl['mean_batch']:.1f})")
    return res


def merge_check(ctx, O):
    """Merge of duplicate candidates (the hit list's consumer): the device pair
    tests at the reference's 50000-candidate cap, and the host merge (16
    threads) beside it on a 5000-candidate sample (its O(n^2) cost makes the
    full cap about 100x the sample); both produce the same clusters."""
    from scipy.spatial.transform import Rotation

    from lauematching_amd import api

    rng = np.random.default_rng(11)
    # candidates as a hit list would give them: members of a few thousand
    # grains (within 1.6 degrees of their grain) in shuffled list order
    n_grains, per = 2500, 20
    centres = Rotation.from_matrix(O[rng.choice(len(O), n_grains, replace=False)].reshape(-1, 3, 3))
    axes = rng.normal(size=(n_grains * (per - 1), 3))
    axes /= np.linalg.norm(axes, axis=1, keepdims=True)
    ang = np.radians(rng.uniform(0, 1.6, len(axes)))
    c = Rotation.from_quat(np.repeat(centres.as_quat(), per - 1, axis=0))
    mats = np.concatenate([centres.as_matrix().reshape(-1, 9),
                           (c * Rotation.from_rotvec(axes * ang[:, None])).as_matrix().reshape(-1, 9)])
    rows = rng.permutation(len(mats)).astype(np.uint64)
    scores = rng.uniform(100, 1000, len(rows))
    out = {"candidates": int(len(rows)), "max_angle": 2.0, "space_group": 225}
    api.merge_candidates(mats, rows[:64], scores[:64], 2.0, 225, ctx=ctx)
    t0 = time.perf_counter()
    g = api.merge_candidates(mats, rows, scores, 2.0, 225, ctx=ctx)
    out["gpu_ms"] = (time.perf_counter() - t0) * 1e3
    out["clusters"] = int(len(g[0]))
    k = 5000
    t0 = time.perf_counter()
    h = api.merge_candidates(mats, rows[:k], scores[:k], 2.0, 225, threads=16)
    out["cpu_sample"] = {"candidates": k, "ms": (time.perf_counter() - t0) * 1e3, "threads": 16}
    t0 = time.perf_counter()
    gk = api.merge_candidates(mats, rows[:k], scores[:k], 2.0, 225, ctx=ctx)
    out["gpu_sample_ms"] = (time.perf_counter() - t0) * 1e3
    out["sample_equal"] = all(np.array_equal(a, b) for a, b in zip(gk, h))
    log(f"merge: {out['candidates']} candidates on the device in {out['gpu_ms']:.0f} ms "
        f"({out['clusters']} clusters); {k}: device {out['gpu_sample_ms']:.1f} ms, host "
        f"{out['cpu_sample']['ms']:.0f} ms (16 threads), equal={out['sample_equal']}")
    return out


def prep_check(ctx, reps=5):
    """Device preprocessing of one raw 2048x2048 frame (background
    subtraction, adaptive threshold, components, blur, uint8 quantization):
    device time of the pipeline and wall time of the call (host copies in
    and out included)."""
    from lauematching_amd import api

    rng = np.random.default_rng(21)
    H = W = 2048
    raw = rng.poisson(50.0, (H, W)).astype(np.float64)
    yy, xx = np.mgrid[0:21, 0:21]
    for _ in range(3000):
        yc, xc = rng.uniform(10, H - 11), rng.uniform(10, W - 11)
        a, sg = rng.uniform(300, 4000), rng.uniform(0.7, 2.5)
        y0, x0 = int(yc) - 10, int(xc) - 10
        raw[y0:y0 + 21, x0:x0 + 21] += a * np.exp(-((yy + y0 - yc) ** 2 + (xx + x0 - xc) ** 2)
                                                  / (2 * sg * sg))
    bg = 45.0 + rng.normal(0, 2.0, (H, W))
    api.preprocess(ctx, raw, bg)
    dev, wall = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        u8, scale, info = api.preprocess(ctx, raw, bg)
        wall.append((time.perf_counter() - t0) * 1e3)
        dev.append(info["ms"])
    out = {"frame": "2048x2048 float64, 3000 spots, adaptive threshold, min_area 10",
           "device_ms": float(np.median(dev)), "wall_ms": float(np.median(wall)),
           "components": info["n_components"], "kept": info["n_kept"], "sigma": info["sigma"]}
    log(f"preprocess: {out['device_ms']:.2f} ms device, {out['wall_ms']:.1f} ms per call "
        f"({out['kept']}/{out['components']} components kept, sigma {out['sigma']:.2f})")
    out["median_background"] = median_check(ctx, raw, reps)
    return out


def median_check(ctx, raw, reps=5, size=101):
    """The median-filter background (compute_background's default size 101,
    one pass) of the 2048x2048 frame as integer counts (a detector's frame)
    and as float64 (every value distinct), the pipeline with that background
    computed, and scipy.ndimage.median_filter (the reference's filter
    without diplib) timed on a 160x160 crop on one host core, with the
    device's result on that crop compared to scipy's; both full frames are
    compared with the C running-median oracle (pinned against scipy)."""
    import scipy.ndimage as ndi
    from lauematching_amd import api
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob

    counts = np.floor(raw)
    out = {"filter_radius": size, "median_passes": 1, "frame": "2048x2048"}
    for name, img in (("integer_counts", counts), ("float", raw)):
        got, _ = api.median_background(ctx, img, size, 1)
        ms = [api.median_background(ctx, img, size, 1)[1] for _ in range(reps)]
        out[f"{name}_device_ms"] = float(np.median(ms))
        t0 = time.perf_counter()
        want = ob.median_background(img, size, 1)
        out[f"{name}_oracle"] = {"equal_full_frame": bool(np.array_equal(got, want)),
                                 "distinct_values": int(np.unique(img).size),
                                 "oracle_s": time.perf_counter() - t0,
                                 "oracle_threads": len(os.sched_getaffinity(0))}
    api.preprocess(ctx, counts, None, filter_radius=size)
    pms = [api.preprocess(ctx, counts, None, filter_radius=size)[2]["ms"] for _ in range(reps)]
    out["pipeline_with_median_device_ms"] = float(np.median(pms))
    crop = counts[:160, :160].copy()
    t0 = time.perf_counter()
    want = ndi.median_filter(crop, size=size)
    cpu_s = time.perf_counter() - t0
    got, _ = api.median_background(ctx, crop, size, 1)
    out["cpu_scipy"] = {"sample": "160x160 crop of the integer frame", "cores": 1,
                        "ms": cpu_s * 1e3, "px_per_s": crop.size / cpu_s,
                        "device_equal": bool(np.array_equal(got, want))}
    out["device_px_per_s"] = counts.size / (out["integer_counts_device_ms"] * 1e-3)
    log(f"median background {size}x{size}: {out['integer_counts_device_ms']:.2f} ms (integer "
        f"counts), {out['float_device_ms']:.2f} ms (float), pipeline with it "
        f"{out['pipeline_with_median_device_ms']:.2f} ms; scipy {crop.size / cpu_s:.3g} px/s "
        f"on one core vs device {out['device_px_per_s']:.3g} px/s, crop equal "
        f"{out['cpu_scipy']['device_equal']}, full frames equal to the oracle "
        f"{out['integer_counts_oracle']['equal_full_frame']}/{out['float_oracle']['equal_full_frame']}")
    return out


def precision_check(db, fp64_ms):
    """Rebuild the table with the FP32 projection + FP64 guard band and check
    it is the FP64 table (64-bit position-tagged checksum over counts and
    spots); report the recompute fraction."""
    from lauematching_amd import api

    c64 = db.checksum()
    st = db.forward(api.LM_FWD_FP32_GUARD)
    c32 = db.checksum()
    out = {"fp64_forward_ms": fp64_ms, "fp32_guard_forward_ms": st.ms,
           "speedup": fp64_ms / st.ms if st.ms else None,
           "pairs": int(st.n_pairs), "recomputed": int(st.n_recomputed),
           "recompute_fraction": st.n_recomputed / st.n_pairs,
           "table_checksum_equal": bool(c64 == c32)}
    log(f"precision: FP32+guard forward {st.ms:.1f} ms vs FP64 {fp64_ms:.1f} ms, "
        f"recomputed {st.n_recomputed / st.n_pairs:.4%}, tables equal={c64 == c32}")
    return out


def config4(ctx, O, n, steps, batch_imgs):
    """Config 4: FCC |hkl| <= 10 (1066 reflections), 5-40 keV; FP64 vs
    FP32+guard forward; scoring a batch of 32 with the table resident versus
    regenerated per batch: every orientation projected (FP32+guard) and
    scored in the same launch with no table (lm_score_fwd_batch_dev), and,
    for comparison, the table rebuilt and then read."""
    from lauematching_amd import api, synth

    p = synth.baseline_params(Ehi=40.0, MaxNrLaueSpots=64)
    hk = synth.fcc_hkls(10, "sphere")
    db = api.OrientationDB(ctx, O, hk, api.Geometry.from_params(p), 64)
    s64 = db.forward(api.LM_FWD_FP64)
    c64 = db.checksum()
    s32 = db.forward(api.LM_FWD_FP32_GUARD)
    c32 = db.checksum()
    res = {"hkls": int(len(hk)), "fp64_forward_ms": s64.ms, "fp32_guard_forward_ms": s32.ms,
           "recompute_fraction": s32.n_recomputed / s32.n_pairs,
           "table_checksum_equal": bool(c64 == c32),
           "spots_per_orientation": s32.n_spots / n}
    if batch_imgs:
        import torch
        B = min(32, len(batch_imgs))
        u8s, scales = zip(*[api.quantize_u8(im) for im, _ in batch_imgs[:B]])
        dev = torch.from_numpy(np.stack(u8s)).to("cuda:0")
        torch.cuda.synchronize()
        db.score_batch_dev(dev.data_ptr(), B, scales, 5, 1000.0)
        t0 = time.perf_counter()
        for _ in range(steps):
            db.score_batch_dev(dev.data_ptr(), B, scales, 5, 1000.0)
        torch.cuda.synchronize()
        ms_res = (time.perf_counter() - t0) / steps * 1e3
        want = db.score_batch_dev(dev.data_ptr(), B, scales, 5, 1000.0)
        t0 = time.perf_counter()
        for _ in range(max(1, steps // 3)):
            db.forward(api.LM_FWD_FP32_GUARD)
            db.score_batch_dev(dev.data_ptr(), B, scales, 5, 1000.0)
        torch.cuda.synchronize()
        ms_rebuild = (time.perf_counter() - t0) / max(1, steps // 3) * 1e3
        # on the fly: every orientation projected (FP32 + guard) and scored in
        # the same launch, no table read or written
        got, fst = db.score_fwd_batch_dev(dev.data_ptr(), B, scales, 5, 1000.0,
                                          api.LM_FWD_FP32_GUARD)
        same = all(g.nr_results == w.nr_results and np.array_equal(g.ids, w.ids) and
                   np.array_equal(g.score, w.score) for g, w in zip(got, want))
        t0 = time.perf_counter()
        for _ in range(max(1, steps // 3)):
            db.score_fwd_batch_dev(dev.data_ptr(), B, scales, 5, 1000.0, api.LM_FWD_FP32_GUARD)
        torch.cuda.synchronize()
        ms_regen = (time.perf_counter() - t0) / max(1, steps // 3) * 1e3
        # the same on the fly in float64 (no guard band needed)
        t0 = time.perf_counter()
        for _ in range(max(1, steps // 3)):
            g64, _ = db.score_fwd_batch_dev(dev.data_ptr(), B, scales, 5, 1000.0, api.LM_FWD_FP64)
        torch.cuda.synchronize()
        ms_regen64 = (time.perf_counter() - t0) / max(1, steps // 3) * 1e3
        same &= all(g.nr_results == w.nr_results and np.array_equal(g.ids, w.ids) and
                    np.array_equal(g.score, w.score) for g, w in zip(g64, want))
        res.update({"batch": B, "resident_ms_per_batch": ms_res,
                    "resident_pairs_per_s": n * B / (ms_res / 1e3),
                    "regenerated_ms_per_batch": ms_regen,
                    "regenerated_pairs_per_s": n * B / (ms_regen / 1e3),
                    "regenerated_kernel_ms": fst.ms,
                    "regenerated_recompute_fraction": fst.n_recomputed / fst.n_pairs,
                    "regenerated_equals_resident": bool(same),
                    "regenerated_fp64_ms_per_batch": ms_regen64,
                    "rebuild_table_then_score_ms": ms_rebuild})
    log(f"config4: {res}")
    db.close()
    return res


def host_cpus():
    """CPUs this process may run on, and the cgroup CPU quota if one is set."""
    out = {"affinity": len(os.sched_getaffinity(0))}
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        out["cgroup_quota_cpus"] = None if q == "max" else int(q) / int(per)
    except (OSError, ValueError):
        pass
    return out


def tcp_sink_rate(items, conns, count):
    """Frames per second that `conns` sender threads (the daemon legs' own
    senders: one queue and thread per connection, sendall of each part of a
    frame) push over loopback TCP into receivers that discard them -- the
    ingest ceiling of a closed loop of `count` such frames.  items: frames as
    tuples of byte parts.  Returns (frames/s, cores busy in this process)."""
    import queue
    import socket
    import threading

    srv = socket.socket()
    srv.bind(("127.0.0.1", 0))
    srv.listen(conns)
    port = srv.getsockname()[1]
    sizes = [sum(len(x) for x in it) for it in items]
    expect = [0] * conns
    for j in range(count):
        expect[j % conns] += sizes[j % len(items)]
    ends = []
    lock = threading.Lock()

    def receiver(c):
        who = bytearray(4)
        c.recv_into(who, 4, socket.MSG_WAITALL)
        left = expect[int.from_bytes(who, "little")]
        buf = bytearray(1 << 24)
        while left:
            n = c.recv_into(buf, min(len(buf), left))
            if n == 0:
                break
            left -= n
        with lock:
            ends.append(time.perf_counter())
        c.close()

    socks = [socket.create_connection(("127.0.0.1", port), timeout=60) for _ in range(conns)]
    rts = []
    for _ in range(conns):
        c, _ = srv.accept()
        rts.append(threading.Thread(target=receiver, args=(c,), daemon=True))
        rts[-1].start()
    for k, s in enumerate(socks):
        s.sendall(k.to_bytes(4, "little"))
    qs = [queue.Queue() for _ in range(conns)]

    def sender(k):
        while True:
            item = qs[k].get()
            if item is None:
                return
            for part in item:
                socks[k].sendall(part)

    ths = [threading.Thread(target=sender, args=(k,), daemon=True) for k in range(conns)]
    for th in ths:
        th.start()
    c0 = os.times()
    t0 = time.perf_counter()
    for j in range(count):
        qs[j % conns].put(items[j % len(items)])
    for q in qs:
        q.put(None)
    for th in ths:
        th.join()
    for th in rts:
        th.join(timeout=120)
    for s in socks:
        s.close()
    srv.close()
    c1 = os.times()
    if len(ends) != conns:
        return None, None
    wall = max(ends) - t0
    return count / wall, (c1.user + c1.system - c0.user - c0.system) / wall


def daemon_leg(O, hk, p, frames, rates=(1, 20, 100, 200), duration=3.0, closed=2048, warm=512,
               conns=8):
    """bin/LaueMatchingGPUStream -- the drop-in daemon the reference's senders
    talk to -- on the same 100M-orientation DB, fed wire-v1 frames (uint16
    image number + float32 pixels, the reference's format) over TCP from
    `conns` connections: Poisson arrivals at each rate, then closed loops
    (`warm` untimed frames, then `closed` timed frames sent back to back) of
    wire-v1 float32 and wire-v2 uint8 frames, then the same senders into
    receivers that discard the frames (the loopback TCP ceiling).  Latencies
    are the daemon's own: to the compact hit list (frame received ->
    quantized, queued, scored) and end to end (frame received -> merge, fit
    and result rows written)."""
    import queue
    import re
    import shutil
    import signal
    import socket
    import struct
    import tempfile
    import threading

    from lauematching_amd import synth

    work = tempfile.mkdtemp(prefix="lm_bench_", dir="/dev/shm" if os.path.isdir("/dev/shm") else None)
    out = {"connections": conns, "frames": "wire v1, float32 2048x2048 (16 MiB)"}
    proc = None
    try:
        t0 = time.time()
        O.tofile(os.path.join(work, "orients.bin"))
        synth.write_hkl_file(os.path.join(work, "hkls.csv"), hk)
        pp = synth.baseline_params(MaxNrLaueSpots=p.MaxNrLaueSpots, MinNrSpots=p.MinNrSpots,
                                   MinIntensity=p.MinIntensity, DoFwd=1, ForwardFile="")
        pp.extra = {"MinGoodSpots": p.MinNrSpots, "ResultDir": os.path.join(work, "res")}
        pp.write(os.path.join(work, "params.txt"))
        with socket.socket() as so:
            so.bind(("127.0.0.1", 0))
            port = so.getsockname()[1]
        exe = os.path.join(ROOT, "bin", "LaueMatchingGPUStream")
        env = dict(os.environ, LM_STREAM_PORT=str(port))
        proc = subprocess.Popen([exe, os.path.join(work, "params.txt"),
                                 os.path.join(work, "orients.bin"), os.path.join(work, "hkls.csv"),
                                 "16"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                text=True, env=env, bufsize=1)
        gpu, total, tail = {}, {}, []
        ready = threading.Event()
        rx_g = re.compile(r"\[Image (\d+)\] GPU: .*hit list ([0-9.]+) ms after arrival")
        rx_t = re.compile(r"\[Image (\d+)\] Total: ([0-9.]+) ms after arrival")
        rx_r = re.compile(r"Received image (\d+) .*receive ([0-9.]+) ms, queued ([0-9.]+) ms, "
                          r"quantize ([0-9.]+) ms, hand-over ([0-9.]+) ms")
        stage = {}
        lock = threading.Lock()

        def reader():
            for ln in proc.stdout:
                now = time.perf_counter()
                tail.append(ln.rstrip())
                del tail[:-40]
                if "Waiting for images" in ln:
                    ready.set()
                elif "Autotuned" in ln:
                    out["autotune_line"] = ln.strip()
                elif "Init time" in ln:
                    out["init_s"] = float(ln.split()[-2])
                m = rx_g.search(ln)
                if m:
                    with lock:
                        gpu[int(m.group(1))] = (float(m.group(2)), now)
                m = rx_t.search(ln)
                if m:
                    with lock:
                        total[int(m.group(1))] = (float(m.group(2)), now)
                m = rx_r.search(ln)
                if m:
                    with lock:
                        stage[int(m.group(1))] = [float(m.group(k)) for k in range(2, 6)]
            ready.set()

        rt = threading.Thread(target=reader, daemon=True)
        rt.start()
        if not ready.wait(600) or proc.poll() is not None:
            raise RuntimeError("daemon did not start: " + " | ".join(tail[-8:]))
        out["startup_s"] = time.time() - t0
        log(f"daemon ready in {out['startup_s']:.1f} s ({out.get('autotune_line', '')})")
        payloads = [memoryview(im.astype(np.float32).tobytes()) for im in frames]
        socks = [socket.create_connection(("127.0.0.1", port), timeout=30) for _ in range(conns)]
        qs = [queue.Queue() for _ in range(conns)]

        def sender(k):
            while True:
                item = qs[k].get()
                if item is None:
                    return
                num, pl = item
                socks[k].sendall(struct.pack("<H", num))
                socks[k].sendall(pl)

        ths = [threading.Thread(target=sender, args=(k,), daemon=True) for k in range(conns)]
        for th in ths:
            th.start()
        nxt = [1]

        def wait_for(nums, limit):
            end = time.time() + limit
            while time.time() < end:
                with lock:
                    if all(x in total for x in nums):
                        return True
                time.sleep(0.01)
            return False

        def stats(nums):
            with lock:
                g = np.array([gpu[x][0] for x in nums if x in gpu])
                t = np.array([total[x][0] for x in nums if x in total])
                st = np.array([stage[x] for x in nums if x in stage]).reshape(-1, 4)
            return {"p50_ms": float(np.percentile(g, 50)), "p99_ms": float(np.percentile(g, 99)),
                    "end_to_end_p50_ms": float(np.percentile(t, 50)),
                    "end_to_end_p99_ms": float(np.percentile(t, 99)),
                    "host_stage_mean_ms": dict(zip(("receive", "queued", "quantize", "hand_over"),
                                                   st.mean(axis=0).tolist()))}

        def cpu_now():
            """CPU seconds so far of the daemon and of this process (the
            senders and the stdout reader)."""
            with open(f"/proc/{proc.pid}/stat") as f:
                fields = f.read().rsplit(")", 1)[1].split()
            tick = os.sysconf("SC_CLK_TCK")
            t = os.times()
            return (int(fields[11]) + int(fields[12])) / tick, t.user + t.system

        def cores_used(c0, wall):
            c1 = cpu_now()
            return {"daemon_cores": (c1[0] - c0[0]) / wall, "client_cores": (c1[1] - c0[1]) / wall}

        out["host_cpus"] = host_cpus()
        rng = np.random.default_rng(77)
        rows = []
        for r in rates:
            nreq = max(8 if r < 5 else 20, int(r * duration))
            arr = np.cumsum(rng.exponential(1.0 / r, nreq))
            nums = list(range(nxt[0], nxt[0] + nreq))
            nxt[0] += nreq
            t0 = time.perf_counter()
            for j, num in enumerate(nums):
                dt = t0 + arr[j] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                qs[j % conns].put((num, payloads[num % len(payloads)]))
            if not wait_for(nums, 120):
                raise RuntimeError("daemon results missing at rate %s" % r)
            row = {"rate_images_per_s": r, "requests": nreq, **stats(nums)}
            rows.append(row)
            log(f"daemon {r}/s: to hit list p50 {row['p50_ms']:.1f} p99 {row['p99_ms']:.1f} ms, "
                f"end to end p50 {row['end_to_end_p50_ms']:.1f} p99 {row['end_to_end_p99_ms']:.1f} ms")
        out["per_rate"] = rows

        def closed_loop(key, label, put):
            """`warm` frames sent back to back (untimed: the daemon's frame
            buffers and the server's batch widths warm up), then `closed`
            frames timed: the rate over the whole loop (including the
            pipeline's fill and drain) and over the middle half of the hit
            lists (the sustained rate)."""
            warm_nums = list(range(nxt[0], nxt[0] + warm))
            nxt[0] += warm
            for j, num in enumerate(warm_nums):
                put(j, num)
            if not wait_for(warm_nums, 300):
                raise RuntimeError(f"daemon {label} warm-up results missing")
            nums = list(range(nxt[0], nxt[0] + closed))
            nxt[0] += closed
            c0 = cpu_now()
            t0 = time.perf_counter()
            for j, num in enumerate(nums):
                put(j, num)
            if not wait_for(nums, 300):
                raise RuntimeError(f"daemon {label} results missing")
            wall = time.perf_counter() - t0
            with lock:
                tg = sorted(gpu[x][1] for x in nums)
                tt = max(total[x][1] for x in nums) - t0
            lo, hi = len(tg) // 4, 3 * len(tg) // 4
            res = {"images": closed, "warmup_images": warm,
                   "images_per_s_to_hit_list": closed / (tg[-1] - t0),
                   "sustained_images_per_s_to_hit_list": (hi - lo) / (tg[hi] - tg[lo]),
                   "images_per_s_end_to_end": closed / tt, **stats(nums),
                   **cores_used(c0, wall)}
            out[key] = res
            log(f"daemon closed loop ({label}): {res['images_per_s_to_hit_list']:.0f} img/s to the "
                f"hit list over {closed} frames, {res['sustained_images_per_s_to_hit_list']:.0f} "
                f"img/s sustained (middle half), {res['images_per_s_end_to_end']:.0f} img/s end "
                f"to end; p50 {res['p50_ms']:.1f} ms; host stages per frame (ms) "
                f"{ {k: round(v, 2) for k, v in res['host_stage_mean_ms'].items()} }; cores busy: "
                f"daemon {res['daemon_cores']:.1f}, senders + reader {res['client_cores']:.1f} "
                f"(host {out['host_cpus']})")

        closed_loop("closed_loop", "wire v1 float32 frames",
                    lambda j, num: qs[j % conns].put((num, payloads[num % len(payloads)])))
        for q in qs:
            q.put(None)
        for th in ths:
            th.join()
        for so in socks:
            so.close()
        # the same closed loop with wire-v2 uint8 frames (4 MiB, a quarter of
        # the float32 bytes; the frames quantize to the same image)
        from lauematching_amd import api
        u8p = []
        for im in frames:
            u8_, _ = api.quantize_u8(im)
            H, W = im.shape
            u8p.append((struct.pack("<HHIII", 0, 3, W, H, 0), memoryview(u8_.tobytes())))
        socks = [socket.create_connection(("127.0.0.1", port + 1), timeout=30)
                 for _ in range(conns)]
        qs = [queue.Queue() for _ in range(conns)]

        def sender2(k):
            while True:
                item = qs[k].get()
                if item is None:
                    return
                num, (hdr, pl) = item
                socks[k].sendall(struct.pack("<II", 0x32464D4C, num) + hdr)
                socks[k].sendall(pl)

        ths = [threading.Thread(target=sender2, args=(k,), daemon=True) for k in range(conns)]
        for th in ths:
            th.start()
        closed_loop("closed_loop_u8_frames", "wire v2 uint8 frames",
                    lambda j, num: qs[j % conns].put((num, u8p[num % len(u8p)])))
        for q in qs:
            q.put(None)
        for th in ths:
            th.join()
        for so in socks:
            so.close()
        # the same senders and frames into receivers that discard them
        ceil_v1, cores_v1 = tcp_sink_rate([(struct.pack("<H", 1), pl) for pl in payloads], conns,
                                          closed)
        ceil_u8, cores_u8 = tcp_sink_rate([(struct.pack("<II", 0x32464D4C, 1) + h, pl)
                                           for h, pl in u8p], conns, closed)
        out["tcp_ingest_ceiling"] = {"v1_float32_images_per_s": ceil_v1,
                                     "v2_uint8_images_per_s": ceil_u8,
                                     "v1_cores": cores_v1, "v2_cores": cores_u8}
        log(f"loopback TCP ingest ceiling ({conns} connections, same senders, frames discarded): "
            f"{ceil_v1:.0f} img/s float32 frames ({cores_v1:.1f} cores), {ceil_u8:.0f} img/s "
            f"uint8 frames ({cores_u8:.1f} cores)")
        proc.send_signal(signal.SIGTERM)
        proc.wait(timeout=120)
        rt.join(timeout=10)
        out["exit_code"] = proc.returncode
        out["final_line"] = next((l for l in tail if l.startswith("Processed")), None)
        out["host_time_line"] = next((l for l in tail if l.startswith("Host time")), None)
        log(f"daemon: {out['final_line']}; {out['host_time_line']}")
    finally:
        if proc is not None and proc.poll() is None:
            proc.kill()
            proc.wait()
        shutil.rmtree(work, ignore_errors=True)
    return out


def kernel_name(db):
    """Name of the single-image kernel the DB's layout options select, with
    every template argument as rocprofv3 prints it (the LDS walk's last
    argument, the slow path for oversized chunks, depends on the table and
    is left as ``*``)."""
    tb, loads = db.banded_layout()
    if tb == 13 and (loads == 3 or loads >= 5):
        loads = 0  # no LDS walk at 2^13 rows: the byte-gather walk runs
    if tb == 0:
        return "score_kernel<1,false>"
    if ((loads == 3 or loads >= 5) and tb >= 14) or tb == 16:
        e1, e2 = (8, 12) if tb == 16 else (4, 8)
        shape = {3: f"{e1},2", 5: f"{e2},2", 6: f"{e1},3", 7: f"{e2},3"}.get(loads, f"{e1},2")
        return f"score_band_lds_kernel<{tb},{12 if tb == 14 else 11},{shape},1024,*,true>"
    threads = 1024 if tb >= 14 else 512
    if loads == 0:
        return f"score_band_kernel<{tb},{threads},8,false>"
    if loads == 4:
        return f"score_band_kernel<{tb},{threads},8,true>"
    return f"score_band_v4_kernel<{tb},{threads},{2 if loads == 1 else 4}>"


def same_kernel(profiled, name):
    """True when a profiled kernel name is the one ``kernel_name`` gave
    (spaces ignored, ``*`` matching one template argument)."""
    import fnmatch
    profiled = profiled.replace(" ", "")
    return fnmatch.fnmatchcase(profiled, name.replace("[", "[[]")) and profiled.count(",") == name.count(",")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--orients", type=int, default=100_000_000)
    ap.add_argument("--grains", type=int, default=5)
    ap.add_argument("--min-spots", type=int, default=5)
    ap.add_argument("--min-intensity", type=float, default=1000.0)
    ap.add_argument("--cpu-sample", type=int, default=10_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tile-bits", type=int, default=None,
                    help="force the single-image super-tile size (log2) of the banded "
                         "table (0 = none) instead of the autotuned one")
    ap.add_argument("--loads", type=int, default=0,
                    help="with --tile-bits: the single-image walk (LM_OPT_BANDED_LOADS)")
    ap.add_argument("--batch", type=str, default="8,16,32,64,128,256",
                    help="comma-separated batch widths to measure as well (one pass over the "
                         "table per batch), e.g. 8,16,32,64,128; empty = none")
    ap.add_argument("--no-autotune", action="store_true",
                    help="score batches with the default layouts instead of the autotuned ones")
    ap.add_argument("--graphs", action="store_true",
                    help="replay the single-image launch sequence as a graph")
    ap.add_argument("--no-merge", action="store_true",
                    help="skip the duplicate-merge and preprocessing timings")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the FP32+guard rebuild of the table")
    ap.add_argument("--no-config4", action="store_true",
                    help="skip the |hkl|<=10, 5-40 keV precision configuration (table "
                         "resident vs regenerated per batch of 32)")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving configuration (two crystals x two detectors "
                         "resident, 70/30 detector mix, Poisson rates 1..200 images/s and "
                         "overload, closed loop)")
    ap.add_argument("--latency", action="store_true",
                    help="also measure serving latency on the single 100M FCC DB alone")
    ap.add_argument("--no-daemon", action="store_true",
                    help="skip bin/LaueMatchingGPUStream fed over TCP")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(every box here has one GPU; across GPUs the path would shard by orientation range)", file=sys.stderr)
        sys.exit(2)
    widths = [int(b) for b in args.batch.split(",") if b]
    batch_imgs = []
    args.config4 = not args.no_config4
    args.serve = not args.no_serve
    args.no_latency = not args.latency
    need = max(widths + ([32] if args.config4 else []), default=0)
    if need:
        # worker processes are started before anything touches the GPU
        t0 = time.time()
        batch_imgs = make_batch_images(need, args.orients, args.seed)
        log(f"generated {len(batch_imgs)} batch images in {time.time() - t0:.1f} s")
    n_hex = args.orients // 2
    serve_pool = []
    if args.serve or not args.no_latency or not args.no_daemon:
        combos = [("fcc", 2048, args.orients, args.seed)]
        if args.serve:
            combos += [("fcc", 1024, args.orients, args.seed),
                       ("hex", 2048, n_hex, args.seed + 7), ("hex", 1024, n_hex, args.seed + 7)]
        t0 = time.time()
        serve_pool = make_serve_pool(combos, 8)
        log(f"generated {len(serve_pool)} serving images in {time.time() - t0:.1f} s")

    import torch
    from lauematching_amd import api, synth

    p = synth.baseline_params(MaxNrLaueSpots=64, MinNrSpots=args.min_spots,
                              MinIntensity=args.min_intensity)
    hk = synth.fcc_hkls(8, "box")
    n = args.orients
    t0 = time.time()
    O = synth.random_orientations(n, seed=args.seed)
    log(f"generated {n} orientations in {time.time() - t0:.1f} s")
    rng = np.random.default_rng(args.seed + 1)
    grains = np.sort(rng.choice(n, args.grains, replace=False))
    img = synth.synth_image(p, hk, O[grains], seed=args.seed + 2)
    ctx = api.Context(0)
    t0 = time.time()
    db = api.OrientationDB(ctx, O, hk, api.Geometry.from_params(p), p.MaxNrLaueSpots)
    log(f"DB upload {time.time() - t0:.1f} s")
    st = db.forward()
    log(f"forward simulation: {st.n_spots} spots ({st.n_spots / n:.2f}/orientation), "
        f"{st.ms:.1f} ms device time")
    u8, im_scale = api.quantize_u8(img)

    # layouts: the autotuner times every walk on this image (and the batch
    # images) before anything is measured; --tile-bits forces the single walk
    tune = None
    if not args.no_autotune:
        samples = [u8] + [api.quantize_u8(im)[0] for im, _ in batch_imgs[:3]]
        t0 = time.time()
        tune = db.autotune(samples, max_batch=256, reps=2)
        tune["seconds"] = time.time() - t0
        log(f"autotune ({tune['seconds']:.1f} s): single {tune['ms_single']} -> "
            f"{tune['banded_tile_bits']}/{tune['banded_loads']}; batch tile bits "
            f"{tune['batch_tile_bits']}, widths mask {tune['batch_widths_mask']:#x}; "
            f"per width {tune['ms_batch']}")
    if args.tile_bits is not None:
        db.set_banded_tile_bits(args.tile_bits)
        db.set_banded_loads(args.loads)
    single_kernel = kernel_name(db)
    d_img = torch.from_numpy(u8).to("cuda:0")
    torch.cuda.synchronize()
    # optionally the launch sequence (counter fill, walk, recount) as one graph
    db.set_use_graphs(args.graphs)

    def step():
        return db.score_dev(d_img.data_ptr(), im_scale, p.MinNrSpots, p.MinIntensity)

    for _ in range(args.warmup):
        h = step()
    torch.cuda.synchronize()
    kms, wms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        h = step()
        kms.append(h.kernel_ms)
        wms.append(h.walk_ms)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = dt / args.steps * 1e3
    value = n / (ms_step / 1e3)
    kernel_ms = float(np.mean(kms))  # counter fill + compare kernel + recount
    walk_ms = float(np.mean(wms))    # the compare kernel alone
    log(f"step {ms_step:.3f} ms, sequence {kernel_ms:.3f} ms, compare kernel {walk_ms:.3f} ms, "
        f"hits {len(h.ids)} "
        f"(nrResults {h.nr_results})")
    found = all(g in set(h.ids.tolist()) for g in grains)
    # every result below is checked against the oracle over the whole table
    parity_cases = [("single", u8, im_scale, p.MinNrSpots, p.MinIntensity, h)]

    # the same frame as a detector would give it (maximum not 255, so
    # imScale != 1): candidate superset from the integer walk, then the
    # float64 in-order re-sum of the candidates
    u8n, scale_n = api.quantize_u8(img * 0.731 + (img > 0) * 0.013)
    d_imgn = torch.from_numpy(u8n).to("cuda:0")
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        db.score_dev(d_imgn.data_ptr(), scale_n, p.MinNrSpots, p.MinIntensity * 0.731)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        hn = db.score_dev(d_imgn.data_ptr(), scale_n, p.MinNrSpots, p.MinIntensity * 0.731)
    torch.cuda.synchronize()
    nonunit = {"im_scale": scale_n, "ms_per_step": (time.perf_counter() - t0) / args.steps * 1e3,
               "kernel_ms": hn.kernel_ms, "hits": int(len(hn.ids))}
    parity_cases.append(("nonunit_scale", u8n, scale_n, p.MinNrSpots, p.MinIntensity * 0.731, hn))
    log(f"imScale {scale_n:.4f} frame: step {nonunit['ms_per_step']:.3f} ms, sequence "
        f"{hn.kernel_ms:.3f} ms, hits {len(hn.ids)}")
    del d_imgn
    # a sparse frame, as preprocessing leaves them (3% of pixels nonzero),
    # with the thresholds of the reference's parameter files (MinNrSpots 5,
    # MinIntensity 50): the 2^16 walk counts spots instead of summing pixels
    sparse = np.where(img > np.percentile(img, 97.0), img, 0.0)
    u8s, scale_s = api.quantize_u8(sparse)
    d_imgs = torch.from_numpy(u8s).to("cuda:0")
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        db.score_dev(d_imgs.data_ptr(), scale_s, 5, 50.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        hs_ = db.score_dev(d_imgs.data_ptr(), scale_s, 5, 50.0)
    torch.cuda.synchronize()
    sparse_line = {"nonzero_fraction": float(np.count_nonzero(u8s)) / u8s.size,
                   "min_nr_spots": 5, "min_intensity": 50.0, "im_scale": scale_s,
                   "ms_per_step": (time.perf_counter() - t0) / args.steps * 1e3,
                   "kernel_ms": hs_.kernel_ms, "walk_ms": hs_.walk_ms,
                   "nr_results": int(hs_.nr_results)}
    parity_cases.append(("sparse", u8s, scale_s, 5, 50.0, hs_))
    log(f"sparse frame ({100 * sparse_line['nonzero_fraction']:.1f}% nonzero, MinNrSpots 5, "
        f"MinIntensity 50): step {sparse_line['ms_per_step']:.3f} ms, sequence "
        f"{hs_.kernel_ms:.3f} ms, {hs_.nr_results} results")
    del d_imgs

    # a plain device copy of known size: the HBM rate reachable on this box,
    # and the reference kernel that scales the FETCH_SIZE / WRITE_SIZE counters
    csrc = torch.empty(1 << 30, dtype=torch.uint8, device="cuda:0")
    cdst = torch.empty_like(csrc)
    cdst.copy_(csrc)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        cdst.copy_(csrc)
    e1.record()
    torch.cuda.synchronize()
    copy_ms = e0.elapsed_time(e1) / 5
    copy_gbps = 2.0 * (1 << 30) / (copy_ms / 1e3) / 1e9
    del csrc, cdst
    log(f"plain 1 GiB device copy: {copy_ms:.3f} ms, {copy_gbps:.0f} GB/s (read + write)")
    # host-to-device rate from pinned memory (the serving closed loop uploads
    # every image), and the headline step with the image handed over in host
    # memory (upload included; reported beside `value`, never as it)
    hsrc = torch.empty(1 << 28, dtype=torch.uint8, pin_memory=True)
    hdst = torch.empty(1 << 28, dtype=torch.uint8, device="cuda:0")
    hdst.copy_(hsrc, non_blocking=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(5):
        hdst.copy_(hsrc, non_blocking=True)
    e1.record()
    torch.cuda.synchronize()
    h2d_gbps = (1 << 28) / (e0.elapsed_time(e1) / 5 / 1e3) / 1e9
    del hsrc, hdst
    for _ in range(args.warmup):
        db.score(u8, im_scale, p.MinNrSpots, p.MinIntensity)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        db.score(u8, im_scale, p.MinNrSpots, p.MinIntensity)
    host_step_ms = (time.perf_counter() - t0) / args.steps * 1e3
    log(f"pinned host-to-device copy {h2d_gbps:.1f} GB/s; step with the image in host memory "
        f"{host_step_ms:.3f} ms (upload included)")

    # algorithmic bytes of one compare launch (per orientation 2 + 4 n_o bytes,
    # the image once, 16 bytes per hit)
    b_alg = 2.0 * n + 4.0 * st.n_spots + float(u8.size) + 16.0 * h.nr_results
    achieved = b_alg / (walk_ms / 1e3)
    traffic = None
    tf = os.path.join(ROOT, "profiles", "compare_traffic.json")
    if os.path.exists(tf):
        try:
            tj = json.load(open(tf))
            if tj.get("orients") == n and same_kernel(tj.get("kernel", ""), single_kernel):
                traffic = tj.get("bytes_per_launch")
        except Exception:
            traffic = None

    batched = None
    if widths:
        db.set_use_graphs(True)
        batched = run_batches(db, batch_imgs, widths, p, max(3, args.steps // 2), 1,
                              parity_cases, st.n_spots)
        db.set_use_graphs(False)

    latency = None
    if not args.no_latency:
        reqs = [(0, *api.quantize_u8(img)) for c, d, img in serve_pool if (c, d) == ("fcc", 2048)]
        latency = run_serving(ctx, [db], reqs, [50, 200, 400, 800, 1600], 2.0, 2048)
    serving = None
    if args.serve:
        from lauematching_amd import synth as sy

        t0 = time.time()
        dbs = {("fcc", 2048): db}
        Ohex = sy.random_orientations_reduced(n_hex, args.seed + 7, sy.hexagonal_ops())
        for cr, det in [("fcc", 1024), ("hex", 2048), ("hex", 1024)]:
            pp = serve_params(cr, det)
            hk2 = sy.fcc_hkls(8, "box") if cr == "fcc" else sy.hcp_hkls(6)
            d2 = api.OrientationDB(ctx, O if cr == "fcc" else Ohex, hk2,
                                   api.Geometry.from_params(pp), 64)
            d2.forward(api.LM_FWD_FP32_GUARD)
            if not args.no_autotune:  # per-DB layouts and per-width costs for batching
                d2.autotune([api.quantize_u8(im)[0] for c, d, im in serve_pool
                             if (c, d) == (cr, det)][:3], max_batch=256, reps=2)
            dbs[(cr, det)] = d2
        keys = list(dbs)
        log(f"serving DBs ready in {time.time() - t0:.1f} s; HBM in use "
            f"{ctx.bytes_in_use() / 1e9:.1f} GB")
        # 70% 2048x2048, 30% 1024x1024; crystals alternate
        reqs = []
        for c, d, img_ in serve_pool:
            u8_, s_ = api.quantize_u8(img_)
            reps = 7 if d == 2048 else 3
            reqs += [(keys.index((c, d)), u8_, s_)] * reps
        serving = run_serving(ctx, [dbs[k] for k in keys], reqs, [1, 20, 100, 200, 800, 1600],
                              2.0, 8192)
        serving["dbs"] = [{"crystal": c, "detector": d, "orientations": dbs[(c, d)].n}
                          for c, d in keys]
        serving["hbm_in_use_gb"] = ctx.bytes_in_use() / 1e9
        for k in keys[1:]:
            dbs[k].close()
    merge = None if args.no_merge else merge_check(ctx, O)
    prep = None if args.no_merge else prep_check(ctx)
    precision = None if args.no_precision else precision_check(db, st.ms)
    c4 = None
    if args.config4:
        c4 = config4(ctx, O, n, max(3, args.steps // 2), batch_imgs)

    # the forward simulation: the reference evaluates about 65 operations per
    # (orientation, hkl) pair and 45 per orientation (divisions and square
    # roots counted as one); this kernel's pre-test skips most of them, so
    # that count over the kernel time is an equivalent rate, not issued FLOPs.
    # The kernel's own bound is VALU issue (busy ~96% in the counters).
    fwd_ops = 65.0 * n * len(hk) + 45.0 * n
    forward = {"ms": st.ms, "pairs": int(st.n_pairs), "projected_pairs": int(st.n_projected),
               "projected_fraction": st.n_projected / st.n_pairs, "spots": int(st.n_spots),
               "pairs_per_s": st.n_pairs / (st.ms / 1e3),
               "ref_op_count": fwd_ops,
               "ref_op_equivalent_tflops": fwd_ops / (st.ms / 1e3) / 1e12,
               "ref_op_equivalent_frac_of_fp64_peak": fwd_ops / (st.ms / 1e3) / 78.6e12,
               "bound": "VALU issue (counters: profiles/counters_100M.md)"}
    log(f"forward simulation: {forward['pairs_per_s']:.3e} (orientation, hkl) pairs/s, "
        f"{forward['ref_op_equivalent_tflops']:.1f} TFLOP/s equivalent by the reference's op "
        f"count; {st.n_projected / st.n_pairs:.3%} of pairs projected past the pre-test")

    cpu, parity = None, None
    if not args.no_cpu:
        sample = min(args.cpu_sample, n)
        cpu = cpu_baseline(db, u8, im_scale, p.MinNrSpots, p.MinIntensity, sample)
        log(f"cpu baseline {cpu['value']:.3e} pairs/s ({cpu['cores']} cores)")
        cpu["forward"] = cpu_forward_baseline(O, hk, p, 1_000_000)
        forward["cpu_port_pairs_per_s"] = cpu["forward"]["pairs_per_s"]
        parity = oracle_parity(db, parity_cases)

    daemon = None
    if not args.no_daemon:
        # the drop-in daemon on its own DB: free this process's table first
        db.close()
        dframes = [img_ for c, d, img_ in serve_pool if (c, d) == ("fcc", 2048)]
        try:
            daemon = daemon_leg(O, hk, p, dframes)
        except Exception as ex:  # reported in the line, the rest stands
            daemon = {"error": str(ex)}
            log(f"daemon leg failed: {ex}")

    line = {
        "metric": "orient×image pairs/s at 100M orients + p50/p99 image latency vs plain GPU",
        "value": value,
        "unit": "orient×image pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "plain GPU drop-in: 100M orientations x 1 image (2048x2048 u8, "
                               "5 grains), FCC |hkl|<=8 (1240), M=64, 5-30 keV",
                   "orientations": n, "images_per_step": 1, "hkls": int(len(hk)),
                   "max_spots": 64, "min_nr_spots": p.MinNrSpots,
                   "min_intensity": p.MinIntensity, "spots_per_orientation": st.n_spots / n,
                   "hits": int(len(h.ids)), "nr_results": int(h.nr_results),
                   "grains_recovered": bool(found),
                   "oracle_parity_rows": parity["rows"] if parity else 0,
                   "oracle_parity_all_equal": parity["all_equal"] if parity else None,
                   "forward_ms": st.ms, "kernel_ms": kernel_ms, "walk_ms": walk_ms,
                   "hbm_copy_GBps": copy_gbps, "h2d_pinned_GBps": h2d_gbps,
                   "image_resident_in_hbm": True,
                   "host_image_ms_per_step": host_step_ms,
                   "single_kernel": single_kernel, "graph_replay": bool(args.graphs)},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic},
        "cpu_baseline": cpu,
    }
    if batched:
        for b in batched:
            b["speedup_vs_plain"] = b["pairs_per_s"] / value
        line["batched"] = batched
    if parity:
        line["oracle_parity"] = parity
    if tune:
        line["autotune"] = tune
    if latency:
        rows = []
        for b, pl in zip(latency["batched"]["poisson"], latency["plain"]["poisson"]):
            rows.append({"rate_images_per_s": b["rate"], "p50_ms": b["p50_ms"],
                         "p99_ms": b["p99_ms"], "plain_p50_ms": pl["p50_ms"],
                         "plain_p99_ms": pl["p99_ms"]})
        line["latency"] = {"per_rate": rows,
                           "closed_loop_images_per_s": latency["batched"]["closed_loop"]["images_per_s"],
                           "plain_closed_loop_images_per_s":
                               latency["plain"]["closed_loop"]["images_per_s"],
                           "detail": latency}
    line["forward"] = forward
    if serving:
        serving["per_rate"] = [
            {"rate_images_per_s": b["rate"], "p50_ms": b["p50_ms"], "p99_ms": b["p99_ms"],
             "plain_p50_ms": pl["p50_ms"], "plain_p99_ms": pl["p99_ms"]}
            for b, pl in zip(serving["batched"]["poisson"], serving["plain"]["poisson"])]
        serving["closed_loop_images_per_s"] = serving["batched"]["closed_loop"]["images_per_s"]
        serving["plain_closed_loop_images_per_s"] = serving["plain"]["closed_loop"]["images_per_s"]
        serving["closed_loop_vs_plain"] = (serving["closed_loop_images_per_s"] /
                                           serving["plain_closed_loop_images_per_s"])
        serving["sustained_vs_plain"] = (
            serving["batched"]["closed_loop"]["sustained_images_per_s"] /
            serving["plain"]["closed_loop"]["sustained_images_per_s"])
        line["serving"] = serving
    if daemon:
        line["daemon"] = daemon
    if precision:
        line["precision"] = precision
    line["nonunit_scale"] = nonunit
    line["sparse_frame"] = sparse_line
    if merge:
        line["merge"] = merge
    if prep:
        line["preprocess"] = prep
    if c4:
        line["config4"] = c4
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
