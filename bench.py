"""Exact-intersect throughput on one MI355X.

Workload (default): 20,000 complex polygons (Pareto(1.2) vertex counts in
[1,000, 200,000], 3-octave radial noise) as sources and 500,000 lake-like
polygons (20..2,000 vertices) as targets in a [0,40]^2 box, rasterised with
-v 256 (vertices per pixel, the reference's meaning of -v), candidate pairs
from the host R-tree.  --scale s keeps the polygon densities and shrinks the
counts by s (box side scaled by sqrt(s)).

A step is one pass of the exact intersect (filter + refinement) over one
batch of candidate pairs already resident in HBM; the batches partition the
candidate list.  value = pairs / s over the K timed steps.

Also reported:
  plain     the one-pair-per-workgroup kernel on the first timed batch, and
            the batched kernel without its exact chunk prune (what batching
            alone gains, and what the prune adds)
  learned   the learned part on the plain intersect workload's distribution
            (star polygons, 2,000 per side per [0,100]^2): trained on 2M pairs
            labeled by the exact path, -a at 99% precision (Clopper-Pearson
            calibration) evaluated on 10M fresh pairs against the exact bits,
            beside the reference heuristic -a and the filter alone
  learned_order  the cost head trained on this workload's warmup pairs
            orders the timed steps' refinement queue
  roofline  the refinement kernel: 42 FP32 ops per executed segment test
            (6 orientations x 7 ops) / its event-timed duration, against the
            MI355X FP32 vector peak.  The kernel's exact prune skips most
            tests whole, so roofline.reference_tests_per_s also counts the
            pruned ones (the reference's own work), and roofline.valu_issue
            gives the VALU lane-ops the kernel issues (rocprofv3
            SQ_INSTS_VALU, per executed test) against the issue peak
  cpu_baseline  the CPU oracle (restatement of the reference intersect_cpu
            query phase) on a bounded random sample of the same pairs,
            all host threads; the sample is also checked bit for bit
  rasterize the set is rasterized on the device (farm_rasterize), checked
            byte for byte against the host builder's arrays and used as the
            device set; both times are reported
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

FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
OPS_PER_TEST = 42


def measured_bytes_per_test(tile: int, chunk: int) -> float | None:
    """HBM bytes per executed segment test of k_refine<tile> on this workload,
    from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (already
    scaled by the copy-kernel calibration); None for a width it has no
    figures for."""
    import re

    path = os.path.join(ROOT, "profiles", "k_refine_hbm_traffic_complex_x_lakes.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    f = re.search(rf"fetch_bytes_per_segment_test W{tile} C{chunk}: ([0-9.eE+-]+)", text)
    w = re.search(rf"write_bytes_per_segment_test W{tile} C{chunk}: ([0-9.eE+-]+)", text)
    return float(f.group(1)) + float(w.group(1)) if f and w else None


# one VALU lane-op per lane slot: 256 CUs x 4 SIMDs x 32 lanes per cycle x 2.4 GHz
# (the 157.3 TFLOP/s FP32 vector peak counts each FMA as 2)
VALU_ISSUE_PEAK_TOPS = 256 * 4 * 32 * 2.4e9 / 1e12


def measured_valu_ops_per_test(tile: int, chunk: int) -> float | None:
    """VALU lane-ops k_refine<tile> issues per executed segment test on this
    workload, from the committed rocprofv3 SQ_INSTS_VALU summary; None for a
    width it has no figure for."""
    import re

    try:
        text = open(os.path.join(ROOT, "profiles", "k_refine_valu_complex_x_lakes.md")).read()
    except OSError:
        return None
    m = re.search(rf"valu_lane_ops_per_segment_test W{tile} C{chunk}: ([0-9.eE+-]+)", text)
    return float(m.group(1)) if m else None


def measured_bytes_per_pair(name: str) -> float | None:
    """HBM bytes per candidate pair of a query kernel, from its committed
    FETCH_SIZE / WRITE_SIZE summary (already scaled by the copy-kernel
    calibration)."""
    import re

    try:
        text = open(os.path.join(ROOT, "profiles", name)).read()
    except OSError:
        return None
    f = re.search(r"fetch_bytes_per_pair: ([0-9.eE+-]+)", text)
    w = re.search(r"write_bytes_per_pair: ([0-9.eE+-]+)", text)
    return float(f.group(1)) + float(w.group(1)) if f and w else None


T_START = time.time()


def log(msg: str) -> None:
    print(f"[bench {time.time() - T_START:6.1f}s] {msg}", file=sys.stderr, flush=True)


CAND_TIMES: dict = {}
WITHIN_TRAFFIC = "k_within_hbm_traffic_lakes_300k_10km.md"
INTERSECTION_TRAFFIC = "k_intersection_hbm_traffic_complex_x_lakes.md"


def device_candidates(dev, mbb_a, mbb_b, pa, pb, b_base: int, **kw) -> dict:
    """The same candidate pairs built on the device (farm_generate_pairs),
    timed and compared with the host R-tree's; they stay as the pair list."""
    dev.generate_pairs(mbb_a, mbb_b, b_base=b_base, **kw)  # first call allocates
    t = time.time()
    n = dev.generate_pairs(mbb_a, mbb_b, b_base=b_base, **kw)
    dt = time.time() - t
    ga, gb = dev.download_pairs()
    same = bool(n == len(pa) and np.array_equal(ga, pa) and np.array_equal(gb, pb.astype(np.uint32) + np.uint32(b_base)))
    log(f"device candidate pairs: {n} in {dt * 1000:.1f} ms, identical to the host R-tree's: {same}")
    if not same:
        raise RuntimeError("device candidate pairs differ from the host R-tree's")
    return {"pairs": int(n), "host_rtree_s": CAND_TIMES.get("host_rtree_s"), "device_s": dt, "identical": same}


def build_workload(scale: float, seed: int, vpr: int, threads: int):
    from farm_amd import capi

    side = 40.0 * np.sqrt(scale)
    n_c = max(1, int(round(20000 * scale)))
    n_l = max(1, int(round(500000 * scale)))
    t = time.time()
    A = capi.HostSet.complex_polygons(seed, n_c, 0.0, side, nthreads=threads)
    B = capi.HostSet.lake_polygons(seed + 1, n_l, 0.0, side, nthreads=threads)
    log(f"generated {n_c} complex + {n_l} lakes on {threads} host threads in {time.time() - t:.1f}s")
    t = time.time()
    pa, pb = capi.candidates(A, B, nthreads=threads)
    CAND_TIMES["host_rtree_s"] = time.time() - t
    log(f"{len(pa)} candidate pairs (host R-tree, {threads} threads) in {CAND_TIMES['host_rtree_s']:.1f}s")
    t = time.time()
    both = capi.HostSet.concat(A, B)
    both.rasterize(vpr, threads)
    RASTER_TIMES["host_s"] = time.time() - t
    bad = both.raster_failures()
    log(f"rasterized {len(both)} polygons on {threads} host threads in {RASTER_TIMES['host_s']:.1f}s "
        f"(first rejected polygon: {bad})")
    return (A, B), A, B, both, pa, pb


RASTER_TIMES: dict = {}


def device_rasterize(dev, H, vpr: int, nlow: int, threads: int) -> dict:
    """The set rasterized on the device (farm_rasterize) becomes the device
    set; its arrays must equal the host builder's byte for byte."""
    t = time.time()
    info = dev.rasterize_hostset(H, vpr, nlow, threads)
    wall = time.time() - t
    D = dev.download_polyset()
    L = H.layout()
    same = D["hdr"].tobytes() == L["hdr"].tobytes()
    for k in ("status", "pixoff", "seqs", "voff", "vnodes"):
        same = same and np.array_equal(D[k], L[k])
    if nlow:
        same = same and D["layers"].tobytes() == L["layers"].tobytes() and np.array_equal(D["lstatus"], L["lstatus"])
        same = same and np.array_equal(D["areas"].view(np.uint64), L["areas"].view(np.uint64))
    if not same:
        raise RuntimeError("the device-built raster set differs from the host builder's")
    log(f"device rasterization: {len(H)} polygons, {info['nverts']} vertices in {wall:.2f}s (streamed upload "
        f"{info['upload_ms']:.0f} ms through pinned staging, build {info['build_ms']:.0f} ms), identical to the "
        f"host's {RASTER_TIMES.get('host_s', float('nan')):.1f}s on {threads} threads")
    return {"polygons": len(H), "vertices": int(info["nverts"]), "vpr": vpr, "nlow": nlow,
            "host_s": RASTER_TIMES.get("host_s"), "host_threads": threads, "device_s": wall,
            "device_build_ms": info["build_ms"], "upload_ms": info["upload_ms"],
            "upload_gb_per_s": 8.0 * info["nverts"] / max(info["upload_ms"], 1e-9) / 1e6,
            "rejected": info["nrejected"], "identical": True}


ONE_THREAD_SECONDS = 4.0


def host_cpu() -> dict:
    """CPU model and core count of this host (the CPU baselines run here)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu": model, "nproc": os.cpu_count()}


def one_thread(run, n_sample: int, dt: float, threads: int) -> dict:
    """The oracle on one thread over the first m pairs of the timed sample, m
    chosen to take about ONE_THREAD_SECONDS; run(m) returns (seconds, whether
    the answers equal the multi-thread run's)."""
    m = int(min(n_sample, max(20, n_sample * ONE_THREAD_SECONDS / max(dt * threads, 1e-3))))
    t1, same = run(m)
    return {"value": m / t1, "unit": "candidate pairs/s", "cores": 1, "pairs": m, "seconds": t1,
            "same_answers": bool(same)}


def cpu_baseline(rings, pa, pb, vpr, threads, target_s, rng, plain_bits_fn):
    """Oracle query-phase throughput on a random sample of the pairs."""
    import oracle

    HA, HB = rings
    n = len(pa)
    sample = np.sort(rng.choice(n, size=min(n, 400), replace=False))
    results = {}
    # grow the sample until the timed oracle query takes target_s
    while True:
        ua, ia = np.unique(pa[sample], return_inverse=True)
        ub, ib = np.unique(pb[sample], return_inverse=True)
        OA = oracle.OracleSet(*HA.subset_rings(ua))
        OB = oracle.OracleSet(*HB.subset_rings(ub))
        OA.rasterize(vpr, nthreads=threads)
        OB.rasterize(vpr, nthreads=threads)
        t = time.time()
        bits, st = oracle.intersect(OA, OB, ia.astype(np.uint32), ib.astype(np.uint32), 0, threads)
        dt = time.time() - t
        results = {"sample": sample, "bits": bits, "seconds": dt, "stats": st}
        if dt >= target_s or len(sample) >= n:
            break
        grow = int(len(sample) * min(8.0, max(1.5, target_s / max(dt, 1e-3))))
        sample = np.sort(rng.choice(n, size=min(n, grow), replace=False))

    def run1(m):
        t1 = time.time()
        b1, _ = oracle.intersect(OA, OB, ia[:m].astype(np.uint32), ib[:m].astype(np.uint32), 0, 1)
        return time.time() - t1, np.array_equal(b1, results["bits"][:m])

    results["one_thread"] = one_thread(run1, len(sample), results["seconds"], threads)
    return results


FP64_PEAK_TFLOPS = 78.6
# algorithmic FP64 operations of the within query (counted from the
# reference formulas): a point-segment distance (geometry_computation.h:49-89)
# is 4 f32->f64 differences, dot and squared length (6), the projection
# parameter (1 division), the foot point (4), the offsets (2), two unit
# divisions and the squared norm (5): 22; a box-pair distance (Box.cpp:89-149)
# is the pixel boxes of both sides (8), the axis gaps (2), two divisions, the
# squared norm and the square root (5): 15
OPS_PER_PT_SEG = 22
OPS_PER_BOX_PAIR = 15


def within_workload(scale: float, seed: int, threads: int):
    from farm_amd import capi

    side = 20.0 * np.sqrt(scale)
    n = max(2, int(round(300000 * scale)))
    t = time.time()
    H = capi.HostSet.lake_polygons(seed + 3, n, 0.0, side, nthreads=threads)
    log(f"within: generated {n} lakes in a [0,{side:.1f}]^2 box in {time.time() - t:.1f}s")
    t = time.time()
    H.rasterize_h(10, 1, threads)
    RASTER_TIMES["host_s"] = time.time() - t
    log(f"within: -h rasterized (-v 10, --NLow 1) on {threads} host threads in {RASTER_TIMES['host_s']:.1f}s")
    t = time.time()
    pa, pb = capi.candidates(H, H, expand_km=10.0, skip_self=True, nthreads=threads)
    CAND_TIMES["host_rtree_s"] = time.time() - t
    log(f"within: {len(pa)} ordered candidate pairs (MBB expanded by 10 km, host R-tree) in "
        f"{CAND_TIMES['host_rtree_s']:.1f}s")
    return H, pa, pb


def within_cpu_baseline(H, pa, pb, threads, target_s, rng):
    """Oracle within throughput on a random sample of the pairs (query phase)."""
    import oracle

    n = len(pa)
    sample = np.sort(rng.choice(n, size=min(n, 2000), replace=False))
    while True:
        u, inv = np.unique(np.concatenate([pa[sample], pb[sample]]), return_inverse=True)
        O = oracle.OracleSet(*H.subset_rings(u))
        O.rasterize(10, hierarchy=True, nlow=1, nthreads=threads)
        ia = inv[: len(sample)].astype(np.uint32)
        ib = inv[len(sample):].astype(np.uint32)
        t = time.time()
        bits, md, st = oracle.within(O, O, ia, ib, 10.0, threads)
        dt = time.time() - t
        res = {"sample": sample, "bits": bits, "md": md, "seconds": dt}
        if dt >= target_s or len(sample) >= n:

            def run1(m):
                t1 = time.time()
                b1, md1, _ = oracle.within(O, O, ia[:m], ib[:m], 10.0, 1)
                return time.time() - t1, (np.array_equal(b1, bits[:m])
                                          and np.array_equal(md1.view(np.uint64), md[:m].view(np.uint64)))

            res["one_thread"] = one_thread(run1, len(sample), dt, threads)
            return res
        grow = int(len(sample) * min(8.0, max(1.5, target_s / max(dt, 1e-3))))
        sample = np.sort(rng.choice(n, size=min(n, grow), replace=False))


def bench_within(args) -> dict:
    """Within-distance self-join: 300,000 lakes in [0,20]^2, 10 km, -h."""
    from farm_amd import capi

    H, pa, pb = within_workload(args.scale, args.seed, args.threads)
    dev = capi.Device(0)
    raster_info = device_rasterize(dev, H, 10, 1, args.threads)
    m = H.mbbs()
    cand_info = device_candidates(dev, m, m, pa, pb, 0, expand_km=10.0, skip_self=True)
    nb = args.steps + args.warmup
    cuts = np.linspace(0, len(pa), nb + 1).astype(np.int64)
    for s in range(args.warmup):
        st = dev.within(10.0, int(cuts[s]), int(cuts[s + 1]), want_results=False)[2]
        log(f"within warmup step {s}: {cuts[s + 1] - cuts[s]} pairs, {st['total_ms']:.1f} ms")
    per = []
    t0 = time.time()
    for s in range(args.warmup, nb):
        st = dev.within(10.0, int(cuts[s]), int(cuts[s + 1]), want_results=False)[2]
        per.append(st)
        log(f"within step {s - args.warmup}: {cuts[s + 1] - cuts[s]} pairs, {st['total_ms']:.1f} ms, "
            f"found {st['found']}, box pairs {st['refine_items']:.3e}, distances {st['segment_tests']:.3e}")
    wall = time.time() - t0
    timed = int(cuts[nb] - cuts[args.warmup])
    kms = sum(p["total_ms"] for p in per)
    ops = sum(p["segment_tests"] * OPS_PER_PT_SEG + p["refine_items"] * OPS_PER_BOX_PAIR for p in per)
    achieved = ops / (kms / 1000.0) / 1e12 if kms > 0 else 0.0
    bpp = measured_bytes_per_pair(WITHIN_TRAFFIC) if args.scale == 1.0 else None
    traffic = bpp * timed / args.steps if bpp else None  # HBM bytes per launch
    out = {
        "metric": "within-distance candidate pairs/s",
        "value": timed / wall,
        "unit": "candidate pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1000.0 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (box and point-segment distances)",
        "data": "synthetic, seeded",
        "config": {"workload": f"within self-join, {len(H)} lakes, 10 km, -h --NLow 1 -v 10",
                   "candidate_pairs": int(len(pa)), "timed_pairs": timed,
                   "found": int(sum(p["found"] for p in per)), "scale": args.scale},
        "point_segment_distances_per_s": sum(p["segment_tests"] for p in per) / wall,
        "box_pair_distances_per_s": sum(p["refine_items"] for p in per) / wall,
        "roofline": {"bound": "valu", "kernel": "k_within", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "traffic_source": WITHIN_TRAFFIC if traffic else None,
                     "ops_per_point_segment": OPS_PER_PT_SEG, "ops_per_box_pair": OPS_PER_BOX_PAIR},
        "candidates": cand_info,
        "rasterize": raster_info,
    }
    if not args.no_cpu:
        bits, md, _ = dev.within(10.0)
        rng = np.random.default_rng(args.seed + 9)
        res = within_cpu_baseline(H, pa, pb, args.threads, args.cpu_seconds, rng)
        samp = res["sample"]
        mism = int((bits[samp] != res["bits"]).sum())
        mdm = int((md[samp].view(np.uint64) != res["md"].view(np.uint64)).sum())
        out["cpu_baseline"] = {"value": len(samp) / res["seconds"], "unit": "candidate pairs/s",
                               "cores": args.threads, "kind": "port",
                               "sample": f"{len(samp)} random candidate pairs of this workload, oracle within "
                                         f"query only (-h rasterization untimed), {res['seconds']:.1f}s",
                               "one_thread": res["one_thread"], "host": host_cpu()}
        out["parity_sample"] = {"pairs": int(len(samp)), "mismatches": mism, "min_dist_mismatches": mdm}
        log(f"within cpu baseline: {len(samp)} pairs in {res['seconds']:.1f}s; mismatches {mism}, "
            f"min-distance mismatches {mdm}")
    dev.close()
    return out


def bench_candidates(args) -> dict:
    """Candidate pairs of the within self-join (300,000 lakes in [0,20]^2,
    boxes expanded by 10 km) built on the device, against the host R-tree."""
    from farm_amd import capi

    side = 20.0 * np.sqrt(args.scale)
    n = max(1, int(round(300000 * args.scale)))
    H = capi.HostSet.lake_polygons(args.seed + 3, n, 0.0, side, nthreads=args.threads)
    t = time.time()
    pa, pb = capi.candidates(H, H, expand_km=10.0, skip_self=True, nthreads=args.threads)
    host_s = time.time() - t
    H.rasterize(1 << 30, args.threads)  # one pixel per polygon: the upload needs a raster, the join does not
    dev = capi.Device(0)
    dev.upload_polygons(H)
    m = H.mbbs()
    for _ in range(args.warmup):
        dev.generate_pairs(m, m, expand_km=10.0, skip_self=True)
    ts = []
    for s in range(args.steps):
        t = time.time()
        npairs = dev.generate_pairs(m, m, expand_km=10.0, skip_self=True)
        ts.append(time.time() - t)
        log(f"candidates step {s}: {npairs} pairs in {ts[-1] * 1000:.1f} ms")
    ga, gb = dev.download_pairs()
    same = bool(np.array_equal(ga, pa) and np.array_equal(gb, pb))
    dev.close()
    wall = sum(ts)
    return {
        "metric": "candidate pairs generated/s (within self-join, 10 km)",
        "value": npairs * args.steps / wall,
        "unit": "candidate pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1000.0 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (box comparisons)",
        "data": "synthetic, seeded",
        "config": {"workload": f"MBB join, {n} lakes, boxes expanded by 10 km, self pairs skipped",
                   "candidate_pairs": int(npairs), "scale": args.scale},
        "host_rtree": {"value": len(pa) / host_s, "unit": "candidate pairs/s", "threads": args.threads,
                       "what": "the whole join on the host R-tree (farm_candidates)"},
        "parity_sample": {"pairs": int(len(pa)), "identical": same},
    }


def bench_intersection(args) -> dict:
    """Intersection (crossing points, boundary pieces, area of A n B):
    5,000 complex x 100,000 lakes in [0,20]^2, -v 256."""
    import oracle
    from farm_amd import capi

    side = 20.0 * np.sqrt(args.scale)
    n_c = max(1, int(round(5000 * args.scale)))
    n_l = max(1, int(round(100000 * args.scale)))
    t = time.time()
    A = capi.HostSet.complex_polygons(args.seed + 5, n_c, 0.0, side, nthreads=args.threads)
    B = capi.HostSet.lake_polygons(args.seed + 6, n_l, 0.0, side, nthreads=args.threads)
    log(f"intersection: generated {n_c} complex + {n_l} lakes in {time.time() - t:.1f}s")
    pa, pb = capi.candidates(A, B, nthreads=args.threads)
    both = capi.HostSet.concat(A, B)
    t = time.time()
    both.rasterize(args.vpr, args.threads)
    log(f"intersection: {len(pa)} candidate pairs; rasterized in {time.time() - t:.1f}s")
    dev = capi.Device(0)
    dev.upload_polygons(both)
    dev.upload_pairs(pa, pb + len(A))
    nb = args.steps + args.warmup
    cuts = np.linspace(0, len(pa), nb + 1).astype(np.int64)
    for s in range(args.warmup):
        st = dev.intersection_area(int(cuts[s]), int(cuts[s + 1]), want_results=False)[2]
        log(f"intersection warmup step {s}: {cuts[s + 1] - cuts[s]} pairs, {st['total_ms']:.1f} ms")
    per = []
    t0 = time.time()
    for s in range(args.warmup, nb):
        st = dev.intersection_area(int(cuts[s]), int(cuts[s + 1]), want_results=False)[2]
        per.append(st)
        log(f"intersection step {s - args.warmup}: {cuts[s + 1] - cuts[s]} pairs, {st['total_ms']:.1f} ms, "
            f"edge pairs {st['segment_tests']:.3e}, crossing records {st['refine_items']}")
    wall = time.time() - t0
    timed = int(cuts[nb] - cuts[args.warmup])
    tests = sum(p["segment_tests"] for p in per)
    kms = sum(p["total_ms"] for p in per)
    # the edge-pair test: 3 f32 differences x 2 coordinates, 3 crosses (3 ops
    # each), the screen (5 compares and 4 products): 24 f32 operations
    ops = 24
    achieved = tests * ops / (kms / 1000.0) / 1e12 if kms > 0 else 0.0
    bpp = measured_bytes_per_pair(INTERSECTION_TRAFFIC) if (args.scale == 1.0 and args.vpr == 256) else None
    traffic = bpp * timed / args.steps if bpp else None  # HBM bytes per launch
    out = {
        "metric": "intersection (area of A n B) candidate pairs/s",
        "value": timed / wall,
        "unit": "candidate pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1000.0 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (crossing tests), f64 (parameters, areas)",
        "data": "synthetic, seeded",
        "config": {"workload": f"intersection, {n_c} complex x {n_l} lakes, -v {args.vpr}",
                   "candidate_pairs": int(len(pa)), "timed_pairs": timed,
                   "found": int(sum(p["found"] for p in per)), "scale": args.scale},
        "edge_pair_tests_per_s": tests / wall,
        "roofline": {"bound": "valu", "kernel": "k_intersection", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic,
                     "traffic_source": INTERSECTION_TRAFFIC if traffic else None, "ops_per_edge_pair": ops},
    }
    if not args.no_cpu:
        counts, area, _ = dev.intersection_area()
        rng = np.random.default_rng(args.seed + 13)
        n = len(pa)
        sample = np.sort(rng.choice(n, size=min(n, 500), replace=False))
        while True:
            ua, ia = np.unique(pa[sample], return_inverse=True)
            ub, ib = np.unique(pb[sample], return_inverse=True)
            OA = oracle.OracleSet(*A.subset_rings(ua))
            OB = oracle.OracleSet(*B.subset_rings(ub))
            OA.rasterize(args.vpr, nthreads=args.threads)
            OB.rasterize(args.vpr, nthreads=args.threads)
            t = time.time()
            wc, wa = oracle.intersection(OA, OB, ia.astype(np.uint32), ib.astype(np.uint32), args.threads)
            dt = time.time() - t
            if dt >= args.cpu_seconds or len(sample) >= n:
                break
            grow = int(len(sample) * min(8.0, max(1.5, args.cpu_seconds / max(dt, 1e-3))))
            sample = np.sort(rng.choice(n, size=min(n, grow), replace=False))
        def run1(m):
            t1 = time.time()
            c1, a1 = oracle.intersection(OA, OB, ia[:m].astype(np.uint32), ib[:m].astype(np.uint32), 1)
            return time.time() - t1, np.array_equal(c1, wc[:m]) and np.array_equal(a1.view(np.uint64),
                                                                                    wa[:m].view(np.uint64))

        single = one_thread(run1, len(sample), dt, args.threads)
        cm = int((counts[sample] != wc).any(1).sum())
        am = int((np.abs(area[sample] - wa) > 1e-9 * np.abs(wa) + 1e-15).sum())
        out["cpu_baseline"] = {"value": len(sample) / dt, "unit": "candidate pairs/s", "cores": args.threads,
                               "kind": "port",
                               "sample": f"{len(sample)} random candidate pairs of this workload, oracle "
                                         f"intersection query only (rasterization untimed), {dt:.1f}s",
                               "one_thread": single, "host": host_cpu()}
        out["parity_sample"] = {"pairs": int(len(sample)), "count_mismatches": cm, "area_mismatches_1e-9": am}
        log(f"intersection cpu baseline: {len(sample)} pairs in {dt:.1f}s; count mismatches {cm}, "
            f"area mismatches {am}")
    dev.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--vpr", type=int, default=256)
    ap.add_argument("--category-bits", type=int, default=8,
                    help="-w: fill classes at 2^b categories for the learned features and -a (exact bits do not depend on it)")
    ap.add_argument("--tile", type=int, default=0, help="source edges per lane (0: autotuned with --chunk)")
    ap.add_argument("--chunk", type=int, default=0, help="target segments per chunk box (with --tile; default 16)")
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the within and intersection sub-benchmarks")
    ap.add_argument("--train-pairs", type=int, default=1_000_000)
    ap.add_argument("--label-pairs", type=int, default=2_000_000)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--c5-train", type=int, default=2_000_000,
                    help="learned part on the plain intersect distribution: training pairs")
    ap.add_argument("--c5-eval", type=int, default=10_000_000,
                    help="learned part on the plain intersect distribution: fresh evaluation pairs (0: skip)")
    ap.add_argument("--workload", choices=["intersect", "within", "intersection", "candidates"], default="intersect")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only time a plain device copy of this many bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        from farm_amd import capi

        dev = capi.Device(0)
        gbps = dev.copy_probe(args.copy_probe, 10)
        dev.close()
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "launches": 11, "gbps_read_plus_write": gbps}))
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; the pair list would be sharded "
                                   "across GPUs by predicted cost, one process per GPU", "n_gpus": args.gpus}))
        return 2
    if args.workload == "within":
        print(json.dumps(bench_within(args)), flush=True)
        return 0
    if args.workload == "intersection":
        print(json.dumps(bench_intersection(args)), flush=True)
        return 0
    if args.workload == "candidates":
        print(json.dumps(bench_candidates(args)), flush=True)
        return 0

    from farm_amd import capi

    rings, A, B, both, pa, pb = build_workload(args.scale, args.seed, args.vpr, args.threads)
    dev = capi.Device(0)
    raster_info = device_rasterize(dev, both, args.vpr, 0, args.threads)
    dev.set_categories(args.category_bits)
    # the pair list is built on the device (and checked against the host's)
    cand_info = device_candidates(dev, A.mbbs(), B.mbbs(), pa, pb, len(A))

    # Every libfarm call returns after its stream has drained (the counters are
    # read back), so host wall-clock brackets exactly the device work.  The
    # timed steps keep results on the device (count only); the bits are read
    # back afterwards for the parity sample.
    nb = args.steps + args.warmup
    cuts = np.linspace(0, len(pa), nb + 1).astype(np.int64)
    if not args.tile:
        # autotune the refinement (tile width, chunk size, workgroups per CU)
        # on part of the warmup window
        t = time.time()
        tun = dev.autotune(0, int(min(cuts[1], 3_000_000)))
        log(f"autotuned tile {tun['tile_a']}, chunk {tun['chunk']}, {tun['blocks_per_cu']} workgroups per CU "
            f"in {time.time() - t:.1f}s")
    else:
        dev.set_tuning(args.tile, args.chunk or 16, 8)
    tun = dev.tuning()
    args.tile, args.chunk, args.blocks_per_cu = tun["tile_a"], tun["chunk"], tun["blocks_per_cu"]

    # Warmup doubles as the learned part's training data: the exact path
    # labels the warmup window (bits + executed tests per pair), the MLP is
    # trained on the device on the pairs the filter leaves undecided.
    order = 0
    learned_info = None
    if not args.no_learned:
        from farm_amd import learned

        t = time.time()
        w0 = int(cuts[0])
        w1 = int(min(cuts[1], cuts[0] + args.label_pairs))
        lbits, ltests, _ = dev.intersect_costs(w0, w1)
        feat, dec = dev.pair_features(w0, w1)
        und = np.nonzero(dec == -1)[0]
        rng = np.random.default_rng(args.seed + 11)
        # training pairs: undecided pairs of the first 80% of the window;
        # the last 20% (all pairs) calibrate the -a threshold
        split = int(len(dec) * 0.8)
        take = np.sort(rng.choice(und[und < split], size=min(int((und < split).sum()), args.train_pairs),
                                  replace=False))
        lab = learned.Labeled(feat[take], dec[take], lbits[take], ltests[take])
        norm, losses = learned.train(dev, lab, seed=args.seed, epochs=args.epochs)
        z_cal = dev.mlp_forward(learned.apply_norm(feat[split:], norm))[:, 0]
        thr = learned.threshold_for_precision(dec[split:], z_cal, lbits[split:].astype(bool), 0.99)
        order = 1
        learned_info = {"train_pairs": int(len(take)), "epochs": args.epochs, "loss_first": float(losses[0]),
                        "loss_last": float(losses[-1]), "threshold": thr, "seconds": time.time() - t}
        log(f"learned: labeled + trained on {len(take)} undecided warmup pairs in {time.time() - t:.1f}s, "
            f"loss {losses[0]:.4f} -> {losses[-1]:.4f}, -a threshold {thr:.3f}")
    for s in range(1 if learned_info else 0, args.warmup):
        st = dev.intersect(int(cuts[s]), int(cuts[s + 1]), capi.MODE_EXACT_BATCHED, args.tile, want_bits=False,
                           order=order).stats
        log(f"warmup step {s}: {cuts[s + 1] - cuts[s]} pairs, {st['total_ms']:.1f} ms")
    # the timed steps leave their bits in device memory (read back after the
    # timed region): they are the ones the parity sample checks
    from farm_amd.hipmem import DeviceBuffer

    t_lo = int(cuts[args.warmup])
    tbuf = DeviceBuffer(int(cuts[nb]) - t_lo)
    per_step = []
    t0 = time.time()
    for s in range(args.warmup, nb):
        st = dev.intersect_dev(tbuf.ptr + int(cuts[s]) - t_lo, int(cuts[s]), int(cuts[s + 1]),
                               capi.MODE_EXACT_BATCHED, args.tile, order=order)
        per_step.append(st)
        log(f"step {s - args.warmup}: {cuts[s + 1] - cuts[s]} pairs, filter {st['filter_ms']:.1f} ms, "
            f"refine {st['refine_ms']:.1f} ms, total {st['total_ms']:.1f} ms, tests {st['segment_tests']:.3e} "
            f"(pruned {st['pruned_tests']:.3e})")
    wall = time.time() - t0
    timed_pairs = int(cuts[nb] - cuts[args.warmup])
    value = timed_pairs / wall
    timed_bits = tbuf.to_host(np.uint8, timed_pairs)
    tbuf.free()
    tests = sum(s["segment_tests"] for s in per_step)
    pruned = sum(s["pruned_tests"] for s in per_step)
    refine_ms = sum(s["refine_ms"] for s in per_step)
    filter_ms = sum(s["filter_ms"] for s in per_step)
    achieved = tests * OPS_PER_TEST / (refine_ms / 1000.0) / 1e12 if refine_ms > 0 else 0.0
    found = sum(s["found"] for s in per_step)
    full_size = args.scale == 1.0 and args.vpr == 256
    bpt = measured_bytes_per_test(args.tile, args.chunk) if full_size else None
    traffic = tests / args.steps * bpt if bpt else None
    opt = measured_valu_ops_per_test(args.tile, args.chunk) if full_size else None
    issued = tests * opt / (refine_ms / 1000.0) / 1e12 if opt and refine_ms > 0 else None

    line = {
        "metric": "exact intersect pairs/s plain vs batched+learned; approx-mode recall at 99% prec",
        "value": value,
        "unit": "candidate pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1000.0 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (segment predicates), f64 (raster windows)",
        "data": "synthetic, seeded",
        "config": {
            "workload": f"exact intersect, {len(A)} complex x {len(B)} lakes, -v {args.vpr} -w {args.category_bits}, "
                        f"batched kernel, {'learned' if order else 'cost-bound'} queue order",
            "category_bits": args.category_bits,
            "candidate_pairs": int(len(pa)),
            "timed_pairs": timed_pairs,
            "found": int(found),
            "scale": args.scale,
            "tile_a": args.tile,
            "chunk": args.chunk,
            "blocks_per_cu": args.blocks_per_cu,
        },
        "segment_tests_per_s": tests / wall,
        "filter_ms": filter_ms,
        "refine_ms": refine_ms,
        "roofline": {
            "bound": "valu",
            "kernel": "k_refine",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": traffic,
            "traffic_source": "profiles/k_refine_hbm_traffic_complex_x_lakes.md (bytes per test x tests per launch)"
            if traffic else None,
            "algorithmic_bytes": sum(s["refine_bytes"] for s in per_step) / args.steps,
            "ops_per_segment_test": OPS_PER_TEST,
            "pruned_tests_per_step": pruned / args.steps,
            "reference_tests_per_s": (tests + pruned) / (refine_ms / 1000.0) if refine_ms > 0 else 0.0,
            "reference_tests_note": "segment tests the reference runs for the same pairs (executed + pruned "
                                    "whole by the exact chunk prune) per second of k_refine: a rate of the "
                                    "reference's work, not a FLOP rate",
            "valu_issue": {"lane_ops_per_test": opt, "achieved": issued, "peak": VALU_ISSUE_PEAK_TOPS,
                           "unit": "tera lane-ops/s", "frac": issued / VALU_ISSUE_PEAK_TOPS,
                           "source": "profiles/k_refine_valu_complex_x_lakes.md (SQ_INSTS_VALU x 64 / tests)"}
            if issued else None,
        },
    }

    # comparisons on the first timed batch
    s = args.warmup
    b0, b1 = int(cuts[s]), int(cuts[s + 1])
    batched_first = per_step[0]["total_ms"] / 1000.0
    if order:
        tq = time.time()
        st0 = dev.intersect(b0, b1, capi.MODE_EXACT_BATCHED, args.tile, want_bits=False, order=0).stats
        t_cost = time.time() - tq
        line["batched_cost_bound_order"] = {"pairs_per_s": (b1 - b0) / t_cost, "batch_pairs": b1 - b0,
                                            "found": int(st0["found"]),
                                            "learned_order_speedup": t_cost / batched_first}
        log(f"cost-bound queue order on step-0 batch: {t_cost * 1000:.1f} ms")
    if not args.no_plain:
        tp = time.time()
        stp = dev.intersect(b0, b1, capi.MODE_EXACT_PLAIN, want_bits=False).stats
        plain_s = time.time() - tp
        plain_rate = (b1 - b0) / plain_s
        # the batched kernel without its exact chunk prune: what batching alone
        # gains over the plain kernel, and what the prune adds on top
        tq = time.time()
        stn = dev.intersect(b0, b1, capi.MODE_EXACT_BATCHED, args.tile, want_bits=False, order=order,
                            flags=capi.FLAG_NO_PRUNE).stats
        noprune_s = time.time() - tq
        line["plain"] = {"pairs_per_s": plain_rate, "batch_pairs": b1 - b0, "found": int(stp["found"]),
                         "found_batched": int(per_step[0]["found"]),
                         "speedup_batched_vs_plain": plain_s / batched_first if batched_first > 0 else None,
                         "batched_no_prune": {"pairs_per_s": (b1 - b0) / noprune_s, "found": int(stn["found"]),
                                              "segment_tests": int(stn["segment_tests"]),
                                              "speedup_vs_plain": plain_s / noprune_s,
                                              "prune_speedup": noprune_s / batched_first}}
        log(f"plain kernel on step-0 batch: {plain_s * 1000:.1f} ms ({plain_rate:.3e} pairs/s); batched without "
            f"the prune {noprune_s * 1000:.1f} ms")

    # exact bits of every pair (untimed, cost-bound queue order): the
    # approximate modes' truth, and a check that the timed steps' bits (the
    # learned queue order) are the same
    bits = dev.intersect(0, len(pa), capi.MODE_EXACT_BATCHED, args.tile).bits
    order_mismatch = int((bits[t_lo:int(cuts[nb])] != timed_bits).sum())
    log(f"timed steps' bits vs an untimed cost-bound-order run: {order_mismatch} mismatches")
    if learned_info is not None:
        from farm_amd import learned

        e1 = int(min(b1, b0 + args.label_pairs))  # evaluation window: fresh pairs of the first timed batch
        truth = bits[b0:e1].astype(bool)
        tq = time.time()
        dev.intersect(b0, e1, capi.MODE_EXACT_BATCHED, args.tile, want_bits=False, order=order)
        t_x = time.time() - tq
        tq = time.time()
        heur = dev.intersect(b0, e1, capi.MODE_APPROX_HEURISTIC).bits.astype(bool)
        t_h = time.time() - tq
        tq = time.time()
        lrn = dev.intersect(b0, e1, capi.MODE_APPROX_LEARNED, accept=learned_info["threshold"]).bits.astype(bool)
        t_l = time.time() - tq
        p_h, r_h = learned.precision_recall(heur, truth)
        p_l, r_l = learned.precision_recall(lrn, truth)
        _, dec = dev.pair_features(b0, e1)
        p_f, r_f = learned.precision_recall(dec == 1, truth)
        learned_info.update({
            "eval_pairs": e1 - b0,
            "exact_pairs_per_s": (e1 - b0) / t_x,
            "learned_a": {"precision": p_l, "recall": r_l, "pairs_per_s": (e1 - b0) / t_l,
                          "speedup_vs_exact": t_x / t_l},
            "heuristic_a": {"precision": p_h, "recall": r_h, "pairs_per_s": (e1 - b0) / t_h,
                            "speedup_vs_exact": t_x / t_h},
            "filter_only": {"precision": p_f, "recall": r_f},
        })
        line["learned_order"] = learned_info
        log(f"this workload's -a: learned P={p_l:.4f} R={r_l:.4f} ({t_l * 1000:.1f} ms) | heuristic P={p_h:.4f} "
            f"R={r_h:.4f} ({t_h * 1000:.1f} ms) | filter-only R={r_f:.4f}")

    line["candidates"] = cand_info
    line["rasterize"] = raster_info

    if not args.no_cpu:
        rng = np.random.default_rng(args.seed + 7)
        res = cpu_baseline(rings, pa, pb, args.vpr, args.threads, args.cpu_seconds, rng, None)
        samp = res["sample"]
        timed = (samp >= t_lo) & (samp < cuts[nb])
        mism = int((timed_bits[samp[timed] - t_lo] != res["bits"][timed]).sum())
        line["cpu_baseline"] = {
            "value": len(samp) / res["seconds"],
            "unit": "candidate pairs/s",
            "cores": args.threads,
            "kind": "port",
            "sample": f"{len(samp)} random candidate pairs of this workload, oracle query phase only "
                      f"(rasterization untimed), {res['seconds']:.1f}s",
            "one_thread": res["one_thread"],
            "host": host_cpu(),
        }
        line["parity_sample"] = {"pairs": int(timed.sum()), "mismatches": mism,
                                 "bits_from": f"the {args.steps} timed steps ({'learned' if order else 'cost-bound'} "
                                              f"queue order, tile {args.tile}, chunk {args.chunk})",
                                 "timed_vs_cost_bound_order_mismatches": order_mismatch}
        log(f"cpu baseline: {len(samp)} pairs in {res['seconds']:.1f}s; sample mismatches vs GPU: {mism}")
    dev.close()
    if not args.no_learned and args.c5_eval > 0:
        # the learned part on the plain intersect workload's distribution at
        # its stated sizes: trained on c5_train pairs labeled by the exact
        # path, -a evaluated on c5_eval fresh pairs against the exact bits
        from farm_amd import learned

        t = time.time()
        d5 = capi.Device(0)
        res5, _, _ = learned.config5(d5, args.c5_train, args.c5_eval, threads=args.threads, log=log)
        d5.close()
        res5["seconds"] = time.time() - t
        line["learned"] = res5
        log(f"config5 -a on {res5['eval_pairs']} pairs: learned P={res5['learned_a']['precision']:.4f} "
            f"R={res5['learned_a']['recall']:.4f} | heuristic P={res5['heuristic_a']['precision']:.4f} "
            f"R={res5['heuristic_a']['recall']:.4f} | filter-only R={res5['filter_only']['recall']:.4f} "
            f"({res5['seconds']:.1f}s)")
    if not args.no_extra:
        # the other two queries at their own full sizes, each with its own
        # throughput, roofline, CPU baseline and parity sample
        import copy

        sub = copy.copy(args)
        sub.steps, sub.warmup, sub.cpu_seconds, sub.scale = 2, 1, 10.0, args.scale
        for name, fn in (("within", bench_within), ("intersection", bench_intersection)):
            r = fn(sub)
            line[name] = {k: r[k] for k in ("metric", "value", "unit", "ms_per_step", "config", "roofline",
                                              "cpu_baseline", "parity_sample", "candidates") if k in r}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
