"""The learned part: an MLP on each candidate pair's fill-ratio features.

Labels come from the exact path itself: the intersect bit and the segment
tests the batched refinement executed for the pair (farm_intersect_costs).
The network (80 -> 128 -> 128 -> 2, bf16 MFMA products, f32 accumulation) is
trained on the GPU by libfarm's own forward/backward/Adam kernels
(farm_mlp_train).  It serves two jobs:

  * cost head  -> the order of the persistent refinement queue in exact mode
                  (results stay bit-exact: intersect is an OR);
  * logit head -> the approximate (-a) decision for pairs the raster filter
                  cannot decide (an IN overlap decides true, no BORDER pixel
                  pair decides false); pairs whose logit falls in a band
                  around the threshold can fall back to exact refinement.

The workload is the plain intersect workload's distribution: star polygons
at its density, as many seeded instances of 2,000 x 2,000 polygons in
[0,100]^2 (build_pairsets) or as one larger box (star_domain).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from farm_amd import capi, datagen

NFEAT = capi.NFEAT
SHAPES = [(80, 128), (128,), (128, 128), (128,), (128, 2), (2,)]


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights, zero biases, flat in the C ABI layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for shp in SHAPES:
        if len(shp) == 2:
            lim = np.sqrt(6.0 / shp[0])
            parts.append(rng.uniform(-lim, lim, size=shp).astype(np.float32).reshape(-1))
        else:
            parts.append(np.zeros(shp, np.float32))
    p = np.concatenate(parts)
    assert p.size == capi.MLP_NPARAMS
    return p


@dataclass
class PairSet:
    """Many seeded star-pair instances in one uploaded polygon set."""
    both: capi.HostSet
    pa: np.ndarray
    pb: np.ndarray  # already offset into the combined set
    n_instances: int
    on_device: bool = False  # rasterized on the device: it is that device's current set


def build_pairsets(seeds: list[int], n_per_side: int = 2000) -> PairSet:
    offs_s, xys_s, offs_t, xys_t, pairs = [], [], [], [], []
    ns_total = 0
    nt_total = 0
    for sd in seeds:
        (sro, sxy), (tro, txy) = datagen.star_pair_set(sd, n_per_side, n_per_side)
        A = capi.HostSet.from_rings(sro, sxy)
        B = capi.HostSet.from_rings(tro, txy)
        if len(A) != n_per_side or len(B) != n_per_side:
            raise RuntimeError("generator produced degenerate rings")
        a, b = capi.candidates(A, B)
        pairs.append((a.astype(np.int64) + ns_total, b.astype(np.int64) + nt_total))
        offs_s.append((sro, sxy))
        offs_t.append((tro, txy))
        ns_total += len(A)
        nt_total += len(B)

    def cat(parts):
        xy = np.concatenate([p[1] for p in parts])
        base = 0
        ro = [np.zeros(1, np.uint64)]
        for r, x in parts:
            ro.append(r[1:] + np.uint64(base))
            base += len(x)
        return np.concatenate(ro), xy

    sro, sxy = cat(offs_s)
    tro, txy = cat(offs_t)
    ro = np.concatenate([sro, tro[1:] + sro[-1]])
    xy = np.concatenate([sxy, txy])
    both = capi.HostSet.from_rings(ro, xy)
    pa = np.concatenate([p[0] for p in pairs]).astype(np.uint32)
    pb = (np.concatenate([p[1] for p in pairs]) + ns_total).astype(np.uint32)
    return PairSet(both, pa, pb, len(seeds))


def fit_norm(X: np.ndarray) -> np.ndarray:
    mu = X.mean(0)
    sd = X.std(0)
    sd = np.where(sd > 1e-6, sd, 1.0)
    return np.concatenate([mu, 1.0 / sd]).astype(np.float32)


def apply_norm(X: np.ndarray, norm: np.ndarray) -> np.ndarray:
    return ((X - norm[:NFEAT]) * norm[NFEAT:]).astype(np.float32)


@dataclass
class Labeled:
    feat: np.ndarray
    decided: np.ndarray
    bits: np.ndarray
    tests: np.ndarray


def label(dev: capi.Device, ps: PairSet, vpr: int = 10) -> Labeled:
    ps.both.rasterize(vpr)
    dev.upload_polygons(ps.both)
    dev.upload_pairs(ps.pa, ps.pb)
    bits, tests, _ = dev.intersect_costs()
    feat, dec = dev.pair_features()
    return Labeled(feat, dec, bits, tests)


def cost_target(tests: np.ndarray) -> np.ndarray:
    return np.log2(1.0 + tests.astype(np.float64)).astype(np.float32)


def train(dev: capi.Device, lab: Labeled, seed: int = 0, epochs: int = 12, batch: int = 2048, lr: float = 2e-3,
          lam: float = 0.05) -> tuple[np.ndarray, np.ndarray]:
    """Trains on the pairs the filter leaves undecided; returns (norm, epoch losses)."""
    und = lab.decided == -1
    X = lab.feat[und]
    norm = fit_norm(X)
    Xn = apply_norm(X, norm)
    y = lab.bits[und].astype(np.float32)
    c = cost_target(lab.tests[und])
    dev.mlp_set_params(init_params(seed), norm)
    losses = dev.mlp_train(Xn, y, c, epochs, batch, lr, lam, seed)
    return norm, losses


def decisions(decided: np.ndarray, logit: np.ndarray, accept: float) -> np.ndarray:
    out = decided == 1
    und = decided == -1
    out[und] = logit[und] > accept
    return out


def precision_recall(pred: np.ndarray, truth: np.ndarray) -> tuple[float, float]:
    tp = int((pred & truth).sum())
    fp = int((pred & ~truth).sum())
    fn = int((~pred & truth).sum())
    return (tp / max(tp + fp, 1), tp / max(tp + fn, 1))


def split_calibration(lab: Labeled, frac: float = 0.2, seed: int = 0) -> tuple[Labeled, Labeled]:
    """Random split of labeled pairs into a training part and a held-out
    calibration part (for the -a threshold)."""
    rng = np.random.default_rng(seed)
    cal = rng.random(len(lab.bits)) < frac
    take = lambda m: Labeled(lab.feat[m], lab.decided[m], lab.bits[m], lab.tests[m])  # noqa: E731
    return take(~cal), take(cal)


def threshold_for_precision(decided, logit, truth, target: float = 0.99, confidence: float = 0.999) -> float:
    """Smallest logit threshold whose overall precision is at least target
    with the given confidence: the one-sided Clopper-Pearson lower bound of
    the accepted pairs' precision (IN-decided pairs count as accepted) reaches
    target.  Pick it on pairs the network was not trained on."""
    from scipy.stats import beta

    und = decided == -1
    base_tp = int(((decided == 1) & truth).sum())
    base_fp = int(((decided == 1) & ~truth).sum())
    z = logit[und]
    t = truth[und]
    order = np.argsort(-z)
    tp = base_tp + np.cumsum(t[order])
    fp = base_fp + np.cumsum(~t[order])
    with np.errstate(all="ignore"):
        lower = np.where(tp > 0, beta.ppf(1.0 - confidence, tp, fp + 1), 0.0)
    ok = np.nonzero(lower >= target)[0]
    if len(ok) == 0:
        return float(np.inf)
    k = ok[-1]  # accept the top k+1 logits
    # a binary32 value (the device compares binary32 logits against a
    # binary32 threshold): the largest one below the (k+1)-th logit
    zk = np.float32(z[order][k]) if k + 1 < len(z) else np.float32(z.min() - 1.0)
    return float(np.nextafter(zk, np.float32(-np.inf)))


MODEL_MAGIC = b"FARMMLP1"


def save_model(path: str, params: np.ndarray, norm: np.ndarray, accept: float, reject: float | None = None) -> None:
    """A trained MLP for the executables' learned -a mode (intersect -a
    --model FILE): magic, u32 parameter and feature counts, f32 accept and
    reject logits, the parameters and the feature normalisation."""
    params = np.ascontiguousarray(params, np.float32)
    norm = np.ascontiguousarray(norm, np.float32)
    assert params.size == capi.MLP_NPARAMS and norm.size == 2 * NFEAT
    with open(path, "wb") as f:
        f.write(MODEL_MAGIC)
        f.write(np.array([params.size, NFEAT], "<u4").tobytes())
        f.write(np.array([accept, accept if reject is None else reject], "<f4").tobytes())
        f.write(params.astype("<f4").tobytes())
        f.write(norm.astype("<f4").tobytes())


def load_model(path: str) -> tuple[np.ndarray, np.ndarray, float, float]:
    with open(path, "rb") as f:
        if f.read(8) != MODEL_MAGIC:
            raise ValueError(f"{path} is not a trained MLP file")
        n, nf = np.frombuffer(f.read(8), "<u4")
        acc, rej = np.frombuffer(f.read(8), "<f4")
        params = np.frombuffer(f.read(4 * int(n)), "<f4").copy()
        norm = np.frombuffer(f.read(8 * int(nf)), "<f4").copy()
    return params, norm, float(acc), float(rej)


# The plain intersect workload's polygon distribution: 50..500 vertices,
# R ~ LogNormal(ln 0.5, 0.5), r = R (1 + 0.4 u), 2,000 polygons per side in
# [0,100]^2.  Larger sets keep its density in a larger box.
STAR = dict(nv_lo=50, nv_hi=500, r_mu=float(np.log(0.5)), r_sigma=0.5, amp=0.4)
PAIRS_PER_INSTANCE = 4240  # candidate pairs per 2,000 x 2,000 polygons at that density


def star_domain(seed: int, instances: int, threads: int | None = None, vpr: int = 10,
                dev: capi.Device | None = None) -> PairSet:
    """instances x 2,000 source and target star polygons in a box of side
    100 sqrt(instances) (the plain intersect workload's density), generated
    by the library in parallel, with their candidate pairs, rasterized on
    the host, or with dev on that device (farm_rasterize, the same bytes),
    where it becomes the current set."""
    side = 100.0 * float(np.sqrt(instances))
    n = 2000 * instances
    A = capi.HostSet.star_polygons(2 * seed, n, 0.0, side, nthreads=threads, **STAR)
    B = capi.HostSet.star_polygons(2 * seed + 1, n, 0.0, side, nthreads=threads, **STAR)
    pa, pb = capi.candidates(A, B, nthreads=threads)
    both = capi.HostSet.concat(A, B)
    if dev is None:
        both.rasterize(vpr, threads)
    else:
        dev.rasterize_hostset(both, vpr, 0, threads)
    return PairSet(both, pa, (pb.astype(np.int64) + len(A)).astype(np.uint32), instances, dev is not None)


def load(dev: capi.Device, ps: PairSet) -> None:
    """Make ps the device's polygon set and pair list."""
    if not ps.on_device:
        dev.upload_polygons(ps.both)
    dev.upload_pairs(ps.pa, ps.pb)


def label_set(dev: capi.Device, ps: PairSet) -> Labeled:
    """Exact labels (bits, executed tests) and features of an uploaded set."""
    load(dev, ps)
    bits, tests, _ = dev.intersect_costs()
    feat, dec = dev.pair_features()
    return Labeled(feat, dec, bits, tests)


def config5(dev: capi.Device, train_pairs: int = 2_000_000, eval_pairs: int = 10_000_000, cal_pairs: int = 500_000,
            batch_pairs: int = 1_000_000, epochs: int = 8, seed: int = 5, threads: int | None = None,
            log=None) -> tuple[dict, np.ndarray, np.ndarray]:
    """The learned part on the plain intersect workload's distribution:
    train on train_pairs pairs labeled by the exact path, choose the -a
    threshold for 99% precision (Clopper-Pearson, 99.9% confidence) on
    cal_pairs held-out pairs, then evaluate on eval_pairs fresh pairs in
    batches: precision and recall of learned -a, of the reference heuristic
    -a and of the filter alone against the exact bits, and each mode's
    pairs/s (every mode over the same uploaded batch).  Returns the results
    and the trained model (parameters, feature normalisation)."""
    import time

    say = log or (lambda m: None)
    t0 = time.time()
    tr = label_set(dev, star_domain(1000 + seed, max(1, round(train_pairs / PAIRS_PER_INSTANCE)), threads, dev=dev))
    norm, losses = train(dev, tr, seed=seed, epochs=epochs)
    params = dev.mlp_get_params()
    cal = label_set(dev, star_domain(2000 + seed, max(1, round(cal_pairs / PAIRS_PER_INSTANCE)), threads, dev=dev))
    dev.mlp_set_params(params, norm)
    z_cal = dev.mlp_forward(apply_norm(cal.feat, norm))[:, 0]
    thr = threshold_for_precision(cal.decided, z_cal, cal.bits.astype(bool), 0.99, 0.999)
    t_train = time.time() - t0
    say(f"config5: trained on {int((tr.decided == -1).sum())} undecided of {len(tr.bits)} pairs, "
        f"calibrated on {len(cal.bits)}, threshold {thr:.3f}, {t_train:.1f}s")
    counts = {k: np.zeros(3, np.int64) for k in ("learned_a", "heuristic_a", "filter_only")}
    secs = {k: 0.0 for k in ("exact", "learned_a", "heuristic_a")}
    n_eval, b = 0, 0
    while n_eval < eval_pairs:
        inst = max(1, round(min(batch_pairs, eval_pairs - n_eval) / PAIRS_PER_INSTANCE))
        ps = star_domain(100_000 + 1000 * seed + b, inst, threads, dev=dev)
        load(dev, ps)
        n = len(ps.pa)
        res = {}
        for name, kw in (("exact", dict(mode=capi.MODE_EXACT_BATCHED)),
                         ("heuristic_a", dict(mode=capi.MODE_APPROX_HEURISTIC)),
                         ("learned_a", dict(mode=capi.MODE_APPROX_LEARNED, accept=thr))):
            dev.intersect(**kw, want_bits=False)  # warm (buffers sized for this batch)
            st = dev.intersect(**kw).stats
            secs[name] += st["total_ms"] / 1000.0
        truth = dev.intersect(mode=capi.MODE_EXACT_BATCHED).bits.astype(bool)
        res["heuristic_a"] = dev.intersect(mode=capi.MODE_APPROX_HEURISTIC).bits.astype(bool)
        res["learned_a"] = dev.intersect(mode=capi.MODE_APPROX_LEARNED, accept=thr).bits.astype(bool)
        _, dec = dev.pair_features()
        res["filter_only"] = dec == 1
        for k, pred in res.items():
            counts[k] += [int((pred & truth).sum()), int((pred & ~truth).sum()), int((~pred & truth).sum())]
        n_eval += n
        b += 1
        say(f"config5: evaluated {n_eval} pairs ({b} batches)")
    out = {"distribution": "plain intersect workload (star polygons, 50..500 vertices, 2,000 per side per "
                           "[0,100]^2), -v 10",
           "train_pairs": int(len(tr.bits)), "train_undecided": int((tr.decided == -1).sum()),
           "calibration_pairs": int(len(cal.bits)), "eval_pairs": int(n_eval), "eval_batches": b,
           "threshold": thr, "calibration": "Clopper-Pearson lower bound >= 0.99 at 99.9% confidence",
           "epochs": epochs, "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
           "train_seconds": t_train}
    for k, (tp, fp, fn) in counts.items():
        out[k] = {"precision": tp / max(tp + fp, 1), "recall": tp / max(tp + fn, 1)}
    for k in ("learned_a", "heuristic_a"):
        out[k]["pairs_per_s"] = n_eval / secs[k]
        out[k]["speedup_vs_exact"] = secs["exact"] / secs[k]
    out["exact_pairs_per_s"] = n_eval / secs["exact"]
    out["learned_vs_heuristic_speed"] = secs["heuristic_a"] / secs["learned_a"]
    return out, params, norm


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description="train the learned -a mode on the plain intersect workload's "
                                             "distribution, evaluate it, and optionally save the model")
    ap.add_argument("--train-pairs", type=int, default=2_000_000)
    ap.add_argument("--eval-pairs", type=int, default=10_000_000)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--save-model", default="", help="write the trained MLP here (for intersect -a --model)")
    a = ap.parse_args()
    dev = capi.Device(0)
    res, params, norm = config5(dev, a.train_pairs, a.eval_pairs, epochs=a.epochs, seed=a.seed,
                                log=lambda m: print(m, flush=True))
    if a.save_model:
        save_model(a.save_model, params, norm, res["threshold"])
    print(json.dumps(res), flush=True)
    dev.close()
