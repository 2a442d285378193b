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

The workload is the star-polygon pair set (many seeded instances of 2,000 x
2,000 polygons in [0,100]^2), so pairs never mix instances.
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


def evaluate(train_pairs: int = 500_000, eval_pairs: int = 2_000_000, target: float = 0.99, epochs: int = 12,
             seed: int = 1) -> dict:
    """Learned -a on the star-pair distribution: train on 80% of
    `train_pairs` labeled pairs of seeded instances, pick the threshold for
    `target` precision (with a 2-sigma margin) on the held-out 20%, evaluate
    on `eval_pairs` pairs of fresh seeds (exact labels, the reference
    heuristic and the filter alone on the same pairs)."""
    import time

    per_instance = 4200  # candidate pairs of one 2,000 x 2,000 instance, about
    n_tr = max(1, train_pairs // per_instance)
    n_ev = max(1, eval_pairs // per_instance)
    dev = capi.Device(0)
    t = time.time()
    tr_ps = build_pairsets([10_000 + seed * 100_000 + i for i in range(n_tr)])
    tr, cal = split_calibration(label(dev, tr_ps), 0.2, seed)
    norm, losses = train(dev, tr, seed=seed, epochs=epochs)
    z_cal = dev.mlp_forward(apply_norm(cal.feat, norm))[:, 0]
    thr = threshold_for_precision(cal.decided, z_cal, cal.bits.astype(bool), target)
    t_train = time.time() - t
    ev_ps = build_pairsets([50_000 + seed * 100_000 + i for i in range(n_ev)])
    ev = label(dev, ev_ps)
    truth = ev.bits.astype(bool)
    res = {"train_pairs": int(len(tr.bits)), "eval_pairs": int(len(ev.bits)), "threshold": thr,
           "train_seconds": t_train, "loss_first": float(losses[0]), "loss_last": float(losses[-1])}
    timings = {}
    for name, kw in (("exact", dict(mode=capi.MODE_EXACT_BATCHED)),
                     ("exact_learned_order", dict(mode=capi.MODE_EXACT_BATCHED, order=1)),
                     ("heuristic_a", dict(mode=capi.MODE_APPROX_HEURISTIC)),
                     ("learned_a", dict(mode=capi.MODE_APPROX_LEARNED, accept=thr))):
        dev.intersect(**kw)  # warm
        tq = time.time()
        bits = dev.intersect(**kw).bits.astype(bool)
        timings[name] = time.time() - tq
        p, r = precision_recall(bits, truth)
        res[name] = {"precision": p, "recall": r, "pairs_per_s": len(truth) / timings[name]}
        if name.startswith("exact"):
            assert np.array_equal(bits, truth)
    for name in ("heuristic_a", "learned_a", "exact_learned_order"):
        res[name]["speedup_vs_exact"] = timings["exact"] / timings[name]
    p, r = precision_recall(ev.decided == 1, truth)
    res["filter_only"] = {"precision": p, "recall": r}
    dev.close()
    return res


def split_calibration(lab: Labeled, frac: float = 0.2, seed: int = 0) -> tuple[Labeled, Labeled]:
    """Random split of labeled pairs into a training part and a held-out
    calibration part (for the -a threshold)."""
    rng = np.random.default_rng(seed)
    cal = rng.random(len(lab.bits)) < frac
    take = lambda m: Labeled(lab.feat[m], lab.decided[m], lab.bits[m], lab.tests[m])  # noqa: E731
    return take(~cal), take(cal)


def threshold_for_precision(decided, logit, truth, target: float = 0.99, zconf: float = 2.0) -> float:
    """Smallest logit threshold whose overall precision, less zconf standard
    errors of its binomial estimate, reaches target (pick it on pairs the
    network was not trained on)."""
    und = decided == -1
    base_tp = int(((decided == 1) & truth).sum())
    base_fp = int(((decided == 1) & ~truth).sum())
    z = logit[und]
    t = truth[und]
    order = np.argsort(-z)
    tp = base_tp + np.cumsum(t[order])
    fp = base_fp + np.cumsum(~t[order])
    n = np.maximum(tp + fp, 1)
    prec = tp / n
    lower = prec - zconf * np.sqrt(prec * (1 - prec) / n)
    ok = np.nonzero(lower >= target)[0]
    if len(ok) == 0:
        return float(np.inf)
    k = ok[-1]  # accept the top k+1 logits
    return float(z[order][k]) - 1e-6 if k + 1 < len(z) else float(z.min()) - 1.0


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description="train and evaluate the learned -a mode on star-polygon pairs")
    ap.add_argument("--train-pairs", type=int, default=500_000)
    ap.add_argument("--eval-pairs", type=int, default=2_000_000)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args()
    print(json.dumps(evaluate(a.train_pairs, a.eval_pairs, epochs=a.epochs, seed=a.seed)), flush=True)
