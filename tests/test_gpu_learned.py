"""GPU: the learned part.  MFMA GEMMs, forward and gradients against the CPU
reference MLP (oracle/mlp_ref.py); training; the learned -a mode against the
exact labels and against the reference heuristic; learned queue order keeps
exact mode bit-exact."""
import numpy as np
import pytest

from farm_amd import capi, learned
from oracle import mlp_ref

pytestmark = pytest.mark.gpu


def _rand_problem(seed, m):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.15, capi.MLP_NPARAMS).astype(np.float32)
    X = rng.normal(0, 1, (m, capi.NFEAT)).astype(np.float32)
    y = (rng.uniform(size=m) < 0.3).astype(np.float32)
    c = rng.normal(8, 3, m).astype(np.float32)
    return p, X, y, c


def test_forward_matches_reference(device):
    # shapes cover partial 64-row tiles and the 2-wide output layer
    for m in (1, 17, 64, 300):
        p, X, _, _ = _rand_problem(m, m)
        device.mlp_set_params(p)
        got = device.mlp_forward(X)
        want, _ = mlp_ref.forward(p.astype(np.float64), X, rounding=True)
        scale = np.abs(want).max() + 1e-6
        assert np.abs(got - want).max() <= 2e-4 * scale, (m, np.abs(got - want).max(), scale)


def test_gradients_match_reference(device):
    for m, lam in ((64, 0.1), (1000, 0.02), (4096, 0.05)):
        p, X, y, c = _rand_problem(7 + m, m)
        device.mlp_set_params(p)
        L, g = device.mlp_loss_grad(X, y, c, lam)
        Lr, gr = mlp_ref.loss_grad(p.astype(np.float64), X, y.astype(np.float64), c.astype(np.float64), lam,
                                   rounding=True)
        assert abs(L - Lr) <= 1e-4 * abs(Lr) + 1e-6
        # per parameter block, relative to the block's largest gradient
        o = 0
        for shp in learned.SHAPES:
            n = int(np.prod(shp))
            blk, ref = g[o:o + n], gr[o:o + n]
            err = np.abs(blk - ref).max() / (np.abs(ref).max() + 1e-12)
            assert err < 3e-3, (m, shp, err)
            o += n


def test_training_reduces_loss(device):
    rng = np.random.default_rng(3)
    X = rng.normal(0, 1, (20000, capi.NFEAT)).astype(np.float32)
    w = rng.normal(0, 1, capi.NFEAT)
    y = (X @ w > 0).astype(np.float32)
    c = (np.abs(X[:, :4]).sum(1) * 2).astype(np.float32)
    device.mlp_set_params(learned.init_params(0))
    losses = device.mlp_train(X, y, c, epochs=6, batch=1024, lr=3e-3, lam=0.05, seed=1)
    assert losses[-1] < 0.3 * losses[0], losses
    z = device.mlp_forward(X[:2000])
    acc = ((z[:, 0] > 0) == (y[:2000] > 0)).mean()
    assert acc > 0.95


@pytest.fixture(scope="module")
def star_model(device):
    # the plain intersect workload's distribution: train on one seeded box,
    # choose the -a threshold on a second, test on a third (fresh seeds)
    train_ps = learned.star_domain(500, 16)
    lab = learned.label_set(device, train_ps)
    norm, losses = learned.train(device, lab, seed=0, epochs=15)
    params = device.mlp_get_params()
    cal_ps = learned.star_domain(600, 40)
    cal = learned.label_set(device, cal_ps)
    test_ps = learned.star_domain(900, 16)
    test_lab = learned.label_set(device, test_ps)
    return {"norm": norm, "params": params, "train": lab, "cal": cal, "test": test_lab, "test_ps": test_ps,
            "losses": losses}


def test_learned_approx_beats_heuristic_at_99_precision(device, star_model):
    M = star_model
    assert M["losses"][-1] < M["losses"][0]
    cal, te = M["cal"], M["test"]
    device.mlp_set_params(M["params"], M["norm"])
    # threshold chosen on held-out calibration pairs, evaluated on fresh seeds
    device.upload_polygons(M["test_ps"].both)  # (label() left the test set uploaded)
    device.upload_pairs(M["test_ps"].pa, M["test_ps"].pb)
    z_te = device.mlp_predict_pairs()
    truth = te.bits.astype(bool)
    z_cal = device.mlp_forward(learned.apply_norm(cal.feat, M["norm"]))
    thr = learned.threshold_for_precision(cal.decided, z_cal[:, 0], cal.bits.astype(bool), 0.99, 0.999)
    pred = learned.decisions(te.decided, z_te[:, 0], thr)
    p_l, r_l = learned.precision_recall(pred, truth)
    # the reference heuristic on the same pairs
    heur = device.intersect(mode=capi.MODE_APPROX_HEURISTIC).bits.astype(bool)
    p_h, r_h = learned.precision_recall(heur, truth)
    # the filter alone (IN overlaps only): the method without refinement or MLP
    p_f, r_f = learned.precision_recall(te.decided == 1, truth)
    print(f"learned P={p_l:.4f} R={r_l:.4f} | heuristic P={p_h:.4f} R={r_h:.4f} | filter-only R={r_f:.4f}")
    assert p_l >= 0.99
    assert r_l > r_h + 0.05 and r_l > r_f + 0.05
    # the device learned mode takes the same decisions as the host rule; a
    # logit exactly at the threshold lies in the (one-point) fallback band
    # [reject, accept] and gets the exact answer (pairs with identical
    # features have identical logits, so such ties do occur)
    r = device.intersect(mode=capi.MODE_APPROX_LEARNED, accept=thr)
    tie = (te.decided == -1) & (z_te[:, 0] == np.float32(thr))
    assert np.array_equal(r.bits.astype(bool)[~tie], pred[~tie])
    assert np.array_equal(r.bits.astype(bool)[tie], truth[tie])


def test_learned_fallback_band_is_exact_inside_band(device, star_model):
    M = star_model
    te = M["test"]
    device.mlp_set_params(M["params"], M["norm"])
    device.upload_polygons(M["test_ps"].both)
    device.upload_pairs(M["test_ps"].pa, M["test_ps"].pb)
    z = device.mlp_predict_pairs()[:, 0]
    lo, hi = -2.0, 2.0
    r = device.intersect(mode=capi.MODE_APPROX_LEARNED, accept=hi, reject=lo)
    band = (te.decided == -1) & (z >= lo) & (z <= hi)
    assert band.sum() > 0
    assert np.array_equal(r.bits[band], te.bits[band])
    dec = te.decided != -1
    assert np.array_equal(r.bits[dec], te.bits[dec])


def test_learned_queue_order_keeps_exact_results(device, star_model):
    M = star_model
    device.mlp_set_params(M["params"], M["norm"])
    device.upload_polygons(M["test_ps"].both)
    device.upload_pairs(M["test_ps"].pa, M["test_ps"].pb)
    a = device.intersect(order=0).bits
    b = device.intersect(order=1).bits
    assert np.array_equal(a, b) and np.array_equal(a, M["test"].bits)


def test_fused_inference_equals_the_three_launch_forward(device, star_model):
    # farm_mlp_predict_pairs runs normalisation and the three layers in one
    # kernel; its logits and cost predictions equal the generic MFMA GEMM
    # path's (farm_mlp_forward on the same normalised features) bit for bit
    M = star_model
    device.mlp_set_params(M["params"], M["norm"])
    device.upload_polygons(M["test_ps"].both)
    device.upload_pairs(M["test_ps"].pa, M["test_ps"].pb)
    fused = device.mlp_predict_pairs()
    feat, _ = device.pair_features()
    three = device.mlp_forward(learned.apply_norm(feat, M["norm"]))
    assert fused.shape == three.shape and len(fused) > 1000
    assert np.array_equal(fused.view(np.uint32), three.view(np.uint32))
