"""GPU parity on the headline workload's shape: complex polygons (Pareto
vertex counts from 1,000 up, 3-octave radial noise) against lake-like
polygons at -v 256, the exact intersect of every kernel variant -- the plain
kernel, the batched kernel at every tile width, both queue orders, with and
without graph capture, host and device outputs -- bit-exact against the
oracle.  The set is the full-size workload's densities in a smaller box
(80 complex x 2,000 lakes, 75,502 candidate pairs), so that the oracle
finishes in seconds; its long target sequences run the refinement's full
chunks (8, 16 and 32 steps) and its exact chunk prune."""
import os

import numpy as np
import pytest

from farm_amd import capi, datagen

pytestmark = pytest.mark.gpu

SCALE = 0.004
SEED = 4242
VPR = 256


@pytest.fixture(scope="module")
def headline():
    import oracle

    side = 40.0 * np.sqrt(SCALE)
    nc, nl = int(round(20000 * SCALE)), int(round(500000 * SCALE))
    a = datagen.complex_set(SEED, nc, 0.0, side)
    b = datagen.lake_set(SEED + 1, nl, 0.0, side)
    A, B = capi.HostSet.from_rings(*a), capi.HostSet.from_rings(*b)
    pa, pb = capi.candidates(A, B)
    both = capi.HostSet.concat(A, B)
    both.rasterize(VPR)
    assert both.raster_failures() == -1
    OA, OB = oracle.OracleSet(*a), oracle.OracleSet(*b)
    OA.rasterize(VPR)
    OB.rasterize(VPR)
    threads = min(16, os.cpu_count() or 1)
    print(f"headline workload: {nc} complex x {nl} lakes, {len(pa)} pairs; oracle on {threads} threads", flush=True)
    want, st = oracle.intersect(OA, OB, pa, pb, 0, threads)
    heur, _ = oracle.intersect(OA, OB, pa, pb, 1, threads)
    assert 0 < want.sum() < len(want) and not (want == 2).any()
    return {"both": both, "nA": len(A), "pa": pa, "pb": pb, "want": want, "heur": heur, "stats": st}


def _upload(dev, w):
    dev.upload_polygons(w["both"])
    dev.upload_pairs(w["pa"], w["pb"] + w["nA"])


def _agree(got, w, what):
    diff = np.nonzero(got != w["want"])[0]
    assert diff.size == 0, f"{what}: {diff.size} of {len(got)} pairs differ from the oracle, first {diff[:10]}"


@pytest.mark.parametrize("chunk", [8, 16, 32])
@pytest.mark.parametrize("tile", [1, 2, 4, 8])
def test_batched_every_tile_width(device, headline, tile, chunk):
    print(f"batched tile={tile} chunk={chunk}", flush=True)
    _upload(device, headline)
    r = device.intersect(mode=capi.MODE_EXACT_BATCHED, tile_a=tile, chunk=chunk)
    _agree(r.bits, headline, f"batched W={tile} C={chunk}")
    st = r.stats
    assert st["found"] == int(headline["want"].sum())
    assert st["bb_pixel_pairs"] > 0 and st["segment_tests"] > 0
    # the exact prune skips tests whole; executed + pruned never exceeds the
    # reference's own count of tests for the full refinement of every pair
    assert st["pruned_tests"] > 0


def test_plain_kernel(device, headline):
    _upload(device, headline)
    r = device.intersect(mode=capi.MODE_EXACT_PLAIN)
    _agree(r.bits, headline, "plain")


def test_learned_queue_order(device, headline):
    # any cost predictions must leave the bits unchanged: the queue order
    # only decides when a pair's crossing is found
    _upload(device, headline)
    rng = np.random.default_rng(17)
    params = (rng.standard_normal(capi.MLP_NPARAMS) * 0.1).astype(np.float32)
    device.mlp_set_params(params)
    for tile in (4, 8):
        r = device.intersect(mode=capi.MODE_EXACT_BATCHED, tile_a=tile, order=1)
        _agree(r.bits, headline, f"learned order W={tile}")


def test_graphs_off_and_device_output(device, headline):
    from farm_amd.hipmem import DeviceBuffer

    _upload(device, headline)
    n = len(headline["pa"])
    buf = DeviceBuffer(n)
    try:
        device.set_graphs(False)
        r = device.intersect(mode=capi.MODE_EXACT_BATCHED, tile_a=8)
        _agree(r.bits, headline, "graphs off")
        device.set_graphs(True)
        st = device.intersect_dev(buf.ptr, mode=capi.MODE_EXACT_BATCHED, tile_a=8)
        _agree(buf.to_host(np.uint8, n), headline, "device output")
        assert st["found"] == int(headline["want"].sum())
    finally:
        device.set_graphs(True)
        buf.free()


def test_batch_windows(device, headline):
    # uneven windows, each its own call (graph replays with new windows)
    _upload(device, headline)
    n = len(headline["pa"])
    cuts = [0, 1, 1000, n // 3, n // 3 + 17, n - 5, n]
    got = np.concatenate([device.intersect(b, e, tile_a=8).bits for b, e in zip(cuts[:-1], cuts[1:])])
    _agree(got, headline, "windows")


def test_reference_heuristic(device, headline):
    _upload(device, headline)
    r = device.intersect(mode=capi.MODE_APPROX_HEURISTIC)
    diff = np.nonzero(r.bits != headline["heur"])[0]
    assert diff.size == 0, f"-a: {diff.size} pairs differ, first {diff[:10]}"


def test_batched_without_the_prune(device, headline):
    # every (source edge, target chunk) refined in full: the same bits
    _upload(device, headline)
    r = device.intersect(mode=capi.MODE_EXACT_BATCHED, tile_a=8, flags=capi.FLAG_NO_PRUNE)
    _agree(r.bits, headline, "no prune")
    assert r.stats["pruned_tests"] == 0


def test_autotuned_configuration(device, headline, tmp_path):
    # the autotuner's choice (tile width, chunk size, workgroups per CU) is
    # exact, persists per shape bucket and is applied again from the file
    _upload(device, headline)
    n = len(headline["pa"])
    cache = str(tmp_path / "tuning.txt")
    t = device.autotune(0, min(n, 20000), cache=cache)
    print(f"tuned {t}", flush=True)
    assert t["tile_a"] in (1, 2, 4, 8) and t["chunk"] in (8, 16, 32) and t["blocks_per_cu"] in (4, 8, 16)
    assert t["ms"] > 0
    r = device.intersect(mode=capi.MODE_EXACT_BATCHED)  # tile_a = tile_b = 0: the tuned configuration
    _agree(r.bits, headline, "autotuned")
    lines = open(cache).read().split("\n")
    assert lines[0] == "farm-tuning 1"
    assert lines[1].split()[:4] == [str(t["shape_bucket"]), str(t["tile_a"]), str(t["chunk"]),
                                    str(t["blocks_per_cu"])]
    # another configuration, then the file's entry applied back on re-upload
    device.set_tuning(1, 8 if t["chunk"] != 8 else 32, 4)
    assert device.tuning()["tile_a"] == 1
    _upload(device, headline)  # the handle's table still holds the set's bucket (the last set_tuning)
    device.tuning_load(cache)
    back = device.tuning()
    assert (back["tile_a"], back["chunk"], back["blocks_per_cu"]) == (t["tile_a"], t["chunk"], t["blocks_per_cu"])
    r = device.intersect(mode=capi.MODE_EXACT_BATCHED)
    _agree(r.bits, headline, "tuning loaded from the file")
    # served from the file without re-timing
    assert device.autotune(0, min(n, 20000), cache=cache)["tile_a"] == t["tile_a"]
