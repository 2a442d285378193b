"""GPU parity: exact intersect (plain and batched kernels, all tile widths)
and the reference -a heuristic, bit-exact against the oracle."""
import numpy as np
import pytest

from farm_amd import capi
from tests.helpers import Workload, rings_from_list, star_pairs

pytestmark = pytest.mark.gpu

MODES = [
    (capi.MODE_EXACT_PLAIN, 0),
    (capi.MODE_EXACT_BATCHED, 1),
    (capi.MODE_EXACT_BATCHED, 2),
    (capi.MODE_EXACT_BATCHED, 4),
    (capi.MODE_EXACT_BATCHED, 8),
]


def _check(dev, w: Workload, modes=MODES, allow_ref_failures=False):
    want, st = w.oracle(0)
    # polygons on which the reference raster builder aborts or never ends:
    # the product must flag exactly the same ones (and answer false on them)
    ok = w.both.raster_ok()
    ref_failed = np.concatenate([w.OA.failed(), w.OB.failed()])
    assert np.array_equal(~ok, ref_failed)
    if not allow_ref_failures:
        assert not (want == 2).any(), "reference would abort on this input"
    done = want != 2
    w.upload(dev)
    for mode, tile in modes:
        print(f"  mode={mode} tile={tile} pairs={len(w.pa)}", flush=True)
        r = dev.intersect(mode=mode, tile_a=tile)
        diff = np.nonzero((r.bits != want) & done)[0]
        assert diff.size == 0, f"mode {mode} tile {tile}: {diff.size} pairs differ, first {diff[:10]}"
        assert not r.bits[~done].any()
        assert r.stats["found"] == int((want == 1).sum())
    return want, st


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_star_pairs_2k_full(device, seed):
    print(f"star pairs 2000 x 2000 seed={seed}", flush=True)
    w = star_pairs(seed)
    want, _ = _check(device, w)
    assert 0 < want.sum() < len(want)


def test_star_pairs_many_seeds(device):
    # 20 smaller seeded instances, every kernel variant
    for seed in range(100, 120):
        print(f"star pairs 400 x 400 seed={seed}", flush=True)
        _check(device, star_pairs(seed, n=400))


def test_heuristic_mode_matches_oracle(device):
    w = star_pairs(5)
    want, _ = w.oracle(1)
    w.upload(device)
    r = device.intersect(mode=capi.MODE_APPROX_HEURISTIC)
    assert np.array_equal(r.bits, want)


def test_autotuned_tile_is_exact(device, tmp_path):
    w = star_pairs(8, n=800)
    want, _ = w.oracle(0)
    w.upload(device)
    cache = str(tmp_path / "tune.txt")
    t = device.autotune(cache=cache)
    assert t["tile_a"] in (1, 2, 4, 8) and t["chunk"] in (8, 16, 32)
    again = device.autotune(cache=cache)  # served from the cache
    assert [again[k] for k in ("tile_a", "chunk", "blocks_per_cu", "shape_bucket")] == \
        [t[k] for k in ("tile_a", "chunk", "blocks_per_cu", "shape_bucket")]
    r = device.intersect()  # tile_a = tile_b = 0: the tuned configuration
    assert np.array_equal(r.bits, want)


@pytest.mark.parametrize("chunk", [8, 32])
def test_chunk_sizes(device, chunk):
    # the refinement with 8- and 32-step target chunks (16 is the default of
    # the tests above) on 10 small seeded instances, every tile width
    for seed in range(130, 140):
        print(f"star pairs 400 x 400 seed={seed} chunk={chunk}", flush=True)
        w = star_pairs(seed, n=400)
        want, _ = w.oracle(0)
        w.upload(device)
        for tile in (1, 2, 4, 8):
            r = device.intersect(mode=capi.MODE_EXACT_BATCHED, tile_a=tile, chunk=chunk)
            diff = np.nonzero(r.bits != want)[0]
            assert diff.size == 0, f"tile {tile} chunk {chunk}: {diff.size} pairs differ, first {diff[:10]}"


def test_tuning_rejects_bad_values(device):
    w = star_pairs(8, n=200)
    w.upload(device)
    with pytest.raises(capi.FarmError):
        device.set_tuning(3, 16, 8)
    with pytest.raises(capi.FarmError):
        device.set_tuning(8, 12, 8)
    with pytest.raises(capi.FarmError):
        device.intersect(mode=capi.MODE_EXACT_BATCHED, tile_a=8, chunk=24)


def test_batch_windows_concatenate(device):
    w = star_pairs(9, n=600)
    want, _ = w.oracle(0)
    w.upload(device)
    n = len(w.pa)
    cuts = [0, 1, 7, n // 3, n // 2 + 5, n]
    got = np.concatenate([device.intersect(b, e).bits for b, e in zip(cuts[:-1], cuts[1:])])
    assert np.array_equal(got, want)
    empty = device.intersect(3, 3)
    assert empty.bits.size == 0 and empty.stats["found"] == 0


def _square(x0, y0, x1, y1):
    return np.array([[x0, y0], [x1, y0], [x1, y1], [x0, y1]], dtype=np.float32)


def test_degenerate_shapes(device):
    # shared edges, touching corners, collinear overlaps, containment,
    # identical polygons, grid-aligned coordinates and one-pixel rasters
    a = [
        _square(10, 10, 20, 20),
        _square(30, 30, 31, 31),
        np.array([[50, 50], [60, 50], [55, 58]], np.float32),
        _square(70, 70, 90, 90),
        _square(1, 1, 1.5, 1.5),
        np.array([[5, 5], [6, 5], [6, 6], [5.5, 5.5], [5, 6]], np.float32),
    ]
    b = [
        _square(20, 12, 25, 18),       # shares the edge x = 20
        _square(31, 31, 32, 32),       # touches at one corner
        np.array([[55, 58], [65, 58], [60, 66]], np.float32),  # touches a vertex
        _square(75, 75, 80, 80),       # strictly inside a[3]
        _square(1, 1, 1.5, 1.5),       # identical
        _square(5.25, 5.25, 5.75, 5.75),
        _square(12, 21, 18, 22),       # near miss
        _square(20.0000001, 10, 21, 20),
    ]
    for vpr in (1, 2, 10):
        print(f"degenerate vpr={vpr}", flush=True)
        w = Workload(rings_from_list(a), rings_from_list(b), vpr=vpr)
        want, _ = _check(device, w, allow_ref_failures=True)
        assert (want != 2).sum() >= 5


def test_zero_length_target_segments(device):
    # Target rings strictly on one side of the line of a source edge, inside
    # that edge's box, each with a segment of length 0: a repeated vertex, or
    # (one-pixel raster) the reference's extra edge onto the repeated closing
    # vertex.  For c == d the reference's segment_intersect takes its
    # collinear branch and answers with the box overlap alone, so both pairs
    # are true; a prune that skipped the chunk by the line side would miss
    # them.
    a = [np.array([[0, 0], [1, 0], [1, 1]], np.float32)]
    b = [np.array([[0.45, 0.5], [0.45, 0.5], [0.40, 0.55], [0.35, 0.50]], np.float32),
         np.array([[0.45, 0.5], [0.40, 0.55], [0.35, 0.50]], np.float32)]
    w = Workload(rings_from_list(a), rings_from_list(b), pairs=([0, 0], [0, 1]))
    want, _ = _check(device, w)
    assert list(want) == [1, 1]


@pytest.mark.parametrize("bits", [3, 8])
def test_category_bits(device, bits):
    # -w b: fill classes at 2^b categories.  Exact bits do not change; the
    # reference -a decodes the finer classes (oracle mode 1 at 2^b)
    w = star_pairs(6, n=600)
    want, _ = w.oracle(0)
    import oracle

    heur, _ = oracle.intersect(w.OA, w.OB, w.pa, w.pb, 1, category_bits=bits)
    heur16, _ = oracle.intersect(w.OA, w.OB, w.pa, w.pb, 1)
    w.both.set_categories(bits)
    try:
        w.upload(device)
        assert np.array_equal(device.intersect(mode=capi.MODE_EXACT_BATCHED).bits, want)
        got = device.intersect(mode=capi.MODE_APPROX_HEURISTIC).bits
        assert np.array_equal(got, heur), int((got != heur).sum())
        assert not np.array_equal(heur, heur16) or bits == 4  # the category count matters to -a
    finally:
        w.both.set_categories(4)
