"""GPU parity: the raster model built on the device (farm_rasterize) is the
host builder's, byte for byte -- headers, fill classes, pixel offsets, edge
sequences, grid-line offsets and nodes, per-pixel areas, and with -h the
coarse layers -- and the rejected polygons are the same.  The host builder
equals the oracle byte for byte (test_oracle_raster.py, test_host_hierarchy.py),
so this pins the device builder to the oracle; one set is also compared with
the oracle directly.  Queries on a device-built set answer as on the
host-built one."""
import numpy as np
import pytest

import oracle
from farm_amd import capi, datagen
from tests.helpers import rings_from_list

pytestmark = pytest.mark.gpu


def _build(dev, H: capi.HostSet, vpr: int, nlow: int = 0):
    off, x, y = H.closed_rings()
    info = dev.rasterize(off, x, y, vpr, nlow)
    return info, dev.download_polyset()


def _same_nodes(a, b):
    # the reference sorts each grid line's nodes with std::sort: -0.0 and
    # 0.0 may come in either order, so nodes compare as numbers
    return a.shape == b.shape and np.array_equal(a, b)


def _check(dev, H: capi.HostSet, vpr: int, nlow: int = 0):
    print(f"  {len(H)} polygons, -v {vpr}" + (f", -h NLow {nlow}" if nlow else ""), flush=True)
    info, D = _build(dev, H, vpr, nlow)
    if nlow:
        H.rasterize_h(vpr, nlow)
    else:
        H.rasterize(vpr)
    L = H.layout()
    ok = H.raster_ok()
    assert info["nrejected"] == int((~ok).sum())
    assert np.array_equal((D["hdr"]["flags"] & capi.HDR_REJECTED) == 0, ok)
    assert D["hdr"].tobytes() == L["hdr"].tobytes(), _first_diff(D["hdr"], L["hdr"])
    for k in ("status", "pixoff", "seqs", "voff"):
        assert np.array_equal(D[k], L[k]), (k, _first_diff(D[k], L[k]))
    assert _same_nodes(D["vnodes"], L["vnodes"])
    areas = np.concatenate([H.areas(i, int(h["dimx"]) * int(h["dimy"])) for i, h in enumerate(L["hdr"])])
    assert np.array_equal(D["areas"].view(np.uint64), areas.view(np.uint64)), _first_diff(D["areas"], areas)
    if nlow:
        assert D["layers"].tobytes() == L["layers"].tobytes()
        assert np.array_equal(D["lstatus"], L["lstatus"])
    return info, D


def _first_diff(a, b):
    a, b = np.asarray(a).reshape(len(a), -1), np.asarray(b).reshape(len(b), -1)
    if a.shape != b.shape or a.dtype != b.dtype:
        return f"shapes {a.shape} {a.dtype} vs {b.shape} {b.dtype}"
    ab = np.ascontiguousarray(a).view(np.uint8).reshape(len(a), -1)
    bb = np.ascontiguousarray(b).view(np.uint8).reshape(len(b), -1)
    d = np.nonzero((ab != bb).any(1))[0]
    return f"{d.size} entries differ, first {d[:8]}"


def test_star_polygons_v10(device):
    (a, b) = datagen.star_pair_set(21, 1500, 1500)
    H = capi.HostSet.concat(capi.HostSet.from_rings(*a), capi.HostSet.from_rings(*b))
    info, _ = _check(device, H, 10)
    assert info["nseqs"] > 0 and info["nvnodes"] > 0


@pytest.mark.parametrize("vpr", [1, 64])
def test_star_polygons_other_resolutions(device, vpr):
    (a, _) = datagen.star_pair_set(22, 400, 1)
    _check(device, capi.HostSet.from_rings(*a), vpr)


def test_complex_and_lakes_v256(device):
    # the headline workload's shapes: long edge sequences, strips of
    # thousands of points (the hash in HBM past 1,024 candidates)
    c = datagen.complex_set(23, 10, 0.0, 4.0, nv_cap=60000)
    lk = datagen.lake_set(24, 600, 0.0, 4.0)
    H = capi.HostSet.concat(capi.HostSet.from_rings(*c), capi.HostSet.from_rings(*lk))
    _check(device, H, 256)
    _check(device, H, 4096)


def test_near_axes_and_negative_coordinates(device):
    # points within 0.0162 of an axis take the tolerant comparison; polygons
    # straddling both axes and tiny ones at the origin
    (ro, xy), _ = datagen.star_pair_set(25, 120, 1, lo=-0.05, hi=0.05)
    xy = (xy * np.float32(0.1)).astype(np.float32)
    _check(device, capi.HostSet.from_rings(ro, xy), 10)
    ro, xy = datagen.lake_set(26, 200, lo=-1.0, hi=1.0)
    _check(device, capi.HostSet.from_rings(ro, xy), 4)


def _square(x0, y0, x1, y1):
    return np.array([[x0, y0], [x1, y0], [x1, y1], [x0, y1]], dtype=np.float32)


def test_degenerate_rings(device):
    # axis-aligned edges on grid lines, one-pixel rings, repeated vertices,
    # a ring revisiting a vertex, collinear runs and rings the reference
    # builder rejects (an edge through a grid corner within one pixel row)
    polys = [
        _square(0, 0, 4, 4),
        _square(10, 10, 10.5, 10.5),
        np.array([[0, 0], [2, 0], [2, 0], [2, 2], [0, 2], [0, 2]], np.float32),
        np.array([[0, 0], [3, 0], [3, 3], [1.5, 1.5], [0, 3], [1.5, 1.5]], np.float32),
        np.array([[0, 0], [1, 0], [2, 0], [3, 0], [3, 1], [0, 1]], np.float32),
        np.array([[5, 5], [6, 5], [6, 6], [5.5, 5.5], [5, 6]], np.float32),
        np.array([[0, 0], [8, 0], [8, 1], [0, 1]], np.float32),
        np.array([[0, 0], [1, 1e-7], [2, 0], [2, 2], [0, 2]], np.float32),
        np.array([[1, 1], [4, 1], [4, 1.0000001], [1, 4]], np.float32),
    ]
    ro, xy = rings_from_list(polys)
    for vpr in (1, 2, 3, 10):
        _check(device, capi.HostSet.from_rings(ro, xy), vpr)


def test_matches_the_oracle_directly(device):
    ro, xy = datagen.lake_set(27, 300, 0.0, 3.0)
    H = capi.HostSet.from_rings(ro, xy)
    O = oracle.OracleSet(ro, xy)
    O.rasterize(16)
    _, D = _build(device, H, 16)
    assert np.array_equal((D["hdr"]["flags"] & capi.HDR_REJECTED) != 0, O.failed())
    for i in range(len(H)):
        h = D["hdr"][i]
        if h["flags"]:
            continue
        r = O.raster(i)
        npix = int(h["dimx"]) * int(h["dimy"])
        ps, po, sq = int(h["pix_start"]), int(h["pixoff_start"]), int(h["seq_start"])
        assert np.array_equal(D["status"][ps:ps + npix], r["status"]), i
        assert np.array_equal(D["pixoff"][po:po + npix + 1], r["pixoff"]), i
        assert np.array_equal(D["seqs"][sq:sq + int(h["nseqs"])], r["seqs"]), i
        vl, vn = int(h["vline_start"]), int(h["vnode_start"])
        assert np.array_equal(D["voff"][vl:vl + int(h["dimx"]) + 2], r["voff"]), i
        assert np.array_equal(D["vnodes"][vn:vn + len(r["vnodes"])], r["vnodes"]), i
        assert np.array_equal(D["areas"][ps:ps + npix], r["areas"]), i


@pytest.mark.parametrize("nlow", [1, 64])
def test_hierarchy(device, nlow):
    ro, xy = datagen.lake_set(28, 500, 0.0, 5.0)
    _check(device, capi.HostSet.from_rings(ro, xy), 10, nlow)


def test_queries_on_a_device_built_set(device):
    # intersect against the oracle and within against the host-built set
    (a, b) = datagen.star_pair_set(29, 600, 600)
    A, B = capi.HostSet.from_rings(*a), capi.HostSet.from_rings(*b)
    pa, pb = capi.candidates(A, B)
    OA, OB = oracle.OracleSet(*a), oracle.OracleSet(*b)
    OA.rasterize(10)
    OB.rasterize(10)
    want, _ = oracle.intersect(OA, OB, pa, pb, 0)
    both = capi.HostSet.concat(A, B)
    _build(device, both, 10)
    device.upload_pairs(pa, pb + len(A))
    for mode in (capi.MODE_EXACT_BATCHED, capi.MODE_EXACT_PLAIN):
        r = device.intersect(mode=mode)
        assert np.array_equal(r.bits, want)
    ro, xy = datagen.lake_set(30, 400, 0.0, 2.0)
    L = capi.HostSet.from_rings(ro, xy)
    qa, qb = capi.candidates(L, L, expand_km=10.0, skip_self=True)
    _build(device, L, 10, 1)
    device.upload_pairs(qa, qb)
    got, gmin, _ = device.within(10.0)
    L.rasterize_h(10, 1)
    device.upload_polygons(L)
    device.upload_pairs(qa, qb)
    want_w, wmin, _ = device.within(10.0)
    assert np.array_equal(got, want_w) and np.array_equal(gmin.view(np.uint64), wmin.view(np.uint64))


@pytest.mark.parametrize("bits", [3, 8])
def test_fill_categories_on_the_device(device, bits):
    # -w b on a device-built set: the host's fill bytes at 2^b categories,
    # and the reference -a at those categories (oracle mode 1)
    (a, b) = datagen.star_pair_set(31, 500, 500)
    A, B = capi.HostSet.from_rings(*a), capi.HostSet.from_rings(*b)
    both = capi.HostSet.concat(A, B)
    _build(device, both, 10)
    device.set_categories(bits)
    D = device.download_polyset()
    both.rasterize(10)
    both.set_categories(bits)
    assert np.array_equal(D["fill"], both.layout()["fill"])
    pa, pb = capi.candidates(A, B)
    OA, OB = oracle.OracleSet(*a), oracle.OracleSet(*b)
    OA.rasterize(10)
    OB.rasterize(10)
    want, _ = oracle.intersect(OA, OB, pa, pb, 1, category_bits=bits)
    device.upload_pairs(pa, pb + len(A))
    assert np.array_equal(device.intersect(mode=capi.MODE_APPROX_HEURISTIC).bits, want)
    device.set_categories(4)


def test_streamed_upload_through_pinned_staging(device):
    # farm_hostset_rasterize_device: the same set as farm_rasterize, with the
    # rings streamed through two pinned buffers -- many chunks, and one ring
    # longer than a buffer (its own pinned buffer)
    c = datagen.complex_set(32, 3, 0.0, 3.0, nv_cap=30000)
    lk = datagen.lake_set(33, 3000, 0.0, 3.0)
    H = capi.HostSet.concat(capi.HostSet.from_rings(*c), capi.HostSet.from_rings(*lk))
    _, want = _build(device, H, 16)
    try:
        device.set_staging(8192)
        info = device.rasterize_hostset(H, 16, nthreads=4)
        got = device.download_polyset()
        assert info["upload_ms"] > 0
        for k in want:
            assert np.asarray(got[k]).tobytes() == np.asarray(want[k]).tobytes(), k
    finally:
        device.set_staging(1 << 22)
