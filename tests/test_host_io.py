"""Host library on the CPU: .dt byte format (dump_polygons / load), WKT
parsing, candidate pairs against brute force, and the C ABI exports."""
import os
import re
import struct
import subprocess

import numpy as np

from farm_amd import capi, datagen
from tests.helpers import rings_from_list

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "farm_amd", "bin")


def _parse_dt(path):
    """Independent reader of the .dt layout (MyPolygon::encode + PolygonMeta)."""
    data = open(path, "rb").read()
    (n,) = struct.unpack_from("<Q", data, len(data) - 8)
    meta = [struct.unpack_from("<IIQ4d", data, len(data) - 8 - 48 * n + 48 * i) for i in range(n)]
    polys = []
    for size, nv_meta, off, *mbr in meta:
        nh, nv = struct.unpack_from("<QQ", data, off)
        ring = np.frombuffer(data, np.float32, 2 * nv, off + 16).reshape(nv, 2)
        o = off + 16 + 8 * nv
        holes = []
        for _ in range(nh):
            (hn,) = struct.unpack_from("<Q", data, o)
            holes.append(np.frombuffer(data, np.float32, 2 * hn, o + 8).reshape(hn, 2))
            o += 8 + 8 * hn
        assert o - off == size and nv == nv_meta
        polys.append((ring, holes, mbr))
    return polys


def test_wkt_to_dt_bytes(tmp_path):
    (ro, xy), _ = datagen.star_pair_set(31, 50, 1)
    lines = datagen.to_wkt_lines(ro, xy)
    lines[3] = lines[3][:-1] + ", (1 1, 1.5 1, 1.5 1.5, 1 1))"  # one polygon with a hole
    wkt = tmp_path / "in.wkt"
    wkt.write_text("\n".join(lines) + "\n")
    dt = tmp_path / "out.dt"
    subprocess.run([os.path.join(BIN, "dump_polygons"), str(wkt), str(dt)], check=True, capture_output=True)
    polys = _parse_dt(dt)
    assert len(polys) == 50
    bias = np.float64(np.float32(1e-9))
    for i, (ring, holes, mbr) in enumerate(polys):
        want = xy[int(ro[i]):int(ro[i + 1])]
        assert np.array_equal(ring, want)  # shortest round-trip decimals parse back exactly
        assert np.allclose(mbr, [want[:, 0].min() - bias, want[:, 1].min() - bias, want[:, 0].max() + bias,
                                 want[:, 1].max() + bias], rtol=0, atol=0)
        assert len(holes) == (1 if i == 3 else 0)
    # loading keeps file order and rings; saving again gives the same bytes
    H = capi.HostSet.load_dt(str(dt))
    assert len(H) == 50 and np.array_equal(H.ring(7), xy[int(ro[7]):int(ro[8])])
    dt2 = tmp_path / "again.dt"
    H.save_dt(str(dt2))
    assert dt.read_bytes() == dt2.read_bytes()


def test_tiny_rings_dropped_on_load(tmp_path):
    ro, xy = rings_from_list([np.array([[0, 0], [1, 0], [1, 1]], np.float32)])
    H = capi.HostSet.from_rings(np.array([0, 1, 5], np.uint64), np.vstack([[[5, 5]], xy]))
    # a ring of 1 stored vertex decodes to 2 vertices (< 3) and is dropped
    assert len(H) == 1


def _brute_pairs(ma, mb):
    ov = ~((mb[None, :, 0] > ma[:, None, 2]) | (mb[None, :, 2] < ma[:, None, 0]) |
           (mb[None, :, 1] > ma[:, None, 3]) | (mb[None, :, 3] < ma[:, None, 1]))
    i, j = np.nonzero(ov)
    order = np.lexsort((i, j))
    return i[order].astype(np.uint32), j[order].astype(np.uint32)


def test_candidates_equal_brute_force():
    (sa, sxy), (ta, txy) = datagen.star_pair_set(32, 1500, 1700)
    A, B = capi.HostSet.from_rings(sa, sxy), capi.HostSet.from_rings(ta, txy)
    pa, pb = capi.candidates(A, B)
    ba, bb = _brute_pairs(A.mbbs(), B.mbbs())
    assert np.array_equal(pa, ba) and np.array_equal(pb, bb)


def test_candidates_inclusive_touching_boxes():
    # identical f32 edges: boxes touch exactly and must overlap inclusively
    sq = [np.array([[x, 0], [x + 1, 0], [x + 1, 1], [x, 1]], np.float32) for x in range(6)]
    A = capi.HostSet.from_rings(*rings_from_list(sq[0::2]))
    B = capi.HostSet.from_rings(*rings_from_list(sq[1::2]))
    pa, pb = capi.candidates(A, B)
    ba, bb = _brute_pairs(A.mbbs(), B.mbbs())
    assert np.array_equal(pa, ba) and np.array_equal(pb, bb) and len(pa) == 5


def test_candidates_self_join_expand():
    ro, xy = datagen.lake_set(33, 3000, hi=10.0)
    L = capi.HostSet.from_rings(ro, xy)
    pa, pb = capi.candidates(L, L, expand_km=10.0, skip_self=True)
    m = L.mbbs()
    e = m.copy()
    lat_deg_per_km = 360.0 / 40076.0
    # expand by 10 km with the longitude table at each box's low latitude
    lon = np.array([360.0 / (np.sin((90 - k / 10.0) * 3.14159265 / 180) * 40076) for k in range(900)])
    lon = np.array([float(f"{v:.6f}") for v in lon])
    sx = 10.0 * lon[(np.abs(m[:, 1]) * 10).astype(int)]
    e[:, 0] -= sx
    e[:, 2] += sx
    e[:, 1] -= 10.0 * lat_deg_per_km
    e[:, 3] += 10.0 * lat_deg_per_km
    ba, bb = _brute_pairs(m, e)
    keep = ba != bb
    assert np.array_equal(pa, ba[keep]) and np.array_equal(pb, bb[keep])


def test_capi_exports_every_declared_symbol():
    names = set()
    for h in ("farm_capi.h", "farm_host.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(farm_\w+)\s*\(", text, re.M))
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(capi.lib, n)]
    assert not missing, missing
    assert set(capi.EXPORTS) == names


def test_device_calls_fail_cleanly_without_gpu():
    # on a host without a usable device, farm_open reports an error code
    import ctypes as C

    h = C.c_void_p()
    rc = capi.lib.farm_open(99, C.byref(h))
    assert rc == capi.FARM_EDEVICE and not h.value


def _write_dt(path, polys):
    # .dt bytes: records (u64 holes, u64 nv, f32x2 ring, then each hole), the
    # 48-byte meta per polygon (size, vertices, offset, biased box), u64 count
    import struct

    body, meta = bytearray(), bytearray()
    bias = np.float64(np.float32(1e-9))
    for ring, holes in polys:
        off = len(body)
        rec = struct.pack("<QQ", len(holes), len(ring)) + np.asarray(ring, np.float32).tobytes()
        for h in holes:
            rec += struct.pack("<Q", len(h)) + np.asarray(h, np.float32).tobytes()
        body += rec
        r = np.asarray(ring, np.float32).astype(np.float64)
        meta += struct.pack("<IIQdddd", len(rec), len(ring), off, r[:, 0].min() - bias, r[:, 1].min() - bias,
                            r[:, 0].max() + bias, r[:, 1].max() + bias)
    path.write_bytes(bytes(body + meta + struct.pack("<Q", len(polys))))


def test_dt_load_holes_and_dropped_rings(tmp_path):
    rng = np.random.default_rng(5)

    def ring(n, c):
        t = np.sort(rng.random(n)) * 2 * np.pi
        p = np.stack([c[0] + np.cos(t), c[1] + np.sin(t)], 1)
        return np.vstack([p, p[:1]]).astype(np.float32)

    polys = []
    for i in range(40):
        holes = [ring(5 + k, (3.0 * i, 0.5 * k)) for k in range(i % 4)]
        polys.append((ring(20 + i, (3.0 * i, 0.0)), holes))
    tiny = (np.array([[9.0, 9.0]], np.float32), [ring(6, (9.0, 9.0)), ring(7, (9.0, 9.5))])  # dropped, holes too
    full, want = tmp_path / "full.dt", tmp_path / "want.dt"
    _write_dt(full, polys[:17] + [tiny] + polys[17:])
    _write_dt(want, polys)
    H = capi.HostSet.load_dt(str(full))
    assert len(H) == 40
    for i in (0, 16, 17, 39):
        assert np.array_equal(H.ring(i), polys[i][0])
    again = tmp_path / "again.dt"
    H.save_dt(str(again))
    assert again.read_bytes() == want.read_bytes()
