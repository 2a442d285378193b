"""Host -h rasterization (grid alignment, coarse layers, merged fill classes)
and within candidate generation, bit-exact against the oracle."""
import numpy as np
import pytest

import oracle
from farm_amd import capi, datagen


def _compare(ro, xy, vpr, nlow):
    O = oracle.OracleSet(ro, xy)
    O.rasterize(vpr, hierarchy=True, nlow=nlow)
    H = capi.HostSet.from_rings(ro, xy)
    H.rasterize_h(vpr, nlow)
    failed = O.failed()
    assert np.array_equal(~failed, H.raster_ok())
    L = H.layout()
    layers = 0
    for i in range(len(O)):
        if failed[i]:
            continue
        nl = O.nlayers(i)
        assert L["hdr"][i]["nlayers"] == nl
        for k in range(nl):
            a, b = O.layer(i, k), H.layer(i, k)
            assert np.array_equal(a["low"], b["low"]) and a["step"] == b["step"], (i, k)
            assert (a["dimx"], a["dimy"]) == (b["dimx"], b["dimy"]), (i, k)
            assert np.array_equal(a["status"], b["status"]), (i, k)
            layers += 1
        r = O.raster(i)
        h = L["hdr"][i]
        npix = r["dimx"] * r["dimy"]
        assert np.array_equal(L["status"][h["pix_start"]:h["pix_start"] + npix], r["status"])
        assert np.array_equal(H.areas(i, npix), r["areas"])
        assert h["step"][0] == h["step"][1]
    return layers


@pytest.mark.parametrize("vpr,nlow", [(10, 1), (3, 20), (1, 1)])
def test_lake_layers_match_oracle(vpr, nlow):
    ro, xy = datagen.lake_set(41, 400, hi=1.0)
    assert _compare(ro, xy, vpr, nlow) > 400


def test_layers_around_the_origin_and_complex_shapes():
    ro, xy = datagen.lake_set(7, 200, lo=-2.0, hi=2.0)
    _compare(ro, xy, 10, 1)
    ro, xy = datagen.complex_set(5, 12, hi=3.0, nv_cap=5000)
    _compare(ro, xy, 10, 1)


def test_within_candidates_match_oracle():
    ro, xy = datagen.lake_set(41, 800, hi=1.0)
    O = oracle.OracleSet(ro, xy)
    H = capi.HostSet.from_rings(ro, xy)
    for d in (2.0, 10.0, 30.0):
        pa, pb = oracle.candidates_within(O, d)
        qa, qb = capi.candidates(H, H, expand_km=d, skip_self=True)
        assert len(pa) == len(qa)
        assert set(zip(pa.tolist(), pb.tolist())) == set(zip(qa.tolist(), qb.tolist()))


def _square_ring(x0, x1, k):
    t = np.arange(k) / k
    w = x1 - x0
    bottom = np.stack([x0 + w * t, np.full(k, x0)], 1)
    right = np.stack([np.full(k, x1), x0 + w * t], 1)
    top = np.stack([x1 - w * t, np.full(k, x1)], 1)
    left = np.stack([np.full(k, x0), x1 - w * t], 1)
    return np.vstack([bottom, right, top, left]).astype(np.float32)


@pytest.mark.parametrize("nlow,want_layers", [(1, 6), (10, 4)])
def test_hierarchy_closed_form_on_a_grid_snapped_square(nlow, want_layers):
    # The universal grid step is 1.1e-5 (360 / 2^25 to two significant
    # digits, UniversalGrid.h:29-35), so aligned steps are s = 1.1e-5 * 2^k.
    # A square with corners at (m + 1/4) s and (M + 3/4) s, m odd, M + 1 odd,
    # rasterized at -v 1 with 4 * 64 + 2 vertices (a 16 x 16 initial grid of
    # step 21.5 s / 16 = 1.34 s) snaps to step s (nearer than 2 s), its box to
    # [m s, (M + 1) s] (floor / ceil), 22 x 22 pixels; its boundary pixels
    # are 3/4 covered (class ceil(14 * 0.75) = 11, corners ceil(14 * 9/16) =
    # 8), the rest IN.  Each coarser layer j has step 2^j s, box
    # [floor(m / 2^j), ceil((M + 1) / 2^j)] * 2^j s, and a pixel's class
    # ceil(14 r), r the square's exact cover of the merged pixels (IN where
    # every merged pixel is IN) over the finest pixels merge_area takes.
    # Layers are halved until dimx * dimy <= NLow.
    from fractions import Fraction

    s = 1.1e-5
    for _ in range(10):
        s *= 2.0
    m, M = 1001, 1022
    ring = _square_ring((m + 0.25) * s, (M + 0.75) * s, 64)
    ro, xy = rings_from_list_local([ring])
    O = oracle.OracleSet(ro, xy)
    assert O.rasterize(1, hierarchy=True, nlow=nlow) == -1
    nl = O.nlayers(0)
    assert nl == want_layers
    lo = Fraction(m) + Fraction(1, 4)
    hi = Fraction(M) + Fraction(3, 4)
    checked = 0
    for k in range(nl):
        L = O.layer(0, k)
        j = nl - 1 - k  # 0 = finest
        sj = s * 2 ** j
        assert L["step"] == sj
        f0 = m // 2 ** j
        c1 = -(-(M + 1) // 2 ** j)
        assert L["low"][0] == f0 * sj and L["low"][1] == f0 * sj, (j, L["low"], f0 * sj)
        assert L["dimx"] == L["dimy"] == c1 - f0
        st = L["status"].reshape(L["dimy"], L["dimx"])
        n = c1 - f0
        dim = M + 1 - m  # finest pixels per axis

        def merged(X0):
            # finest pixels a coarse box [X0, X0 + 2^j] merges (Ideal::merge_area,
            # Ideal.cpp:1008-1038): from its clamped low offset to its clamped high
            # offset minus one -- a box reaching past the grid's far edge thus
            # leaves out the last pixel -- and at least one
            c0 = min(max(X0 - m, 0), dim - 1)
            c1 = max(min(max(X0 + 2 ** j - m, 0), dim - 1) - 1, c0)
            return c0, c1

        def cover(c):  # the square's cover of finest pixel c along one axis
            return max(Fraction(0), min(hi, m + c + 1) - max(lo, Fraction(m + c)))

        for y in range(n):
            for x in range(n):
                X0, Y0 = (f0 + x) * 2 ** j, (f0 + y) * 2 ** j  # box in units of s
                cx0, cx1 = merged(X0)
                cy0, cy1 = merged(Y0)
                r = sum(cover(c) for c in range(cx0, cx1 + 1)) * sum(cover(c) for c in range(cy0, cy1 + 1)) / 4 ** j
                # all IN when the merged pixels avoid the boundary ring
                inner = cx0 >= 1 and cy0 >= 1 and cx1 <= dim - 2 and cy1 <= dim - 2
                if inner:
                    want = 15
                else:
                    t = 14 * r
                    if abs(t - round(t)) < Fraction(1, 1000):
                        continue  # class decided by the f32 clipping's last digits
                    want = -(-t.numerator // t.denominator)  # ceil(14 r)
                assert st[y, x] == want, (j, x, y, st[y, x], want, float(r))
                checked += 1
    assert checked > 500


def rings_from_list_local(polys):
    from tests.helpers import rings_from_list

    return rings_from_list(polys)
