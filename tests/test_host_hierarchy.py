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
