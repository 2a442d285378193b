"""Shared builders for parity tests: one polygon workload run through the
product (host raster + HIP kernels via the C ABI) and through the oracle."""
from __future__ import annotations

import numpy as np

from farm_amd import capi, datagen


def rings_from_list(polys: list[np.ndarray]) -> tuple[np.ndarray, np.ndarray]:
    """Closed rings (first vertex repeated) from a list of (n, 2) vertex arrays."""
    parts, offs = [], [0]
    for p in polys:
        p = np.asarray(p, dtype=np.float32)
        if not np.array_equal(p[0], p[-1]):
            p = np.vstack([p, p[:1]])
        parts.append(p)
        offs.append(offs[-1] + len(p))
    return np.asarray(offs, dtype=np.uint64), np.vstack(parts).astype(np.float32)


class Workload:
    """Sources A, targets B, candidate pairs, product layout and oracle sets."""

    def __init__(self, a_rings, b_rings, vpr: int = 10, pairs=None):
        import oracle

        (self.sa, self.sxy), (self.ta, self.txy) = a_rings, b_rings
        self.A = capi.HostSet.from_rings(self.sa, self.sxy)
        self.B = capi.HostSet.from_rings(self.ta, self.txy)
        if pairs is None:
            self.pa, self.pb = capi.candidates(self.A, self.B)
        else:
            self.pa, self.pb = (np.asarray(pairs[0], np.uint32), np.asarray(pairs[1], np.uint32))
        self.both = capi.HostSet.concat(self.A, self.B)
        self.both.rasterize(vpr)
        self.OA = oracle.OracleSet(self.sa, self.sxy)
        self.OB = oracle.OracleSet(self.ta, self.txy)
        self.bad_a = self.OA.rasterize(vpr)
        self.bad_b = self.OB.rasterize(vpr)

    def oracle(self, mode: int = 0):
        import oracle

        return oracle.intersect(self.OA, self.OB, self.pa, self.pb, mode)

    def upload(self, dev: capi.Device) -> None:
        dev.upload_polygons(self.both)
        dev.upload_pairs(self.pa, self.pb + len(self.A))


def star_pairs(seed: int, n: int = 2000) -> Workload:
    """n x n seeded star polygons (50..500 vertices) and their candidate pairs."""
    a, b = datagen.star_pair_set(seed, n, n)
    return Workload(a, b)


def check_wkt_areas(got: dict, want: dict) -> None:
    """Shoelace areas of the assembled WKT rings against the Green's-theorem
    areas of the same IN pieces.  Where two crossing points lie a few ulps
    apart with a BORDER piece between them, the IN pieces leave a gap: the
    rings bridge it, while the Green's sum over the open chain moves with its
    origin by about gap x distance (up to ~1e-6 in a [0, 20]^2 box).  Every
    other pair agrees to 1e-9."""
    assert set(want) <= set(got)
    close = [abs(got[k] - x) <= 1e-9 * x + 1e-15 for k, x in want.items()]
    assert sum(close) >= 0.97 * len(close), f"{len(close) - sum(close)} of {len(close)} areas differ"
    for k, x in want.items():
        assert abs(got[k] - x) <= 2e-5 * x + 2e-6, (k, got[k], x)
