"""Seeded synthetic polygon sets for the polygon-join workloads.

All sets are simple star-shaped polygons: sorted angles around a centre and a
positive radius per vertex, stored counter-clockwise as closed rings (the
first vertex repeated at the end, as WKT rings are).  Coordinates are drawn in
float64 and rounded to float32, the precision of the .dt format; writing them
as WKT with shortest round-trip decimals and converting with dump_polygons
gives the same .dt bytes.

Sorted uniform angles are drawn as normalised partial sums of exponentials
(the order statistics of uniforms), which avoids a sort per polygon.

Workloads (sizes are the full ones; callers may scale the counts down):
  star_pair_set      2,000 + 2,000 polygons in [0,100]^2, 50..500 vertices,
                     R ~ LogNormal(ln 0.5, 0.5), r = R (1 + 0.4 u)
  complex_set        Pareto(1.2) vertex counts clipped to [1000, 200000],
                     3-octave radial noise of amplitude 0.3,
                     R ~ LogNormal(ln 0.8, 0.6)
  lake_set           20..2000 vertices, R ~ LogNormal(ln 0.05, 0.8),
                     r = R (1 + 0.4 u)
"""
from __future__ import annotations

import numpy as np

TWO_PI = 2.0 * np.pi


def _sorted_angles(rng: np.random.Generator, nv: np.ndarray) -> np.ndarray:
    """Concatenated sorted U[0, 2pi) samples, nv[i] of them per polygon."""
    total = int(nv.sum())
    seg_start = np.concatenate([[0], np.cumsum(nv)[:-1]])
    e = rng.exponential(1.0, size=total + len(nv))
    # per polygon: nv[i] + 1 exponentials; angle_k = 2 pi S_k / S_{n+1}
    ext_start = seg_start + np.arange(len(nv))
    cs = np.cumsum(e)
    base = np.concatenate([[0.0], cs])[ext_start]  # cumsum before the segment
    owner = np.repeat(np.arange(len(nv)), nv)
    idx_in_ext = np.arange(total) - seg_start[owner] + ext_start[owner]
    partial = cs[idx_in_ext] - base[owner]
    norm = cs[ext_start + nv] - base
    u = partial / norm[owner]
    # the first sample is the smallest; shift so angles start in [0, 2pi)
    return (u * TWO_PI) % TWO_PI


def star_polygons(rng: np.random.Generator, n: int, lo: float, hi: float, nv: np.ndarray, r_mu: float,
                  r_sigma: float, noise: str = "uniform", amp: float = 0.4) -> tuple[np.ndarray, np.ndarray]:
    """n star polygons; returns (ring_off uint64[n+1], xy float32[total, 2]) with closed rings."""
    nv = np.asarray(nv, dtype=np.int64)
    cx = rng.uniform(lo, hi, size=n)
    cy = rng.uniform(lo, hi, size=n)
    R = rng.lognormal(mean=r_mu, sigma=r_sigma, size=n)
    th = _sorted_angles(rng, nv)
    owner = np.repeat(np.arange(n), nv)
    if noise == "uniform":
        u = rng.uniform(-1.0, 1.0, size=th.shape[0])
        r = R[owner] * (1.0 + amp * u)
    elif noise == "octave":
        f = rng.integers(2, 7, size=n).astype(np.float64)
        ph = rng.uniform(0.0, TWO_PI, size=(n, 3))
        s = np.zeros_like(th)
        for o in range(3):
            s += 0.5 ** o * np.sin(f[owner] * (2 ** o) * th + ph[owner, o])
        r = R[owner] * (1.0 + amp * s / 1.75)
    else:
        raise ValueError(noise)
    x = cx[owner] + r * np.cos(th)
    y = cy[owner] + r * np.sin(th)
    # closed rings: nv[i] vertices + the first one again
    ring_len = nv + 1
    ring_off = np.concatenate([[0], np.cumsum(ring_len)]).astype(np.uint64)
    xy = np.empty((int(ring_len.sum()), 2), dtype=np.float32)
    pos = np.arange(th.shape[0]) + owner  # skip one slot per previous ring
    xy[pos, 0] = x.astype(np.float32)
    xy[pos, 1] = y.astype(np.float32)
    first = ring_off[:-1].astype(np.int64)
    last = ring_off[1:].astype(np.int64) - 1
    xy[last] = xy[first]
    return ring_off, xy


def star_pair_set(seed: int, n_src: int = 2000, n_tgt: int = 2000, lo: float = 0.0, hi: float = 100.0):
    """Source and target sets of the plain intersect workload."""
    rng = np.random.default_rng(seed)
    out = []
    for n in (n_src, n_tgt):
        nv = rng.integers(50, 501, size=n)
        out.append(star_polygons(rng, n, lo, hi, nv, np.log(0.5), 0.5, "uniform", 0.4))
    return out[0], out[1]


def complex_set(seed: int, n: int, lo: float = 0.0, hi: float = 40.0, nv_cap: int = 200_000):
    rng = np.random.default_rng(seed)
    nv = np.clip(np.floor((rng.pareto(1.2, size=n) + 1.0) * 1000.0), 1000, nv_cap).astype(np.int64)
    return star_polygons(rng, n, lo, hi, nv, np.log(0.8), 0.6, "octave", 0.3)


def lake_set(seed: int, n: int, lo: float = 0.0, hi: float = 40.0, chunk: int = 50_000):
    """Lake-like polygons, generated in chunks to bound temporaries."""
    rng = np.random.default_rng(seed)
    offs, xys = [], []
    base = 0
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        nv = rng.integers(20, 2001, size=m)
        ro, xy = star_polygons(rng, m, lo, hi, nv, np.log(0.05), 0.8, "uniform", 0.4)
        offs.append(ro[:-1] + base)
        xys.append(xy)
        base += int(ro[-1])
    ring_off = np.concatenate(offs + [np.array([base], dtype=np.uint64)]).astype(np.uint64)
    return ring_off, np.concatenate(xys)


def to_wkt_lines(ring_off: np.ndarray, xy: np.ndarray) -> list[str]:
    """POLYGON WKT with shortest round-trip float32 decimals."""
    lines = []
    for i in range(len(ring_off) - 1):
        pts = xy[int(ring_off[i]):int(ring_off[i + 1])]
        body = ", ".join(f"{repr(float(np.float32(x)))} {repr(float(np.float32(y)))}" for x, y in
                         zip(pts[:, 0].tolist(), pts[:, 1].tolist()))
        lines.append(f"POLYGON (({body}))")
    return lines


def subset(ring_off: np.ndarray, xy: np.ndarray, idx: np.ndarray):
    """Rings idx of a set, as a new (ring_off, xy)."""
    idx = np.asarray(idx, dtype=np.int64)
    lens = (ring_off[idx + 1] - ring_off[idx]).astype(np.int64)
    new_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    parts = [xy[int(ring_off[i]):int(ring_off[i + 1])] for i in idx]
    return new_off, (np.concatenate(parts) if parts else np.zeros((0, 2), np.float32))
