"""The drop-in executables (intersect, intersection, within_polygon) on .dt
files written by dump_polygons: their count-found lines against the oracle
on the same inputs."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from farm_amd import capi, datagen
from tests.helpers import check_wkt_areas

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "farm_amd", "bin")


def _dt(tmp_path, name, ro, xy):
    wkt = tmp_path / f"{name}.wkt"
    wkt.write_text("\n".join(datagen.to_wkt_lines(ro, xy)) + "\n")
    dt = tmp_path / f"{name}.dt"
    subprocess.run([os.path.join(BIN, "dump_polygons"), str(wkt), str(dt)], check=True, capture_output=True,
                   timeout=60)
    return str(dt)


def _found(args):
    print("  run", " ".join(os.path.basename(a) for a in args), flush=True)
    p = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr[-2000:]
    m = re.search(r"count-found:\s+(\d+)", p.stderr)
    assert m, p.stderr[-2000:]
    return int(m.group(1))


def test_intersect_and_intersection_executables(tmp_path):
    (ra, xa), (rb, xb) = datagen.star_pair_set(21, 300, 300, 0.0, 30.0)
    s, t = _dt(tmp_path, "src", ra, xa), _dt(tmp_path, "tgt", rb, xb)
    A, B = capi.HostSet.load_dt(s), capi.HostSet.load_dt(t)
    pa, pb = capi.candidates(A, B)
    OA = oracle.OracleSet(ra, xa)
    OB = oracle.OracleSet(rb, xb)
    assert OA.rasterize(10) == -1 and OB.rasterize(10) == -1
    bits, _ = oracle.intersect(OA, OB, pa, pb)
    _, area = oracle.intersection(OA, OB, pa, pb)
    tune = str(tmp_path / "tuning.txt")
    for extra in ([], ["-b", "1000"], ["--gpu-candidates"], ["--gpu-candidates", "-b", "777"],
                  ["--tile", "2", "--chunk", "32"], ["--tune", tune], ["--tune", tune, "-b", "900"]):
        assert _found([os.path.join(BIN, "intersect"), "-s", s, "-t", t, "-r", "-g", "-n", "8"] + extra) == int(
            (bits == 1).sum())
    assert open(tune).readline().strip() == "farm-tuning 1"  # written by the first --tune run
    # -a: the reference's heuristic (IN overlap or a BORDER pair with prob > 1)
    abits, _ = oracle.intersect(OA, OB, pa, pb, 1)
    assert _found([os.path.join(BIN, "intersect"), "-s", s, "-t", t, "-r", "-g", "-a"]) == int((abits == 1).sum())
    for extra in ([], ["-b", "1000"], ["--gpu-candidates"]):
        assert _found([os.path.join(BIN, "intersection"), "-s", s, "-t", t, "-r", "-g", "-n", "8"] + extra) == int(
            (area > 0).sum())
    # --wkt: the polygons of A n B, one line per intersecting pair, their
    # shoelace areas equal to the oracle's areas (helpers.check_wkt_areas)
    want = {(int(i), int(j)): float(x) for i, j, x in zip(pa, pb, area) if x > 0}
    for extra in ([], ["-b", "777", "--gpu-candidates"]):
        out = tmp_path / "ab.wkt"
        _found([os.path.join(BIN, "intersection"), "-s", s, "-t", t, "-r", "-g", "--wkt", str(out)] + extra)
        got = {}
        for line in out.read_text().splitlines():
            i, j, body = line.split("\t", 2)
            got[(int(i), int(j))] = capi.wkt_multipolygon_area(body)
        check_wkt_areas(got, want)


def test_within_polygon_executable(tmp_path):
    ro, xy = datagen.lake_set(23, 600, 0.0, 1.0)
    s = _dt(tmp_path, "lakes", ro, xy)
    O = oracle.OracleSet(ro, xy)
    assert O.rasterize(10, hierarchy=True) == -1
    pa, pb = oracle.candidates_within(O, 10.0)
    bits, _, _ = oracle.within(O, O, pa, pb, 10.0)
    want = int((bits == 1).sum())
    exe = os.path.join(BIN, "within_polygon")
    assert _found([exe, "-s", s, "-r", "-g", "-h", "-n", "8"]) == want
    assert _found([exe, "-s", s, "-r", "-g", "-h", "-n", "8", "-b", "5000"]) == want
    assert _found([exe, "-s", s, "-r", "-g", "-h", "-n", "8", "--gpu-candidates"]) == want
    # without -h the reference dereferences an empty layer list; here it is an error
    p = subprocess.run([exe, "-s", s, "-r", "-g"], capture_output=True, text=True, timeout=60)
    assert p.returncode != 0 and "-h" in p.stderr


def test_intersect_learned_a_from_a_model_file(tmp_path):
    # intersect -a --model FILE runs the learned -a mode with a saved MLP:
    # its count equals the learned mode's on the same pairs through the C ABI
    from farm_amd import learned

    (ra, xa), (rb, xb) = datagen.star_pair_set(24, 300, 300, 0.0, 30.0)
    s, t = _dt(tmp_path, "src", ra, xa), _dt(tmp_path, "tgt", rb, xb)
    A, B = capi.HostSet.load_dt(s), capi.HostSet.load_dt(t)
    pa, pb = capi.candidates(A, B)
    both = capi.HostSet.concat(A, B)
    both.rasterize(10)
    dev = capi.Device(0)
    try:
        dev.upload_polygons(both)
        dev.upload_pairs(pa, pb + len(A))
        feat, dec = dev.pair_features()
        rng = np.random.default_rng(5)
        params = (rng.standard_normal(capi.MLP_NPARAMS) * 0.05).astype(np.float32)
        norm = learned.fit_norm(feat)
        dev.mlp_set_params(params, norm)
        z = dev.mlp_predict_pairs()[:, 0]
        thr = float(np.median(z[dec == -1])) if (dec == -1).any() else 0.0
        want = int(dev.intersect(mode=capi.MODE_APPROX_LEARNED, accept=thr).bits.sum())
    finally:
        dev.close()
    path = str(tmp_path / "mlp.bin")
    learned.save_model(path, params, norm, thr)
    p2, n2, acc, _ = learned.load_model(path)
    assert np.array_equal(p2, params) and np.array_equal(n2, norm) and acc == np.float32(thr)
    exe = os.path.join(BIN, "intersect")
    assert _found([exe, "-s", s, "-t", t, "-r", "-g", "-a", "--model", path]) == want
    # a model without -a, or a file that is not a model, is an error
    p = subprocess.run([exe, "-s", s, "-t", t, "--model", path], capture_output=True, text=True, timeout=60)
    assert p.returncode != 0
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"not a model")
    p = subprocess.run([exe, "-s", s, "-t", t, "-a", "--model", str(bad)], capture_output=True, text=True, timeout=60)
    assert p.returncode != 0
