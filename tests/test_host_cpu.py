"""Host-side C++ API on the CPU (through the ccs_selftest binary): .npz/.npy
files as numpy writes them, PNG/PNM/PFM files, option precedence and the label
count rule of the facade (stereo/stereo.cpp:64-73, stereo/stereo_img.cpp:47-103)."""
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cnn-crf-stereo_amd", "ccs_selftest")


def run(*args):
    if not os.path.exists(BIN):
        pytest.fail("ccs_selftest is not built (run __graft_entry__.build())")
    r = subprocess.run([BIN, *map(str, args)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout


def parse_dump(text):
    out = {}
    for line in text.strip().splitlines():
        t = line.split()
        i = t.index("n")
        out[t[0]] = dict(shape=tuple(int(v) for v in t[2:i]), n=int(t[i + 1]), sum=float(t[i + 3]),
                         sq=float(t[i + 5]), first=float(t[i + 7]), last=float(t[i + 9]))
    return out


@pytest.mark.parametrize("compressed", [False, True])
def test_npz_written_by_numpy(tmp_path, compressed):
    rng = np.random.default_rng(0)
    arrays = [rng.standard_normal((100, 3, 3, 3)).astype(np.float32), rng.standard_normal(100).astype(np.float32),
              rng.standard_normal((4, 100, 2, 2)), np.float32([0.99]), np.arange(7, dtype=np.float32)]
    p = tmp_path / "params.npz"
    (np.savez_compressed if compressed else np.savez)(p, *arrays)
    d = parse_dump(run("npz-dump", p))
    assert sorted(d) == ["arr_%d" % i for i in range(len(arrays))]
    for i, a in enumerate(arrays):
        e = d["arr_%d" % i]
        f = a.astype(np.float32).ravel()
        assert e["shape"] == a.shape and e["n"] == f.size
        np.testing.assert_allclose(e["sum"], f.astype(np.float64).sum(), rtol=1e-6, atol=1e-6)
        np.testing.assert_allclose(e["sq"], (f.astype(np.float64) ** 2).sum(), rtol=1e-6)
        assert e["first"] == pytest.approx(float(f[0]), rel=1e-7) and e["last"] == pytest.approx(float(f[-1]), rel=1e-7)


def test_npz_and_npy_written_by_the_library(tmp_path):
    p = tmp_path / "w.npz"
    run("npz-write", p)
    z = np.load(p, allow_pickle=False)
    np.testing.assert_array_equal(z["arr_0"], (0.5 * np.arange(24) - 3).reshape(2, 3, 2, 2).astype(np.float32))
    np.testing.assert_array_equal(z["arr_1"], np.float32([1, -2, 3.5, 1e-3, 7]))
    assert z["arr_2"].shape == () and z["arr_2"] == 42
    q = tmp_path / "v.npy"
    run("npy-write", q)
    np.testing.assert_array_equal(np.load(q), z["arr_0"])
    assert parse_dump(run("npy-dump", q))["npy"]["shape"] == (2, 3, 2, 2)


def test_npz_errors(tmp_path):
    p = tmp_path / "bad.npz"
    p.write_bytes(b"not a zip archive at all")
    r = subprocess.run([BIN, "npz-dump", str(p)], capture_output=True, text=True)
    assert r.returncode == 2 and "ZIP" in r.stderr
    q = tmp_path / "i.npz"
    np.savez(q, np.arange(3, dtype=np.int64))
    r = subprocess.run([BIN, "npz-dump", str(q)], capture_output=True, text=True)
    assert r.returncode == 2 and "dtype" in r.stderr


@pytest.mark.parametrize("lo,step,hi,n,mx", [(0, 1, 127, 128, 127), (0, 1, 128, 256, 255), (0, 1, 255, 256, 255),
                                             (0, 1, 300, 256, 255), (1, 1, 128, 128, 128), (0, 2, 127, 64, 126),
                                             (0, 1, 63, 64, 63), (0, 1, 31, 32, 31), (-10, 1, 10, 32, 21)])
def test_label_count(lo, step, hi, n, mx):
    got = run("labels", lo, step, hi).split()
    assert int(got[0]) == n and float(got[1]) == mx


def options(*args):
    return dict(line.split("=", 1) for line in run("options", *args).strip().splitlines())


def test_option_defaults_and_cli_over_file(tmp_path):
    d = options("a.png", "b.png")
    assert (d["disp-min"], d["disp-max"], d["matching"], d["inference"], d["lambda"], d["L1"], d["L2"]) == \
        ("1", "128", "CENSUS", "CRF", "0.4", "1", "12")
    assert d["precision"] == "fp32"  # the reference's arithmetic unless bf16 is asked for
    cfg = tmp_path / "m.cfg"
    cfg.write_text("matching = COLORCNN\nnum-layers = 7\nlambda = 0.3\n# comment\nL1 = 0.9986288 # learned\n"
                   "L2 = 11.95378876\ncrf-iterations = 5\n")
    d = options("a.png", "b.png", "--config-file", cfg, "--lambda", "0.7", "-o", "res")
    assert d["matching"] == "COLORCNN" and d["L1"] == "0.9986288" and d["L2"] == "11.95378876"
    assert d["lambda"] == "0.7"  # the command line wins over the file
    assert d["output-file"] == "res" and d["crf-iterations"] == "5"
    d = options("--im0=x.png", "--im1", "y.png", "--disp-max=63")
    assert (d["im0"], d["im1"], d["disp-max"]) == ("x.png", "y.png", "63")


def test_option_errors(tmp_path):
    r = subprocess.run([BIN, "options", "--no-such-key", "1"], capture_output=True, text=True)
    assert r.returncode == 2 and "unknown option" in r.stderr
    cfg = tmp_path / "bad.cfg"
    cfg.write_text("bogus = 1\n")
    r = subprocess.run([BIN, "options", "--config-file", str(cfg)], capture_output=True, text=True)
    assert r.returncode == 2 and "unknown option" in r.stderr


def _png(path, img, bitdepth, ctype, palette=None, filters=(0, 1, 2, 3, 4)):
    """Minimal PNG encoder: every row uses the next filter type of `filters`."""
    H, W = img.shape[:2]
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    bpp = ch * bitdepth // 8
    raw = img.reshape(H, W * ch).astype(">u2" if bitdepth == 16 else np.uint8).view(np.uint8).reshape(H, -1)
    rows, prev = [], np.zeros(raw.shape[1], dtype=np.int64)
    for y in range(H):
        f = filters[y % len(filters)]
        cur = raw[y].astype(np.int64)
        a = np.concatenate([np.zeros(bpp, np.int64), cur[:-bpp]])
        c = np.concatenate([np.zeros(bpp, np.int64), prev[:-bpp]])
        if f == 0:
            out = cur
        elif f == 1:
            out = cur - a
        elif f == 2:
            out = cur - prev
        elif f == 3:
            out = cur - (a + prev) // 2
        else:
            p = a + prev - c
            pa, pb, pc = np.abs(p - a), np.abs(p - prev), np.abs(p - c)
            pred = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, prev, c))
            out = cur - pred
        rows.append(bytes([f]) + (out % 256).astype(np.uint8).tobytes())
        prev = cur

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xffffffff)

    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, bitdepth, ctype, 0, 0, 0))
    if palette is not None:
        data += chunk(b"PLTE", palette.astype(np.uint8).tobytes())
    data += chunk(b"IDAT", zlib.compress(b"".join(rows))) + chunk(b"IEND", b"")
    path.write_bytes(data)


def imread(path):
    lines = run("imread", path).strip().splitlines()
    W, H = map(int, lines[0].split())
    return np.array([[float(v) for v in l.split()] for l in lines[1:]]).reshape(H, W, 4)


@pytest.mark.parametrize("bitdepth", [8, 16])
def test_png_rgb_and_gray_all_filters(tmp_path, bitdepth):
    rng = np.random.default_rng(3)
    top = 255 if bitdepth == 8 else 65535
    rgb = rng.integers(0, top + 1, (7, 9, 3))
    _png(tmp_path / "c.png", rgb, bitdepth, 2)
    got = imread(tmp_path / "c.png")
    np.testing.assert_allclose(got[..., :3], rgb / top, rtol=1e-6)
    np.testing.assert_array_equal(got[..., 3], 1.0)
    g = rng.integers(0, top + 1, (6, 5))
    _png(tmp_path / "g.png", g, bitdepth, 0)
    got = imread(tmp_path / "g.png")
    for c in range(3):
        np.testing.assert_allclose(got[..., c], g / top, rtol=1e-6)
    rgba = rng.integers(0, top + 1, (4, 6, 4))
    _png(tmp_path / "a.png", rgba, bitdepth, 6)
    np.testing.assert_allclose(imread(tmp_path / "a.png"), rgba / top, rtol=1e-6)


def test_png_palette_and_pnm(tmp_path):
    rng = np.random.default_rng(4)
    pal = rng.integers(0, 256, (5, 3))
    idx = rng.integers(0, 5, (3, 4))
    _png(tmp_path / "p.png", idx, 8, 3, palette=pal)
    np.testing.assert_allclose(imread(tmp_path / "p.png")[..., :3], pal[idx] / 255.0, rtol=1e-6)
    img = rng.integers(0, 256, (3, 4, 3)).astype(np.uint8)
    (tmp_path / "x.ppm").write_bytes(b"P6\n# made by a test\n4 3\n255\n" + img.tobytes())
    np.testing.assert_allclose(imread(tmp_path / "x.ppm")[..., :3], img / 255.0, rtol=1e-6)


def test_pfm_and_png_writers(tmp_path):
    want = np.array([[0.25 * (x + 10 * y) - 1 for x in range(5)] for y in range(3)], dtype=np.float32)
    run("pfm-write", tmp_path / "d.pfm")
    b = (tmp_path / "d.pfm").read_bytes()
    assert b.startswith(b"Pf\n5 3\n-1.0\n")
    rows = np.frombuffer(b[len(b"Pf\n5 3\n-1.0\n"):], dtype="<f4").reshape(3, 5)
    np.testing.assert_array_equal(rows[::-1], want)  # stored bottom-up
    np.testing.assert_allclose(imread(tmp_path / "d.pfm")[..., 0], want)
    run("png-write", tmp_path / "d.png")
    p = (tmp_path / "d.png").read_bytes()
    assert p[:8] == b"\x89PNG\r\n\x1a\n" and p[12:16] == b"IHDR"
    W, H, depth, ctype = struct.unpack(">IIBB", p[16:26])
    assert (W, H, depth, ctype) == (5, 3, 16, 0)
    idat = p[p.index(b"IDAT") + 4:p.index(b"IEND") - 8]
    raw = np.frombuffer(zlib.decompress(idat), dtype=np.uint8).reshape(3, 11)
    assert np.all(raw[:, 0] == 0)
    vals = raw[:, 1:].copy().view(">u2").reshape(3, 5)
    np.testing.assert_array_equal(vals, np.floor(np.maximum(want * 256.0, 0)).astype(np.uint16))
