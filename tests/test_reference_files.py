"""The reference's own model files through the host-side C++ API, on the CPU.

Every params.npz under the reference's data/parameters tree is read in place
by the library's npz reader and walked by StereoNet's parameter walk
(StereoNet::checkParameterFile: the shapes and checks of setAllParams and
Operator::setParams, stereo/stereonet.cpp:184-227 and
stereo/operators/operator.cpp:55-85) for the architecture its .cfg names
(matching / num-layers / with-xy, stereo/stereo_img.cpp:94-103, 135-212).  The
learned L1, L2 that CRF models append as (1,) arrays must be left unread, as
the reference leaves them.  Every .cfg is parsed by Options::parseFile.

The files are the reference's data (AGPL-3.0): they are read where they lie,
never copied, and the pickled checkpoints beside them are never opened.  The
tests skip where the reference tree is absent.
"""
import glob
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cnn-crf-stereo_amd", "ccs_selftest")
PARAMS = "/root/reference/data/parameters"

pytestmark = pytest.mark.skipif(not os.path.isdir(PARAMS), reason="the reference's data tree is not present")

# relative path -> (num-layers, input channels, pairwise network, arrays the walk reads).
# Where a .cfg sits beside the file, the model is the one it names (checked
# below); the two middlebury cnn+crf+pw models have none: the 7-layer one holds
# a COLORPAIR network without the trailing L1, L2, the 3-layer ones a 3-layer
# tower followed by arrays that no StereoNet architecture reads.
MODELS = {
    "kitti-2015/3-layer/cnn+crf+full/params.npz": (3, 3, True, 11),
    "kitti-2015/3-layer/cnn+crf/params.npz": (3, 3, False, 6),
    "kitti-2015/3-layer/cnn/params.npz": (3, 3, False, 6),
    "kitti-2015/7-layer/cnn+crf+full+xy/params.npz": (7, 5, True, 19),
    "kitti-2015/7-layer/cnn+crf+full/params.npz": (7, 3, True, 19),
    "kitti-2015/7-layer/cnn+crf/params.npz": (7, 3, False, 14),
    "kitti-2015/7-layer/cnn/params.npz": (7, 3, False, 14),
    "middlebury-2014/3-layer/cnn+crf+full/params.npz": (3, 3, True, 11),
    "middlebury-2014/3-layer/cnn+crf+pw/params.npz": (3, 3, False, 6),
    "middlebury-2014/3-layer/cnn+crf+pw/params3.npz": (3, 3, False, 6),
    "middlebury-2014/3-layer/cnn+crf/params.npz": (3, 3, False, 6),
    "middlebury-2014/3-layer/cnn/params.npz": (3, 3, False, 6),
    "middlebury-2014/7-layer/cnn+crf+full/params.npz": (7, 3, True, 19),
    "middlebury-2014/7-layer/cnn+crf+pw/params.npz": (7, 3, True, 19),
    "middlebury-2014/7-layer/cnn+crf/params.npz": (7, 3, False, 14),
    "middlebury-2014/7-layer/cnn/params.npz": (7, 3, False, 14),
}


def selftest(*args, ok=True):
    if not os.path.exists(BIN):
        pytest.fail("ccs_selftest is not built (run __graft_entry__.build())")
    r = subprocess.run([BIN, *map(str, args)], capture_output=True, text=True, timeout=120)
    if ok:
        assert r.returncode == 0, r.stderr
    return r


def walk(path, layers, ic, pairwise):
    """(number of parameters the walk reads, keys it leaves unread)."""
    out = selftest("params", path, layers, ic, int(pairwise)).stdout.split()
    assert out[0] == "ok"
    return int(out[1]), out[2:]


def cfg_values(path):
    vals = {}
    for line in open(path):
        line = line.split("#")[0].strip()
        if line and not line.startswith("[") and "=" in line:
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    return vals


def test_every_parameter_file_is_listed():
    found = sorted(os.path.relpath(p, PARAMS) for p in glob.glob(PARAMS + "/**/*.npz", recursive=True))
    assert found == sorted(MODELS)


@pytest.mark.parametrize("rel", sorted(MODELS))
def test_parameter_file_loads(rel):
    path = os.path.join(PARAMS, rel)
    layers, ic, pairwise, nread = MODELS[rel]
    # The model the .cfg files beside it name.
    for cfg in glob.glob(os.path.join(os.path.dirname(path), "*.cfg")):
        v = cfg_values(cfg)
        assert int(v["num-layers"]) == layers, cfg
        assert (v["matching"] == "COLORPAIR") == pairwise, cfg
        assert v["matching"] in ("COLORCNN", "COLORPAIR"), cfg  # colour input: 3 planes (+ 2 with with-xy)
        assert (3 + 2 * int(v.get("with-xy", "0") in ("1", "true"))) == ic, cfg
    n, unread = walk(path, layers, ic, pairwise)
    assert n == nread
    z = np.load(path, allow_pickle=False)
    assert sorted(unread, key=lambda k: int(k[4:])) == ["arr_%d" % i for i in range(nread, len(z.files))]
    # What the walk read has exactly the shapes of the architecture.
    shapes = [z["arr_%d" % i].shape for i in range(nread)]
    conv = [s for s in shapes if len(s) == 4]
    assert conv[0] == (100, ic, 3, 3) and all(s == (100, 100, 2, 2) for s in conv[1:layers])
    if pairwise:
        assert conv[layers:] == [(64, 3, 3, 3), (64, 64, 3, 3), (2, 64, 1, 1)]
    # Left unread: the learned L1, L2 of CRF models (shape (1,)), or arrays of
    # a network no StereoNet builds (the 3-layer cnn+crf+pw files).
    if "cnn+crf+pw" not in rel:
        assert all(z[k].shape == (1,) for k in unread)


@pytest.mark.parametrize("rel,layers,ic,pairwise", [
    ("kitti-2015/3-layer/cnn+crf/params.npz", 7, 3, False),      # too few layers in the file
    ("kitti-2015/7-layer/cnn/params.npz", 7, 1, False),          # colour weights for a gray network
    ("kitti-2015/7-layer/cnn+crf/params.npz", 7, 3, True),       # no pairwise network in the file
    ("kitti-2015/7-layer/cnn+crf+full+xy/params.npz", 7, 3, True),  # 5-plane first layer
])
def test_wrong_architecture_is_rejected(rel, layers, ic, pairwise):
    r = selftest("params", os.path.join(PARAMS, rel), layers, ic, int(pairwise), ok=False)
    assert r.returncode != 0 and "arr_" in r.stderr


@pytest.mark.parametrize("cfg", sorted(os.path.relpath(p, PARAMS)
                                       for p in glob.glob(PARAMS + "/**/*.cfg", recursive=True)))
def test_config_file_parses(cfg):
    path = os.path.join(PARAMS, cfg)
    out = dict(line.split("=", 1) for line in selftest("config", path).stdout.strip().splitlines())
    want = cfg_values(path)
    for k, v in want.items():
        assert out.get(k) == v, (k, v, out.get(k))
