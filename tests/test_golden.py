"""The committed golden fixtures (tests/golden/*.npz) against the oracle: the
oracle must still produce them, and they must satisfy the invariants their
generator does not check by itself."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden  # noqa: E402
import oracle as O  # noqa: E402


def load(name):
    with np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(make_golden.FIXTURES))
def test_oracle_reproduces_fixture(name):
    want = load(name)
    got = make_golden.FIXTURES[name]()
    assert sorted(got) == sorted(want)
    for k in want:
        g, w = np.asarray(got[k]), want[k]
        assert g.shape == w.shape, k
        if w.dtype.kind in "iu":
            np.testing.assert_array_equal(g, w, err_msg=k)
        else:
            # OpenMP reductions may add the chain sums in another order
            np.testing.assert_allclose(g, w, rtol=1e-9, atol=1e-9, err_msg=k)


def test_frame_fixture_invariants():
    f = load("frame_small")
    D = int(f["D"])
    np.testing.assert_allclose(f["cost"].sum(-1), -1.0, rtol=1e-12)
    assert f["labels"].min() >= 0 and f["labels"].max() < D
    e = O.energy(f["cost"].astype(np.float32), f["wx"].astype(np.float32), f["wy"].astype(np.float32), f["labels"],
                 float(f["L1"]), float(f["L2"]), int(f["delta"]))
    assert f["lb"] <= e + 1e-9
    seq = O.dual_sequence(f["sums"], f["kept"])
    assert all(b >= a - 1e-9 for a, b in zip(seq, seq[1:]))
    # the left image is matched: most pixels away from the border land near the true offsets
    H, W = f["labels"].shape
    d = np.where(np.arange(W) < W // 2, D // 3, (2 * D) // 3)
    inner = f["labels"][3:-3, 3:W // 2 - 3]
    assert np.mean(np.abs(inner - d[None, 3:W // 2 - 3]) <= 1) > 0.8


def test_crf_fixture_invariants():
    f = load("crf_small")
    assert f["lb"] <= f["energy"] + 1e-9
    np.testing.assert_allclose(O.energy(f["cost"], f["wx"], f["wy"], f["labels"], float(f["L1"]), float(f["L2"]),
                                        int(f["delta"])), f["energy"], rtol=1e-12)
    assert f["gap"].min() >= 0
