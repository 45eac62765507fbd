"""Run-to-run determinism of the SlackProp sweeps at sizes where the column
sweep recomputes its messages from checkpoints (volumes of 64 MiB and more).

Each case runs rows, columns, rows, columns and the decode on fresh handles
(so the first call of a handle, which allocates its message scratch, is
included) and on a handle's second call, and requires the unaries, the float64
chain/kept sums and the labels to be bit-identical.  At D = 256 every lane
holds four labels; these sweeps once stored 128-bit vectors whose data
registers the next instruction overwrote, which showed up only here, as
results that changed from call to call.  A full-size case at 1488 x 992 x 256
and the 1242 x 375 x 128 benchmark volume are included.
"""
import pytest
import torch

from cnn_crf_stereo_amd import capi, ops

pytestmark = pytest.mark.gpu

SHAPES = [(992, 300, 256), (200, 1488, 256), (375, 1242, 128), (992, 1488, 256)]


def run_sequence(h, cost, wx, wy):
    H, W, D = cost.shape
    A = cost.clone()
    sums = torch.zeros(8, dtype=torch.float64, device="cuda")
    k = 0
    for _ in range(2):
        for d in (0, 1):
            ops.slackprop_sweep(h, cost, A, wx, wy, 0.5, 2.0, 2, d, chain_sum=sums[k:k + 1],
                                kept_sum=sums[k + 4:k + 5])
            k += 1
    labels = torch.empty((H, W), dtype=torch.int32, device="cuda")
    ops.slackprop_sweep(h, cost, A, wx, wy, 0.5, 2.0, 2, 2, labels=labels)
    torch.cuda.synchronize()
    return A, sums, labels


@pytest.mark.parametrize("H,W,D", SHAPES)
def test_sweeps_bit_identical_across_handles_and_calls(H, W, D):
    g = torch.Generator(device="cuda")
    g.manual_seed(H * 7 + W + D)
    cost = -torch.softmax(torch.randn(H, W, D, device="cuda", generator=g) * 3, dim=-1)
    wx = torch.rand(H, W, device="cuda", generator=g) * 0.8 + 0.2
    wy = torch.rand(H, W, device="cuda", generator=g) * 0.8 + 0.2
    runs = []
    for _ in range(2):
        h = capi.Handle()
        runs.append(run_sequence(h, cost, wx, wy))  # first call of the handle
        runs.append(run_sequence(h, cost, wx, wy))
        h.close()
    A0, s0, l0 = runs[0]
    assert bool(torch.isfinite(A0).all())
    for A, s, lab in runs[1:]:
        assert torch.equal(A.view(torch.int32), A0.view(torch.int32))
        assert torch.equal(s, s0)
        assert torch.equal(lab, l0)


@pytest.mark.parametrize("H,W,D", [(375, 1242, 128), (61, 97, 100), (40, 300, 256)])
def test_first_row_sweep_reads_cost(handle, H, W, D):
    """Direction 3 (unaries read from cost, written to A) equals a copy A = cost
    followed by direction 0, bit for bit, with A holding garbage beforehand."""
    g = torch.Generator(device="cuda")
    g.manual_seed(H + W + D)
    cost = -torch.softmax(torch.randn(H, W, D, device="cuda", generator=g) * 3, dim=-1)
    wx = torch.rand(H, W, device="cuda", generator=g) * 0.8 + 0.2
    wy = torch.rand(H, W, device="cuda", generator=g) * 0.8 + 0.2
    sums = torch.zeros(4, dtype=torch.float64, device="cuda")
    A0 = cost.clone()
    ops.slackprop_sweep(handle, cost, A0, wx, wy, 0.5, 2.0, 2, 0, chain_sum=sums[0:1], kept_sum=sums[1:2])
    A3 = torch.full_like(cost, float("nan"))
    ops.slackprop_sweep(handle, cost, A3, wx, wy, 0.5, 2.0, 2, 3, chain_sum=sums[2:3], kept_sum=sums[3:4])
    torch.cuda.synchronize()
    assert torch.equal(A3.view(torch.int32), A0.view(torch.int32))
    assert torch.equal(sums[2:4], sums[0:2])
    with pytest.raises(capi.CCSError):
        ops.slackprop_sweep(handle, cost, cost, wx, wy, 0.5, 2.0, 2, 3)


# Shapes below the checkpointing threshold, covering the kernel variants that the
# large benchmark volumes do not: labels that do not fill a lane's slots
# (D = 37, 100, 200), delta != 2, and 1, 2 and 4 labels per lane.  The row
# sweeps' message wave splits a chain of W - 1 steps into segments of 12 steps
# (6 at 4 labels per lane), end-aligned: the widths give every remainder class
# used below, chains of one node, of less than one segment and of one to two
# segments.
SMALL = [(97, 131, 100, 2), (120, 90, 37, 3), (64, 70, 256, 4), (150, 257, 200, 2), (61, 97, 64, 2),
         (30, 1, 32, 2), (40, 9, 64, 2), (33, 20, 128, 2), (17, 25, 128, 2), (9, 38, 256, 2), (12, 14, 64, 3)]
# Per-edge penalties (two levels) and a whole ccs_slackprop run, whose first row
# sweep also checks the costs for NaN.  Rows of 8000 nodes need more LDS per
# chain for the message wave (169 KB at D = 32) than a CU has: they keep the
# plain row scheme even where the threshold asks for checkpointing.
SMALL_PE = [(23, 131, 128), (11, 40, 256), (7, 18, 64)]
SMALL_RUN = [(41, 157, 128), (9, 26, 64), (3, 8000, 32)]

_CHILD = r"""
import sys, numpy as np, torch
sys.path.insert(0, sys.argv[1])
import conftest  # loads the package
from cnn_crf_stereo_amd import capi
from test_gpu_crf_determinism import SMALL, small_case, run_small
h = capi.Handle()
out = {}
from test_gpu_crf_determinism import SMALL_PE, SMALL_RUN, pe_case, run_pe, run_full
for (H, W, D, delta) in SMALL:
    A, sums, labels = run_small(h, *small_case(H, W, D), delta)
    out["%d_%d_%d_%d_A" % (H, W, D, delta)] = A.cpu().numpy()
    out["%d_%d_%d_%d_s" % (H, W, D, delta)] = sums.cpu().numpy()
    out["%d_%d_%d_%d_l" % (H, W, D, delta)] = labels.cpu().numpy()
for (H, W, D) in SMALL_PE:
    for k, v in run_pe(h, *pe_case(H, W, D)).items():
        out["pe_%d_%d_%d_%s" % (H, W, D, k)] = v
for (H, W, D) in SMALL_RUN:
    for k, v in run_full(h, *small_case(H, W, D)).items():
        out["run_%d_%d_%d_%s" % (H, W, D, k)] = v
np.savez(sys.argv[2], **out)
print("child ok")
"""


def small_case(H, W, D):
    g = torch.Generator(device="cuda")
    g.manual_seed(H * 3 + W + D)
    cost = -torch.softmax(torch.randn(H, W, D, device="cuda", generator=g) * 3, dim=-1)
    wx = torch.rand(H, W, device="cuda", generator=g) * 0.8 + 0.2
    wy = torch.rand(H, W, device="cuda", generator=g) * 0.8 + 0.2
    return cost, wx, wy


def run_small(h, cost, wx, wy, delta):
    H, W, D = cost.shape
    A = torch.empty_like(cost)
    sums = torch.zeros(8, dtype=torch.float64, device="cuda")
    for it in range(2):
        ops.slackprop_sweep(h, cost, A, wx, wy, 0.5, 2.0, delta, 3 if it == 0 else 0,
                            chain_sum=sums[2 * it:2 * it + 1], kept_sum=sums[4 + 2 * it:5 + 2 * it])
        ops.slackprop_sweep(h, cost, A, wx, wy, 0.5, 2.0, delta, 1,
                            chain_sum=sums[2 * it + 1:2 * it + 2], kept_sum=sums[5 + 2 * it:6 + 2 * it])
    labels = torch.empty((H, W), dtype=torch.int32, device="cuda")
    ops.slackprop_sweep(h, cost, A, wx, wy, 0.5, 2.0, delta, 2, labels=labels)
    torch.cuda.synchronize()
    return A, sums, labels


def pe_case(H, W, D):
    g = torch.Generator(device="cuda")
    g.manual_seed(H * 5 + W + D)
    cost = -torch.softmax(torch.randn(H, W, D, device="cuda", generator=g) * 3, dim=-1)
    p1 = torch.rand(2, H, W, device="cuda", generator=g) * 0.7 + 0.1
    p2 = p1 + torch.rand(2, H, W, device="cuda", generator=g) * 2.0 + 0.2
    return cost, torch.cat([p1, p2]).contiguous()


def run_pe(h, cost, pw):
    pen, bad = ops.pairwise_penalties(h, pw)
    assert not bad.any()
    g = ops.slackprop_pw(h, cost, pen, 2, 3, want_unaries=True)
    torch.cuda.synchronize()
    return {k: g[k].cpu().numpy() for k in ("labels", "sums", "lb", "unaries")}


def run_full(h, cost, wx, wy):
    g = ops.slackprop(h, cost, wx, wy, 0.5, 2.0, 2, 3, want_unaries=True)
    ops.slackprop_status(h)
    torch.cuda.synchronize()
    return {k: g[k].cpu().numpy() for k in ("labels", "sums", "lb", "unaries")}


def test_checkpointed_sweeps_equal_stored_messages(handle, tmp_path):
    """The checkpointed schemes (forced on by the threshold override in a child
    process: columns recompute their messages in the forward pass, rows in a
    second wave) give the store-everything schemes' unaries, sums and labels
    bit for bit, on shapes only the latter run by default -- plain sweeps,
    per-edge penalties and whole runs."""
    import os
    import subprocess
    import sys

    import numpy as np

    here = os.path.dirname(os.path.abspath(__file__))
    path = str(tmp_path / "seg.npz")
    env = dict(os.environ, CCS_CRF_SEG_MIN_BYTES="0")
    r = subprocess.run([sys.executable, "-c", _CHILD, here, path], env=env, cwd=here, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    got = np.load(path)
    for (H, W, D, delta) in SMALL:
        A, sums, labels = run_small(handle, *small_case(H, W, D), delta)
        key = "%d_%d_%d_%d_" % (H, W, D, delta)
        assert np.array_equal(got[key + "A"].view(np.int32), A.cpu().numpy().view(np.int32)), key
        assert np.array_equal(got[key + "s"], sums.cpu().numpy()), key
        assert np.array_equal(got[key + "l"], labels.cpu().numpy()), key
    for (H, W, D) in SMALL_PE:
        ref = run_pe(handle, *pe_case(H, W, D))
        for k, v in ref.items():
            key = "pe_%d_%d_%d_%s" % (H, W, D, k)
            assert np.array_equal(got[key].view(np.uint8), v.view(np.uint8)), key
    for (H, W, D) in SMALL_RUN:
        ref = run_full(handle, *small_case(H, W, D))
        for k, v in ref.items():
            key = "run_%d_%d_%d_%s" % (H, W, D, k)
            assert np.array_equal(got[key].view(np.uint8), v.view(np.uint8)), key


_CHILD_FULL = r"""
import sys, numpy as np, torch
sys.path.insert(0, sys.argv[1])
import conftest  # loads the package
from cnn_crf_stereo_amd import capi
from test_gpu_crf_determinism import full_case, run_sequence
h = capi.Handle()
A, sums, labels = run_sequence(h, *full_case())
np.savez(sys.argv[2], A=A.cpu().numpy(), s=sums.cpu().numpy(), l=labels.cpu().numpy())
print("child ok")
"""


def full_case():
    g = torch.Generator(device="cuda")
    g.manual_seed(4242)
    H, W, D = 375, 1242, 128
    cost = -torch.softmax(torch.randn(H, W, D, device="cuda", generator=g) * 3, dim=-1)
    wx = torch.rand(H, W, device="cuda", generator=g) * 0.8 + 0.2
    wy = torch.rand(H, W, device="cuda", generator=g) * 0.8 + 0.2
    return cost, wx, wy


def test_full_size_checkpointed_equal_stored_messages(handle, tmp_path):
    """At 1242 x 375 x 128 (checkpointed rows and columns by default) the
    sweeps equal the store-everything schemes, forced on in a child process by
    a threshold above the volume, bit for bit."""
    import os
    import subprocess
    import sys

    import numpy as np

    here = os.path.dirname(os.path.abspath(__file__))
    path = str(tmp_path / "full.npz")
    env = dict(os.environ, CCS_CRF_SEG_MIN_BYTES=str(1 << 40))
    r = subprocess.run([sys.executable, "-c", _CHILD_FULL, here, path], env=env, cwd=here, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    got = np.load(path)
    A, sums, labels = run_sequence(handle, *full_case())
    assert np.array_equal(got["A"].view(np.int32), A.cpu().numpy().view(np.int32))
    assert np.array_equal(got["s"], sums.cpu().numpy())
    assert np.array_equal(got["l"], labels.cpu().numpy())
