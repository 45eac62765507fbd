"""The whole frame (normalise, both towers, correlation, edge weights,
SlackProp with its float64 sums and bound, remap) captured once in a HIP graph
and replayed: every replay must reproduce the eager frame bit for bit.

Replays are interleaved with eager frames of a different image pair on the same
handle, so a replay that relied on state left by the previous replay (scratch
contents, counters) would show up.  The frame path records no memset nodes: the
only zero fills are torch fill kernels.
"""
import pytest
import torch

from cnn_crf_stereo_amd import capi, pipeline, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("workload,precision", [("rds-320x240-d32", "bf16"), ("rds-320x240-d32", "fp32")])
def test_frame_graph_replay_matches_eager(workload, precision):
    spec = dict(synth.WORKLOADS[workload])
    data = synth.stereo_pair(workload)
    other = synth.stereo_pair(spec, seed=4321)
    layers, _ = synth.architecture(7, spec["C"], "all3x3")
    params = synth.glorot_weights(layers, seed=5)
    stream = torch.cuda.Stream()  # default priority
    h = capi.Handle(stream)
    frame = pipeline.Frame(h, params, precision, spec["D"], iters=spec["iters"])
    left = torch.as_tensor(data["left"], device="cuda")
    right = torch.as_tensor(data["right"], device="cuda")
    l2 = torch.as_tensor(other["left"], device="cuda")
    r2 = torch.as_tensor(other["right"], device="cuda")
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        eager = frame(left, right)
        keep = {k: eager[k].clone() for k in ("cost", "labels", "sums", "lb", "disparity")}
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        out = frame(left, right)
    torch.cuda.synchronize()
    for _ in range(3):
        # Eager work on the same handle between replays, ordered with them on the
        # handle's stream (a handle's scratch must not be used from two streams
        # at once).
        with torch.cuda.stream(stream):
            frame(l2, r2)
            g.replay()
        torch.cuda.synchronize()
        for k, v in keep.items():
            assert torch.equal(out[k], v), k
    del g
    h.close()
