"""One stereo frame through the hot path, stage by stage, on the handle's stream:

  pad + normalise (left, right)      stereo/stereo.cu:1022-1056, stereo/colorstereonet.cpp:38-77
  feature towers (shared weights)    stereo/stereonet.cpp:250-291
  correlation + -softmax             stereo/stereonet.cpp:301-318
  edge weights                       stereo/stereo.cpp:320-326
  SlackProp CRF                      stereo/stereo.cpp:129-162
  label -> disparity                 stereo/stereo.cpp:155-158

Inputs stay resident in HBM; every stage is a HIP kernel of libccs.so.  An
optional `timer` records HIP events around each stage on the same stream.
"""
import torch

from . import ops


class StageTimer:
    """HIP events around named stages (recorded on torch's current stream, which
    the ccs handle shares)."""

    def __init__(self):
        self.events = []

    def mark(self, name):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        self.events.append((name, ev))

    def durations(self):
        torch.cuda.synchronize()
        out = {}
        for (n0, e0), (_, e1) in zip(self.events[:-1], self.events[1:]):
            out[n0] = out.get(n0, 0.0) + e0.elapsed_time(e1)
        return out


class Frame:
    """Weights and CRF parameters of one configuration, ready to run frames."""

    def __init__(self, h, params, precision, D, disp_min=0, disp_step=1, L1=0.5, L2=2.0, delta=2, iters=8,
                 sigma=0.7, alpha=10.0, beta=1.0, lam=1.0, color=False):
        self.h = h
        self.tower = ops.Tower(h, params, precision)
        self.pad = self.tower.pad
        self.C = self.tower.layers[-1].K
        self.D, self.disp_min, self.disp_step = D, disp_min, disp_step
        self.L1, self.L2, self.delta, self.iters = L1, L2, delta, iters
        self.sigma, self.alpha, self.beta, self.lam = sigma, alpha, beta, lam
        self.color = color

    def __call__(self, left, right, timer=None, crf=True):
        """left/right: C x H x W float32 cuda tensors in [0, 1].  Returns dict."""
        h = self.h
        mark = timer.mark if timer is not None else (lambda name: None)
        mark("normalize")
        xl = ops.pad_normalize(h, left, self.pad)
        xr = ops.pad_normalize(h, right, self.pad)
        mark("cnn")
        fl = self.tower(h, xl)
        fr = self.tower(h, xr)
        mark("correlation")
        cost = ops.correlation_volume(h, fl, fr, self.C, self.disp_min, self.disp_step, self.D)
        out = dict(cost=cost, features=(fl, fr))
        if crf:
            mark("edge_weights")
            gray = ops.rgb_to_gray(h, left) if self.color else left[0]
            wx, wy = ops.edge_weights(h, gray, self.sigma, self.alpha, self.beta, self.lam)
            mark("crf")
            res = ops.slackprop(h, cost, wx, wy, self.L1, self.L2, self.delta, self.iters)
            mark("remap")
            disp = ops.remap(h, res["labels"], self.disp_min, self.disp_step)
            out.update(wx=wx, wy=wy, labels=res["labels"], lb=res["lb"], sums=res["sums"], disparity=disp)
        mark("end")
        return out
