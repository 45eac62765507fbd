#!/usr/bin/env python3
"""Benchmark of the CNN-CRF stereo hot path on one MI355X.

A step is one stereo frame through the whole path with the inputs already in
HBM: pad + normalise, the two 7-layer feature towers, correlation + -softmax,
edge weights, SlackProp CRF and the label -> disparity map.  The workload is
the 1242 x 375 grayscale random-dot pair with D = 128, a 7-layer 100-channel
CNN of 3x3 layers and 8 CRF iterations; the BF16-MFMA form is the measured
value and the FP32 form runs beside it on the same GPU.

Prints one JSON line (see --help).  The headline frame time is the whole
frame captured once in a HIP graph and replayed (config.timing says which;
eager_ms_per_frame is the same frame launched kernel by kernel).  Per-launch
HIP events bracket every kernel launch of the eager frames on the handle's
stream; per-launch figures (the
roofline entry, stages_ms, kernels) come from a pass with both towers on one
stream, so that no event interval spans the other tower's launches.  The
roofline entry is the dominant kernel class's algorithmic work per launch over
its mean duration.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "cnn-crf-stereo_amd")

METRIC = "ms/frame + pixel·disp/s at 1242x375 D=128, BF16-MFMA vs FP32, same GPU"
HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"bf16": 2500.0, "fp32": 157.3}


def _load_package():
    if "cnn_crf_stereo_amd" in sys.modules:
        return sys.modules["cnn_crf_stereo_amd"]
    spec = importlib.util.spec_from_file_location("cnn_crf_stereo_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["cnn_crf_stereo_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


class Launches:
    """HIP events bracketing kernel launches, grouped by kernel class."""

    def __init__(self, enabled):
        import torch

        self.torch = torch
        self.enabled = enabled
        self.records = []  # (cls, work dict, ev0, ev1)

    def run(self, cls, work, fn):
        if not self.enabled:
            return fn()
        e0 = self.torch.cuda.Event(enable_timing=True)
        e1 = self.torch.cuda.Event(enable_timing=True)
        e0.record()
        r = fn()
        e1.record()
        self.records.append((cls, work, e0, e1))
        return r

    def summary(self):
        self.torch.cuda.synchronize()
        out = {}
        for cls, work, e0, e1 in self.records:
            d = out.setdefault(cls, dict(ms=0.0, launches=0, flops=0.0, bytes=0.0))
            d["ms"] += e0.elapsed_time(e1)
            d["launches"] += 1
            d["flops"] += work.get("flops", 0.0)
            d["bytes"] += work.get("bytes", 0.0)
        return out


def run_frame(ops, h, frame, left, right, L, A, side=None, record=True):
    """One frame, every kernel launch bracketed by events (when enabled).

    side = (handle, stream): the right image's normalisation and tower run on
    that second stream, concurrently with the left tower (the weights are
    shared read-only; each stream has its own handle), and the correlation
    waits for both.  Concurrent towers fill the tail of each other's layers."""
    import torch

    H, W = left.shape[1:]
    D = frame.D
    out = {}
    pad = frame.pad
    esize = 2 if frame.tower.precision == "bf16" else 4

    def tower(hh, img):
        y = L.run("normalize", {"bytes": 4.0 * img.numel() * 3}, lambda: ops.pad_normalize(hh, img, pad))
        for li, layer in enumerate(frame.tower.layers):
            if li == 0:
                Hi, Wi = y.shape[1:]
            else:
                Hi, Wi = y.shape[:2]
            Ho, Wo = Hi - layer.KH + 1, Wi - layer.KW + 1
            flops = 2.0 * layer.K * layer.C * layer.KH * layer.KW * Ho * Wo
            cls = "conv_first" if li == 0 else "conv"
            lay = "nchw" if li == 0 else "nhwc"
            y = L.run(cls, {"flops": flops}, lambda y=y, layer=layer, lay=lay: layer(hh, y, layout=lay))
        return y

    if side is None:
        feats = [tower(h, left), tower(h, right)]
    else:
        h2, s2 = side
        main = torch.cuda.current_stream()
        s2.wait_stream(main)
        with torch.cuda.stream(s2):
            fr = tower(h2, right)
        fl = tower(h, left)
        main.wait_stream(s2)
        if record:  # (not under graph capture: the graph's memory pool keeps fr)
            fr.record_stream(main)
        feats = [fl, fr]
    C = frame.C
    cs = feats[0].shape[2]
    corr_bytes = 2.0 * H * W * cs * esize + 4.0 * H * W * D
    cost = L.run("correlation", {"bytes": corr_bytes, "flops": 2.0 * C * H * W * D},
                 lambda: ops.correlation_volume(h, feats[0], feats[1], C, frame.disp_min, frame.disp_step, D))
    wx, wy = L.run("edge_weights", {"bytes": 12.0 * H * W},
                   lambda: ops.edge_weights(h, left[0], frame.sigma, frame.alpha, frame.beta, frame.lam))
    V = 4.0 * H * W * D
    labels = torch.empty((H, W), dtype=torch.int32, device="cuda")
    for it in range(frame.iters):
        # The first row sweep reads its unaries from the cost volume (direction 3).
        L.run("crf_rows", {"bytes": 3 * V},
              lambda it=it: ops.slackprop_sweep(h, cost, A, wx, wy, frame.L1, frame.L2, frame.delta,
                                                3 if it == 0 else 0))
        L.run("crf_cols", {"bytes": 3 * V}, lambda: ops.slackprop_sweep(h, cost, A, wx, wy, frame.L1, frame.L2,
                                                                        frame.delta, 1))
    L.run("crf_decode", {"bytes": V + 4.0 * H * W},
          lambda: ops.slackprop_sweep(h, cost, A, wx, wy, frame.L1, frame.L2, frame.delta, 2, labels=labels))
    disp = L.run("remap", {"bytes": 8.0 * H * W}, lambda: ops.remap(h, labels, frame.disp_min, frame.disp_step))
    out.update(cost=cost, labels=labels, disparity=disp)
    return out


def time_form(pkg, h, data, spec, params, precision, steps, warmup, side=None):
    import torch
    from cnn_crf_stereo_amd import ops, pipeline

    frame = pipeline.Frame(h, params, precision, spec["D"], iters=spec["iters"])
    left = torch.as_tensor(data["left"], device="cuda")
    right = torch.as_tensor(data["right"], device="cuda")
    H, W = left.shape[1:]
    A = torch.empty((H, W, spec["D"]), dtype=torch.float32, device="cuda")
    for _ in range(warmup):
        run_frame(ops, h, frame, left, right, Launches(False), A, side)
    torch.cuda.synchronize()
    L = Launches(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = run_frame(ops, h, frame, left, right, L, A, side)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / steps
    # Host images handed over per frame (pageable copies + the same frame).
    hl, hr = data["left"], data["right"]
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(max(1, steps // 2)):
        left.copy_(torch.from_numpy(hl))
        right.copy_(torch.from_numpy(hr))
        run_frame(ops, h, frame, left, right, Launches(False), A, side)
    torch.cuda.synchronize()
    ms_pcie = (time.perf_counter() - t2) * 1e3 / max(1, steps // 2)
    summary = L.summary()
    summary["_pcie_inclusive_ms"] = ms_pcie
    return ms, summary, out


def time_crf_forms(h, data, spec, cost, steps):
    """ms per SlackProp run (all iterations, rows + columns + decode) on the
    frame's cost volume with edge weights (ccs_slackprop: w L1, w L2) and with
    the same penalties given per edge (ccs_slackprop_pw), on the handle's stream."""
    import torch
    from cnn_crf_stereo_amd import ops

    left = torch.as_tensor(data["left"], device="cuda")
    wx, wy = ops.edge_weights(h, left[0], 0.7, 10.0, 1.0, 1.0)
    L1, L2, iters = 0.5, 2.0, spec["iters"]
    pw = torch.stack([wx * L1, wy * L1, wx * L2, wy * L2]).contiguous()
    pen, _ = ops.pairwise_penalties(h, pw)
    out = {}
    for name, fn in (("edge_weights", lambda: ops.slackprop(h, cost, wx, wy, L1, L2, 2, iters)),
                     ("per_edge", lambda: ops.slackprop_pw(h, cost, pen, 2, iters))):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        out[name] = round(e0.elapsed_time(e1) / steps, 4)
    out["per_edge_over_edge_weights"] = round(out["per_edge"] / out["edge_weights"], 4)
    return out


def time_pipelined(pkg, lanes, data, spec, params, precision, steps, warmup):
    """Steady-state ms per frame with len(lanes) frames in flight: frame i runs
    entirely on lane i % len(lanes), a (handle, stream) pair with its own CRF
    state, so consecutive frames overlap wherever the GPU has room (the
    latency-bound CRF row sweeps leave most SIMDs idle)."""
    import torch
    from cnn_crf_stereo_amd import ops, pipeline

    left = torch.as_tensor(data["left"], device="cuda")
    right = torch.as_tensor(data["right"], device="cuda")
    H, W = left.shape[1:]
    frames = [pipeline.Frame(h, params, precision, spec["D"], iters=spec["iters"]) for (h, _) in lanes]
    As = [torch.empty((H, W, spec["D"]), dtype=torch.float32, device="cuda") for _ in lanes]
    main = torch.cuda.current_stream()

    def go(i):
        h, st = lanes[i % len(lanes)]
        with torch.cuda.stream(st):
            return run_frame(ops, h, frames[i % len(lanes)], left, right, Launches(False), As[i % len(lanes)])

    torch.cuda.synchronize()
    for st in [st for (_, st) in lanes]:
        st.wait_stream(main)
    for i in range(warmup * len(lanes)):
        go(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        out = go(i)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps, out


def time_graph(pkg, data, spec, params, precision, steps, two_streams=True):
    """ms per frame when the whole frame (towers on two streams unless
    two_streams is False) is captured once in a HIP graph and replayed (no
    per-launch host work or launch gaps), and whether every replay's labels
    equal an eager run's bit for bit."""
    import torch
    from cnn_crf_stereo_amd import capi, ops, pipeline

    s_cap, s2 = torch.cuda.Stream(), torch.cuda.Stream()  # default priority
    h, h2 = capi.Handle(s_cap), capi.Handle(s2)
    frame = pipeline.Frame(h, params, precision, spec["D"], iters=spec["iters"])
    left = torch.as_tensor(data["left"], device="cuda")
    right = torch.as_tensor(data["right"], device="cuda")
    H, W = left.shape[1:]
    A = torch.empty((H, W, spec["D"]), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    with torch.cuda.stream(s_cap):  # eager warm-up: scratch reserved, kernels configured
        for _ in range(2):
            ref = run_frame(ops, h, frame, left, right, Launches(False), A, (h2, s2) if two_streams else None)
        ref_labels = ref["labels"].clone()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s_cap):
        out = run_frame(ops, h, frame, left, right, Launches(False), A, (h2, s2) if two_streams else None,
                        record=False)
    torch.cuda.synchronize()
    same = True
    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
        same = same and bool(torch.equal(out["labels"], ref_labels))
    t0 = time.perf_counter()
    for _ in range(steps):
        g.replay()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    g.replay()
    torch.cuda.synchronize()
    same = same and bool(torch.equal(out["labels"], ref_labels))
    del g
    h.close()
    h2.close()
    return ms, same


def measured_traffic(workload):
    """HBM bytes per launch by kernel class, from the committed counter summary
    (rocprofv3 FETCH_SIZE + WRITE_SIZE passes, scaled by a streaming kernel of
    known size); empty when no summary exists for this workload."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_%s.md" % workload)
    if not os.path.exists(path):
        return {}
    text = open(path).read()
    start = text.find("```json")
    if start < 0:
        return {}
    body = text[start + len("```json"):text.find("```", start + 7)]
    return json.loads(body)


def bad_pixel_rate(labels, gt, D, border=0):
    import numpy as np

    lab = labels.astype(np.float64)
    err = np.abs(lab - gt) > 1.0
    if border:
        err = err[:, border:]
    return float(err.mean())


def oracle_frame(O, left, right, spec, params, pad):
    """The FP64 oracle of the whole pipeline on planar float64 images."""
    import numpy as np

    D = spec["D"]
    feats = []
    for img in (left, right):
        x, _, _ = O.normalize(O.pad_symmetric(img, pad))
        for (w, b) in params:
            x = O.conv2d(x, w, b, True)
        feats.append(x)
    S = O.correlation(feats[0], feats[1], -np.arange(D))
    cost = O.neg_softmax(S)
    gray = O.rgb2gray(left) if left.shape[0] == 3 else left[0]
    wx, wy = O.edge_weights(gray, 0.7, 10.0, 1.0, 1.0)
    r = O.slackprop(cost, wx, wy, 0.5, 2.0, 2, spec["iters"])
    return feats, cost, wx, wy, r


def cpu_baseline(data, spec, params, pad, rows, repeats=3, time_rows=0):
    """The FP64 oracle of the same pipeline on a band of `rows` image rows
    (run once: its outputs are the reference of the GPU error figures), then
    timed as the median of `repeats` runs on a band of `time_rows` rows (the
    first run is the warm-up)."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    H = data["left"].shape[1]

    def band(n):
        n = H if n <= 0 or n > H else n
        y0 = max(0, H // 2 - n // 2)
        sl = slice(y0, y0 + n)
        return n, y0, data["left"][:, sl, :].astype(np.float64), data["right"][:, sl, :].astype(np.float64)

    rows, y0, left, right = band(rows)
    feats, cost, wx, wy, r = oracle_frame(O, left, right, spec, params, pad)
    trows, ty0, tl, tr = band(time_rows if time_rows > 0 else rows)
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        oracle_frame(O, tl, tr, spec, params, pad)
        times.append(time.perf_counter() - t0)
    Wd = left.shape[2]
    energy = O.energy(cost, wx, wy, r["labels"], 0.5, 2.0, 2)
    return dict(seconds=float(np.median(times)), times=times, time_rows=trows, time_y0=ty0,
                pd_timed=trows * Wd * spec["D"], rows=rows, y0=y0, pd=rows * Wd * spec["D"], labels=r["labels"],
                lb=r["lb"], energy=energy, feat_left=feats[0], cost=cost, left=left.astype(np.float32),
                right=right.astype(np.float32))


def time_config(pkg, h, workload, precision, steps, cpu_repeats=0, arch="all3x3", layers_n=7):
    """One workload's frame (FP32 or BF16) on the GPU: median ms over `steps`
    frames (HIP events, after a warm-up frame), and optionally the FP64 oracle
    of the same frame (median of cpu_repeats runs after the GPU one) with the
    agreement of the labels."""
    import numpy as np
    import torch
    from cnn_crf_stereo_amd import pipeline, synth

    spec = dict(synth.WORKLOADS[workload])
    data = synth.stereo_pair(workload)
    layers, pad = synth.architecture(layers_n, spec["C"], arch)
    params = synth.glorot_weights(layers)
    fr = pipeline.Frame(h, params, precision, spec["D"], iters=spec["iters"], color=spec["C"] == 3)
    left = torch.as_tensor(data["left"], device="cuda")
    right = torch.as_tensor(data["right"], device="cuda")
    out = fr(left, right)
    ms = []
    for _ in range(steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        out = fr(left, right)
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    H, W, D = spec["H"], spec["W"], spec["D"]
    res = {"precision": precision, "ms_per_frame": round(float(np.median(ms)), 4),
           "pixel_disp_per_s": H * W * D / (float(np.median(ms)) / 1e3),
           "bad_px_gt_gt1": bad_pixel_rate(out["labels"].cpu().numpy(), data["disp"], D)}
    if cpu_repeats:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        l64, r64 = data["left"].astype(np.float64), data["right"].astype(np.float64)
        times, r = [], None
        for _ in range(cpu_repeats + 1):  # the first run is the warm-up
            t0 = time.perf_counter()
            _, _, _, _, r = oracle_frame(O, l64, r64, spec, params, pad)
            times.append(time.perf_counter() - t0)
        cpu_s = float(np.median(times[1:]))
        lab = out["labels"].cpu().numpy()
        res["cpu_oracle"] = {"ms_per_frame": round(cpu_s * 1e3, 1), "runs": cpu_repeats, "cores": cpu_cores(),
                             "pixel_disp_per_s": H * W * D / cpu_s,
                             "bad_px_gt_gt1": bad_pixel_rate(r["labels"], data["disp"], D),
                             "labels_equal_gpu": float((lab == r["labels"]).mean())}
    return res


def cpu_cores():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def stage_errors(ops, h, fr, out, cb):
    """Per-stage error of one GPU form against the FP64 oracle on the same band."""
    import numpy as np
    import torch

    C = fr.C
    fl = out["features"][0][:, :, :C].float().cpu().numpy().transpose(2, 0, 1).astype(np.float64)
    ref_f = cb["feat_left"]
    cost = out["cost"].cpu().numpy().astype(np.float64)
    ref_c = cb["cost"]
    e = ops.crf_energy(h, out["cost"], out["wx"], out["wy"], out["labels"], fr.L1, fr.L2, fr.delta)
    torch.cuda.synchronize()
    energy = float(e.item())

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    return {
        "features_max_abs": float(np.abs(fl - ref_f).max()), "features_rel_l2": rel(fl, ref_f),
        "cost_max_abs": float(np.abs(cost - ref_c).max()), "cost_rel_l2": rel(cost, ref_c),
        "crf_energy_rel": abs(energy - cb["energy"]) / max(abs(cb["energy"]), 1e-300),
        "crf_lb_rel": abs(float(out["lb"].item()) - cb["lb"]) / max(abs(cb["lb"]), 1e-300),
    }


def train_bench(pkg, h, workload, layers_n, arch, batch, steps, warmup, iters, precision="fp32", checkpoint=True,
                grad_check=True, seed=11):
    """Training steps (forward, CRF hinge, backward through the hand-written
    adjoints, SGD) on `batch` seeded frames of `workload` with FP32 master
    weights, the towers and the correlation in `precision`, checkpointed towers
    (only the padded images kept; layer outputs recomputed in the backward) or
    store-everything layers.  Reports ms/step, the peak of torch's device
    allocations over the timed steps (all live tensors of the process, and the
    part above what was allocated when the timed steps began: the step's
    activations and gradients; the handle's own scratch is not a torch
    allocation), and the relative error (L2 over all
    tower and CRF parameters) of the first step's gradient against the FP32
    store-everything backward at the same weights."""
    import numpy as np
    import torch
    from cnn_crf_stereo_amd import autograd as AG
    from cnn_crf_stereo_amd import synth

    spec = dict(synth.WORKLOADS[workload])
    D = spec["D"]
    layers, _ = synth.architecture(layers_n, spec["C"], arch)
    params = synth.glorot_weights(layers, seed=seed)
    frames = []
    for i in range(batch):
        d = synth.stereo_pair(spec, seed=2000 + i)
        tgt = np.clip(np.rint(d["disp"]), 0, D - 1).astype(np.int32)
        frames.append((torch.as_tensor(d["left"], device="cuda"), torch.as_tensor(d["right"], device="cuda"),
                       torch.as_tensor(tgt, device="cuda")))
    out = {"workload": workload, "batch": batch, "layers": "%d x %s, 100 channels" % (layers_n, arch),
           "crf_iterations": iters, "dtype": precision,
           "towers": "checkpointed (padded image kept, layer outputs recomputed)" if checkpoint else
                     "store-everything"}
    if grad_check:
        ref = AG.StereoModel(params, precision="fp32", checkpoint=False)
        _, g_ref = AG.gradients(ref, h, frames, D, iters=iters)
        del ref
        model = AG.StereoModel(params, precision=precision, checkpoint=checkpoint)
        _, g = AG.gradients(model, h, frames, D, iters=iters)
        num = sum(float(((a - b).double() ** 2).sum()) for a, b in zip(g, g_ref))
        den = sum(float((b.double() ** 2).sum()) for b in g_ref)
        per = [float((a - b).double().norm() / max(float(b.double().norm()), 1e-300)) for a, b in zip(g, g_ref)]
        out["grad_rel_err_vs_fp32_store_everything"] = float(np.sqrt(num / max(den, 1e-300)))
        out["grad_rel_err_max_per_tensor"] = max(per)
        del g, g_ref
    model = AG.StereoModel(params, precision=precision, checkpoint=checkpoint)
    for _ in range(warmup):
        AG.training_step(model, h, frames, D, lr=1e-4, iters=iters)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()
    base = torch.cuda.memory_allocated()
    t0 = time.perf_counter()
    losses = [AG.training_step(model, h, frames, D, lr=1e-4, iters=iters) for _ in range(steps)]
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    peak = torch.cuda.max_memory_allocated()
    out.update({"ms_per_step": round(ms, 2), "loss": [round(v, 5) for v in losses],
                "peak_torch_allocated_gb": round(peak / 1e9, 3),
                "step_peak_above_start_gb": round((peak - base) / 1e9, 3)})
    return out


def train_steps(pkg, h, steps):
    """The training measurements: 4 x 256x512 (FP32 checkpointed, FP32
    store-everything, BF16 checkpointed) and 1488x992 D = 256 (BF16 checkpointed)."""
    return [
        train_bench(pkg, h, "rds-512x256-d128", 7, "all3x3", batch=4, steps=steps, warmup=1, iters=8),
        train_bench(pkg, h, "rds-512x256-d128", 7, "all3x3", batch=4, steps=steps, warmup=1, iters=8,
                    checkpoint=False, grad_check=False),
        train_bench(pkg, h, "rds-512x256-d128", 7, "all3x3", batch=4, steps=steps, warmup=1, iters=8,
                    precision="bf16"),
        train_bench(pkg, h, "rds-1488x992-d256", 7, "all3x3", batch=1, steps=max(1, steps - 1), warmup=1, iters=10,
                    precision="bf16"),
    ]


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="rds-1242x375-d128")
    ap.add_argument("--arch", default="all3x3", choices=["all3x3", "reference"])
    ap.add_argument("--layers", type=int, default=7)
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 form")
    ap.add_argument("--cpu-rows", type=int, default=0,
                    help="rows of the CPU-baseline band (0 = the whole frame)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-time-rows", type=int, default=94,
                    help="rows of the band the CPU baseline is timed on (median of 3 runs)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 320x240 and 744x496 RGB configurations")
    ap.add_argument("--no-train", action="store_true", help="skip the training-step measurements")
    ap.add_argument("--train-only", action="store_true",
                    help="run only the training-step measurements and print them as one JSON line")
    ap.add_argument("--streams", type=int, default=2, choices=[1, 2],
                    help="2: the right tower runs on a second stream beside the left one")
    ap.add_argument("--inflight", type=int, default=2,
                    help="frames in flight for the pipelined throughput figure (0: skip it)")
    ap.add_argument("--graph", type=int, default=1, choices=[0, 1],
                    help="1: also time the BF16 frame captured in a HIP graph and replayed")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (requested --gpus %d, WORLD_SIZE %d)"
              % (args.gpus, world), file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    pkg = _load_package()
    from cnn_crf_stereo_amd import capi, synth

    torch.cuda.set_device(0)
    if args.train_only:
        h = capi.Handle()
        print(json.dumps({"train_step": train_steps(pkg, h, args.steps if args.steps < 10 else 3)}))
        return
    spec = dict(synth.WORKLOADS[args.workload])
    data = synth.stereo_pair(args.workload)
    layers, pad = synth.architecture(args.layers, spec["C"], args.arch)
    params = synth.glorot_weights(layers)
    h = capi.Handle()
    H, W, D = spec["H"], spec["W"], spec["D"]
    pd_per_frame = float(H * W * D)

    side = None
    if args.streams == 2:
        s2 = torch.cuda.Stream()  # default priority
        side = (capi.Handle(s2), s2)
    ms_bf16, launches_bf16, out_bf16 = time_form(pkg, h, data, spec, params, "bf16", args.steps, args.warmup, side)
    result_extra = {"pcie_inclusive_ms_per_frame": round(launches_bf16.pop("_pcie_inclusive_ms"), 4)}
    result_extra["crf_ms_per_run"] = time_crf_forms(h, data, spec, out_bf16["cost"], max(3, args.steps // 2))
    if side is not None:
        # Per-launch figures come from a one-stream pass: with the towers on two
        # streams each event interval also spans the other tower's launches.
        ms_1s, launches_bf16, _ = time_form(pkg, h, data, spec, params, "bf16", args.steps, 1, None)
        launches_bf16.pop("_pcie_inclusive_ms")
        result_extra["ms_per_frame_one_stream"] = round(ms_1s, 4)
    if args.inflight > 1:
        lanes = []
        for _ in range(args.inflight):
            st = torch.cuda.Stream()  # default priority
            lanes.append((capi.Handle(st), st))
        pipe = {}
        for prec in ("bf16", "fp32") if not args.no_fp32 else ("bf16",):
            ms_p, out_p = time_pipelined(pkg, lanes, data, spec, params, prec, args.steps, 1)
            pipe[prec] = round(ms_p, 4)
            if prec == "bf16":
                assert bool((out_p["labels"] == out_bf16["labels"]).all()), "pipelined frames must match"
        result_extra["pipelined"] = {"frames_in_flight": args.inflight, "ms_per_frame": pipe,
                                     "pixel_disp_per_s": {k: pd_per_frame / (v / 1e3) for k, v in pipe.items()}}
    timing = "eager launches"
    ms_eager = ms_bf16
    if args.graph:
        # The headline frame time is the captured two-stream frame replayed; the
        # eager figure stays beside it.  A capture that fails or disagrees with
        # the eager labels leaves the eager figure as the headline.
        try:
            g = {}
            for prec in ("bf16",) + (() if args.no_fp32 else ("fp32",)):
                ms_g, same_g = time_graph(pkg, data, spec, params, prec, args.steps, args.streams == 2)
                g[prec] = {"ms_per_frame": round(ms_g, 4), "labels_equal_eager": same_g}
            result_extra["graph"] = g
            if g["bf16"]["labels_equal_eager"]:
                ms_bf16 = g["bf16"]["ms_per_frame"]
                timing = "HIP graph replay of the captured two-stream frame"
        except Exception as e:  # recorded, not fatal
            result_extra["graph"] = {"error": repr(e)[:300]}
        result_extra["eager_ms_per_frame"] = round(ms_eager, 4)
    if not args.no_fp32:
        ms_fp32, launches_fp32, out_fp32 = time_form(pkg, h, data, spec, params, "fp32", args.steps, args.warmup,
                                                     side)
        launches_fp32.pop("_pcie_inclusive_ms")
        if side is not None:
            _, launches_fp32, _ = time_form(pkg, h, data, spec, params, "fp32", max(2, args.steps // 2), 1, None)
            launches_fp32.pop("_pcie_inclusive_ms")
        g = result_extra.get("graph", {})
        if timing != "eager launches" and "fp32" in g and g["fp32"]["labels_equal_eager"]:
            result_extra["fp32_eager_ms_per_frame"] = round(ms_fp32, 4)
            ms_fp32 = g["fp32"]["ms_per_frame"]
        result_extra["fp32_ms_per_frame"] = round(ms_fp32, 4)
        result_extra["fp32_pixel_disp_per_s"] = pd_per_frame / (ms_fp32 / 1e3)
        result_extra["speedup_bf16_vs_fp32"] = round(ms_fp32 / ms_bf16, 3)
        nf = max(2, args.steps // 2) if side is not None else args.steps
        result_extra["stages_ms_fp32"] = {k: round(v["ms"] / nf, 4) for k, v in launches_fp32.items()}

    gt = data["disp"]
    bad = {"bf16": bad_pixel_rate(out_bf16["labels"].cpu().numpy(), gt, D)}
    if not args.no_fp32:
        bad["fp32"] = bad_pixel_rate(out_fp32["labels"].cpu().numpy(), gt, D)
        lab_b = out_bf16["labels"].cpu().numpy()
        lab_f = out_fp32["labels"].cpu().numpy()
        result_extra["labels_differ_bf16_vs_fp32"] = float((lab_b != lab_f).mean())
    result_extra["bad_px_gt_gt1"] = bad

    # Dominant kernel class of the BF16 form and its roofline.
    stages = {k: v["ms"] / args.steps for k, v in launches_bf16.items()}
    dom = max(launches_bf16, key=lambda k: launches_bf16[k]["ms"])
    d = launches_bf16[dom]
    mean_ms = d["ms"] / d["launches"]
    if d["flops"] > 0 and dom.startswith("conv"):
        per_launch = d["flops"] / d["launches"]
        achieved = per_launch / (mean_ms / 1e3) / 1e12
        roof = dict(kernel=dom, bound="mfma", achieved=round(achieved, 2), peak=MFMA_PEAK_TFLOPS["bf16"],
                    unit="TFLOP/s", frac=round(achieved / MFMA_PEAK_TFLOPS["bf16"], 4), traffic=None,
                    work_per_launch=per_launch, mean_launch_ms=round(mean_ms, 5))
    else:
        per_launch = d["bytes"] / d["launches"]
        achieved = per_launch / (mean_ms / 1e3) / 1e9
        roof = dict(kernel=dom, bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None, work_per_launch=per_launch,
                    mean_launch_ms=round(mean_ms, 5))
    traffic = measured_traffic(args.workload).get(dom)
    if traffic is not None and args.arch == "all3x3" and args.layers == 7:
        roof["traffic"] = traffic
    kernels = {}
    for k, v in launches_bf16.items():
        m = v["ms"] / v["launches"]
        e = dict(launches_per_frame=v["launches"] // args.steps, mean_launch_ms=round(m, 5))
        if v["flops"] and k.startswith("conv"):
            e["tflops"] = round(v["flops"] / v["launches"] / (m / 1e3) / 1e12, 2)
        if v["bytes"]:
            e["gbs"] = round(v["bytes"] / v["launches"] / (m / 1e3) / 1e9, 1)
        kernels[k] = e

    cpu = None
    if not args.no_cpu_baseline:
        cb = cpu_baseline(data, spec, params, pad, args.cpu_rows, repeats=3, time_rows=args.cpu_time_rows)
        cpu = dict(value=cb["pd_timed"] / cb["seconds"], unit="pixel*disp/s", cores=cpu_cores(), kind="port",
                   sample="FP64 oracle of every stage on a %dx%d band (rows %d-%d) of the same pair, D=%d, %d CRF "
                          "iterations: median of 3 runs after a warm-up run (the whole-frame oracle run), "
                          "%.2f s (runs %s s)" % (W, cb["time_rows"], cb["time_y0"],
                                                  cb["time_y0"] + cb["time_rows"] - 1, D, spec["iters"],
                                                  cb["seconds"], ", ".join("%.2f" % t for t in cb["times"])))
        # The same band through both GPU forms, compared with the oracle's labels.
        from cnn_crf_stereo_amd import ops, pipeline

        band_l = torch.as_tensor(cb["left"], device="cuda")
        band_r = torch.as_tensor(cb["right"], device="cuda")
        gt_band = gt[cb["y0"]:cb["y0"] + cb["rows"]]
        band = {"fp64_oracle": bad_pixel_rate(cb["labels"], gt_band, D)}
        for prec in (["bf16"] + ([] if args.no_fp32 else ["fp32"])):
            fr = pipeline.Frame(h, params, prec, D, iters=spec["iters"])
            o = fr(band_l, band_r)
            lab = o["labels"].cpu().numpy()
            band[prec] = bad_pixel_rate(lab, gt_band, D)
            band[prec + "_labels_equal_oracle"] = float((lab == cb["labels"]).mean())
            result_extra.setdefault("errors_vs_fp64_oracle", {})[prec] = stage_errors(ops, h, fr, o, cb)
        result_extra["bad_px_band_vs_oracle"] = band

    if not args.no_train:
        result_extra["train_step"] = train_steps(pkg, h, 3)

    if not args.no_configs:
        result_extra["configs"] = {
            "rds-320x240-d32": [time_config(pkg, h, "rds-320x240-d32", "fp32", 10, cpu_repeats=3),
                                time_config(pkg, h, "rds-320x240-d32", "bf16", 10)],
            "rds-744x496-rgb-d64": [time_config(pkg, h, "rds-744x496-rgb-d64", "fp32", 10),
                                    time_config(pkg, h, "rds-744x496-rgb-d64", "bf16", 10)],
        }

    # Each stage's floor at the chip's peak (dense BF16 MFMA 2.5 PFLOP/s, HBM
    # 8 TB/s) with its algorithmic work: the conv FLOPs, the correlation's
    # feature reads and volume write, 3 volume passes per CRF half-step (read C,
    # read the other family's dual, write the dual), one volume read for the
    # decode; the frame time as a fraction of their sum.
    tower_flops = 0.0
    hh, ww = H + 2 * pad, W + 2 * pad
    for (K, C, KH, KW) in layers:
        hh, ww = hh - KH + 1, ww - KW + 1
        tower_flops += 2.0 * K * C * KH * KW * hh * ww
    V = 4.0 * H * W * D
    cs = (layers[-1][0] + 7) // 8 * 8
    floors = {
        "conv": 2 * tower_flops / (MFMA_PEAK_TFLOPS["bf16"] * 1e12) * 1e3,
        "correlation": (2.0 * H * W * cs * 2 + V) / (HBM_PEAK_GBS * 1e9) * 1e3,
        "crf_sweeps": spec["iters"] * 2 * 3 * V / (HBM_PEAK_GBS * 1e9) * 1e3,
        "crf_decode": V / (HBM_PEAK_GBS * 1e9) * 1e3,
        "normalize_edges_remap": (2 * 3 * 4.0 * (H + 2 * pad) * (W + 2 * pad) + 12.0 * H * W + 8.0 * H * W)
                                 / (HBM_PEAK_GBS * 1e9) * 1e3,
    }
    result_extra["frame_roofline"] = {
        "floor_ms": {k: round(v, 4) for k, v in floors.items()},
        "sum_floor_ms": round(sum(floors.values()), 4),
        "frame_ms": round(ms_bf16, 4),
        "frac": round(sum(floors.values()) / ms_bf16, 4),
    }

    line = {
        "metric": METRIC,
        "value": pd_per_frame / (ms_bf16 / 1e3),
        "unit": "pixel*disp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_bf16, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": "%s: grayscale random-dot pair %dx%d, D=%d, 7-layer 100-channel CNN (%s), "
                               "%d SlackProp iterations" % (args.workload, W, H, D, args.arch, spec["iters"]),
                   "form": "BF16-MFMA CNN + correlation (FP32 accumulate), FP32 CRF min-sum, FP64 bound",
                   "streams": "towers on %d HIP stream%s" % (args.streams, "s" if args.streams > 1 else ""),
                   "timing": timing},
        "roofline": roof,
        "cpu_baseline": cpu,
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "kernels": kernels,
    }
    line.update(result_extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
