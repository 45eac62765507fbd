"""Seeded synthetic random-dot stereo pairs and network weights.

Left image: uniform noise blurred with a Gaussian (sigma 1), min-max scaled to
[0, 1] (independently per channel for colour).  Disparity: piecewise-planar,
one random plane per cell of a Voronoi partition, clipped to [0, dmax].
Right image: R(x - d, y) = L(x, y) -- the matching convention of the
reference's correlation (offset -d, stereo/operators/stereocorrelation.cpp:83-85)
-- built by forward warping with linear interpolation and a depth buffer
(nearest surface wins), plus N(0, 2/255) noise; pixels no left pixel reaches
are filled with fresh blurred noise.
"""
import numpy as np

# Workloads named by size (width x height, channels, disparities, planes, CRF iterations).
WORKLOADS = {
    "rds-320x240-d32": dict(W=320, H=240, C=1, D=32, planes=6, iters=5),
    "rds-744x496-rgb-d64": dict(W=744, H=496, C=3, D=64, planes=10, iters=8),
    "rds-1242x375-d128": dict(W=1242, H=375, C=1, D=128, planes=12, iters=8),
    "rds-512x256-d128": dict(W=512, H=256, C=1, D=128, planes=12, iters=8),
    "rds-1488x992-d256": dict(W=1488, H=992, C=1, D=256, planes=16, iters=10),
}


def _reflect(i, n):
    if n == 1:
        return np.zeros_like(i)
    period = 2 * n
    i = np.mod(i, period)
    return np.where(i < n, i, period - 1 - i)


def gaussian_blur(img, sigma):
    """Separable Gaussian (radius ceil(3 sigma)), symmetric border, last two axes."""
    r = int(np.ceil(3 * sigma))
    t = np.arange(-r, r + 1)
    k = np.exp(-(t * t) / (2.0 * sigma * sigma))
    k /= k.sum()
    H, W = img.shape[-2:]
    out = np.zeros_like(img, dtype=np.float64)
    xi = np.arange(W)
    for j, kv in zip(t, k):
        out += kv * img[..., _reflect(xi + j, W)]
    tmp = out
    out = np.zeros_like(tmp)
    yi = np.arange(H)
    for j, kv in zip(t, k):
        out += kv * tmp[..., _reflect(yi + j, H), :]
    return out


def _blurred_noise(rng, C, H, W):
    n = rng.uniform(0.0, 1.0, size=(C, H, W))
    b = gaussian_blur(n, 1.0)
    lo = b.min(axis=(1, 2), keepdims=True)
    hi = b.max(axis=(1, 2), keepdims=True)
    return (b - lo) / np.maximum(hi - lo, 1e-12)


def disparity_field(rng, H, W, planes, dmax):
    """Piecewise-planar disparity over a Voronoi partition, values in [0, dmax]."""
    sx = rng.uniform(0, W, planes)
    sy = rng.uniform(0, H, planes)
    c = rng.uniform(0.1 * dmax, 0.9 * dmax, planes)
    a = rng.uniform(-0.15, 0.15, planes) * dmax / W
    b = rng.uniform(-0.15, 0.15, planes) * dmax / H
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    dist = (xx[None] - sx[:, None, None]) ** 2 + (yy[None] - sy[:, None, None]) ** 2
    cell = np.argmin(dist, axis=0)
    d = a[cell] * (xx - sx[cell]) + b[cell] * (yy - sy[cell]) + c[cell]
    return np.clip(d, 0.0, float(dmax))


def warp_right(left, disp, fill):
    """R(x - d(x, y), y) = L(x, y) by forward warping consecutive left pixel pairs
    of one surface (|d1 - d0| < 1) with linear interpolation and a depth buffer."""
    C, H, W = left.shape
    x = np.arange(W, dtype=np.float64)
    xr = x[None, :] - disp
    x0, x1 = xr[:, :-1], xr[:, 1:]
    d0, d1 = disp[:, :-1], disp[:, 1:]
    cont = np.abs(d1 - d0) < 1.0
    ys_all, xs_all, dd_all, t_all, src_all = [], [], [], [], []
    yy = np.broadcast_to(np.arange(H)[:, None], x0.shape)
    src = np.broadcast_to(np.arange(W - 1)[None, :], x0.shape)
    start = np.ceil(x0)
    for off in range(3):
        xt = start + off
        ok = cont & (xt <= x1) & (xt >= 0) & (xt < W)
        t = np.where(ok, (xt - x0) / np.where(cont, x1 - x0, 1.0), 0.0)
        ys_all.append(yy[ok])
        xs_all.append(xt[ok].astype(np.int64))
        dd_all.append((d0 + t * (d1 - d0))[ok])
        t_all.append(t[ok])
        src_all.append(src[ok])
    ys = np.concatenate(ys_all)
    xs = np.concatenate(xs_all)
    dd = np.concatenate(dd_all)
    tt = np.concatenate(t_all)
    ss = np.concatenate(src_all)
    order = np.argsort(dd, kind="stable")[::-1]  # nearest surface (largest d) first
    lin = ys[order] * W + xs[order]
    _, first = np.unique(lin, return_index=True)
    sel = order[first]
    right = fill.copy()
    covered = np.zeros((H, W), dtype=bool)
    yv, xv, tv, sv = ys[sel], xs[sel], tt[sel], ss[sel]
    for c in range(C):
        right[c, yv, xv] = (1.0 - tv) * left[c, yv, sv] + tv * left[c, yv, sv + 1]
    covered[yv, xv] = True
    return right, covered


def stereo_pair(workload, seed=None):
    """Returns dict(left, right: float32 C x H x W in ~[0,1]; disp: float64 H x W;
    D, iters, covered)."""
    spec = WORKLOADS[workload] if isinstance(workload, str) else dict(workload)
    if seed is None:
        seed = 1000 + sorted(WORKLOADS).index(workload) if isinstance(workload, str) else 1000
    rng = np.random.default_rng(seed)
    C, H, W, D = spec["C"], spec["H"], spec["W"], spec["D"]
    left = _blurred_noise(rng, C, H, W)
    disp = disparity_field(rng, H, W, spec["planes"], D - 1)
    fill = _blurred_noise(rng, C, H, W)
    right, covered = warp_right(left, disp, fill)
    right = right + rng.normal(0.0, 2.0 / 255.0, size=right.shape)
    return dict(left=left.astype(np.float32), right=right.astype(np.float32), disp=disp,
                D=D, iters=spec.get("iters", 5), covered=covered)


def architecture(num_layers, in_channels, kind="reference", channels=100):
    """Layer list [(K, C, KH, KW)] of one tower.

    kind "reference": 3x3 first layer then 2x2 layers (stereo/stereonet.cpp:69-129);
    kind "all3x3": every layer 3x3.  Returns (layers, pad) where pad is the
    mirror padding that keeps the output at the input size
    (stereo/stereo.cpp:204-217 for the reference kind)."""
    layers = [(channels, in_channels, 3, 3)]
    k = 2 if kind == "reference" else 3
    for _ in range(num_layers - 1):
        layers.append((channels, channels, k, k))
    shrink = sum(kh - 1 for (_, _, kh, _) in layers)
    assert shrink % 2 == 0
    return layers, shrink // 2


def glorot_weights(layers, seed=7):
    """Glorot-uniform filters a = sqrt(6 / (fan_in + fan_out)), fan = channels * KH * KW; zero bias."""
    rng = np.random.default_rng(seed)
    params = []
    for (K, C, KH, KW) in layers:
        a = np.sqrt(6.0 / (C * KH * KW + K * KH * KW))
        params.append((rng.uniform(-a, a, size=(K, C, KH, KW)).astype(np.float32),
                       np.zeros(K, dtype=np.float32)))
    return params
