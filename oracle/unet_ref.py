"""CPU oracle of the learned velocity-preconditioner correction (test infrastructure).

Only tests, smoke() and bench.py's CPU-baseline leg use this module, and only
as the checker.  The network has no counterpart in the reference (it is an
addition beside the drop-in path), so the oracle is pinned by mathematics
instead of golden vectors:

* the vectorised 3x3 convolution agrees with a brute-force loop
  (``conv3x3_loops``);
* in exact mode (FP64, no BF16 rounding) the backward pass agrees with
  central finite differences of the forward pass along random directions;
* Adam's first step from zero moments reduces to lr * g / (|g| + eps).

Network (same layer table as the HIP kernels): 3x3 convolutions with zero
padding and ReLU, 2x2 average pooling, nearest upsampling, channel concat.

    e1a = conv(in[8] -> 16)         e1 = conv(e1a -> 16)          H x W
    e2a = conv(pool(e1) -> 32)      e2 = conv(e2a -> 32)          H/2 x W/2
    e3a = conv(pool(e2) -> 32)      e3 = conv(e3a -> 32)          H/4 x W/4
    d2a = conv(up(e3) || e2 -> 32)  d2 = conv(d2a -> 32)          H/2 x W/2
    d1a = conv(up(d2) || e1 -> 16)  d1 = conv(d1a -> 16)          H x W
    out = conv1x1(d1 -> 2)

Parameters (flat FP32): for each conv W[cout][9*cin] (k = tap*cin + ci,
tap = 3*(dy+1) + (dx+1)), then b[cout]; then Wo[2][16], bo[2].
Loss = 0.5 * sum((out - t)^2) / (2 H W).

In ``bf16=True`` mode the oracle rounds where the GPU kernels store BF16:
weights of the 3x3 convolutions, every activation (after bias and ReLU),
pooled tensors, the masked output gradients dz and the data gradients dx.
Sums are FP64 (the GPU's are FP32), so the two agree to the BF16 rounding
of values that fall near a rounding boundary.
"""
from __future__ import annotations

import numpy as np

# (c1, up1, c2, cout, level) -- first input channels (upsampled by 2**up1),
# skip channels, outputs, pyramid level
SPEC = [
    (8, 0, 0, 16, 0), (16, 0, 0, 16, 0),
    (16, 0, 0, 32, 1), (32, 0, 0, 32, 1),
    (32, 0, 0, 32, 2), (32, 0, 0, 32, 2),
    (32, 1, 32, 32, 1), (32, 0, 0, 32, 1),
    (32, 1, 16, 16, 0), (16, 0, 0, 16, 0),
]
N_IN, N_OUT = 8, 2


def param_count() -> int:
    n = sum(co * 9 * (c1 + c2) + co for c1, _, c2, co, _ in SPEC)
    return n + N_OUT * 16 + N_OUT


def split_params(p: np.ndarray):
    """Views (W_l as (cout, 9*cin), b_l) for each conv, then Wo (2,16), bo (2,)."""
    off = 0
    Ws, bs = [], []
    for c1, _, c2, co, _ in SPEC:
        k9 = 9 * (c1 + c2)
        Ws.append(p[off:off + co * k9].reshape(co, k9))
        off += co * k9
        bs.append(p[off:off + co])
        off += co
    Wo = p[off:off + N_OUT * 16].reshape(N_OUT, 16)
    off += N_OUT * 16
    bo = p[off:off + N_OUT]
    off += N_OUT
    assert off == p.size
    return Ws, bs, Wo, bo


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 (ties to even), returned as float32 values."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(f.shape)


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """BF16 bit patterns (uint16) of x (round to nearest even)."""
    return (bf16(x).view(np.uint32) >> 16).astype(np.uint16)


def from_bf16_bits(b: np.ndarray) -> np.ndarray:
    return (b.astype(np.uint32) << 16).view(np.float32)


def _taps():
    return [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def _shift(x: np.ndarray, dy: int, dx: int) -> np.ndarray:
    """s[y, x] = x[y + dy, x + dx] with zeros outside (x is H x W x C)."""
    H, W = x.shape[:2]
    s = np.zeros_like(x)
    ys, yd = (slice(dy, H), slice(0, H - dy)) if dy >= 0 else (slice(0, H + dy), slice(-dy, H))
    xs, xd = (slice(dx, W), slice(0, W - dx)) if dx >= 0 else (slice(0, W + dx), slice(-dx, W))
    s[yd, xd] = x[ys, xs]
    return s


def im2col(x: np.ndarray) -> np.ndarray:
    """(H, W, cin) -> (H*W, 9*cin), column k = tap*cin + ci."""
    H, W, c = x.shape
    return np.concatenate([_shift(x, dy, dx) for dy, dx in _taps()], axis=2).reshape(H * W, 9 * c)


def conv3x3(x: np.ndarray, Wm: np.ndarray, b: np.ndarray | None) -> np.ndarray:
    H, W, _ = x.shape
    y = im2col(x.astype(np.float64)) @ Wm.astype(np.float64).T
    if b is not None:
        y = y + b.astype(np.float64)
    return y.reshape(H, W, Wm.shape[0])


def conv3x3_loops(x: np.ndarray, Wm: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Brute-force restatement of conv3x3 (for tiny inputs)."""
    H, W, cin = x.shape
    cout = Wm.shape[0]
    y = np.zeros((H, W, cout))
    for yy in range(H):
        for xx in range(W):
            for co in range(cout):
                s = float(b[co])
                for t, (dy, dx) in enumerate(_taps()):
                    sy, sx = yy + dy, xx + dx
                    if 0 <= sy < H and 0 <= sx < W:
                        for ci in range(cin):
                            s += float(Wm[co, t * cin + ci]) * float(x[sy, sx, ci])
                y[yy, xx, co] = s
    return y


def conv3x3_data_grad(dz: np.ndarray, Wm: np.ndarray, cin: int) -> np.ndarray:
    """dx[q, ci] = sum_{co, tap} W[co, tap*cin+ci] dz[q - off(tap), co]."""
    H, W, cout = dz.shape
    dx = np.zeros((H, W, cin))
    dz64 = dz.astype(np.float64)
    for t, (dy, ddx) in enumerate(_taps()):
        Wt = Wm[:, t * cin:(t + 1) * cin].astype(np.float64)  # (cout, cin)
        dx += _shift(dz64, -dy, -ddx) @ Wt
    return dx


def pool2(x: np.ndarray) -> np.ndarray:
    H, W, c = x.shape
    return 0.25 * x.reshape(H // 2, 2, W // 2, 2, c).sum(axis=(1, 3))


def up2(x: np.ndarray) -> np.ndarray:
    return x.repeat(2, axis=0).repeat(2, axis=1)


def sum2x2(x: np.ndarray) -> np.ndarray:
    H, W, c = x.shape
    return x.reshape(H // 2, 2, W // 2, 2, c).sum(axis=(1, 3))


def forward(p: np.ndarray, img: np.ndarray, bf: bool = True):
    """img: (H, W, 8) input values.  Returns (out (H, W, 2), cache)."""
    Ws, bs, Wo, bo = split_params(np.asarray(p, dtype=np.float64 if not bf else np.float32))
    rnd = bf16 if bf else (lambda a: a)
    Wc = [rnd(w) for w in Ws]
    x0 = rnd(img)
    act, ins = [], []
    p1 = p2 = None
    for l, (c1, up1, c2, co, lev) in enumerate(SPEC):
        if l == 0:
            xin = x0
        elif l == 2:
            p1 = rnd(pool2(act[1]))
            xin = p1
        elif l == 4:
            p2 = rnd(pool2(act[3]))
            xin = p2
        elif l == 6:
            xin = np.concatenate([up2(act[5]), act[3]], axis=2)
        elif l == 8:
            xin = np.concatenate([up2(act[7]), act[1]], axis=2)
        else:
            xin = act[l - 1]
        y = np.maximum(conv3x3(xin, Wc[l], bs[l]), 0.0)
        ins.append(xin)
        act.append(rnd(y))
    d1 = act[9].astype(np.float64)
    out = d1 @ Wo.astype(np.float64).T + bo.astype(np.float64)
    return out, dict(Wc=Wc, Wo=Wo, act=act, ins=ins, rnd=rnd)


def loss_and_grads(p: np.ndarray, img: np.ndarray, target: np.ndarray, bf: bool = True):
    """Loss 0.5*sum((out-t)^2)/(2HW) and its gradient (flat, FP64)."""
    out, cache = forward(p, img, bf)
    H, W, _ = img.shape
    n = H * W * N_OUT
    d = out - target.astype(np.float64)
    loss = 0.5 * float((d * d).sum()) / n
    go = d / n
    act, ins, Wc, Wo, rnd = cache["act"], cache["ins"], cache["Wc"], cache["Wo"], cache["rnd"]
    gWs = [None] * len(SPEC)
    gbs = [None] * len(SPEC)
    d1 = act[9].astype(np.float64).reshape(H * W, 16)
    gWo = go.reshape(H * W, N_OUT).T @ d1
    gbo = go.reshape(H * W, N_OUT).sum(axis=0)
    gact = [np.zeros(a.shape) for a in act]
    gact[9] += go @ Wo.astype(np.float64)
    for l in range(len(SPEC) - 1, -1, -1):
        c1, up1, c2, co, lev = SPEC[l]
        cin = c1 + c2
        Hl, Wl = act[l].shape[:2]
        dz = rnd(gact[l] * (act[l] > 0))
        dz2 = dz.astype(np.float64).reshape(Hl * Wl, co)
        gWs[l] = dz2.T @ im2col(ins[l].astype(np.float64))
        gbs[l] = dz2.sum(axis=0)
        if l == 0:
            break
        dx = rnd(conv3x3_data_grad(dz, Wc[l], cin)).astype(np.float64)
        if l == 2:
            gact[1] += 0.25 * up2(dx)
        elif l == 4:
            gact[3] += 0.25 * up2(dx)
        elif l == 6:
            gact[5] += sum2x2(dx[..., :32])
            gact[3] += dx[..., 32:]
        elif l == 8:
            gact[7] += sum2x2(dx[..., :32])
            gact[1] += dx[..., 32:]
        else:
            gact[l - 1] += dx
    parts = []
    for gW, gb in zip(gWs, gbs):
        parts += [gW.ravel(), gb.ravel()]
    parts += [gWo.ravel(), gbo.ravel()]
    return loss, out, np.concatenate(parts)


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015), bias-corrected; returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return p - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v


def init_params(seed: int, scale_out: float = 1e-3) -> np.ndarray:
    """He-normal weights, zero biases, small output layer (numpy RNG; the GPU
    library draws its own initial weights)."""
    rng = np.random.default_rng(seed)
    parts = []
    for c1, _, c2, co, _ in SPEC:
        k9 = 9 * (c1 + c2)
        parts += [rng.normal(0, np.sqrt(2.0 / k9), co * k9), np.zeros(co)]
    parts += [rng.normal(0, scale_out, N_OUT * 16), np.zeros(N_OUT)]
    return np.concatenate(parts).astype(np.float32)
