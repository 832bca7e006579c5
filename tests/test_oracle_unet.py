"""Pins the CPU oracle of the learned correction's network (oracle/unet_ref.py).

The network has no reference counterpart, so the oracle is pinned by
independent implementations and mathematics: torch's BF16 conversion, a
brute-force convolution loop, torch.nn.functional.conv2d, torch autograd of
the same network in FP64, central finite differences, and Adam's closed-form
first step.
"""
import os
import sys

import numpy as np
import pytest
import torch
import torch.nn.functional as TF

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
from oracle import unet_ref as U  # noqa: E402


def test_param_count_by_hand():
    # 3x3 convs: cout*9*cin + cout; then the 1x1 layer 2*16 + 2
    convs = [(8, 16), (16, 16), (16, 32), (32, 32), (32, 32), (32, 32), (64, 32), (32, 32), (48, 16), (16, 16)]
    n = sum(co * 9 * ci + co for ci, co in convs) + 34
    assert n == 72866 == U.param_count()


def test_param_count_matches_library():
    lib_path = os.path.join(ROOT, "drd_application_amd", "libdrd_gpu.so")
    if not os.path.exists(lib_path):
        pytest.skip("library not built")
    from drd_application_amd import solver
    assert solver.unet_param_count() == U.param_count()


def test_bf16_rounding_matches_torch_and_ties_to_even():
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.normal(0, 1, 4000), rng.normal(0, 1e-3, 1000), rng.normal(0, 1e4, 1000)])
    x = x.astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    assert np.array_equal(U.bf16(x), want)
    ties = np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -(1.0 + 2 ** -8)], np.float32)
    assert U.bf16(ties).tolist() == [1.0, 1.0 + 2 ** -6, -1.0]
    bits = U.bf16_bits(x)
    assert np.array_equal(U.from_bf16_bits(bits), want)


def test_conv_matches_brute_force_loop():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(5, 6, 3))
    W = rng.normal(size=(4, 27))
    b = rng.normal(size=4)
    assert np.allclose(U.conv3x3(x, W, b), U.conv3x3_loops(x, W, b), rtol=0, atol=1e-12)


def test_conv_matches_torch_conv2d():
    rng = np.random.default_rng(2)
    H, W, cin, cout = 7, 9, 5, 6
    x = rng.normal(size=(H, W, cin))
    Wm = rng.normal(size=(cout, 9 * cin))
    b = rng.normal(size=cout)
    # W[co][tap*cin + ci], tap = 3*ky + kx (ky = dy + 1) -> torch weight[co][ci][ky][kx]
    wt = torch.from_numpy(Wm.reshape(cout, 3, 3, cin).transpose(0, 3, 1, 2).copy())
    xt = torch.from_numpy(x.transpose(2, 0, 1)[None].copy())
    yt = TF.conv2d(xt, wt, torch.from_numpy(b), padding=1)[0].numpy().transpose(1, 2, 0)
    assert np.allclose(U.conv3x3(x, Wm, b), yt, rtol=0, atol=1e-12)


def _torch_net_loss(p: torch.Tensor, img: torch.Tensor, target: torch.Tensor):
    """The same network written with torch primitives (FP64, NCHW)."""
    off = 0
    Ws, bs = [], []
    for c1, _, c2, co, _ in U.SPEC:
        cin = c1 + c2
        w = p[off:off + co * 9 * cin].reshape(co, 3, 3, cin).permute(0, 3, 1, 2)
        off += co * 9 * cin
        Ws.append(w)
        bs.append(p[off:off + co])
        off += co
    Wo = p[off:off + 32].reshape(2, 16)
    bo = p[off + 32:off + 34]

    def conv(x, l):
        return torch.relu(TF.conv2d(x, Ws[l], bs[l], padding=1))

    def up(x):
        return x.repeat_interleave(2, dim=2).repeat_interleave(2, dim=3)

    x = img.permute(2, 0, 1)[None]
    e1 = conv(conv(x, 0), 1)
    e2 = conv(conv(TF.avg_pool2d(e1, 2), 2), 3)
    e3 = conv(conv(TF.avg_pool2d(e2, 2), 4), 5)
    d2 = conv(conv(torch.cat([up(e3), e2], 1), 6), 7)
    d1 = conv(conv(torch.cat([up(d2), e1], 1), 8), 9)
    out = torch.einsum("oc,nchw->nhwo", Wo, d1)[0] + bo
    n = out.numel()
    return 0.5 * ((out - target) ** 2).sum() / n, out


def _case(seed, H=8, W=12, scale_out=0.3):
    rng = np.random.default_rng(seed)
    p = U.init_params(seed, scale_out=scale_out).astype(np.float64)
    # nonzero biases so every ReLU pattern is exercised
    Ws, bs, _, bo = U.split_params(p)
    for b in bs:
        b[:] = rng.normal(0, 0.1, b.shape)
    bo[:] = rng.normal(0, 0.1, bo.shape)
    img = rng.normal(size=(H, W, 8))
    target = rng.normal(size=(H, W, 2))
    return p, img, target


def test_exact_backward_matches_torch_autograd():
    p, img, target = _case(3)
    loss, out, g = U.loss_and_grads(p, img, target, bf=False)
    pt = torch.from_numpy(p.copy()).requires_grad_(True)
    lt, ot = _torch_net_loss(pt, torch.from_numpy(img), torch.from_numpy(target))
    lt.backward()
    assert np.allclose(out, ot.detach().numpy(), rtol=1e-12, atol=1e-12)
    assert abs(loss - lt.item()) <= 1e-12 * abs(loss)
    gt = pt.grad.numpy()
    assert np.linalg.norm(g - gt) <= 1e-10 * np.linalg.norm(gt)


def test_exact_backward_matches_finite_differences_per_layer():
    p, img, target = _case(4)
    _, _, g = U.loss_and_grads(p, img, target, bf=False)
    rng = np.random.default_rng(9)
    # one random direction inside each parameter block (conv weights, biases, output layer)
    blocks, off = [], 0
    for c1, _, c2, co, _ in U.SPEC:
        blocks += [(off, off + co * 9 * (c1 + c2)), (off + co * 9 * (c1 + c2), off + co * 9 * (c1 + c2) + co)]
        off += co * 9 * (c1 + c2) + co
    blocks += [(off, off + 32), (off + 32, off + 34)]
    h = 1e-6
    for a, b in blocks:
        v = np.zeros_like(p)
        v[a:b] = rng.normal(size=b - a)
        lp = U.loss_and_grads(p + h * v, img, target, bf=False)[0]
        lm = U.loss_and_grads(p - h * v, img, target, bf=False)[0]
        fd = (lp - lm) / (2 * h)
        an = float(g @ v)
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-12) + 1e-12, (a, b, fd, an)


def test_bf16_mode_stays_close_to_exact():
    # the rounding hooks change values by BF16 rounding only
    p, img, target = _case(6)
    l0, o0, g0 = U.loss_and_grads(p, img, target, bf=False)
    l1, o1, g1 = U.loss_and_grads(p, img, target, bf=True)
    assert np.linalg.norm(o1 - o0) <= 0.05 * np.linalg.norm(o0)
    assert np.linalg.norm(g1 - g0) <= 0.05 * np.linalg.norm(g0)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(0)
    p = rng.normal(size=100)
    g = rng.normal(size=100)
    lr, eps = 1e-3, 1e-8
    p1, m, v = U.adam_step(p, g, np.zeros(100), np.zeros(100), 1, lr, eps=eps)
    assert np.allclose(p1, p - lr * g / (np.abs(g) + eps), rtol=0, atol=1e-15)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)
