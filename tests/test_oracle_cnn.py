"""The network oracle (oracle/oracle_cnn.py) pinned against library routines
and calculus on the CPU: torch's conv3d, autograd, GELU, Adam and bfloat16
conversion, and central finite differences.  The network has no counterpart
in the reference, so these checks are what fixes it."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import oracle_cnn as oc

torch.set_num_threads(4)


def rnd(seed, *shape, scale=1.0):
    return np.random.default_rng(seed).standard_normal(shape) * scale


def torch_weight(W):
    """W[co][tap][ci] -> torch [co][ci][kd][kh][kw]"""
    co, _, ci = W.shape
    return torch.from_numpy(W.reshape(co, 3, 3, 3, ci).transpose(0, 4, 1, 2, 3).copy())


def torch_conv(x, W, b, periodic):
    xt = torch.from_numpy(np.ascontiguousarray(x.transpose(0, 4, 1, 2, 3)))
    if periodic:
        xt = F.pad(xt, (1, 1, 0, 0, 1, 1), mode="circular")
        xt = F.pad(xt, (0, 0, 1, 1, 0, 0))
        y = F.conv3d(xt, torch_weight(W), torch.from_numpy(b) if b is not None else None)
    else:
        y = F.conv3d(xt, torch_weight(W), torch.from_numpy(b) if b is not None else None, padding=1)
    return y.numpy().transpose(0, 2, 3, 4, 1)


def random_params(seed, scale=1.0):
    layers = []
    rng = np.random.default_rng(seed)
    for cin, _, cout, _, _, _ in oc.LAYERS:
        layers.append((rng.standard_normal((cout, 27, cin)) * scale / np.sqrt(27 * cin),
                       rng.standard_normal(cout) * 0.1))
    return oc.pack(layers)


@pytest.mark.parametrize("periodic", [False, True])
def test_conv_matches_torch(periodic):
    x = rnd(1, 2, 5, 4, 6, 7)
    W = rnd(2, 3, 27, 7)
    b = rnd(3, 3)
    y = oc.conv(x, W, b, periodic)
    ref = torch_conv(x, W, b, periodic)
    assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_conv_transpose_and_wgrad_are_adjoints():
    # <conv(x, W), dz> = <x, conv_t(dz, W)> = <W, wgrad(dz, x)>
    x = rnd(4, 1, 4, 5, 6, 3)
    W = rnd(5, 2, 27, 3)
    dz = rnd(6, 1, 4, 5, 6, 2)
    for per in (False, True):
        a = np.sum(oc.conv(x, W, None, per) * dz)
        b = np.sum(x * oc.conv_t(dz, W, per))
        c = np.sum(W * oc.wgrad(dz, x, per))
        assert abs(a - b) <= 1e-12 * abs(a) and abs(a - c) <= 1e-12 * abs(a)


def torch_net(params, x, periodic):
    layers = oc.unpack(params)
    ts = []
    a = torch.from_numpy(np.ascontiguousarray(x.transpose(0, 4, 1, 2, 3)))
    for l, (W, b) in enumerate(layers):
        Wt = torch_weight(W).requires_grad_(True)
        bt = torch.from_numpy(b).requires_grad_(True)
        ts.append((Wt, bt))
        if periodic:
            a = F.conv3d(F.pad(F.pad(a, (1, 1, 0, 0, 1, 1), mode="circular"), (0, 0, 1, 1, 0, 0)), Wt, bt)
        else:
            a = F.conv3d(a, Wt, bt, padding=1)
        if l < 3:
            a = F.gelu(a)
    return a, ts


@pytest.mark.parametrize("periodic", [False, True])
def test_network_gradient_matches_autograd(periodic):
    params = random_params(7)
    x = rnd(8, 2, 4, 6, 5, 15)
    t = rnd(9, 2, 4, 6, 5, 5, scale=0.3)
    mask = (np.random.default_rng(10).random((2, 4, 6, 5)) < 0.7).astype(np.uint8)
    loss, y, g = oc.loss_grad(params, x, t, mask, periodic)
    yt, ts = torch_net(params, x, periodic)
    mt = torch.from_numpy(mask.astype(np.float64))[:, None]
    e = yt - torch.from_numpy(np.ascontiguousarray(t.transpose(0, 4, 1, 2, 3)))
    lt = torch.sum(mt * e * e) / (5.0 * mask.sum())
    lt.backward()
    assert abs(loss - lt.item()) <= 1e-13 * abs(loss)
    assert np.max(np.abs(y - yt.detach().numpy().transpose(0, 2, 3, 4, 1))) <= 1e-12
    gl = oc.unpack(g)
    for (gW, gb), (Wt, bt) in zip(gl, ts):
        ref = Wt.grad.numpy().transpose(0, 2, 3, 4, 1).reshape(gW.shape)
        assert np.max(np.abs(gW - ref)) <= 1e-12 * np.max(np.abs(ref))
        assert np.max(np.abs(gb - bt.grad.numpy())) <= 1e-12 * max(np.max(np.abs(bt.grad.numpy())), 1e-30)


def test_network_gradient_finite_differences():
    params = random_params(11)
    x = rnd(12, 1, 3, 4, 4, 15)
    t = rnd(13, 1, 3, 4, 4, 5, scale=0.3)
    mask = np.ones((1, 3, 4, 4), np.uint8)
    _, _, g = oc.loss_grad(params, x, t, mask)
    rng = np.random.default_rng(14)
    idx = []
    for cin, cinp, cout, coutp, wo, bo in oc.LAYERS:
        for _ in range(3):
            co, tap, ci = rng.integers(cout), rng.integers(27), rng.integers(cin)
            idx.append(wo + (co * 27 + tap) * cinp + ci)
        idx.append(bo + rng.integers(cout))
    h = 1e-6
    for i in idx:
        p1 = params.copy(); p1[i] += h
        p2 = params.copy(); p2[i] -= h
        fd = (oc.loss_grad(p1, x, t, mask)[0] - oc.loss_grad(p2, x, t, mask)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(g[i]), 1e-3), (i, fd, g[i])
    # padded entries carry no gradient
    full = oc.pack(oc.unpack(np.ones(oc.NPARAM)))
    assert np.all(g[full == 0] == 0)


def test_gelu_matches_torch():
    z = np.linspace(-8, 8, 2001)
    zt = torch.from_numpy(z).requires_grad_(True)
    gt = F.gelu(zt)
    gt.sum().backward()
    assert np.max(np.abs(oc.gelu(z) - gt.detach().numpy())) <= 1e-15 * 8
    assert np.max(np.abs(oc.gelu_grad(z) - zt.grad.numpy())) <= 1e-14


def test_adam_matches_torch():
    rng = np.random.default_rng(15)
    w0 = rng.standard_normal(50)
    wt = torch.tensor(w0.copy(), requires_grad=True)
    opt = torch.optim.Adam([wt], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    w, m1, m2 = w0.copy(), np.zeros(50), np.zeros(50)
    for t in range(1, 6):
        g = rng.standard_normal(50)
        wt.grad = torch.from_numpy(g.copy())
        opt.step()
        w, m1, m2 = oc.adam(w, m1, m2, g, t, 1e-3)
    assert np.max(np.abs(w - wt.detach().numpy())) <= 1e-14


def test_bf16_matches_torch():
    rng = np.random.default_rng(16)
    x = np.concatenate([rng.standard_normal(100000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 100000),
                        np.array([0.0, -0.0, np.inf, -np.inf, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8,
                                  3.0e38, 1e-40, -1e-41], np.float32)]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(oc.bf16(x).view(np.uint32), ref.view(np.uint32))
    assert np.isnan(oc.bf16(np.array([np.nan], np.float32)))[0]


def test_emulation_stays_near_fp64():
    # BF16 storage costs about 2^-8 relative per rounding; the emulated
    # network stays within a few percent of the FP64 one
    params = random_params(17)
    x = rnd(18, 1, 4, 4, 16, 15)
    y64, _ = oc.forward(params, x)
    y16, _ = oc.forward(params, x, emulate=True)
    err = np.max(np.abs(y16 - y64)) / np.max(np.abs(y64))
    assert 1e-5 < err < 5e-2


def test_init_params_layout():
    p = oc.init_params(1234)
    layers = oc.unpack(p)
    assert np.all(oc.pack(layers) == p)
    for l, (W, b) in enumerate(layers):
        cin = oc.LAYERS[l][0]
        bound = np.sqrt(6 / (27 * cin)) if l < 3 else 0.1 * np.sqrt(3 / (27 * cin))
        assert np.max(np.abs(W)) <= bound and np.max(np.abs(W)) > 0.9 * bound and np.all(b == 0)
