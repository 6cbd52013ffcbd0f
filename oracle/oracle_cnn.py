"""CPU restatement of the learned initial guess's network -- TEST
INFRASTRUCTURE ONLY (tests, the smoke check and the CPU baseline may use it;
the library never does).

The network has no counterpart in the reference, whose physical steps start
from U_ = U^n (main.cpp:378-382).  It is 15 -> 32 -> 32 -> 32 -> 5 channels,
3x3x3 kernels with zero padding (or periodic in D and W), GELU after the
first three layers, channel-last tensors [B][D][H][W][C], weights
W[co][tap][ci] with tap = (dd+1)*9 + (dh+1)*3 + (dw+1), and the parameter
layout of cgli_net_params (padded per layer, W then b).

Two modes:
  emulate=False  plain FP64 throughout (the mathematical network);
  emulate=True   the GPU's rounding points: BF16 inputs, weights and GELU
                 outputs, FP32 bias, output and output gradient, BF16 copies
                 of the back-propagated gradients for the convolutions, FP32
                 for the bias sums; products and sums themselves in FP64.

Pinned (tests/test_oracle_cnn.py) against torch on the CPU: conv3d (zero and
circular padding), autograd gradients of the whole network and loss, GELU and
its derivative, Adam, and the bfloat16 conversion; and against central
finite differences.
"""
import numpy as np
from scipy.special import erf

# (cin, cin_pad, cout, cout_pad, W offset, b offset)
LAYERS = [(15, 16, 32, 32, 0, 13824), (32, 32, 32, 32, 13856, 41504), (32, 32, 32, 32, 41536, 69184),
          (32, 32, 5, 16, 69216, 83040)]
NPARAM = 83056
TAPS = [(dd, dh, dw) for dd in (-1, 0, 1) for dh in (-1, 0, 1) for dw in (-1, 0, 1)]


def bf16(x):
    """FP32 -> nearest-even BF16, returned as float32 values (NaN stays NaN)."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    nan = ((u & 0x7F800000) == 0x7F800000) & ((u & 0x7FFFFF) != 0)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000) & 0xFFFFFFFF
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r)
    return r.astype(np.uint32).view(np.float32)


def unpack(params):
    """flat padded parameters -> [(W[cout][27][cin], b[cout])] as float64"""
    p = np.asarray(params, dtype=np.float64)
    out = []
    for cin, cinp, cout, coutp, wo, bo in LAYERS:
        W = p[wo:wo + coutp * 27 * cinp].reshape(coutp, 27, cinp)[:cout, :, :cin]
        out.append((W.copy(), p[bo:bo + cout].copy()))
    return out


def pack(layers):
    p = np.zeros(NPARAM, dtype=np.float64)
    for (cin, cinp, cout, coutp, wo, bo), (W, b) in zip(LAYERS, layers):
        Wp = np.zeros((coutp, 27, cinp))
        Wp[:cout, :, :cin] = W
        p[wo:wo + coutp * 27 * cinp] = Wp.ravel()
        p[bo:bo + cout] = b
    return p


def shift(x, dd, dh, dw, periodic=False):
    """s[b,d,h,w] = x[b,d+dd,h+dh,w+dw]; zero outside, or periodic in d, w"""
    mode_dw = "wrap" if periodic else "constant"
    xp = np.pad(x, ((0, 0), (1, 1), (0, 0), (1, 1), (0, 0)), mode=mode_dw)
    xp = np.pad(xp, ((0, 0), (0, 0), (1, 1), (0, 0), (0, 0)), mode="constant")
    D, H, W = x.shape[1:4]
    return xp[:, 1 + dd:1 + dd + D, 1 + dh:1 + dh + H, 1 + dw:1 + dw + W, :]


def conv(x, W, b=None, periodic=False):
    """y[v][co] = sum_tap sum_ci W[co][tap][ci] x[v + off(tap)][ci] (+ b)"""
    y = np.zeros(x.shape[:4] + (W.shape[0],))
    for t, (dd, dh, dw) in enumerate(TAPS):
        y += shift(x, dd, dh, dw, periodic) @ W[:, t, :].T
    if b is not None:
        y += b
    return y


def conv_t(dz, W, periodic=False):
    """dx[v][ci] = sum_tap sum_co W[co][tap][ci] dz[v - off(tap)][co]"""
    dx = np.zeros(dz.shape[:4] + (W.shape[2],))
    for t, (dd, dh, dw) in enumerate(TAPS):
        dx += shift(dz, -dd, -dh, -dw, periodic) @ W[:, t, :]
    return dx


def wgrad(dz, x, periodic=False):
    """dW[co][tap][ci] = sum_v dz[v][co] x[v + off(tap)][ci]"""
    dW = np.zeros((dz.shape[-1], 27, x.shape[-1]))
    dz2 = dz.reshape(-1, dz.shape[-1])
    for t, (dd, dh, dw) in enumerate(TAPS):
        dW[:, t, :] = dz2.T @ shift(x, dd, dh, dw, periodic).reshape(-1, x.shape[-1])
    return dW


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def _f32(a, emulate):
    return np.asarray(a, np.float32).astype(np.float64) if emulate else a


def forward(params, x, periodic=False, emulate=False):
    """y [B,D,H,W,5] and the activations the backward pass needs"""
    layers = unpack(params)
    a = bf16(x).astype(np.float64) if emulate else np.asarray(x, np.float64)
    acts, zs = [a], []
    for l, (W, b) in enumerate(layers):
        if emulate:
            W = bf16(W).astype(np.float64)
            b = _f32(b, True)
        z = conv(a, W, b, periodic)
        zs.append(z)
        if l < 3:
            a = gelu(z)
            a = bf16(a).astype(np.float64) if emulate else a
            acts.append(a)
    y = _f32(zs[-1], emulate)
    return y, (layers, acts, zs)


def loss_grad(params, x, t, mask, periodic=False, emulate=False):
    """masked MSE over the 5 outputs and its gradient in the flat layout"""
    y, (layers, acts, zs) = forward(params, x, periodic, emulate)
    m = np.asarray(mask, np.float64)[..., None]
    cnt = float(np.sum(mask))
    inv = 1.0 / (5.0 * cnt)
    e = _f32(y - np.asarray(t, np.float64), emulate)
    loss = float(np.sum(m * e * e) * inv)
    dy = _f32(2.0 * e * _f32(inv, emulate) * m, emulate)
    grads = [None] * 4
    dz_mm = bf16(dy).astype(np.float64) if emulate else dy   # operand of the convolutions
    dz_b = dy                                                 # operand of the bias sums
    for l in range(3, -1, -1):
        W, _ = layers[l]
        if emulate:
            W = bf16(W).astype(np.float64)
        grads[l] = (wgrad(dz_mm, acts[l], periodic), dz_b.reshape(-1, dz_b.shape[-1]).sum(axis=0))
        if l == 0:
            break
        da = _f32(conv_t(dz_mm, W, periodic), emulate)
        dz = _f32(da * gelu_grad(zs[l - 1]), emulate)
        dz_mm = bf16(dz).astype(np.float64) if emulate else dz
        dz_b = dz
    return loss, y, pack(grads)


def adam(w, m1, m2, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """one Adam step (Kingma & Ba 2015, Algorithm 1) at step count t >= 1"""
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    mh = m1 / (1 - b1 ** t)
    vh = m2 / (1 - b2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m1, m2


def record_scales(dq, dq1=None, dq2=None):
    """inverse scales of record t from dq_t, dq_{t-1}, dq_{t-2} (FP32
    differences d1 = dq_t - dq_{t-1}, d2 = d1 - (dq_{t-1} - dq_{t-2}), FP64
    sums): 1 / field RMS of dq_t, d1, d2, or 1 where an RMS is 0 or the
    earlier records are missing"""
    a = np.asarray(dq, np.float32).reshape(-1, 5)
    sums = [np.mean(a.astype(np.float64) ** 2, axis=0), np.zeros(5), np.zeros(5)]
    if dq1 is not None:
        b = np.asarray(dq1, np.float32).reshape(-1, 5)
        d1 = (a - b).astype(np.float32)
        sums[1] = np.mean(d1.astype(np.float64) ** 2, axis=0)
        if dq2 is not None:
            c = np.asarray(dq2, np.float32).reshape(-1, 5)
            d2 = (d1 - (b - c)).astype(np.float32)
            sums[2] = np.mean(d2.astype(np.float64) ** 2, axis=0)
    out = []
    for sm in sums:
        rms = np.sqrt(sm)
        with np.errstate(divide="ignore"):
            out.append(np.where(rms > 0, 1.0 / rms, 1.0).astype(np.float32))
    return tuple(out)


def features(dq0, dq1, dq2, i1, idd, i2):
    """network input from the last three increments (newest first) and the
    newest record's inverse scales, FP32 arithmetic as on the device, then
    BF16: [..., 15] = dq0 i1, d1 idd, d2 i2"""
    i1, idd, i2 = (np.asarray(v, np.float32) for v in (i1, idd, i2))
    a, b, c = (np.asarray(v, np.float32) for v in (dq0, dq1, dq2))
    d1 = (a - b).astype(np.float32)
    d1p = (b - c).astype(np.float32)
    x = np.concatenate([a * i1, d1 * idd, (d1 - d1p) * i2], axis=-1).astype(np.float32)
    return bf16(x)


def target(dq_next, dq0, dq1, i2):
    """the next second difference, scaled: ((dq_next - dq0) - (dq0 - dq1)) i2"""
    a, b, c = (np.asarray(v, np.float32) for v in (dq_next, dq0, dq1))
    return (((a - b).astype(np.float32) - (b - c).astype(np.float32)) * np.asarray(i2, np.float32)).astype(
        np.float32)


def init_params(seed):
    """the library's initialisation (uniform +-sqrt(6/fan_in) for the hidden
    layers, +-0.1 sqrt(3/fan_in) for the output layer, zero biases) drawn
    from splitmix64 in layer, co, tap, ci order"""
    st = [seed & 0xFFFFFFFFFFFFFFFF]
    M = 0xFFFFFFFFFFFFFFFF

    def unit():
        st[0] = (st[0] + 0x9E3779B97F4A7C15) & M
        z = st[0]
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        return (z >> 11) * 2.0 ** -53

    layers = []
    for l, (cin, _, cout, _, _, _) in enumerate(LAYERS):
        fan = 27.0 * cin
        bound = np.sqrt(6.0 / fan) if l < 3 else 0.1 * np.sqrt(3.0 / fan)
        W = np.zeros((cout, 27, cin))
        for co in range(cout):
            for tap in range(27):
                for ci in range(cin):
                    W[co, tap, ci] = np.float32(bound * (2 * unit() - 1))
        layers.append((W, np.zeros(cout)))
    return pack(layers)
