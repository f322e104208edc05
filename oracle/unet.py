"""FP64 restatement of the learned warm-start network (test infrastructure).

There is no reference counterpart: the reference has no learned component.
This module defines the network's semantics for the parity tests of the HIP
training and inference kernels (waves_amd/csrc/wv_net.hip):

* tensors are channels-last, cell index i = x + nx (y + ny z) within a sample,
  samples stacked after one another;
* conv3: 3x3x3, zero padding at the grid edge, tap t = (dz+1) 9 + (dy+1) 3 +
  (dx+1); weights W[co][k] with k = t Cin + ci and the bias stored in column
  27 Cin (an input that is always 1);
* pool: mean over the existing children of a 2x2x2 block (coarse dims are
  ceil(n/2));
* up: nearest (coarse cell of fine cell i is i >> 1 per axis), and decoder
  inputs are concat(skip, up(coarser));
* two heads: a fine field qf at level 0 and a coarse field qc at level 2
  (prolonged by i >> 2 per axis);
* loss = mean over samples of sum_wet (qf + up4 qc - t)^2 / sum_wet t^2.

With `bf16=True` every stored activation, weight and input is rounded to
bfloat16 (round to nearest even) where the device stores it, so the device's
results differ from this oracle only by accumulation order.

Parity pins: the conv is checked against a brute-force triple loop on tiny
grids and the gradients against central finite differences (tests/
test_net_oracle.py).
"""
from __future__ import annotations

import numpy as np

#: (name, level, inputs, cin_skip, cin_up, cout, relu)
LAYERS = [
    ("e0a", 0, "x", 8, 0, 8, True),
    ("e0b", 0, "e0a", 8, 0, 8, True),
    ("e1a", 1, "pool:e0b", 8, 0, 16, True),
    ("e1b", 1, "e1a", 16, 0, 16, True),
    ("e2a", 2, "pool:e1b", 16, 0, 32, True),
    ("e2b", 2, "e2a", 32, 0, 32, True),
    ("b3a", 3, "pool:e2b", 32, 0, 64, True),
    ("b3b", 3, "b3a", 64, 0, 64, True),
    ("d2a", 2, "e2b+b3b", 32, 64, 32, True),
    ("d2b", 2, "d2a", 32, 0, 32, True),
    ("hc", 2, "d2b", 32, 0, 1, False),
    ("d1a", 1, "e1b+d2b", 16, 32, 16, True),
    ("d1b", 1, "d1a", 16, 0, 16, True),
    ("d0a", 0, "e0b+d1b", 8, 16, 8, True),
    ("d0b", 0, "d0a", 8, 0, 8, True),
    ("hf", 0, "d0b", 8, 0, 1, False),
]
IN_CHANNELS = 8


def kpad(cin: int) -> int:
    """Row length of a weight matrix: 27 cin taps + bias, padded to 16."""
    return (27 * cin + 1 + 15) // 16 * 16


def layer_shapes():
    out = []
    for name, lev, src, ca, cb, co, relu in LAYERS:
        out.append((name, co, ca + cb))
    return out


def param_count() -> int:
    return sum(co * kpad(ci) for _, co, ci in layer_shapes())


def split_params(flat: np.ndarray):
    """Flat parameter vector -> {name: W[co][kpad]} views."""
    out, off = {}, 0
    for name, co, ci in layer_shapes():
        n = co * kpad(ci)
        out[name] = flat[off:off + n].reshape(co, kpad(ci))
        off += n
    return out


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero bias and zero padding columns (float32)."""
    rng = np.random.default_rng(seed)
    flat = np.zeros(param_count(), np.float32)
    for name, w in split_params(flat).items():
        co, kp = w.shape
        ci = [c for n, _, c in layer_shapes() if n == name][0]
        fan = 27 * ci
        w[:, :27 * ci] = rng.normal(0, np.sqrt(2.0 / fan), (co, 27 * ci)).astype(np.float32)
    return flat


def level_dims(n, levels: int = 4):
    dims = [tuple(int(v) for v in n)]
    for _ in range(levels - 1):
        dims.append(tuple((v + 1) // 2 for v in dims[-1]))
    return dims


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float64/float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isnan(f), np.nan, out)


def _shift(x: np.ndarray, dz: int, dy: int, dx: int) -> np.ndarray:
    """y[z, y, x] = x[z + dz, y + dy, x + dx] with zeros outside; x is (B, nz, ny, nx, C)."""
    out = np.zeros_like(x)
    B, nz, ny, nx, C = x.shape
    zs, ys, xs = slice(max(0, -dz), nz - max(0, dz)), slice(max(0, -dy), ny - max(0, dy)), slice(max(0, -dx), nx - max(0, dx))
    zt, yt, xt = slice(max(0, dz), nz - max(0, -dz)), slice(max(0, dy), ny - max(0, -dy)), slice(max(0, dx), nx - max(0, -dx))
    out[:, zs, ys, xs] = x[:, zt, yt, xt]
    return out


def taps():
    return [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def conv3(x: np.ndarray, w: np.ndarray) -> np.ndarray:
    """x (B, nz, ny, nx, Cin), w (Cout, kpad) -> (B, nz, ny, nx, Cout)."""
    cin = x.shape[-1]
    y = np.zeros(x.shape[:-1] + (w.shape[0],))
    for t, (dz, dy, dx) in enumerate(taps()):
        y += _shift(x, dz, dy, dx) @ w[:, t * cin:(t + 1) * cin].T
    return y + w[:, 27 * cin]


def conv3_backward(x: np.ndarray, w: np.ndarray, dy: np.ndarray):
    """Gradients of conv3 w.r.t. x and w (same shapes)."""
    cin = x.shape[-1]
    dx = np.zeros_like(x)
    dw = np.zeros_like(w)
    flat_dy = dy.reshape(-1, dy.shape[-1])
    for t, (a, b, c) in enumerate(taps()):
        xs = _shift(x, a, b, c).reshape(-1, cin)
        dw[:, t * cin:(t + 1) * cin] = flat_dy.T @ xs
        dx += _shift(dy @ w[:, t * cin:(t + 1) * cin], -a, -b, -c)
    dw[:, 27 * cin] = flat_dy.sum(axis=0)
    return dx, dw


def pool(x: np.ndarray) -> np.ndarray:
    B, nz, ny, nx, C = x.shape
    cz, cy, cx = (nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2
    s = np.zeros((B, cz, cy, cx, C))
    cnt = np.zeros((1, cz, cy, cx, 1))
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                part = x[:, a::2, b::2, c::2]
                s[:, :part.shape[1], :part.shape[2], :part.shape[3]] += part
                cnt[:, :part.shape[1], :part.shape[2], :part.shape[3]] += 1
    return s / cnt


def pool_backward(dy: np.ndarray, fine_shape) -> np.ndarray:
    B, nz, ny, nx, C = fine_shape
    cz, cy, cx = dy.shape[1:4]
    cnt = np.zeros((1, cz, cy, cx, 1))
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                sh = (len(range(a, nz, 2)), len(range(b, ny, 2)), len(range(c, nx, 2)))
                cnt[:, :sh[0], :sh[1], :sh[2]] += 1
    g = dy / cnt
    dx = np.zeros(fine_shape)
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                sub = dx[:, a::2, b::2, c::2]
                sub += g[:, :sub.shape[1], :sub.shape[2], :sub.shape[3]]
    return dx


def up(x: np.ndarray, fine_shape, factor: int = 2) -> np.ndarray:
    B, nz, ny, nx = fine_shape[:4]
    iz, iy, ix = np.arange(nz) // factor, np.arange(ny) // factor, np.arange(nx) // factor
    return x[:, iz][:, :, iy][:, :, :, ix]


def up_backward(dy: np.ndarray, coarse_shape, factor: int = 2) -> np.ndarray:
    dx = np.zeros(coarse_shape)
    B, nz, ny, nx, C = dy.shape
    iz, iy, ix = np.arange(nz) // factor, np.arange(ny) // factor, np.arange(nx) // factor
    np.add.at(dx, (slice(None), iz[:, None, None], iy[None, :, None], ix[None, None, :]), dy)
    return dx


def forward(params: np.ndarray, x: np.ndarray, use_bf16: bool = False):
    """x: (B, nz, ny, nx, 8).  Returns (qf (B,nz,ny,nx), qc (B, level-2 dims), cache)."""
    P = split_params(np.asarray(params, np.float64))
    rnd = bf16 if use_bf16 else (lambda a: a)
    if use_bf16:
        P = {k: bf16(v) for k, v in P.items()}
    acts = {"x": rnd(np.asarray(x, np.float64))}
    cache = {}
    for name, lev, src, ca, cb, co, relu in LAYERS:
        if src.startswith("pool:"):
            inp = rnd(pool(acts[src[5:]]))
            cache[name + ":pooled"] = inp
        elif "+" in src:
            sk, cr = src.split("+")
            inp = np.concatenate([acts[sk], up(acts[cr], acts[sk].shape)], axis=-1)
        else:
            inp = acts[src]
        pre = conv3(inp, P[name])
        cache[name + ":in"] = inp
        if relu:
            acts[name] = rnd(np.maximum(pre, 0))
        else:
            acts[name] = pre[..., 0]  # heads stay FP32 on the device; float64 here
    return acts["hf"], acts["hc"], (P, acts, cache)


def combine(qf: np.ndarray, qc: np.ndarray) -> np.ndarray:
    """qf + nearest prolongation of qc by 4 per axis."""
    return qf + up(qc[..., None], qf.shape + (1,), 4)[..., 0]


def loss_and_grad(params: np.ndarray, x: np.ndarray, target: np.ndarray, mask: np.ndarray, use_bf16: bool = False):
    """Loss and d loss / d params (flat, float64)."""
    qf, qc, (P, acts, cache) = forward(params, x, use_bf16)
    B = qf.shape[0]
    q = combine(qf, qc)
    m = mask.astype(np.float64)
    den = (target ** 2 * m).reshape(B, -1).sum(axis=1) + 1e-30
    diff = (q - target) * m
    loss = float(np.mean((diff ** 2).reshape(B, -1).sum(axis=1) / den))
    dq = 2 * diff / den[:, None, None, None] / B
    grads = {}
    dacts = {name: None for name, *_ in LAYERS}

    def add(name, g):
        dacts[name] = g if dacts[name] is None else dacts[name] + g

    add("hf", dq[..., None])
    add("hc", up_backward(dq[..., None], qc.shape + (1,), 4))
    for name, lev, src, ca, cb, co, relu in reversed(LAYERS):
        g = dacts[name]
        if g is None:
            g = np.zeros(acts[name].shape if not name.startswith("h") else acts[name].shape + (1,))
        if relu:
            g = g * (acts[name] > 0)
        if use_bf16:
            g = bf16(g)
        inp = cache[name + ":in"]
        dx, dw = conv3_backward(inp, P[name], g)
        grads[name] = dw
        if src.startswith("pool:"):
            add(src[5:], pool_backward(dx, acts[src[5:]].shape))
        elif "+" in src:
            sk, cr = src.split("+")
            add(sk, dx[..., :ca])
            add(cr, up_backward(dx[..., ca:], acts[cr].shape))
        elif src != "x":
            add(src, dx)
    flat = np.concatenate([grads[name].ravel() for name, *_ in LAYERS])
    return loss, flat


def features(F, r, p, wet):
    """Network input (cells, 8) from volume fraction, residual, previous
    pressure and the wet mask: F, r / rms_wet(r), p / rms_wet(p), wet, 0..."""
    w = wet.astype(np.float64)
    cnt = max(w.sum(), 1.0)
    sr = np.sqrt((np.asarray(r, np.float64) ** 2 * w).sum() / cnt)
    sp = np.sqrt((np.asarray(p, np.float64) ** 2 * w).sum() / cnt)
    x = np.zeros((len(F), IN_CHANNELS))
    x[:, 0] = F
    x[:, 1] = np.asarray(r, np.float64) / sr * w if sr > 0 else 0
    x[:, 2] = np.asarray(p, np.float64) / sp * w if sp > 0 else 0
    x[:, 3] = w
    return x
