"""CPU oracle of the learned Transform-PAL 3D mask (test infrastructure only).

Nothing here is part of the decode path.  Only the tests, the smoke check
and bench.py's CPU comparison may import this module; the library computes
the same quantities with its own HIP kernels (libchromadec_amd/csrc/
mask_net_kernels.hip) and these functions are the yardstick they are checked
against.

The learned mask replaces the hard keep/discard rule of
TransformPal3D::applyFilter (src/decoders/palcolour/transform_pal_3d.cpp:
249-333) at the same insertion point.  It is defined on the rule's own
index set, the 768 visits v = (z*32 + y)*3 + (x-2) of the thresholds file:
a visit (z, y, x), x = 2..4, looks at bin a = F[z][y][x] and its carrier
mirror b = F[zr][yr][8-x], zr = (10-z) % 8, yr = (40-y) % 32
(transform_pal_3d.cpp:284-300).  The network maps per-visit features to a
soft mask m in [0, 1]; bins x = 2..4 are multiplied by their own visit's m
and bins x = 5, 6 by the m of the visit whose partner they are.

Network: three 3x3x3 convolutions over the visit grid (Z=8, Y=32, X=3),
circular in z and y (the DFT bins wrap) and zero-padded in x:
4 -> 32 (ReLU), 32 -> 32 (ReLU), 32 -> 1 (sigmoid).

Features per visit: f0 = (ln(|a|+1) - 10)/3, f1 = the same of |b|, f2 = the
tile mean of f0, f3 = 1 on the four self-mirror visits.

Training target per visit: with the truth chroma spectrum C of the same
tile, the squared error of the bins a visit owns, sum |m F - C|^2, is
w m^2 - 2 num m + const with w = sum |F|^2 and num = sum Re(conj(C) F)
over those bins (own bin, plus the partner bin for x = 2, 3).  The loss is
scale * sum (w m^2 - 2 num m).

Tap index t = (dz+1)*9 + (dy+1)*3 + (dx+1); weights W[t][cin][cout].
"""
from __future__ import annotations

import numpy as np

Z, Y, X = 8, 32, 3
P = Z * Y * X
CH = 32
N_PARAMS = 27 * 4 * CH + CH + 27 * CH * CH + CH + 27 * CH + 1

_zr = (10 - np.arange(Z)) % Z
_yr = (40 - np.arange(Y)) % Y


def _nbr_table():
    """NBR[t][p]: index of the neighbour of position p at tap t, or P (a zero row)."""
    nb = np.full((27, P), P, dtype=np.int64)
    for t in range(27):
        dz, dy, dx = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
        for z in range(Z):
            for y in range(Y):
                for x in range(X):
                    xs = x + dx
                    if 0 <= xs < X:
                        nb[t, (z * Y + y) * X + x] = (((z + dz) % Z) * Y + (y + dy) % Y) * X + xs
    return nb


NBR = _nbr_table()
NBR_INV = NBR[::-1].copy()  # tap -t: (-dz, -dy, -dx) is tap 26 - t

SELF_MIRROR = np.zeros(P, bool)
for _z in (1, 5):
    for _y in (4, 20):
        SELF_MIRROR[(_z * Y + _y) * X + 2] = True


def split_params(theta):
    """Flat parameter vector -> (W1, b1, W2, b2, W3, b3) views."""
    theta = np.asarray(theta)
    o = 0
    out = []
    for shape in ((27, 4, CH), (CH,), (27, CH, CH), (CH,), (27, CH, 1), (1,)):
        n = int(np.prod(shape))
        out.append(theta[o:o + n].reshape(shape))
        o += n
    return out


def init_params(seed: int) -> np.ndarray:
    """He-style initialisation; the last bias starts the mask near 0.5."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2.0 / (27 * 4)), 27 * 4 * CH), np.zeros(CH),
             rng.normal(0, np.sqrt(2.0 / (27 * CH)), 27 * CH * CH), np.zeros(CH),
             rng.normal(0, np.sqrt(1.0 / (27 * CH)), 27 * CH), np.zeros(1)]
    return np.concatenate(parts).astype(np.float32)


def pair_view(spec):
    """spec [T, 5, 32, 8] complex (bin x = k+2, ky, kz) -> visit grids a, b [T, P]."""
    spec = np.asarray(spec)
    t = spec.shape[0]
    a = np.empty((t, Z, Y, X), spec.dtype)
    b = np.empty((t, Z, Y, X), spec.dtype)
    for x in range(X):
        a[:, :, :, x] = np.transpose(spec[:, x], (0, 2, 1))                       # [T, z, y]
        part = np.transpose(spec[:, 4 - x], (0, 2, 1))                          # bin 8-(x+2)
        b[:, :, :, x] = part[:, _zr][:, :, _yr]
    return a.reshape(t, P), b.reshape(t, P)


def features(spec):
    a, b = pair_view(spec)
    f0 = (np.log(np.abs(a) + 1.0) - 10.0) / 3.0
    f1 = (np.log(np.abs(b) + 1.0) - 10.0) / 3.0
    f2 = np.broadcast_to(f0.mean(axis=1, keepdims=True), f0.shape)
    f3 = np.broadcast_to(SELF_MIRROR.astype(np.float64), f0.shape)
    return np.stack([f0, f1, f2, f3], axis=-1)  # [T, P, 4]


def targets(spec, spec_truth):
    a, b = pair_view(spec)
    ca, cb = pair_view(spec_truth)
    own2 = (np.arange(P) % X) < 2  # x = 2, 3 also own their partner bin
    w = np.abs(a) ** 2 + np.where(own2, np.abs(b) ** 2, 0.0)
    num = (np.conj(ca) * a).real + np.where(own2, (np.conj(cb) * b).real, 0.0)
    return w, num


def visit_to_bins(m):
    """m [T, P] -> bin mask [T, 5, 32, 8] (x = k+2, ky, kz)."""
    t = m.shape[0]
    mg = m.reshape(t, Z, Y, X)
    out = np.empty((t, 5, Y, Z), m.dtype)
    for x in range(X):
        out[:, x] = np.transpose(mg[:, :, :, x], (0, 2, 1))
    mp = mg[:, _zr][:, :, _yr]  # at (z', y'): the visit whose partner is (z', y')
    out[:, 3] = np.transpose(mp[:, :, :, 1], (0, 2, 1))
    out[:, 4] = np.transpose(mp[:, :, :, 0], (0, 2, 1))
    return out


def bf16_round(x):
    """Round float values to bfloat16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _conv(inp, W, b):
    """inp [T, P, Cin], W [27, Cin, Cout] -> [T, P, Cout]."""
    t, _, cin = inp.shape
    pad = np.concatenate([inp, np.zeros((t, 1, cin), inp.dtype)], axis=1)
    out = np.zeros((t, P, W.shape[2]), np.float64) + b
    for k in range(27):
        out += pad[:, NBR[k]] @ W[k]
    return out


def forward(theta, feat, bf16: bool = False):
    """Returns (m, cache).  With bf16=True, inputs, weights and stored
    activations are rounded to bfloat16 as the GPU kernels hold them."""
    W1, b1, W2, b2, W3, b3 = [np.asarray(v, np.float64) for v in split_params(theta)]
    r = (lambda v: bf16_round(v).astype(np.float64)) if bf16 else (lambda v: np.asarray(v, np.float64))
    f = r(feat)
    act1 = r(np.maximum(_conv(f, r(W1), b1), 0.0))
    act2 = r(np.maximum(_conv(act1, r(W2), b2), 0.0))
    logit = _conv(act2, r(W3), b3)[..., 0]
    m = 1.0 / (1.0 + np.exp(-logit))
    return m, dict(f=f, act1=act1, act2=act2, logit=logit, W=(r(W1), r(W2), r(W3)))


def loss(m, w, num, scale):
    return float(scale * np.sum(w * m * m - 2.0 * num * m))


def backward(theta, feat, w, num, scale, bf16: bool = False):
    """Gradient of loss(forward(theta, feat)) with respect to theta."""
    m, c = forward(theta, feat, bf16)
    W1, W2, W3 = c["W"]
    r = (lambda v: bf16_round(v).astype(np.float64)) if bf16 else (lambda v: v)
    dlogit = 2.0 * scale * (w * m - num) * m * (1.0 - m)   # [T, P]

    def wgrad(inp, dout):
        t, _, cin = inp.shape
        pad = np.concatenate([inp, np.zeros((t, 1, cin))], axis=1)
        return np.stack([np.einsum('tpc,tpo->co', pad[:, NBR[k]], dout) for k in range(27)])

    def dgrad(dout, W):
        # din[q] = sum_t dout[nbr(q, -t)] W[t]^T
        t, _, cout = dout.shape
        pad = np.concatenate([dout, np.zeros((t, 1, cout))], axis=1)
        return sum(pad[:, NBR_INV[k]] @ W[k].T for k in range(27))

    d3 = dlogit[..., None]
    gW3 = wgrad(c["act2"], r(d3))
    gb3 = d3.sum(axis=(0, 1))
    dpre2 = r(dgrad(r(d3), W3) * (c["act2"] > 0))
    gW2 = wgrad(c["act1"], dpre2)
    gb2 = dpre2.sum(axis=(0, 1))
    dpre1 = r(dgrad(dpre2, W2) * (c["act1"] > 0))
    gW1 = wgrad(c["f"], dpre1)
    gb1 = dpre1.sum(axis=(0, 1))
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
