"""CPU oracle of the learned cut proposal (3D CNN) -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of the network the library runs on the matrix cores (cugip_hip.h,
"learned cut proposal"): four 3x3x3 convolutions with 16 channels and ReLU, a 1x1x1 sigmoid head,
the mean binary cross-entropy against S_min, and its gradient. With bf16=True it rounds to bfloat16
(round to nearest even) exactly where the GPU kernels store bfloat16: the network input, the
weights, each layer's activations, and each backward gradient at a pre-activation; sums are
float64 here and float32 on the GPU. With bf16=False everything is float64, which the finite-
difference test uses to pin the backward pass.

The reference has no learned component; this oracle is pinned by finite differences (test_cnn_oracle.py)
and by closed forms (a single-tap network reduces to a per-voxel logistic model).
"""
import numpy as np

C = 16
TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
N_W1 = 27 * 4 * C
N_W = 27 * C * C
N_PARAMS = N_W1 + C + 3 * (N_W + C) + C + 1


def bf16(a):
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(p):
    """Parameter vector -> (W list of (27, cin, 16), b list of (16,), w_head (16,), b_head)."""
    p = np.asarray(p, np.float64)
    W, b = [], []
    W.append(p[:N_W1].reshape(27, 4, C))
    off = N_W1
    b.append(p[off:off + C])
    off += C
    for _ in range(3):
        W.append(p[off:off + N_W].reshape(27, C, C))
        off += N_W
        b.append(p[off:off + C])
        off += C
    wh = p[off:off + C]
    bh = p[off + C]
    return W, b, wh, bh


def pack(W, b, wh, bh):
    parts = []
    for w, bb in zip(W, b):
        parts += [np.asarray(w).reshape(-1), np.asarray(bb).reshape(-1)]
    parts += [np.asarray(wh).reshape(-1), np.array([bh])]
    return np.concatenate(parts)


def _shift(a, dz, dy, dx):
    """out[z, y, x] = a[z + dz, y + dy, x + dx] (zero outside); a is (nz, ny, nx, c)."""
    nz, ny, nx = a.shape[:3]
    pad = np.zeros((nz + 2, ny + 2, nx + 2) + a.shape[3:], a.dtype)
    pad[1:-1, 1:-1, 1:-1] = a
    return pad[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]


def network_input(volume, cs, ct, bf=True):
    vol = np.asarray(volume, np.float32)
    x = np.zeros(vol.shape + (4,), np.float32)
    x[..., 0] = vol * np.float32(1.0 / 255.0)
    x[..., 1] = np.asarray(cs, np.float32) * np.float32(1e-3)
    x[..., 2] = np.asarray(ct, np.float32) * np.float32(1e-3)
    return bf16(x).astype(np.float64) if bf else x.astype(np.float64)


def forward(params, volume, cs, ct, bf=True):
    """Returns (logits (nz, ny, nx), activations [x0, a1, a2, a3, a4])."""
    W, b, wh, bh = unpack(params)
    rnd = (lambda a: bf16(a).astype(np.float64)) if bf else (lambda a: a)
    acts = [network_input(volume, cs, ct, bf)]
    for layer in range(4):
        w = rnd(W[layer])
        x = acts[-1]
        out = np.zeros(x.shape[:3] + (C,), np.float64)
        for t, (dz, dy, dx) in enumerate(TAPS):
            out += _shift(x, dz, dy, dx) @ w[t]
        out += b[layer]
        acts.append(rnd(np.maximum(out, 0.0)))
    logits = acts[-1] @ wh + bh
    return logits, acts


def loss_and_grad(params, volume, cs, ct, label, bf=True):
    """Mean binary cross-entropy of sigmoid(logits) against `label` (0/1) and its gradient."""
    W, b, wh, bh = unpack(params)
    rnd = (lambda a: bf16(a).astype(np.float64)) if bf else (lambda a: a)
    z, acts = forward(params, volume, cs, ct, bf)
    y = np.asarray(label, np.float64).reshape(z.shape)
    n = z.size
    loss = float(np.mean(np.maximum(z, 0) + np.log1p(np.exp(-np.abs(z))) - y * z))
    dz = (1.0 / (1.0 + np.exp(-z)) - y) / n
    gW = [None] * 4
    gb = [None] * 4
    gwh = np.tensordot(dz, acts[4], axes=([0, 1, 2], [0, 1, 2]))
    gbh = float(dz.sum())
    g = rnd(dz[..., None] * wh[None, None, None, :] * (acts[4] > 0))
    for layer in range(3, -1, -1):
        x = acts[layer]
        cin = x.shape[3]
        gw = np.zeros((27, cin, C))
        for t, (dz_, dy_, dx_) in enumerate(TAPS):
            gw[t] = np.tensordot(_shift(x, dz_, dy_, dx_), g, axes=([0, 1, 2], [0, 1, 2]))
        gW[layer] = gw
        gb[layer] = g.sum(axis=(0, 1, 2))
        if layer == 0:
            break
        w = rnd(W[layer])
        gin = np.zeros(x.shape, np.float64)
        for t, (dz_, dy_, dx_) in enumerate(TAPS):
            gin += _shift(g, -dz_, -dy_, -dx_) @ w[t].T
        g = rnd(gin * (x > 0))
    gW[0][:, 3, :] = 0.0  # the padding input channel
    return loss, pack(gW, gb, gwh, gbh)
