"""The CNN oracle on the CPU, pinned independently of its own formulas: its float64 gradient
against central finite differences of its loss, closed forms of degenerate networks, and the
bfloat16 rounding it shares with the GPU kernels."""
import numpy as np

from oracle import cnn


def small_case(seed, shape=(3, 4, 5)):
    rng = np.random.default_rng(seed)
    vol = rng.integers(0, 256, shape).astype(np.uint8)
    cs = rng.integers(0, 1000, shape)
    ct = rng.integers(0, 1000, shape)
    label = (rng.random(shape) < 0.4).astype(np.uint8)
    p = rng.normal(0, 0.3, cnn.N_PARAMS)
    p[:cnn.N_W1].reshape(27, 4, 16)[:, 3, :] = 0
    return p, vol, cs, ct, label


def test_gradient_matches_finite_differences():
    p, vol, cs, ct, label = small_case(0)
    loss, grad = cnn.loss_and_grad(p, vol, cs, ct, label, bf=False)
    rng = np.random.default_rng(1)
    groups = [(0, cnn.N_W1), (cnn.N_W1, cnn.N_W1 + 16)]
    off = cnn.N_W1 + 16
    for _ in range(3):
        groups += [(off, off + cnn.N_W), (off + cnn.N_W, off + cnn.N_W + 16)]
        off += cnn.N_W + 16
    groups += [(off, off + 16), (off + 16, off + 17)]
    eps = 1e-6
    for lo, hi in groups:
        for i in rng.integers(lo, hi, 4):
            if i < cnn.N_W1 and (i // 16) % 4 == 3:
                continue  # padding channel: gradient defined as 0
            pp, pm = p.copy(), p.copy()
            pp[i] += eps
            pm[i] -= eps
            lp, _ = cnn.loss_and_grad(pp, vol, cs, ct, label, bf=False)
            lm, _ = cnn.loss_and_grad(pm, vol, cs, ct, label, bf=False)
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - grad[i]) <= 1e-6 + 1e-4 * abs(fd), (lo, i, fd, grad[i])


def test_closed_form_bias_only_network():
    """All weights zero: every logit is b_head, the loss is mean softplus(b) - y b and only the
    head bias has a gradient, mean(sigmoid(b) - y)."""
    _, vol, cs, ct, label = small_case(2)
    p = np.zeros(cnn.N_PARAMS)
    p[-1] = 0.7
    loss, grad = cnn.loss_and_grad(p, vol, cs, ct, label, bf=False)
    y = label.astype(float)
    assert np.isclose(loss, np.mean(np.log1p(np.exp(0.7)) - 0.7 * y))
    assert np.isclose(grad[-1], np.mean(1 / (1 + np.exp(-0.7)) - y))
    assert np.count_nonzero(grad[:-1]) == 0


def test_closed_form_single_tap_is_per_voxel():
    """Only the centre tap of every layer, identity-like channel 0 paths: the logit of a voxel
    depends on that voxel alone, z = w * relu(relu(relu(relu(a * x0)))) + b for x0 = I/255 (float64)."""
    _, vol, cs, ct, label = small_case(3)
    W, b, wh, bh = cnn.unpack(np.zeros(cnn.N_PARAMS))
    W[0][13, 0, 0] = 2.0
    for layer in (1, 2, 3):
        W[layer][13, 0, 0] = 1.5
    wh = np.zeros(16)
    wh[0] = 0.25
    p = cnn.pack(W, b, wh, -0.3)
    z, _ = cnn.forward(p, vol, cs, ct, bf=False)
    x0 = vol.astype(np.float32) * np.float32(1 / 255)
    assert np.allclose(z, 0.25 * (2.0 * 1.5 ** 3) * x0.astype(np.float64) - 0.3)


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1e-3, 0.0], np.float32)
    out = cnn.bf16(vals)
    assert out[0] == 1.0
    assert out[1] == 1.0                    # tie -> even mantissa
    assert out[2] == 1.0 + 4 * 2 ** -8      # tie -> even (up)
    assert out[3] == -2.5 and out[5] == 0.0
    assert abs(out[4] - 1e-3) <= 1e-3 * 2 ** -8


def test_forward_matches_torch_conv3d():
    """The float64 forward equals torch.nn.functional.conv3d (zero padding 1, cross-correlation)
    with the tap order t = (dz+1)*9 + (dy+1)*3 + (dx+1) mapped to kernel index [dz+1, dy+1, dx+1],
    on random weights where every tap of every layer is nonzero and a grid small enough that every
    voxel touches a border: a transposed tap order, a flipped tap or missing padding all fail it."""
    import torch
    import torch.nn.functional as F
    rng = np.random.default_rng(7)
    shape = (4, 5, 6)
    vol = rng.integers(0, 256, shape).astype(np.uint8)
    cs = rng.integers(0, 1000, shape)
    ct = rng.integers(0, 1000, shape)
    p = rng.normal(0, 0.4, cnn.N_PARAMS)
    p[:cnn.N_W1].reshape(27, 4, 16)[:, 3, :] = 0
    z, _ = cnn.forward(p, vol, cs, ct, bf=False)
    W, b, wh, bh = cnn.unpack(p)
    x = torch.from_numpy(cnn.network_input(vol, cs, ct, bf=False)).permute(3, 0, 1, 2)[None]  # (1, 4, z, y, x)
    for layer in range(4):
        w = np.asarray(W[layer])  # (27, cin, cout)
        k = torch.from_numpy(w.reshape(3, 3, 3, w.shape[1], w.shape[2]).transpose(4, 3, 0, 1, 2).copy())
        x = F.relu(F.conv3d(x, k, torch.from_numpy(np.asarray(b[layer])), padding=1))
    ref = (x[0].permute(1, 2, 3, 0).numpy() @ wh) + bh
    assert np.allclose(z, ref, rtol=1e-12, atol=1e-12)
    # the same network with the tap order read as (dx, dy, dz) instead differs
    Wt = [np.asarray(w).reshape(3, 3, 3, *np.asarray(w).shape[1:]).transpose(2, 1, 0, 3, 4).reshape(np.asarray(w).shape)
          for w in W]
    zt, _ = cnn.forward(cnn.pack(Wt, b, wh, bh), vol, cs, ct, bf=False)
    assert not np.allclose(zt, ref)
