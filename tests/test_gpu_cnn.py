"""Learned cut proposal on the GPU (cgp_cnn_*, cgp_max_flow_certified): the matrix-core forward and
backward passes against the numpy oracle (oracle/cnn.py, same bfloat16 rounding points, float64
sums), training on the solver's own cuts, and the certified solve, whose results are exact whether
or not the certificate stops it early."""
import time

import numpy as np
import pytest

import cugip_amd
from oracle import cnn as ocnn
from tests.gpu_util import grid_cut_capacity, say

pytestmark = pytest.mark.gpu


def params(seed):
    p = cugip_amd.Cnn.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    W, b, wh, bh = ocnn.unpack(p)
    b = [rng.normal(0, 0.05, 16) for _ in range(4)]
    return ocnn.pack(W, b, wh, 0.1).astype(np.float32)


def test_forward_matches_oracle():
    """Odd dimensions (partial bricks on every axis). Tolerance: the GPU sums in float32 and the
    oracle in float64, so a stored bfloat16 can differ by one unit in the last place."""
    nx, ny, nz = 21, 13, 11
    say(f"cnn forward {nx}x{ny}x{nz}")
    vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, 5)
    p = params(3)
    with cugip_amd.GridGraph() as g, cugip_amd.Cnn() as net:
        g.set_grid(cs, ct, cn)
        net.set_params(p)
        side, prob, cert = net.predict(g, vol)
    z, _ = ocnn.forward(p, vol, cs, ct, bf=True)
    ref = 1.0 / (1.0 + np.exp(-z.reshape(-1)))
    err = np.abs(prob - ref)
    assert err.max() < 2e-2, err.max()
    assert np.quantile(err, 0.999) < 2e-3, np.quantile(err, 0.999)
    zf = z.reshape(-1)
    differ = side != (zf > 0)
    assert not differ[np.abs(zf) > 2e-2].any()
    assert cert == grid_cut_capacity(side, cs, ct, cn)


def test_gradients_match_oracle():
    nx, ny, nz = 19, 14, 10
    say(f"cnn gradients {nx}x{ny}x{nz}")
    vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, 8)
    p = params(4)
    with cugip_amd.GridGraph() as g, cugip_amd.Cnn() as net:
        g.set_grid(cs, ct, cn)
        g.max_flow()
        label = g.source_set()
        net.set_params(p)
        loss, grad = net.gradients(g, vol)
    ref_loss, ref_grad = ocnn.loss_and_grad(p, vol, cs, ct, label.reshape(nz, ny, nx), bf=True)
    assert abs(loss - ref_loss) <= 1e-3 * abs(ref_loss)
    bounds = [0, ocnn.N_W1, ocnn.N_W1 + 16]
    for _ in range(3):
        bounds += [bounds[-1] + ocnn.N_W, bounds[-1] + ocnn.N_W + 16]
    bounds += [bounds[-1] + 16, bounds[-1] + 17]
    for lo, hi in zip(bounds[:-1], bounds[1:]):
        a, b = grad[lo:hi].astype(np.float64), ref_grad[lo:hi]
        nb = np.linalg.norm(b)
        if nb == 0:
            assert np.linalg.norm(a) == 0
            continue
        rel = np.linalg.norm(a - b) / nb
        cos = float(a @ b / (np.linalg.norm(a) * nb))
        assert rel < 3e-2 and cos > 0.999, (lo, hi, rel, cos)


def test_training_tightens_the_certificate_and_certified_solves_are_exact():
    """Training on the solver's S_min lowers the loss and brings the certificate (capacity of the
    predicted cut) closer to the max flow than the untrained network's. The certified solve
    returns the plain solve's flow, S_min and T_min, and stops early when the certificate is the
    max flow itself."""
    say("cnn training and certificate")
    n = 40
    train = [cugip_amd.make_instance(n, n, n, 300 + i) for i in range(6)]
    test = cugip_amd.make_instance(n, n, n, 400)
    vol_t, cs_t, ct_t, cn_t = test
    with cugip_amd.Cnn() as net, cugip_amd.GridGraph() as gt:
        gt.set_grid(cs_t, ct_t, cn_t)
        flow = gt.max_flow()
        smin = gt.source_set()
        tmin = gt.fill_segments(1, 0)
        net.set_params(cugip_amd.Cnn.init_params(0))
        _, _, cert0 = net.predict(gt, vol_t)
        graphs = []
        for vol, cs, ct, cn in train:
            g = cugip_amd.GridGraph()
            g.set_grid(cs, ct, cn)
            g.max_flow()
            graphs.append((g, vol))
        losses = []
        for step in range(90):
            g, vol = graphs[step % len(graphs)]
            losses.append(net.train_step(g, vol, 3e-3))
        for g, _ in graphs:
            g.close()
        side, prob, cert1 = net.predict(gt, vol_t)
        acc = float((side == smin).mean())
        say(f"  loss {np.mean(losses[:6]):.4f} -> {np.mean(losses[-6:]):.4f}, certificate gap "
            f"{cert0 - flow} -> {cert1 - flow} (flow {flow}), voxel accuracy {acc:.4f}")
        assert np.mean(losses[-6:]) < 0.7 * np.mean(losses[:6])
        assert flow <= cert1 < cert0
        for certificate in (cert1, flow):
            t0 = time.perf_counter()
            f, early = gt.max_flow_certified(certificate)
            ms = (time.perf_counter() - t0) * 1e3
            say(f"  certificate {certificate}: flow {f}, stopped early {early}, {ms:.2f} ms")
            assert f == flow
            assert np.array_equal(gt.source_set(), smin)
            assert np.array_equal(gt.fill_segments(1, 0), tmin)
            assert early == (certificate == flow) or certificate != flow
        assert early  # the last certificate was the max flow itself


def test_resident_volume_and_trained_weights():
    """With a grid built on the device from its volume, the network reads that volume (no host copy)
    and gives the same certificate as from the host volume; the committed trained weights give a
    certificate within 1 % of the max flow on a held-out 96^3 grid, and the certified solve is exact."""
    import os
    say("cnn resident volume, trained weights")
    vol = cugip_amd.generate_volume(96, 96, 96, 777001)
    weights = os.path.join(os.path.dirname(cugip_amd.__file__), "cnn_weights.npy")
    with cugip_amd.GridGraph() as g, cugip_amd.Cnn() as net:
        g.set_grid_volume(vol)
        net.set_params(np.load(weights))
        c_dev = net.certificate(g)
        c_host = net.certificate(g, vol)
        assert c_dev == c_host
        flow = g.max_flow()
        smin = g.source_set()
        assert flow <= c_dev <= flow * 1.01, (flow, c_dev)
        f, _ = g.max_flow_certified(c_dev)
        assert f == flow and np.array_equal(g.source_set(), smin)


def test_learned_certificate_on_held_out_192_and_256_grids():
    """The committed weights on held-out grids of the learned workload's shapes (8 of 192^3 and 2
    of 256^3, seeds outside the training range 100000-101999): every certified solve returns the
    plain solve's flow and S_min bit for bit; the trained network's certificate is an upper bound
    on the max flow and tighter than the untrained network's on every grid (what training gains);
    early stops, and the solve times of plain, trained-certificate, untrained-certificate and
    exact-certificate solves (the last is the most any proposal could save), are reported."""
    import os
    weights = os.path.join(os.path.dirname(cugip_amd.__file__), "cnn_weights.npy")
    grids = [(192, 900000 + i) for i in range(8)] + [(256, 900100 + i) for i in range(2)]
    trained, untrained = np.load(weights), cugip_amd.Cnn.init_params(0)
    early = 0
    times = {"plain": 0.0, "trained": 0.0, "untrained": 0.0, "exact": 0.0}

    def timed(fn):
        best = None
        for _ in range(2):
            t0 = time.perf_counter()
            out = fn()
            dt = time.perf_counter() - t0
            best = dt if best is None or dt < best else best
        return out, best

    with cugip_amd.Cnn() as net:
        for n, seed in grids:
            say(f"learned held-out {n}^3 seed {seed}")
            vol = cugip_amd.generate_volume(n, n, n, seed)
            with cugip_amd.GridGraph() as g:
                g.set_grid_volume(vol)
                flow, dt = timed(g.max_flow)
                times["plain"] += dt
                smin = g.source_set()
                net.set_params(untrained)
                cert0 = net.certificate(g)
                net.set_params(trained)
                cert1 = net.certificate(g)
                say(f"  flow {flow}, certificate gap trained {cert1 - flow}, untrained {cert0 - flow}")
                assert flow <= cert1 < cert0
                for name, cert in (("trained", cert1), ("untrained", cert0), ("exact", flow)):
                    (f, stopped), dt = timed(lambda: g.max_flow_certified(cert))
                    times[name] += dt
                    assert f == flow, name
                    assert np.array_equal(g.source_set(), smin), name
                    if name == "trained":
                        early += int(stopped)
                    if name == "exact":
                        assert stopped
    say(f"  {len(grids)} grids: trained certificate stopped {early} solves early; total ms " +
        ", ".join(f"{k} {1e3 * v:.2f}" for k, v in times.items()))
