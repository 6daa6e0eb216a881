"""The synthetic instance generator (host code of libcugip_hip.so) pinned by an independent
numpy restatement and by closed-form table values. CPU only."""
import math
import os
import re

import numpy as np
import pytest

import cugip_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def load_tables():
    text = open(os.path.join(ROOT, "cugip_amd", "csrc", "gen_tables.h")).read()

    def arr(name):
        body = re.search(name + r"\[[^\]]*\](?:\[[^\]]*\])? = \{(.*?)\};", text, re.S).group(1)
        return np.array([int(v.rstrip("ull")) for v in re.findall(r"\d+(?:ull)?", body)], dtype=np.int64)

    return dict(src=arr("CGP_TLINK_SOURCE"), snk=arr("CGP_TLINK_SINK"), w=arr("CGP_NLINK_WEIGHT").reshape(3, 256),
                cdf=arr("CGP_NOISE_CDF"))


def splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def numpy_volume(nx, ny, nz, seed, instance, tables):
    with np.errstate(over="ignore"):
        key = splitmix64(splitmix64(np.uint64(seed)) ^ (np.uint64(0x5851F42D4C957F2D) * np.uint64(instance + 1)))

        def draw(i):
            return int(splitmix64(key + np.uint64(i)))

    m = min(nx, ny, nz)
    amin = max(1, (8 * m) // 100)
    amax = max(amin, (25 * m) // 100)
    count = 3 + draw(0) % 6
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    vol = np.full((nz, ny, nx), 60, np.int64)
    inside_any = np.zeros((nz, ny, nx), bool)
    for b in range(count):
        base = 1 + 7 * b
        cx, cy, cz = draw(base) % nx, draw(base + 1) % ny, draw(base + 2) % nz
        a, bb, c = (amin + draw(base + k) % (amax - amin + 1) for k in (3, 4, 5))
        a2, b2, c2 = a * a, bb * bb, c * c
        inside = ((x - cx) ** 2 * b2 * c2 + (y - cy) ** 2 * a2 * c2 + (z - cz) ** 2 * a2 * b2) <= a2 * b2 * c2
        inside_any |= inside
    vol[inside_any] = 180
    with np.errstate(over="ignore"):
        nkey = splitmix64(key ^ np.uint64(0xD1B54A32D192ED03))
        idx = np.arange(nx * ny * nz, dtype=np.uint64)
        u = (splitmix64(nkey + idx) >> np.uint64(32)).astype(np.int64)
    noise = np.searchsorted(tables["cdf"], u, side="right") - 150
    vol = vol + noise.reshape(nz, ny, nx)
    return np.clip(vol, 0, 255).astype(np.uint8)


def numpy_caps(vol, conn, tables):
    nz, ny, nx = vol.shape
    iv = vol.astype(np.int64)
    cs = tables["src"][iv]
    ct = tables["snk"][iv]
    if conn == 6:
        offs = [(-1, 0, 0), (0, -1, 0), (0, 0, -1)]
    else:
        offs = [(j % 3 - 1, (j // 3) % 3 - 1, j // 9 - 1) for j in range(13)]
    cn = np.zeros((len(offs), nz, ny, nx), np.int64)
    pad = np.pad(iv, 1, constant_values=-1)
    for j, (dx, dy, dz) in enumerate(offs):
        nb = pad[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
        cls = (dx != 0) + (dy != 0) + (dz != 0) - 1
        w = tables["w"][cls][np.abs(iv - nb).clip(0, 255)]
        cn[j] = np.where(nb >= 0, w, 0)
    return cs, ct, cn


@pytest.fixture(scope="module")
def tables():
    return load_tables()


@pytest.mark.parametrize("dims,seed,instance", [((17, 13, 11), 0, 0), ((32, 24, 20), 5, 3), ((9, 40, 7), 123, 7),
                                                 ((64, 64, 64), 0, 0)])
def test_volume_matches_numpy_restatement(dims, seed, instance, tables):
    nx, ny, nz = dims
    got = cugip_amd.generate_volume(nx, ny, nz, seed, instance)
    want = numpy_volume(nx, ny, nz, seed, instance, tables)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("conn", [6, 26])
def test_caps_match_numpy_restatement(conn, tables):
    vol = cugip_amd.generate_volume(21, 14, 9, 3, 1)
    cs, ct, cn = cugip_amd.caps_from_volume(vol, conn)
    ws, wt, wn = numpy_caps(vol, conn, tables)
    assert np.array_equal(cs, ws) and np.array_equal(ct, wt) and np.array_equal(cn, wn)


def test_tables_closed_forms(tables):
    # -ln N(mu; mu, 30) = ln(30 sqrt(2 pi)) = 4.3204 -> 432; 60 units away adds 2 -> 632
    assert tables["src"][60] == 432 and tables["snk"][180] == 432
    assert tables["src"][0] == 632 and tables["src"][120] == 632 and tables["snk"][120] == 632
    assert tables["src"][180] == 1000 and tables["snk"][60] == 1000  # capped
    assert np.all(tables["src"][:61] == tables["src"][120:59:-1])   # symmetric about 60
    # n-links: 500 exp(-d^2/800) / sqrt(k)
    assert tables["w"][0][0] == 500 and tables["w"][1][0] == 354 and tables["w"][2][0] == 289
    assert tables["w"][0][20] == round(500 * math.exp(-0.5))  # one sigma: 303.27 -> 303
    assert np.all(np.diff(tables["w"][0]) <= 0) and tables["w"][0][255] == 0
    # noise CDF of round(25 Z): increasing, symmetric, P(<= 0) = Phi(0.02)
    cdf = tables["cdf"]
    assert len(cdf) == 301 and cdf[-1] == 2 ** 32 and np.all(np.diff(cdf) >= 0)
    assert abs(cdf[150] / 2 ** 32 - 0.5079783137169019) < 1e-9
    assert np.all(np.abs(cdf[:149] + cdf[299:150:-1] - 2 ** 32) <= 2)


def test_noise_statistics():
    vol = cugip_amd.generate_volume(128, 128, 64, 11, 0).astype(np.float64)
    bg = vol[(vol < 130) & (vol > 0)]
    assert abs(bg.mean() - 60) < 1.5 and abs(bg.std() - 25) < 1.5


def test_generator_rejects_bad_arguments():
    with pytest.raises(cugip_amd.CgpError):
        cugip_amd.generate_volume(0, 4, 4, 1)
    with pytest.raises(cugip_amd.CgpError):
        cugip_amd.caps_from_volume(np.zeros((3, 3, 3), np.uint8), conn=8)


def test_golden_volume_digests(golden):
    import hashlib
    for rec in golden["instances"][:4]:
        nx, ny, nz = rec["dims"]
        vol = cugip_amd.generate_volume(nx, ny, nz, rec["seed"], rec["instance"])
        assert hashlib.sha256(vol.tobytes()).hexdigest() == rec["volume_sha256"]
