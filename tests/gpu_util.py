"""Helpers shared by the GPU tests: grid cut capacity in numpy and case logging."""
import hashlib
import sys

import numpy as np

HALF6 = [(-1, 0, 0), (0, -1, 0), (0, 0, -1)]
HALF26 = [(j % 3 - 1, (j // 3) % 3 - 1, j // 9 - 1) for j in range(13)]


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def digest(mask):
    return hashlib.sha256(np.packbits(np.asarray(mask, np.uint8).reshape(-1))).hexdigest()


def _shift(a, dx, dy, dz, fill):
    """b[z, y, x] = a[z + dz, y + dy, x + dx] (fill outside)."""
    nz, ny, nx = a.shape
    out = np.full(a.shape, fill, dtype=a.dtype)
    zs = slice(max(0, -dz), min(nz, nz - dz))
    ys = slice(max(0, -dy), min(ny, ny - dy))
    xs = slice(max(0, -dx), min(nx, nx - dx))
    zs2 = slice(zs.start + dz, zs.stop + dz)
    ys2 = slice(ys.start + dy, ys.stop + dy)
    xs2 = slice(xs.start + dx, xs.stop + dx)
    out[zs, ys, xs] = a[zs2, ys2, xs2]
    return out


def grid_cut_capacity(source_side, cs, ct, cn, cnb=None, conn=6):
    """Capacity of the cut whose source side is {s} + source_side, on a grid instance; int."""
    S = np.asarray(source_side, bool).reshape(cs.shape)
    total = int(cs[~S].astype(np.int64).sum()) + int(ct[S].astype(np.int64).sum())
    offs = HALF6 if conn == 6 else HALF26
    for j, (dx, dy, dz) in enumerate(offs):
        nbS = _shift(S, dx, dy, dz, False)
        inside = _shift(np.ones(S.shape, bool), dx, dy, dz, False)
        fwd = cn[j].astype(np.int64)
        bwd = (cnb[j] if cnb is not None else cn[j]).astype(np.int64)
        total += int(fwd[S & ~nbS & inside].sum())    # v in S -> v + o_j outside S
        total += int(bwd[~S & nbS & inside].sum())    # v + o_j in S -> v outside S
    return total
