"""Global-position expectations for cubes cut into blocks glued by interfaces
(shared by the oracle pins and the GPU tests; not a test module).

Every DOF of block b at local index idx stands for the global position
origin_b + idx, taken modulo the cube size N."""
import itertools

import numpy as np

DELTA = {"cell": (1, 1, 1), "fx": (0, 1, 1), "fy": (1, 0, 1), "fz": (1, 1, 0),
         "ex": (1, 0, 0), "ey": (0, 1, 0), "ez": (0, 0, 1)}

N = (10, 9, 8)
SPLIT = (2, 2, 2)


def _origin(c, ib):
    sx, sy, sz = c.meta["split"]
    pos = (ib // (sy * sz), (ib // sz) % sy, ib % sz)
    return [c.meta["origins"][a][pos[a]] for a in range(3)]


def _pos_code(P, comp=0):
    return 1e7 * (comp + 1) + 1e4 * P[0] + 1e2 * P[1] + P[2]


def _box_codes(c, ib, delta, comp=0, owned_only=True, block_tag=False):
    """Array over the block's box: code of the global (wrapped) position on the
    owned DOFs (all DOFs when not owned_only), NaN elsewhere."""
    g, d = c.g, c.blocks[ib].dims
    org = _origin(c, ib)
    shape = c.box(ib)
    I, J, K = np.meshgrid(*[np.arange(s) - g for s in shape], indexing="ij")
    P = [np.mod(org[0] + I, N[0]), np.mod(org[1] + J, N[1]), np.mod(org[2] + K, N[2])]
    code = _pos_code(P, comp)
    if block_tag:
        code = code + 1e9 * (ib + 1)
    if not owned_only:
        return code
    f = np.full(shape, np.nan)
    sl = tuple(slice(g, g + m + 1 - dd) for m, dd in zip(d, delta))
    f[sl] = code[sl]
    return f


def _out_axes(c, ib, delta):
    g, d = c.g, c.blocks[ib].dims
    shape = c.box(ib)
    idx = np.meshgrid(*[np.arange(s) - g for s in shape], indexing="ij")
    n = np.zeros(shape, dtype=int)
    for a in range(3):
        n += (idx[a] < 0) | (idx[a] >= d[a] + 1 - delta[a])
    return n


def _classes(c, delta):
    """Global position -> sorted list of (block, i, j, k) of its owned copies."""
    cls = {}
    for ib in range(c.nb):
        org, d = _origin(c, ib), c.blocks[ib].dims
        for idx in itertools.product(*[range(m + 1 - dd) for m, dd in zip(d, delta)]):
            P = tuple(int((org[a] + idx[a]) % N[a]) for a in range(3))
            cls.setdefault(P, []).append((ib,) + idx)
    return {P: sorted(v) for P, v in cls.items() if len(v) > 1}
