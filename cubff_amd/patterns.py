"""Interaction patterns for 2-D neighbour pairing.

`grid_pattern(w, h, radius)` builds, as CSR arrays for
`Soup.set_interactions`, the neighbour lists that the reference's
make_2d_pattern.py:17-29 writes as `a b` lines: for each cell in row-major
order, every other cell within `radius` in both coordinates, scanned by row
offset then column offset.
"""
from __future__ import annotations

import numpy as np


def grid_pattern(w: int, h: int, radius: int = 2):
    """(row_off uint64[w*h + 1], nbr uint32[E]) of a w x h grid."""
    ii, jj = np.divmod(np.arange(w * h, dtype=np.int64), w)
    cols = []
    for dy in range(-radius, radius + 1):
        for dx in range(-radius, radius + 1):
            if dy == 0 and dx == 0:
                continue
            y, x = ii + dy, jj + dx
            ok = (y >= 0) & (y < h) & (x >= 0) & (x < w)
            cols.append(np.where(ok, y * w + x, -1))
    nb = np.stack(cols, axis=1)
    deg = (nb >= 0).sum(axis=1)
    row_off = np.zeros(w * h + 1, dtype=np.uint64)
    row_off[1:] = np.cumsum(deg)
    return row_off, nb[nb >= 0].astype(np.uint32)


def read_pattern_file(path: str, num_programs: int):
    """CSR arrays of an `a b` per line file (main.cc:281-299): program a may
    interact with b; lines keep their order within each a."""
    data = np.loadtxt(path, dtype=np.int64, ndmin=2)
    if data.size and (data.min() < 0 or data.max() >= num_programs):
        raise ValueError("invalid interaction pattern: programs not in [0, n) range.")
    a, b = (data[:, 0], data[:, 1]) if data.size else (np.zeros(0, np.int64),) * 2
    order = np.argsort(a, kind="stable")
    deg = np.bincount(a, minlength=num_programs)
    row_off = np.zeros(num_programs + 1, dtype=np.uint64)
    row_off[1:] = np.cumsum(deg)
    return row_off, b[order].astype(np.uint32)
