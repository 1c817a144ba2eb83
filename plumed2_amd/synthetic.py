"""Seeded synthetic inputs: uniform random atoms in periodic cells.

``batch_systems`` draws the many-systems workload: atom counts log-uniform in
[n_lo, n_hi]; per system, edges a, b, c uniform in [4, 12] nm and tilts
uniform in +-0.1 of the edge of their row (rows a = (a,0,0),
b = (t_xy b, b, 0), c = (t_xz c, t_yz c, c)), then the cell is rescaled
isotropically so that its volume is n / density (tilt fractions kept);
positions are uniform in the cell.
"""
from __future__ import annotations

import numpy as np


def batch_systems(nsys=256, n_lo=2000, n_hi=64000, density=33.4, seed=4):
    rng = np.random.default_rng(seed)
    natoms = np.round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), nsys))).astype(np.int64)
    boxes = np.empty((nsys, 3, 3))
    pos = []
    for k in range(nsys):
        e = rng.uniform(4.0, 12.0, 3)
        t = rng.uniform(-0.1, 0.1, 3)
        box = np.array([[e[0], 0.0, 0.0], [t[0] * e[1], e[1], 0.0], [t[1] * e[2], t[2] * e[2], e[2]]])
        box *= (natoms[k] / density / abs(np.linalg.det(box))) ** (1.0 / 3.0)
        boxes[k] = box
        pos.append(rng.random((natoms[k], 3)) @ box)
    return natoms, boxes, np.concatenate(pos)
