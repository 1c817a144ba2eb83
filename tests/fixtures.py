"""Readers for the regtest data files kept under tests/golden/."""
from __future__ import annotations

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def path(*p):
    return os.path.join(GOLDEN, *p)


def read_xyz(fname):
    """plumed driver --ixyz format: natoms, box line (3 or 9 numbers), atoms.
    Returns a list of (positions[n,3], box[3,3])."""
    frames = []
    with open(fname) as f:
        lines = f.read().splitlines()
    i = 0
    while i < len(lines):
        if not lines[i].strip():
            i += 1
            continue
        n = int(lines[i].split()[0])
        b = [float(v) for v in lines[i + 1].split()]
        box = np.diag(b[:3]) if len(b) == 3 else np.array(b[:9]).reshape(3, 3)
        pos = np.array([[float(v) for v in lines[i + 2 + k].split()[1:4]] for k in range(n)])
        frames.append((pos, box))
        i += 2 + n
    return frames


def read_colvar(fname):
    """PLUMED COLVAR-like file -> (field names, array)."""
    names = None
    rows = []
    with open(fname) as f:
        for line in f:
            if line.startswith("#! FIELDS"):
                names = line.split()[2:]
            elif line.startswith("#"):
                continue
            elif line.strip():
                rows.append([float(v) for v in line.split()])
    return names, np.array(rows)


def read_deriv(fname):
    """DUMPDERIVATIVES file -> dict label -> array[frame, parameter]."""
    names, arr = read_colvar(fname)
    labels = names[2:]
    nframes = len(np.unique(arr[:, 0]))
    npar = int(arr[:, 1].max()) + 1
    out = {}
    for k, lab in enumerate(labels):
        out[lab] = arr[:, 2 + k].reshape(nframes, npar)
    return out


def read_hills(fname):
    names, arr = read_colvar(fname)
    return names, arr


def rt_backup_grid_cvs():
    """d1 = DISTANCE ATOMS=1,2 and d2 = DISTANCE ATOMS=3,4 (minimum image in
    the cubic box) over the frames of regtest/basic/rt-backup-grid."""
    out = []
    for pos, box in read_xyz(path("rt-backup-grid", "trajectory.xyz")):
        L = np.diag(box)
        c = []
        for a, b in [(0, 1), (2, 3)]:
            d = pos[b] - pos[a]
            d -= L * np.round(d / L)
            c.append(float(np.sqrt(d @ d)))
        out.append(c)
    return out


def read_forces(fname):
    """plumed driver --dump-forces --dump-full-virial output -> list of
    (virial[3,3], forces[n,3]) per frame."""
    with open(fname) as f:
        lines = [l for l in f.read().splitlines() if l.strip()]
    out = []
    i = 0
    while i < len(lines):
        n = int(lines[i].split()[0])
        vir = np.array([float(v) for v in lines[i + 1].split()]).reshape(3, 3)
        frc = np.array([[float(v) for v in lines[i + 2 + k].split()[1:4]] for k in range(n)])
        out.append((vir, frc))
        i += 2 + n
    return out
