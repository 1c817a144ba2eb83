"""Oracle of the DISTANCE between two centres of mass (COM made whole, then
DISTANCE made whole) and of the METAD chain rule onto the atoms.

regtest/basic/rt18 pins the centre distance: `c1: COM ATOMS=g1` with g1 the
GROUP `ATOMS=1-10,10,9,8 REMOVE=8-10,20 SORT` (atoms 1-10: REMOVE drops the
first occurrences only), `d2: DISTANCE ATOMS=c1,50`, `RESTRAINT ARG=d2 AT=0.9
KAPPA=3.0`, masses from the `mc` file (all 1), on the 5 frames of
regtest/trajectories/trajectory.xyz (cubic box 5.0388).  COLVAR prints d2
with %6.3f; --dump-forces prints the virial diagonal and every atom's force
with %8.4f.  The restraint's force on d2 is f = -KAPPA (d2 - AT)
(src/bias/Restraint.cpp).
"""
import numpy as np
import pytest

import oracle as O
from fixtures import path, read_colvar, read_xyz


def _dumped_forces(fname):
    frames = []
    lines = open(fname).read().split("\n")
    k = 0
    while k < len(lines) and lines[k].strip():
        n = int(lines[k])
        vir = np.array([float(x) for x in lines[k + 1].split()])
        F = np.array([[float(x) for x in lines[k + 2 + i].split()[1:4]] for i in range(n)])
        frames.append((vir, F))
        k += n + 2
    return frames


def test_rt18_golden_value_forces_virial():
    frames = read_xyz(path("rt18", "trajectory.xyz"))
    names, col = read_colvar(path("rt18", "COLVAR.reference"))
    assert names == ["time", "d1", "d2", "d3"]
    mc = np.loadtxt(path("rt18", "mc"))
    masses = mc[:, 1]
    dumped = _dumped_forces(path("rt18", "forces.reference"))
    assert len(frames) == len(col) == len(dumped) == 5
    g1, g2 = list(range(10)), [49]
    for (pos, box), row, (vir_ref, F_ref) in zip(frames, col, dumped):
        v, der, w, _, _ = O.center_distance(box, pos, g1, g2, masses=masses[g1 + g2])
        assert abs(v - row[2]) <= 5.1e-4
        f = -3.0 * (v - 0.9)
        F = np.zeros_like(F_ref)
        F[g1] += f * der[:10]
        F[g2] += f * der[10:]
        assert np.max(np.abs(F - F_ref)) <= 5.1e-5
        np.testing.assert_allclose(np.diag(f * w), vir_ref, rtol=0, atol=5.1e-5)


def test_whole_group_across_the_boundary_closed_form():
    """A compact group wrapped into the cell: its centre is the mean of the
    unwrapped positions (up to a lattice vector), and the distance is the
    minimum-image distance of the two unwrapped means (triclinic cell)."""
    rng = np.random.default_rng(4)
    box = np.array([[4.0, 0.0, 0.0], [0.7, 3.8, 0.0], [-0.5, 0.6, 4.1]])
    inv = np.linalg.inv(box)
    c_a, c_b = np.array([0.1, 0.2, 0.05]), np.array([1.1, 1.0, 0.65])
    ua = c_a + rng.normal(0, 0.35, (40, 3))
    ub = c_b + rng.normal(0, 0.35, (30, 3))
    wrap = lambda x: x - np.floor(x @ inv) @ box  # noqa: E731
    pos = np.concatenate([wrap(ua), wrap(ub)])
    m = rng.uniform(1.0, 16.0, 70)
    v, der, vir, c1, c2 = O.center_distance(box, pos, range(40), range(40, 70), masses=m)
    ma, mb = (m[:40] @ ua) / m[:40].sum(), (m[40:] @ ub) / m[40:].sum()
    n = (c1 - ma) @ inv
    np.testing.assert_allclose(n, np.round(n), atol=1e-12)
    # the second centre is made whole against the first
    np.testing.assert_allclose(c2 - c1, mb - ma, atol=1e-12)
    assert v == pytest.approx(np.linalg.norm(mb - ma), rel=1e-12)
    # the derivative of a centre is w_i times the identity
    d = (mb - ma) / np.linalg.norm(mb - ma)
    np.testing.assert_allclose(der[:40], -np.outer(m[:40] / m[:40].sum(), d), atol=1e-14)
    np.testing.assert_allclose(der[40:], np.outer(m[40:] / m[40:].sum(), d), atol=1e-14)
    np.testing.assert_allclose(der.sum(0), 0.0, atol=1e-14)


def test_center_distance_finite_differences():
    rng = np.random.default_rng(6)
    box = np.diag([3.0, 3.3, 2.8])
    pos = rng.uniform(0.0, 3.0, (25, 3))
    g1, g2 = [0, 3, 5, 7, 11, 12], [2, 4, 20, 21]
    v, der, _, _, _ = O.center_distance(box, pos, g1, g2)
    h = 1e-6
    for k, i in enumerate(g1 + g2):
        for c in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[i, c] += h
            p2[i, c] -= h
            fd = (O.center_distance(box, p1, g1, g2)[0] - O.center_distance(box, p2, g1, g2)[0]) / (2 * h)
            assert fd == pytest.approx(der[k, c], abs=1e-7)
