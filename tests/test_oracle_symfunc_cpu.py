"""Oracle of COORDINATIONNUMBER (per-atom coordination numbers) against the
reference's own golden outputs.

regtest/symfunc/rt-single-groupb: `c1: COORDINATIONNUMBER SPECIESA=1-5
SPECIESB=6 SWITCH={RATIONAL R_0=1.0}`, the five per-atom values printed with
%8.4f for every frame of regtest/trajectories/trajectory.xyz (108 atoms, cubic
box 5.0388).
regtest/symfunc/rt20: `c2: COORDINATIONNUMBER SPECIES=1-5 SWITCH={RATIONAL
R_0=0.5} MORE_THAN={EXP R_0=4.0}`, the sum over atoms of 1 - exp(-cn_i/4),
printed with %8.4f, on the same trajectory.
"""
import numpy as np
import pytest

import oracle as O
from fixtures import path, read_colvar, read_xyz

# both goldens are printed with %8.4f
TOL = 5.1e-5


def _frames():
    return read_xyz(path("symfunc-rt-single-groupb", "trajectory.xyz"))


def test_coordination_numbers_speciesa_speciesb_golden():
    names, ref = read_colvar(path("symfunc-rt-single-groupb", "colvar.reference"))
    assert names[1:] == ["c1.1", "c1.2", "c1.3", "c1.4", "c1.5"]
    frames = _frames()
    assert len(frames) == len(ref)
    sw = O.Switch("RATIONAL R_0=1.0")
    for (pos, box), row in zip(frames, ref):
        cn = O.coordination_numbers(sw, box, pos, species_a=range(0, 5), species_b=[5])
        assert np.max(np.abs(cn - row[1:])) <= TOL


def test_coordination_numbers_species_more_than_golden():
    names, ref = read_colvar(path("symfunc-rt20", "colv2.reference"))
    assert names[1:] == ["c2_morethan"]
    frames = _frames()
    assert len(frames) == len(ref)
    sw = O.Switch("RATIONAL R_0=0.5")
    more = O.Switch("EXP R_0=4.0")
    for (pos, box), row in zip(frames, ref):
        cn = O.coordination_numbers(sw, box, pos, species=range(0, 5))
        total = sum(1.0 - more.dist(float(x))[0][0] for x in cn)
        assert abs(total - row[1]) <= TOL


def test_coordination_numbers_brute_force_and_symmetry():
    """Every row against a direct minimum-image sum; with SPECIES the contact
    matrix is symmetric, so sum_i cn_i is twice COORDINATION of the group."""
    rng = np.random.default_rng(3)
    box = np.diag([2.1, 2.4, 1.9])
    pos = rng.uniform(0.0, 2.5, size=(40, 3))
    sw = O.Switch("RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=0.9")
    cn = O.coordination_numbers(sw, box, pos, species=range(40))
    L = np.diag(box)
    for i in range(40):
        d = pos - pos[i]
        d -= L * np.round(d / L)
        r = np.sqrt((d * d).sum(1))
        expect = sum(sw.dist(float(r[j]))[0][0] for j in range(40) if j != i)
        assert cn[i] == pytest.approx(expect, abs=1e-12)
    total = O.coordination(sw, box, pos, 40)[0]
    assert cn.sum() == pytest.approx(2.0 * total, rel=1e-12)
    # SPECIESA/SPECIESB rows equal the corresponding SPECIES rows minus the A-A contacts
    cab = O.coordination_numbers(sw, box, pos, species_a=range(10), species_b=range(10, 40))
    caa = O.coordination_numbers(sw, box, pos, species=range(10))
    assert np.allclose(cab + caa, cn[:10], rtol=0, atol=1e-12)


def test_coincident_atoms_are_not_in_contact():
    """ContactMatrix::calculateWeight returns nothing for |r|^2 < DBL_EPSILON:
    two distinct atoms at one place do not count s(0) = 1 for each other."""
    box = np.diag([3.0, 3.0, 3.0])
    pos = np.array([[1.0, 1.0, 1.0], [1.0, 1.0, 1.0], [1.2, 1.0, 1.0]])
    sw = O.Switch("RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=0.9")
    cn = O.coordination_numbers(sw, box, pos, species=range(3))
    s02 = sw.dist(0.2)[0][0]
    np.testing.assert_allclose(cn, [s02, s02, 2 * s02], rtol=1e-12)


def test_values_below_epsilon_are_zeroed():
    """A contact whose switch value lies below DBL_EPSILON counts nothing
    (and carries no derivative): an EXP switch far out gives ~1e-20."""
    box = np.diag([30.0, 30.0, 30.0])
    pos = np.array([[1.0, 1.0, 1.0], [1.0 + 0.2 * 46.0, 1.0, 1.0], [1.5, 1.0, 1.0]])
    sw = O.Switch("EXP R_0=0.2")
    assert 0.0 < sw.dist(0.2 * 46.0)[0][0] < np.finfo(float).eps
    cn, der, _ = O.coordination_numbers(sw, box, pos, species=range(3), weights=np.ones(3))
    np.testing.assert_allclose(cn, [sw.dist(0.5)[0][0], 0.0, sw.dist(0.5)[0][0]], rtol=1e-12)
    assert np.all(der[1] == 0.0)


def test_vector_jacobian_product_against_finite_differences():
    """sum_k w_k d cn_k / dx by central differences (h = 1e-6 nm) on a small
    triclinic system, and the box derivative against -sum w_k dfunc d (x) d
    through translation invariance (the columns of the derivative sum to 0)."""
    rng = np.random.default_rng(5)
    box = np.array([[2.3, 0.0, 0.0], [0.4, 2.1, 0.0], [-0.3, 0.5, 2.2]])
    pos = rng.uniform(0.0, 2.2, size=(30, 3))
    sw = O.Switch("RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=0.9")
    a, b = list(range(12)), list(range(8, 30))
    w = rng.normal(size=12)
    cn, der, vir = O.coordination_numbers(sw, box, pos, species_a=a, species_b=b, weights=w)
    h = 1e-6
    for i in range(30):
        for c in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[i, c] += h
            p2[i, c] -= h
            f1 = w @ O.coordination_numbers(sw, box, p1, species_a=a, species_b=b)
            f2 = w @ O.coordination_numbers(sw, box, p2, species_a=a, species_b=b)
            assert (f1 - f2) / (2 * h) == pytest.approx(der[i, c], abs=1e-6 * np.max(np.abs(der)))
    np.testing.assert_allclose(der.sum(0), 0.0, atol=1e-12 * np.max(np.abs(der)))
    np.testing.assert_allclose(vir, vir.T, atol=1e-13 * np.max(np.abs(vir)))
