"""COORDINATIONNUMBER on the cluster pair list (p2a_coord_rows_device through
plumed2_amd.CoordinationNumbers): per-atom values and the vector-Jacobian
product sum_k w_k dcn_k/dx with its box derivative, against
oracle.coordination_numbers (ContactMatrix rows times ONES,
symfunc/CoordinationNumbers.cpp:326-372) and the symfunc goldens.

Tolerances (FP32 pair arithmetic, FP64 sums, as COORDINATION): values
1e-6 relative per row (1e-6 of the largest row for rows near zero);
derivatives and box derivative 1e-5 of their largest entry.  At 50,000 atoms
the oracle's O(rows x environment) loop takes a sample of rows, and the
weighted product is checked through two identities: with all weights 1 it is
twice the COORDINATION derivative (every pair sits in two rows), and it is
linear in the weights."""
import numpy as np
import pytest

import oracle as O
from fixtures import path, read_colvar, read_xyz

pytestmark = pytest.mark.gpu

ROW_TOL, F_TOL = 1e-6, 1e-5
GOLD_TOL = 5.1e-5


def _t(x):
    import torch
    return torch.tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda:0")


def _np(t):
    return t.detach().cpu().numpy()


def _check_rows(got, want):
    scale = np.maximum(np.abs(want), np.max(np.abs(want)) * 1e-3 + 1e-12)
    assert np.max(np.abs(got - want) / scale) <= ROW_TOL, np.max(np.abs(got - want) / scale)


def _check_vec(got, want, tag=""):
    m = max(np.max(np.abs(want)), 1e-12)
    assert np.max(np.abs(got - want)) <= F_TOL * m, (tag, np.max(np.abs(got - want)) / m)


def uniform(n, rho=33.4, seed=0):
    L = (n / rho) ** (1 / 3)
    return np.random.default_rng(seed).uniform(0, L, (n, 3)), L


@pytest.mark.parametrize("defn", ["RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2", "RATIONAL R_0=0.35 NN=8 MM=16 D_MAX=1.0",
                                  "RATIONAL R_0=0.3 NN=10 MM=20 D_MAX=1.2", "RATIONAL R_0=0.3 NN=4 MM=8 D_MAX=1.2",
                                  "RATIONAL R_0=0.25 D_0=0.05 NN=5 MM=9 D_MAX=1.1", "EXP R_0=0.2 D_MAX=1.2"])
def test_species_rows_and_product_match_oracle(defn):
    import plumed2_amd as P
    n = 3000
    x, L = uniform(n, seed=3)
    box = np.diag([L] * 3)
    rng = np.random.default_rng(4)
    species = rng.permutation(n)[:2200]
    cn = P.CoordinationNumbers(defn, species=species)
    w = rng.normal(size=len(species))
    got, vjp, vir = cn(_t(x), box, weights=w)
    assert cn.action.stats()["path"] == 0
    want, owant, ovir = O.coordination_numbers(O.Switch(defn), box, x, species=species, weights=w)
    _check_rows(_np(got), want)
    _check_vec(_np(vjp), owant, "vjp")
    _check_vec(_np(vir), ovir, "box derivative")
    # values alone: the same rows
    _check_rows(_np(cn(_t(x), box)), want)


@pytest.mark.parametrize("overlap", [False, True])
def test_species_a_b_match_oracle(overlap):
    """SPECIESA rows over SPECIESB; with an overlap the same atom in both lists
    is not paired with itself."""
    import plumed2_amd as P
    n = 4000
    x, L = uniform(n, seed=5)
    H = np.array([[L, 0, 0], [0.1 * L, L, 0], [-0.05 * L, 0.08 * L, L]])
    xt = x @ np.linalg.inv(np.diag([L] * 3)) @ H
    a = np.arange(0, 1500)
    b = np.arange(1000, 4000) if overlap else np.arange(1500, 4000)
    defn = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2"
    cn = P.CoordinationNumbers(defn, species_a=a, species_b=b)
    w = np.random.default_rng(6).uniform(0.0, 2.0, size=len(a))
    got, vjp, vir = cn(_t(xt), H, weights=w)
    assert cn.action.stats()["path"] == 0
    want, owant, ovir = O.coordination_numbers(O.Switch(defn), H, xt, species_a=a, species_b=b, weights=w)
    _check_rows(_np(got), want)
    _check_vec(_np(vjp), owant, "vjp")
    _check_vec(_np(vir), ovir, "box derivative")


def test_full_size_50k():
    """SPECIES = all 50,000 atoms of the benchmark system: a sample of 600
    rows against the oracle; sum of the rows = 2 COORDINATION; the product
    with all weights 1 = 2 x the COORDINATION derivative and box derivative
    (each from the oracle over the reference cell-list candidates); linear
    in the weights."""
    import plumed2_amd as P
    n = 50000
    x, L = uniform(n, seed=2)
    box = np.diag([L] * 3)
    defn = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2"
    cn = P.CoordinationNumbers(defn, species=np.arange(n))
    xt = _t(x)
    rows = _np(cn(xt, box))
    sample = np.random.default_rng(7).choice(n, 600, replace=False)
    want = O.coordination_numbers(O.Switch(defn), box, x, species_a=sample, species_b=np.arange(n))
    _check_rows(rows[sample], want)
    pairs = O.nl_cells(O.Pbc(box), 1.2001, x, n, 0)
    ov, od, ow, _ = O.coordination(O.Switch(defn), box, x, n, 0, pairs=pairs)
    assert abs(rows.sum() - 2 * ov) <= ROW_TOL * 2 * ov
    _, vjp1, vir1 = cn(xt, box, weights=np.ones(n))
    _check_vec(_np(vjp1), 2 * od, "ones")
    _check_vec(_np(vir1), 2 * ow, "ones box derivative")
    rng = np.random.default_rng(8)
    w1, w2 = rng.normal(size=n), rng.uniform(size=n)
    _, va, ra = cn(xt, box, weights=w1)
    _, vb, rb = cn(xt, box, weights=w2)
    _, vc, rc = cn(xt, box, weights=w1 + w2)
    _check_vec(_np(vc), _np(va) + _np(vb), "linearity")
    _check_vec(_np(rc), _np(ra) + _np(rb), "linearity box derivative")


def test_symfunc_goldens_through_cluster_rows():
    """regtest/symfunc/rt-single-groupb (SPECIESA=1-5 SPECIESB=6) and
    regtest/symfunc/rt20 (SPECIES=1-5, MORE_THAN summed) through
    CoordinationNumbers; these small cells take the FP64 path."""
    import plumed2_amd as P
    frames = read_xyz(path("symfunc-rt-single-groupb", "trajectory.xyz"))
    _, ref1 = read_colvar(path("symfunc-rt-single-groupb", "colvar.reference"))
    _, ref2 = read_colvar(path("symfunc-rt20", "colv2.reference"))
    more = O.Switch("EXP R_0=4.0")
    c1 = P.CoordinationNumbers("RATIONAL R_0=1.0", species_a=range(5), species_b=[5])
    c2 = P.CoordinationNumbers("RATIONAL R_0=0.5", species=range(5))
    for k, ((pos, box), r1, r2) in enumerate(zip(frames, ref1, ref2)):
        cn1 = _np(c1(_t(pos), box, step=k))
        assert np.max(np.abs(cn1 - r1[1:])) <= GOLD_TOL
        cn2 = _np(c2(_t(pos), box, step=k))
        total = sum(1.0 - more.dist(float(v))[0][0] for v in cn2)
        assert abs(total - r2[1]) <= GOLD_TOL


def test_coincident_atoms_are_not_in_contact():
    """Two distinct atoms at one place contribute nothing to each other's rows
    (ContactMatrix.h:98-106), on the cluster path."""
    import plumed2_amd as P
    x, L = uniform(2000, seed=9)
    x[1] = x[0]
    box = np.diag([L] * 3)
    defn = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2"
    cn = P.CoordinationNumbers(defn, species=np.arange(2000))
    got = _np(cn(_t(x), box))
    assert cn.action.stats()["path"] == 0
    want = O.coordination_numbers(O.Switch(defn), box, x, species=np.arange(2000))
    _check_rows(got, want)


def test_weights_refused_on_fp64_path():
    import plumed2_amd as P
    x = np.random.default_rng(10).uniform(0, 2.0, (200, 3))
    cn = P.CoordinationNumbers("RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2", species=np.arange(200))
    with pytest.raises(P.Plumed2AmdError, match="cluster path"):
        cn(_t(x), np.diag([2.0] * 3), weights=np.ones(200))
