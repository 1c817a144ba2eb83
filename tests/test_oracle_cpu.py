"""Pins the FP64 oracle (CPU, no GPU needed).

Three independent anchors:
  * the reference's own tool classes compiled from its sources
    (oracle/_ref/libplumedref.so; skipped where it was not built),
  * the reference regtest golden files under tests/golden (rt20-switch-*,
    rt42, rt42c),
  * closed forms / invariants chosen so that a dropped term, a wrong sign or
    a transposed operand fails (s(r0)=1/2, finite differences of dfunc, brute-
    force minimum image, translation invariance, virial identity).
"""
import math

import numpy as np
import pytest

import oracle as O
from fixtures import path, read_colvar, read_deriv, read_forces, read_xyz

needs_ref = pytest.mark.skipif(O.ref() is None, reason="compiled reference subset (oracle/_ref) not built here")

SWITCH_DEFS = [
    "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.0",
    "RATIONAL R_0=0.35 NN=8 MM=16 D_MAX=1.0",
    "RATIONAL R_0=1.3 NN=5 MM=11 D_MAX=2.6",
    "RATIONAL R_0=1.3 NN=5 MM=11 D_MAX=2.6 NOSTRETCH",
    "RATIONAL R_0=1.3 NN=5 D_MAX=2.6",
    "RATIONAL R_0=1.3 NN=6 MM=10 D_MAX=2.6",
    "RATIONAL R_0=1 NN=7 MM=14 D_0=0.1 D_MAX=3",
    "RATIONAL R_0=1 NN=14 MM=28 D_MAX=3",
    "RATIONAL R_0=1",
    "EXP R_0=0.8 D_0=0.5 D_MAX=2.6",
    "GAUSSIAN R_0=1.0 D_0=0.0 D_MAX=2.6",
    "GAUSSIAN R_0=1.0 D_0=0.3 D_MAX=2.6",
    "SMAP R_0=1.3 A=3 B=2 D_MAX=2.6",
    "CUBIC D_0=0 D_MAX=2.6",
    "TANH R_0=1.3 D_MAX=2.6",
    "COSINUS R_0=2.6",
    "Q R_0=1.0 D_0=0.3 BETA=1.0 LAMBDA=1.0 REF=1.3 D_MAX=2.6",
]


@needs_ref
@pytest.mark.parametrize("definition", SWITCH_DEFS)
def test_switch_matches_compiled_reference(definition):
    rng = np.random.default_rng(0)
    d = np.concatenate([rng.uniform(0, 3.2, 3000), [0.0, 0.3, 1.0, 1.3, 2.6, 1.3 * (1 + 1e-6)]])
    a, b = O.Switch(definition), O.RefSwitch(definition)
    assert a.dmax == b.dmax
    for fa, fb, x in ((a.sqr, b.sqr, d * d), (a.dist, b.dist, d)):
        va, da = fa(x)
        vb, db = fb(x)
        np.testing.assert_array_equal(va, vb)
        np.testing.assert_array_equal(da, db)


@needs_ref
@pytest.mark.parametrize("nn,mm,r0,d0", [(6, 0, 1.0, 0.0), (7, 0, 1.0, 0.1), (6, 12, 0.3, 0.0), (8, 16, 0.35, 0.0)])
def test_switch_nnmm_matches_compiled_reference(nn, mm, r0, d0):
    d = np.linspace(0, 3 * r0 * 10, 4001)
    a, b = O.Switch(nn=nn, mm=mm, r0=r0, d0=d0), O.RefSwitch(nn=nn, mm=mm, r0=r0, d0=d0)
    assert a.dmax == b.dmax
    np.testing.assert_array_equal(a.sqr(d * d)[0], b.sqr(d * d)[0])
    np.testing.assert_array_equal(a.sqr(d * d)[1], b.sqr(d * d)[1])


def test_rational_closed_form():
    # without stretch s(r0) = 1/2 and s = 1/(1+(r/r0)^6) for NN=6, MM=12
    sw = O.Switch("RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=10.0 NOSTRETCH")
    r = np.linspace(0.05, 2.0, 200)
    v, df = sw.dist(r)
    np.testing.assert_allclose(v, 1.0 / (1.0 + (r / 0.3) ** 6), rtol=1e-14)
    # dfunc = (1/r) ds/dr
    ds = -6 * (r / 0.3) ** 5 / 0.3 / (1 + (r / 0.3) ** 6) ** 2
    np.testing.assert_allclose(df, ds / r, rtol=1e-12)
    assert abs(sw.dist(0.3)[0][0] - 0.5) < 1e-15


@pytest.mark.parametrize("definition", SWITCH_DEFS)
def test_switch_dfunc_finite_difference(definition):
    sw = O.Switch(definition)
    r = np.linspace(0.07, min(sw.dmax, 3.0) * 0.97, 60)
    h = 1e-6
    v, df = sw.dist(r)
    vp, _ = sw.dist(r + h)
    vm, _ = sw.dist(r - h)
    fd = (vp - vm) / (2 * h) / r
    np.testing.assert_allclose(df, fd, rtol=2e-5, atol=2e-6)


def test_stretch_makes_switch_vanish_at_dmax():
    for definition in SWITCH_DEFS:
        sw = O.Switch(definition)
        if sw.dmax < 1e300 and "NOSTRETCH" not in definition:
            assert abs(sw.dist(sw.dmax)[0][0]) < 1e-12, definition
            assert abs(sw.dist(0.0)[0][0] - 1.0) < 1e-12, definition


RT20 = {
    "rational": "RATIONAL R_0=1.3 NN=5 MM=11 D_MAX=2.6",
    "rational-MMeq2NN": "RATIONAL R_0=1.3 NN=5 D_MAX=2.6",
    "fastrational": "RATIONAL R_0=1.3 NN=6 MM=10 D_MAX=2.6",
    "fastrational-MMeq2NN": "RATIONAL R_0=1.3 D_MAX=2.6",
    "exp": "EXP R_0=0.8 D_0=0.5 D_MAX=2.6",
    "gaussian": "GAUSSIAN R_0=1.0 D_0=0.3 D_MAX=2.6",
    "fastgaussian": "GAUSSIAN R_0=1.0 D_0=0.0 D_MAX=2.6",
    "smap": "SMAP R_0=1.3 A=3 B=2 D_MAX=2.6",
    "cubic": "CUBIC D_0=0 D_MAX=2.6",
    "tanh": "TANH R_0=1.3 D_MAX=2.6",
    "cosinus": "COSINUS R_0=2.6",
    "q": "Q R_0=1.0 D_0=0.3 BETA=1.0 LAMBDA=1.0 REF=1.3 D_MAX=2.6",
}


def rt20_expected(name):
    names, colvar = read_colvar(path(f"rt20-switch-{name}", "COLVAR.reference"))
    deriv = read_deriv(path(f"rt20-switch-{name}", "deriv.reference"))
    return names, colvar, deriv


@pytest.mark.parametrize("name", sorted(RT20))
def test_oracle_rt20_golden(name):
    """2-atom distance sweep of regtest/basic/rt20-switch-*: values (%f) and
    atom + box derivatives (%8.4f), stretched (cs) and NOSTRETCH (c)."""
    frames = read_xyz(path("switchtraj", "switchtraj.xyz"))
    names, colvar, deriv = rt20_expected(name)
    definition = RT20[name]
    for label, defn in (("c", definition + " NOSTRETCH"), ("cs", definition)):
        sw = O.Switch(defn)
        col = names.index(label)
        for k, (pos, box) in enumerate(frames):
            v, d, w, _ = O.coordination(sw, box, pos, 1, 1)
            assert abs(v - colvar[k, col]) <= 5.01e-7, (label, k, v, colvar[k, col])
            params = np.concatenate([d.reshape(-1), w.reshape(-1)])
            np.testing.assert_allclose(params, deriv[label][k, :15], atol=5.01e-5, err_msg=f"{label} frame {k}")


def test_oracle_rt42_golden():
    """regtest/basic/rt42: 108 atoms in a triclinic cell, GROUPA=GROUPB (each pair
    twice, self pairs skipped). check_c: R_0=1 (NN/MM form, automatic D_MAX);
    check_e: R_0=1 NN=7 D_0=0.1 (general rational); check_g: EXP R_0=1 without
    D_MAX (every minimum-image pair)."""
    frames = read_xyz(path("rt42", "trajectory.xyz"))
    cases = {"check_c": O.Switch(nn=6, mm=0, r0=1.0), "check_e": O.Switch(nn=7, mm=0, r0=1.0, d0=0.1),
             "check_g": O.Switch("EXP R_0=1")}
    absidx = np.concatenate([np.arange(108), np.arange(108)])
    for key, sw in cases.items():
        _, ref = read_colvar(path("rt42", key + ".reference"))
        for k in range(0, len(frames), 9):
            pos, box = frames[k]
            p2 = np.concatenate([pos, pos])
            v, _, _, _ = O.coordination(sw, box, p2, 108, 108, absidx=absidx)
            assert abs(v - ref[k, 1]) <= 5.01e-5, (key, k, v, ref[k, 1])


def test_oracle_rt42c_golden():
    frames = read_xyz(path("rt42c", "trajectory.xyz"))
    _, ref = read_colvar(path("rt42c", "colvar.reference"))
    sw = O.Switch(nn=6, mm=0, r0=1.0)
    for k in range(0, len(frames), 7):
        pos, box = frames[k]
        v, _, _, _ = O.coordination(sw, box, pos, 108, 0)
        assert abs(v - ref[k, 1]) <= 5.01e-5


# regtest/basic/rt42 biases: RESTRAINT SLOPE=1 on s = c + d and SLOPE=0.001 on
# c - 0.5 d - 0.5 dfast + e - f + g - h.  d and dfast are CUSTOM 1/(1+x^6)
# (the rational form without D_MAX), f is CUSTOM 1/(1+x^7) with D_0=0.1, and
# h = exp(-x) equals g, so the total bias is
# 1.001 c + 0.999 d + 0.001 (e - f).
RT42_BIAS_TERMS = [(1.001, dict(nn=6, mm=0, r0=1.0)), (0.999, "RATIONAL R_0=1 NN=6 MM=12"),
                   (0.001, dict(nn=7, mm=0, r0=1.0, d0=0.1)), (-0.001, "RATIONAL R_0=1 NN=7 MM=14 D_0=0.1")]


def rt42_forces_virial(coords):
    """Forces on the 108 atoms and the virial the driver dumps, from a list of
    (coefficient, (value, deriv[216,3], virial[3,3])) of the four terms."""
    F = np.zeros((108, 3))
    W = np.zeros((3, 3))
    for c, (_, d, w) in coords:
        F -= c * (d[:108] + d[108:])
        W -= c * w
    return F, W


def test_oracle_rt42_forces_full_virial():
    """regtest/basic/rt42 ff.reference (--dump-forces --dump-full-virial,
    %10.5f): forces -dV/dx and the full virial -dV/ds * (box derivative of s)
    of every frame, from the oracle's derivatives and virials."""
    frames = read_xyz(path("rt42", "trajectory.xyz"))
    ff = read_forces(path("rt42", "ff.reference"))
    assert len(ff) == len(frames)
    absidx = np.concatenate([np.arange(108), np.arange(108)])
    sws = [(c, O.Switch(**a) if isinstance(a, dict) else O.Switch(a)) for c, a in RT42_BIAS_TERMS]
    for k in range(0, len(frames), 3):
        pos, box = frames[k]
        p2 = np.concatenate([pos, pos])
        terms = [(c, O.coordination(sw, box, p2, 108, 108, absidx=absidx)[:3]) for c, sw in sws]
        F, W = rt42_forces_virial(terms)
        vir, frc = ff[k]
        np.testing.assert_allclose(F, frc, atol=5.01e-6, rtol=0, err_msg=f"forces frame {k}")
        np.testing.assert_allclose(W, vir, atol=5.01e-6, rtol=0, err_msg=f"virial frame {k}")


def test_oracle_rt42_cells_forces_zero():
    """regtest/basic/rt42-cells ff.reference: the bias 0.001 (d - c) of the
    NLISTCELLS and NLIST forms puts zero force and virial on every frame."""
    frames = read_xyz(path("rt42-cells", "trajectory.xyz"))
    ff = read_forces(path("rt42-cells", "ff.reference"))
    assert len(ff) == len(frames)
    for vir, frc in ff:
        assert np.all(vir == 0) and np.all(frc == 0)


def test_oracle_rt42_cells_golden():
    """regtest/basic/rt42-cells: NLIST and NLISTCELLS (D_MAX=1.5 < NL_CUTOFF=2,
    NL_STRIDE=4) give the same value on every frame (check_diff = 0)."""
    frames = read_xyz(path("rt42-cells", "trajectory.xyz"))
    _, ref = read_colvar(path("rt42-cells", "check_diff.reference"))
    sw = O.Switch("RATIONAL R_0=1 D_MAX=1.5")
    absidx = np.concatenate([np.arange(108), np.arange(108)])
    pairs_c = pairs_k = None
    for k, (pos, box) in enumerate(frames[:24]):
        p2 = np.concatenate([pos, pos])
        pbc = O.Pbc(box)
        if k % 4 == 0:
            pairs_c = O.nl_classic(pbc, True, 2.0, p2, 108, 108)
            pairs_k = O.nl_cells(pbc, 2.0, p2, 108, 108)
        vc = O.coordination(sw, box, p2, 108, 108, absidx=absidx, pairs=pairs_c)[0]
        vk = O.coordination(sw, box, p2, 108, 108, absidx=absidx, pairs=pairs_k)[0]
        assert abs((vk - vc) - ref[k, 1]) <= 5.01e-5


_R = 7
_IJK = np.array([(i, j, k) for i in range(-_R, _R + 1) for j in range(-_R, _R + 1) for k in range(-_R, _R + 1)], float)


def _brute_min_image(box, d):
    t = d[None, :] + _IJK @ box
    return t[np.argmin((t * t).sum(1))]


@pytest.mark.parametrize("box", [np.diag([5.0, 6.0, 7.0]),
                                 np.array([[6, -6, 0], [0, 6, -6], [-6, 6, 6]], float),
                                 np.array([[4.0, 0, 0], [0.4, 5, 0], [-0.3, 0.5, 6]]),
                                 np.array([[3.0, 0, 0], [1.4, 3, 0], [1.3, 1.2, 3.5]])])
def test_pbc_minimum_image_brute_force(box):
    rng = np.random.default_rng(5)
    v1 = rng.uniform(-8, 8, (300, 3))
    v2 = rng.uniform(-8, 8, (300, 3))
    d = O.Pbc(box).distance(v1, v2)
    for k in range(len(v1)):
        b = _brute_min_image(box, v2[k] - v1[k])
        assert abs(np.linalg.norm(d[k]) - np.linalg.norm(b)) < 1e-9


@needs_ref
@pytest.mark.parametrize("box", [np.diag([5.0, 6.0, 7.0]),
                                 np.array([[6, -6, 0], [0, 6, -6], [-6, 6, 6]], float),
                                 np.array([[4.0, 0, 0], [0.4, 5, 0], [-0.3, 0.5, 6]])])
def test_pbc_matches_compiled_reference(box):
    rng = np.random.default_rng(6)
    v1 = rng.uniform(-10, 10, (2000, 3))
    v2 = rng.uniform(-10, 10, (2000, 3))
    np.testing.assert_array_equal(O.Pbc(box).distance(v1, v2), O.ref_pbc_distance(box, v1, v2))


@needs_ref
@pytest.mark.parametrize("nA,nB,box,defn,nlmode", [
    (300, 0, np.diag([3.0, 3, 3]), "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.0", 0),
    (200, 250, np.diag([3.0, 3.5, 4.0]), "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.0", 0),
    (400, 0, np.array([[4.0, 0, 0], [0.4, 5, 0], [-0.3, 0.5, 6]]), "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2", 2),
    (400, 0, np.diag([4.0, 5, 6]), "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2", 1),
    (250, 300, np.diag([4.0, 4.5, 5]), "EXP R_0=0.2 D_MAX=1.0", 2),
])
def test_coordination_matches_compiled_reference(nA, nB, box, defn, nlmode):
    rng = np.random.default_rng(nA + nB)
    pos = rng.uniform(0, 4, (nA + nB, 3))
    pbc = O.Pbc(box)
    pairs = None
    if nlmode == 2:
        pairs = O.nl_cells(pbc, 1.5, pos, nA, nB)
    elif nlmode == 1:
        pairs = O.nl_classic(pbc, True, 1.5, pos, nA, nB)
    v, d, w, _ = O.coordination(O.Switch(defn), box, pos, nA, nB, pairs=pairs)
    rv, rd, rw, rn = O.ref_coordination(O.RefSwitch(defn), box, pos, nA, nB, nlmode=nlmode, nl_cut=1.5)
    assert v == rv
    np.testing.assert_array_equal(d, rd)
    np.testing.assert_array_equal(w, rw)
    if pairs is not None:
        assert len(pairs) == rn


def test_coordination_invariants():
    rng = np.random.default_rng(9)
    n = 500
    L = (n / 33.4) ** (1 / 3)
    box = np.diag([L, L, L])
    x = rng.uniform(0, L, (n, 3))
    sw = O.Switch("RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2")
    v, d, w, nc = O.coordination(sw, box, x, n, 0)
    # translation invariance: sum of derivatives vanishes
    assert np.max(np.abs(d.sum(0))) < 1e-10
    # symmetric virial
    np.testing.assert_allclose(w, w.T, atol=1e-10)
    # permutation of the atoms permutes the derivatives only
    perm = rng.permutation(n)
    v2, d2, w2, _ = O.coordination(sw, box, x[perm], n, 0)
    assert abs(v - v2) < 1e-9 * abs(v)
    np.testing.assert_allclose(d2, d[perm], atol=1e-10)
    # GROUPA/GROUPB split: all A-B pairs counted once
    a, b = 200, 300
    vab, dab, wab, _ = O.coordination(sw, box, x, a, b)
    vaa = O.coordination(sw, box, x[:a], a, 0)[0]
    vbb = O.coordination(sw, box, x[a:], b, 0)[0]
    assert abs(v - (vab + vaa + vbb)) < 1e-9 * abs(v)
    # central finite differences of the value against the derivatives
    h = 1e-6
    for i in (0, 17, 311):
        for k in range(3):
            xp = x.copy()
            xp[i, k] += h
            xm = x.copy()
            xm[i, k] -= h
            fd = (O.coordination(sw, box, xp, n, 0)[0] - O.coordination(sw, box, xm, n, 0)[0]) / (2 * h)
            assert abs(fd - d[i, k]) < 1e-5 * max(1.0, abs(d[i, k]))
    # virial = box derivative: scale the box and positions by (1+eps) along a
    eps = 1e-6
    for a_ in range(3):
        S = np.eye(3)
        S[a_, a_] += eps
        vp = O.coordination(sw, box @ S, x @ S, n, 0)[0]
        S[a_, a_] -= 2 * eps
        vm = O.coordination(sw, box @ S, x @ S, n, 0)[0]
        # d value / d(strain_aa) = sum deriv_a * x_a - ... = -virial_aa (box derivative convention)
        assert abs((vp - vm) / (2 * eps) + w[a_, a_]) < 1e-4 * max(1, abs(w[a_, a_]))


def test_cells_list_contains_all_pairs_within_cutoff():
    rng = np.random.default_rng(4)
    n = 600
    box = np.array([[5.0, 0, 0], [0.5, 5.0, 0], [-0.4, 0.3, 5.5]])
    x = rng.uniform(0, 5, (n, 3))
    pbc = O.Pbc(box)
    cells = {tuple(p) for p in O.nl_cells(pbc, 1.3, x, n, 0)}
    classic = {tuple(p) for p in O.nl_classic(pbc, True, 1.3, x, n, 0)}
    assert classic <= cells
