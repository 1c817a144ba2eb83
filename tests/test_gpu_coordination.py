"""GPU parity of COORDINATION (HIP library through the C ABI) against the FP64
oracle and the reference regtest golden files.

Tolerances (FP32 pair arithmetic, FP64 accumulation):
  CV      |dCV| <= 1e-6 |CV|
  forces  max |dF| <= 1e-5 max |F|
  virial  max |dW| <= 1e-5 max |W|
FP64 paths (NOPBC, PAIR, unbounded D_MAX) are held to 1e-10 relative.
"""
import numpy as np
import pytest

import oracle as O
from fixtures import path, read_colvar, read_deriv, read_xyz

pytestmark = pytest.mark.gpu

CV_TOL, F_TOL, W_TOL = 1e-6, 1e-5, 1e-5


def _P():
    import plumed2_amd as P
    return P


def check(v, d, w, ov, od, ow, cv_tol=CV_TOL, f_tol=F_TOL, w_tol=W_TOL, tag=""):
    assert abs(v - ov) <= cv_tol * max(abs(ov), 1e-12), (tag, v, ov)
    fmax = max(np.max(np.abs(od)), 1e-12)
    assert np.max(np.abs(d - od)) <= f_tol * fmax, (tag, np.max(np.abs(d - od)) / fmax)
    wmax = max(np.max(np.abs(ow)), 1e-12)
    assert np.max(np.abs(w - ow)) <= w_tol * wmax, (tag, np.max(np.abs(w - ow)) / wmax)


def uniform(n, rho=33.4, seed=0):
    L = (n / rho) ** (1 / 3)
    return np.random.default_rng(seed).uniform(0, L, (n, 3)), L


def test_two_group_c1_shape():
    """GROUPA=1000, GROUPB=1000 atoms in a 5 nm cube, no neighbour list."""
    P = _P()
    x = np.random.default_rng(1).uniform(0, 5, (2000, 3))
    box = np.diag([5.0, 5, 5])
    defn = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.0"
    act = P.Coordination(f"c: COORDINATION GROUPA=1-1000 GROUPB=1001-2000 SWITCH={{{defn}}}")
    v, d, w = act.calculate(x, box)
    ov, od, ow, npair = O.coordination(O.Switch(defn), box, x, 1000, 1000)
    check(v, d, w, ov, od, ow)
    st = act.stats()
    assert st["path"] == 0 and st["contributing_pairs"] == npair


@pytest.mark.parametrize("n", [2, 9, 64, 513, 4000])
def test_single_group_sizes(n):
    P = _P()
    x, L = uniform(max(n, 64), seed=n)
    x = x[:n] if n < 64 else x
    box = np.diag([max(L, 3.0)] * 3)
    defn = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2"
    act = P.Coordination(f"c: COORDINATION GROUPA=1-{n} SWITCH={{{defn}}}")
    v, d, w = act.calculate(x, box)
    ov, od, ow, npair = O.coordination(O.Switch(defn), box, x, n, 0)
    if npair == 0:
        assert v == 0.0 and np.all(d == 0)
    else:
        check(v, d, w, ov, od, ow, tag=n)
    assert act.stats()["contributing_pairs"] == npair


def test_neighbour_list_stride_random_walk():
    """NLISTCELLS NL_CUTOFF=1.5 NL_STRIDE=10 over a random walk: every frame
    against the oracle driven by the reference cell-list candidate set rebuilt
    on the same steps."""
    P = _P()
    n = 4000
    x, L = uniform(n, seed=2)
    box = np.diag([L] * 3)
    defn = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2"
    act = P.Coordination(f"c: COORDINATION GROUPA=1-{n} SWITCH={{{defn}}} NLISTCELLS NL_CUTOFF=1.5 NL_STRIDE=10")
    sw = O.Switch(defn)
    rng = np.random.default_rng(3)
    pairs = None
    for step in range(23):
        if step % 10 == 0:
            pairs = O.nl_cells(O.Pbc(box), 1.5, x, n, 0)
        v, d, w = act.calculate(x, box, step=step)
        ov, od, ow, npair = O.coordination(sw, box, x, n, 0, pairs=pairs)
        check(v, d, w, ov, od, ow, tag=step)
        assert act.stats()["contributing_pairs"] == npair
        x = x + rng.normal(0, 0.005, x.shape)
    assert act.stats()["rebuilds"] == 3


@pytest.mark.parametrize("tilt", [(0.1, -0.05, 0.08), (-0.1, 0.1, -0.1)])
def test_triclinic(tilt):
    P = _P()
    n = 3000
    x, L = uniform(n, seed=4)
    box = np.array([[L, 0, 0], [tilt[0] * L, L, 0], [tilt[1] * L, tilt[2] * L, L]])
    defn = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2"
    act = P.Coordination(f"c: COORDINATION GROUPA=1-{n} SWITCH={{{defn}}}")
    v, d, w = act.calculate(x, box)
    ov, od, ow, npair = O.coordination(O.Switch(defn), box, x, n, 0)
    check(v, d, w, ov, od, ow)
    assert act.stats()["path"] == 0 and act.stats()["contributing_pairs"] == npair


def test_long_list_radius_searches_overflowing_clusters():
    """NLISTCELLS NL_CUTOFF=2.8 with D_MAX 1.2 on 12,000 atoms: ~800
    candidate j clusters per i cluster, more than the count pass keeps (384),
    so the fill pass searches again; CV, forces, virial and pair count match
    the oracle."""
    P = _P()
    n = 12000
    x, L = uniform(n, seed=13)
    box = np.diag([L] * 3)
    defn = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2"
    act = P.Coordination(f"c: COORDINATION GROUPA=1-{n} SWITCH={{{defn}}} NLIST NL_CUTOFF=2.8 NL_STRIDE=10")
    v, d, w = act.calculate(x, box, step=0)
    ov, od, ow, npair = O.coordination(O.Switch(defn), box, x, n, 0)
    check(v, d, w, ov, od, ow)
    assert act.stats()["path"] == 0 and act.stats()["contributing_pairs"] == npair


def test_tall_cell_columns_ranked_in_global_memory():
    """A 3 x 3 x 60 nm cell: 16 columns of ~1,130 atoms, more than a column
    ranked in LDS holds (1,024), so the rebuild ranks them in global memory;
    CV, forces, virial and pair count match the oracle."""
    P = _P()
    box = np.diag([3.0, 3.0, 60.0])
    n = int(33.4 * 3.0 * 3.0 * 60.0)
    x = np.random.default_rng(12).uniform(0.0, 1.0, (n, 3)) * np.array([3.0, 3.0, 60.0])
    defn = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2"
    act = P.Coordination(f"c: COORDINATION GROUPA=1-{n} SWITCH={{{defn}}}")
    v, d, w = act.calculate(x, box)
    ov, od, ow, npair = O.coordination(O.Switch(defn), box, x, n, 0)
    check(v, d, w, ov, od, ow)
    assert act.stats()["path"] == 0 and act.stats()["contributing_pairs"] == npair


@pytest.mark.parametrize("defn", ["EXP R_0=0.2 D_MAX=1.0", "GAUSSIAN R_0=0.3 D_0=0.1 D_MAX=1.0",
                                  "RATIONAL R_0=0.3 NN=5 MM=11 D_MAX=1.0", "RATIONAL R_0=0.35 NN=8 MM=16 D_MAX=1.0",
                                  "TANH R_0=0.3 D_MAX=1.0", "COSINUS R_0=1.0", "SMAP R_0=0.3 A=3 B=2 D_MAX=1.0",
                                  "CUBIC D_0=0.1 D_MAX=1.0", "Q R_0=1.0 D_0=0.1 BETA=5.0 LAMBDA=1.0 REF=0.4 D_MAX=1.0",
                                  "GAUSSIAN R_0=1.0 D_MAX=1.0"])
def test_switch_types(defn):
    P = _P()
    n = 2500
    x, L = uniform(n, seed=5)
    box = np.diag([L] * 3)
    act = P.Coordination(f"c: COORDINATION GROUPA=1-{n} SWITCH={{{defn}}}")
    v, d, w = act.calculate(x, box)
    ov, od, ow, npair = O.coordination(O.Switch(defn), box, x, n, 0)
    check(v, d, w, ov, od, ow, tag=defn)


def test_overlapping_groups_skip_self_pairs():
    P = _P()
    x, L = uniform(900, seed=6)
    box = np.diag([L] * 3)
    ga = np.arange(600)
    gb = np.arange(300, 900)
    req = np.concatenate([ga, gb])
    defn = "RATIONAL R_0=0.3 D_MAX=1.0"
    act = P.Coordination(f"c: COORDINATION GROUPA=1-600 GROUPB=301-900 SWITCH={{{defn}}}")
    v, d, w = act.calculate(x[req], box)
    ov, od, ow, _ = O.coordination(O.Switch(defn), box, x[req], 600, 600, absidx=req)
    check(v, d, w, ov, od, ow)


def test_groupb_larger_than_groupa():
    P = _P()
    x, L = uniform(3000, seed=7)
    box = np.diag([L] * 3)
    defn = "RATIONAL R_0=0.35 NN=8 MM=16 D_MAX=1.0"
    act = P.Coordination(f"c: COORDINATION GROUPA=1-200 GROUPB=201-3000 SWITCH={{{defn}}}")
    v, d, w = act.calculate(x, box)
    ov, od, ow, npair = O.coordination(O.Switch(defn), box, x, 200, 2800)
    check(v, d, w, ov, od, ow)
    assert act.stats()["contributing_pairs"] == npair


def test_fp64_paths():
    """NOPBC, PAIR and unbounded D_MAX run the FP64 kernels."""
    P = _P()
    x, L = uniform(1200, seed=8)
    box = np.diag([L] * 3)
    sw = "RATIONAL R_0=0.3 D_MAX=1.0"
    act = P.Coordination(f"c: COORDINATION GROUPA=1-600 GROUPB=601-1200 SWITCH={{{sw}}} NOPBC")
    v, d, w = act.calculate(x, box)
    ov, od, ow, _ = O.coordination(O.Switch(sw), box, x, 600, 600, use_pbc=False)
    assert act.stats()["path"] == 1
    check(v, d, w, ov, od, ow, 1e-10, 1e-10, 1e-10)
    act = P.Coordination("c: COORDINATION GROUPA=1-600 GROUPB=601-1200 PAIR R_0=0.3")
    v, d, w = act.calculate(x, box)
    ov, od, ow, _ = O.coordination(O.Switch(nn=6, mm=0, r0=0.3), box, x, 600, 600, dopair=True)
    assert act.stats()["path"] == 2
    check(v, d, w, ov, od, ow, 1e-10, 1e-10, 1e-10)
    act = P.Coordination("c: COORDINATION GROUPA=1-400 SWITCH={EXP R_0=0.3}")
    v, d, w = act.calculate(x[:400], box)
    ov, od, ow, _ = O.coordination(O.Switch("EXP R_0=0.3"), box, x[:400], 400, 0)
    check(v, d, w, ov, od, ow, 1e-10, 1e-10, 1e-10)


RT20 = {
    "rational": "RATIONAL R_0=1.3 NN=5 MM=11 D_MAX=2.6",
    "fastrational": "RATIONAL R_0=1.3 NN=6 MM=10 D_MAX=2.6",
    "fastrational-MMeq2NN": "RATIONAL R_0=1.3 D_MAX=2.6",
    "exp": "EXP R_0=0.8 D_0=0.5 D_MAX=2.6",
    "gaussian": "GAUSSIAN R_0=1.0 D_0=0.3 D_MAX=2.6",
    "fastgaussian": "GAUSSIAN R_0=1.0 D_0=0.0 D_MAX=2.6",
    "smap": "SMAP R_0=1.3 A=3 B=2 D_MAX=2.6",
    "cubic": "CUBIC D_0=0 D_MAX=2.6",
    "tanh": "TANH R_0=1.3 D_MAX=2.6",
    "q": "Q R_0=1.0 D_0=0.3 BETA=1.0 LAMBDA=1.0 REF=1.3 D_MAX=2.6",
}


@pytest.mark.parametrize("name", sorted(RT20))
def test_rt20_golden(name):
    """regtest/basic/rt20-switch-*: the 2-atom sweep through the GPU (1-atom
    clusters on both sides), against the golden COLVAR (%f) and derivatives (%8.4f)."""
    P = _P()
    frames = read_xyz(path("switchtraj", "switchtraj.xyz"))
    names, colvar = read_colvar(path(f"rt20-switch-{name}", "COLVAR.reference"))
    deriv = read_deriv(path(f"rt20-switch-{name}", "deriv.reference"))
    for label, defn in (("c", RT20[name] + " NOSTRETCH"), ("cs", RT20[name])):
        act = P.Coordination(f"{label}: COORDINATION GROUPA=1 GROUPB=2 SWITCH={{{defn}}}")
        col = names.index(label)
        for k, (pos, box) in enumerate(frames):
            v, d, w = act.calculate(pos, box, step=k)
            assert abs(v - colvar[k, col]) <= 5.01e-7 + 1e-6 * abs(colvar[k, col]), (label, k)
            params = np.concatenate([d.reshape(-1), w.reshape(-1)])
            np.testing.assert_allclose(params, deriv[label][k, :15], atol=5.1e-5, err_msg=f"{label} frame {k}")


def test_rt42_golden_triclinic():
    """regtest/basic/rt42 check_c / check_e / check_g on the triclinic 108-atom frames."""
    P = _P()
    frames = read_xyz(path("rt42", "trajectory.xyz"))
    lines = {"check_c": "c: COORDINATION GROUPA=1-108 GROUPB=1-108 R_0=1",
             "check_e": "e: COORDINATION GROUPA=1-108 GROUPB=1-108 R_0=1 NN=7 D_0=0.1",
             "check_g": "g: COORDINATION GROUPA=1-108 GROUPB=1-108 SWITCH={EXP R_0=1}"}
    for key, line in lines.items():
        _, ref = read_colvar(path("rt42", key + ".reference"))
        act = P.Coordination(line)
        for k in range(0, len(frames), 5):
            pos, box = frames[k]
            v, _, _ = act.calculate(np.concatenate([pos, pos]), box, step=k)
            assert abs(v - ref[k, 1]) <= 5.01e-5 + 1e-6 * abs(ref[k, 1]), (key, k, v, ref[k, 1])


def test_rt42c_golden():
    P = _P()
    frames = read_xyz(path("rt42c", "trajectory.xyz"))
    _, ref = read_colvar(path("rt42c", "colvar.reference"))
    act = P.Coordination("c: COORDINATION GROUPA=1-108 R_0=1")
    for k in range(0, len(frames), 3):
        pos, box = frames[k]
        v, _, _ = act.calculate(pos, box, step=k)
        assert abs(v - ref[k, 1]) <= 5.01e-5 + 1e-6 * abs(ref[k, 1])


def test_repeatable_value_and_virial():
    """CV and virial are reduced in a fixed order: bitwise repeatable."""
    P = _P()
    x, L = uniform(6000, seed=9)
    box = np.diag([L] * 3)
    act = P.Coordination("c: COORDINATION GROUPA=1-6000 SWITCH={RATIONAL R_0=0.3 D_MAX=1.2}")
    v1, d1, w1 = act.calculate(x, box, 0)
    v2, d2, w2 = act.calculate(x, box, 1)
    assert v1 == v2
    np.testing.assert_array_equal(w1, w2)
    np.testing.assert_allclose(d1, d2, rtol=0, atol=1e-12 * np.max(np.abs(d1)))


def test_full_size_properties():
    """50,000 atoms (the benchmark workload, frame 0): value, derivatives and
    virial against the oracle driven by the reference cell-list candidates."""
    P = _P()
    n = 50000
    x, L = uniform(n, seed=2)
    box = np.diag([L] * 3)
    defn = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2"
    act = P.Coordination(f"c: COORDINATION GROUPA=1-{n} SWITCH={{{defn}}} NLISTCELLS NL_CUTOFF=1.5 NL_STRIDE=10")
    v, d, w = act.calculate(x, box, 0)
    pairs = O.nl_cells(O.Pbc(box), 1.5, x, n, 0)
    ov, od, ow, npair = O.coordination(O.Switch(defn), box, x, n, 0, pairs=pairs)
    check(v, d, w, ov, od, ow)
    assert act.stats()["contributing_pairs"] == npair
    assert np.max(np.abs(d.sum(0))) < 1e-9 * np.max(np.abs(d)) * n


def test_device_entry_point_torch_stream():
    import torch

    P = _P()
    n = 3000
    x, L = uniform(n, seed=10)
    box = np.diag([L] * 3)
    act = P.Coordination(f"c: COORDINATION GROUPA=1-{n} SWITCH={{RATIONAL R_0=0.3 D_MAX=1.2}}")
    xt = torch.from_numpy(x).cuda()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        val, der, vir = act.calculate_torch(xt, box, step=0)
    torch.cuda.synchronize()
    v, d, w = act.calculate(x, box, step=1)
    assert float(val.item()) == v
    np.testing.assert_allclose(der.cpu().numpy(), d, atol=1e-12 * np.max(np.abs(d)))
    np.testing.assert_array_equal(vir.cpu().numpy(), w)


def test_errors_are_raised():
    P = _P()
    with pytest.raises(P.Plumed2AmdError, match="R_0 should be explicitly specified"):
        P.Coordination("c: COORDINATION GROUPA=1-10")
    with pytest.raises(P.Plumed2AmdError, match="D_MAX larger than NL_CUTOFF"):
        P.Coordination("c: COORDINATION GROUPA=1-10 SWITCH={RATIONAL R_0=1 D_MAX=3} NLIST NL_CUTOFF=2 NL_STRIDE=5")
    act = P.Coordination("c: COORDINATION GROUPA=1-10 R_0=1")
    with pytest.raises(ValueError):
        act.calculate(np.zeros((9, 3)), np.diag([3.0] * 3))


@pytest.mark.parametrize("chunk,bpc", [(32, 0), (64, 2), (32, 8)])
def test_pair_kernel_decomposition(chunk, bpc):
    """Entries per wave and persistent grid change the work split, not the
    pairs: exact pair count, CV and forces within the FP32 tolerance."""
    P = _P()
    x, L = uniform(20000, seed=21)
    box = np.diag([L] * 3)
    defn = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2"
    line = f"c: COORDINATION GROUPA=1-20000 SWITCH={{{defn}}} NLISTCELLS NL_CUTOFF=1.5 NL_STRIDE=10"
    a0 = P.Coordination(line)
    v0, d0, w0 = a0.calculate(x, box, 0)
    n0 = a0.stats()["contributing_pairs"]
    a = P.Coordination(line)
    a.set_tuning(chunk, bpc)
    v, d, w = a.calculate(x, box, 0)
    assert a.stats()["contributing_pairs"] == n0
    ov, od, ow, npair = O.coordination(O.Switch(defn), box, x, 20000, 0)
    assert npair == n0
    check(v, d, w, ov, od, ow)
    a.close()
    a0.close()


@pytest.mark.parametrize("chunk", [32, 64])
def test_lattice_pairs_exactly_at_dmax(chunk):
    """Simple cubic lattice with spacing D_MAX/2: shells 1-3 inside D_MAX,
    shell 4 (six partners at 2a) exactly at D_MAX, so every work item needs
    the FP64 cut decision (inclusive cut on the distance Pbc::distance
    gives).  Exact pair count, CV, forces."""
    P = _P()
    a, m = 0.2, 20
    g = np.arange(m) * a
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) + 0.013
    # every other atom leaves its site (non-zero forces); pairs of the atoms
    # left on sites stay exactly at D_MAX
    x[::2] += np.random.default_rng(chunk).normal(0.0, 0.005, x[::2].shape)
    n = len(x)
    box = np.diag([m * a] * 3)
    defn = f"RATIONAL R_0=0.15 NN=6 MM=12 D_MAX={2 * a}"
    act = P.Coordination(f"c: COORDINATION GROUPA=1-{n} SWITCH={{{defn}}}")
    act.set_tuning(chunk, 0)
    v, d, w = act.calculate(x, box, 0)
    ov, od, ow, npair = O.coordination(O.Switch(defn), box, x, n, 0)
    assert act.stats()["contributing_pairs"] == npair
    check(v, d, w, ov, od, ow)
    act.close()


@pytest.mark.parametrize("chunk", [32, 64])
def test_lattice_shell_of_thirty_at_dmax(chunk):
    """Simple cubic lattice with D_MAX = 3 a: the 30 partners of every atom
    at (3,0,0) a and (2,2,1) a sit exactly at D_MAX, so the items hold far
    more band pairs than one round of band records (64) and the band pairs
    are collected again for the later rounds.  Exact pair count, CV, virial
    against the oracle; the forces vanish by symmetry, so they are checked
    against the derivative of one nearest-neighbour pair."""
    P = _P()
    a, m = 0.1, 15
    g = np.arange(m) * a
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) + 0.0137
    n = len(x)
    box = np.diag([m * a] * 3)
    defn = f"RATIONAL R_0=0.12 NN=6 MM=12 D_MAX={3 * a}"
    act = P.Coordination(f"c: COORDINATION GROUPA=1-{n} SWITCH={{{defn}}}")
    act.set_tuning(chunk, 0)
    v, d, w = act.calculate(x, box, 0)
    assert act.stats()["path"] == 0
    ov, od, ow, npair = O.coordination(O.Switch(defn), box, x, n, 0)
    assert act.stats()["contributing_pairs"] == npair
    assert abs(v - ov) <= CV_TOL * abs(ov)
    fpair = abs(O.Switch(defn).dist(a)[1][0]) * a
    assert np.max(np.abs(d - od)) <= 2 * F_TOL * fpair, np.max(np.abs(d - od)) / fpair
    assert np.max(np.abs(w - ow)) <= W_TOL * np.max(np.abs(ow))
    act.close()


def test_dense_long_list_radius():
    """100 atoms/nm^3 with NL_CUTOFF=1.6: most i clusters have more list
    candidates than the count pass keeps, so the fill pass searches again
    for them; exact pair count, CV, forces against the oracle."""
    P = _P()
    x, L = uniform(20000, rho=100.0, seed=31)
    box = np.diag([L] * 3)
    defn = "RATIONAL R_0=0.35 NN=8 MM=16 D_MAX=1.0"
    act = P.Coordination(f"c: COORDINATION GROUPA=1-20000 SWITCH={{{defn}}} NLISTCELLS NL_CUTOFF=1.6 NL_STRIDE=5")
    v, d, w = act.calculate(x, box, 0)
    assert act.stats()["list_entries"] > 400 * act.stats()["clusters_i"]
    ov, od, ow, npair = O.coordination(O.Switch(defn), box, x, 20000, 0)
    assert act.stats()["contributing_pairs"] == npair
    check(v, d, w, ov, od, ow)
    act.close()


@pytest.mark.parametrize("defn", ["EXP R_0=0.2 D_0=0.1 D_MAX=1.0", "RATIONAL R_0=0.3 NN=7 MM=13 D_MAX=1.0"])
def test_value_only_generic_switch(defn):
    """Value-only kernel (no derivative and no virial pointers) with switches
    outside fixedRational: the value equals the full kernel's and the oracle's."""
    import torch

    P = _P()
    x, L = uniform(8000, seed=41)
    box = np.diag([L] * 3)
    act = P.Coordination(f"c: COORDINATION GROUPA=1-8000 SWITCH={{{defn}}}")
    v_full, _, _ = act.calculate(x, box, 0)
    xt = torch.tensor(x, dtype=torch.float64, device="cuda")
    val = torch.empty(1, dtype=torch.float64, device="cuda")
    act.set_stream(torch.cuda.current_stream().cuda_stream)
    act.calculate_device(xt.data_ptr(), box, 1, val.data_ptr(), 0, 0)
    torch.cuda.synchronize()
    ov, _, _, npair = O.coordination(O.Switch(defn), box, x, 8000, 0)
    assert abs(float(val.item()) - ov) <= 1e-6 * abs(ov)
    assert abs(v_full - ov) <= 1e-6 * abs(ov)
    assert act.stats()["contributing_pairs"] == npair
    act.close()


@pytest.mark.parametrize("groups", ["one", "two"])
def test_skin_keeps_the_list_and_matches_rebuilding(groups):
    """Without a neighbour list, a skin keeps the pair list while no atom has
    moved more than skin/2.  Along a random walk every step equals a context
    that rebuilds at every call (same contributing pairs; CV, forces and virial
    within the FP32 tolerances), the list is rebuilt only now and then, and
    the last step matches the oracle."""
    P = _P()
    n = 12000
    x, L = uniform(n, seed=7)
    box = np.diag([L] * 3)
    defn = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2"
    grp = f"GROUPA=1-{n}" if groups == "one" else f"GROUPA=1-2000 GROUPB=2001-{n}"
    line = f"c: COORDINATION {grp} SWITCH={{{defn}}}"
    a = P.Coordination(line)
    a.set_skin(0.1)
    b = P.Coordination(line)
    rng = np.random.default_rng(8)
    steps = 30
    for s in range(steps):
        va, da, wa = a.calculate(x, box, step=s)
        vb, db, wb = b.calculate(x, box, step=s)
        check(va, da, wa, vb, db, wb, tag=s)
        assert a.stats()["contributing_pairs"] == b.stats()["contributing_pairs"], s
        last = x
        x = x + rng.normal(0.0, 0.003, x.shape)
    na, nb = (n, 0) if groups == "one" else (2000, n - 2000)
    ov, od, ow, npair = O.coordination(O.Switch(defn), box, last, na, nb)
    check(va, da, wa, ov, od, ow)
    assert a.stats()["contributing_pairs"] == npair
    # the walk moves some atom beyond skin/2 = 0.05 nm within ~15 steps
    assert b.stats()["rebuilds"] == steps
    assert 2 <= a.stats()["rebuilds"] <= steps // 3, a.stats()["rebuilds"]
    a.close()
    b.close()


def test_rebuild_then_new_positions_in_the_same_buffer():
    """A rebuild on device positions packs those positions for the pair pass;
    a later call on the same buffer holding other positions (step 1 of an
    NLIST stride keeps the list) must pack its own: CV, forces and virial
    match the oracle at the new positions."""
    import torch
    P = _P()
    n = 6000
    x, L = uniform(n, seed=14)
    box = np.diag([L] * 3)
    defn = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2"
    act = P.Coordination(f"c: COORDINATION GROUPA=1-{n} SWITCH={{{defn}}} NLIST NL_CUTOFF=1.5 NL_STRIDE=10")
    xt = torch.tensor(x, dtype=torch.float64, device="cuda")
    act.calculate_torch(xt, box, step=0)
    act.rebuild_device(xt.data_ptr(), box)
    # displacements of ~0.05 nm at most: the list built at x (1.5 nm) holds
    # every pair within D_MAX at x2
    x2 = x + np.random.default_rng(15).normal(0.0, 0.01, x.shape)
    xt.copy_(torch.tensor(x2, dtype=torch.float64))
    v, d, w = (t.cpu().numpy() for t in act.calculate_torch(xt, box, step=1))
    ov, od, ow, npair = O.coordination(O.Switch(defn), box, x2, n, 0)
    check(float(v[0]), d, w, ov, od, ow)
    act.close()


def test_skin_in_a_cell_too_narrow_for_it():
    """A cell wide enough for D_MAX but not for D_MAX + skin rebuilds at every
    call and stays correct; a negative skin is refused."""
    P = _P()
    L = 2.8
    n = int(33.4 * L ** 3)
    x = np.random.default_rng(4).uniform(0, L, (n, 3))
    box = np.diag([L] * 3)
    defn = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2"
    act = P.Coordination(f"c: COORDINATION GROUPA=1-{n} SWITCH={{{defn}}}")
    with pytest.raises(P.Plumed2AmdError):
        act.set_skin(-0.1)
    act.set_skin(0.5)
    rng = np.random.default_rng(5)
    for s in range(4):
        v, d, w = act.calculate(x, box, step=s)
        ov, od, ow, npair = O.coordination(O.Switch(defn), box, x, n, 0)
        check(v, d, w, ov, od, ow, tag=s)
        st = act.stats()
        assert st["path"] == 0 and st["contributing_pairs"] == npair
        x = x + rng.normal(0.0, 0.001, x.shape)
    assert act.stats()["rebuilds"] == 4
    act.close()


def test_profiling_samples_every_kth_call():
    """p2a_coord_set_profiling(ctx, k): pair-kernel events on calls 0, k, 2k, ...;
    the contributing pairs keep summing at every call."""
    P = _P()
    x, L = uniform(4000, seed=12)
    box = np.diag([L] * 3)
    act = P.Coordination("c: COORDINATION GROUPA=1-4000 SWITCH={RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2}")
    act.calculate(x, box)
    act.set_profiling(True, every=3)
    for s in range(7):
        act.calculate(x, box, step=s)
    st = act.stats()
    assert st["pair_kernel_launches"] == 3 and st["pair_kernel_ms"] > 0.0
    assert st["total_contributing_pairs"] == 7 * st["contributing_pairs"]
    with pytest.raises(P.Plumed2AmdError):
        act.set_profiling(True, every=-1)
    act.set_profiling(False)
    act.close()
