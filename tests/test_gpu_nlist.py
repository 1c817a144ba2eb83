"""GPU parity of COORDINATION with neighbour lists kept over NL_STRIDE steps,
with the cell changing between rebuilds, and of the forces and full virial of
regtest/basic/rt42.

The reference keeps the pair indices of its neighbour list from one rebuild to
the next (NeighborList::prepare, NeighborList.cpp:433-456) whatever the cell
does in between, and measures each listed pair with the current cell
(Pbc::distance).  The oracle reproduces that by evaluating the candidate pairs
of the rebuild step (nl_classic: pairs within NL_CUTOFF, NeighborList.cpp:
237-308; nl_cells: the 27-cell stencil, NeighborList.cpp:177-236) with the
positions and cell of every step.

Tolerances as in test_gpu_coordination (FP32 pair arithmetic, FP64 sums):
  CV |dCV| <= 1e-6 |CV|, forces and virial 1e-5 of their largest entry;
the contributing-pair count is exact.
"""
import numpy as np
import pytest

import oracle as O
from fixtures import path, read_forces, read_xyz
from test_gpu_coordination import check, uniform
from test_oracle_cpu import RT42_BIAS_TERMS, rt42_forces_virial

pytestmark = pytest.mark.gpu

DEFN = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2"


def _P():
    import plumed2_amd as P
    return P


def _stale_list_run(act, sw, mode, x, box_of_step, nA, nB, stride, nsteps, seed, scale_positions=True):
    """Drive `act` and the oracle over nsteps frames of a random walk whose
    cell follows box_of_step(step); atoms follow the cell affinely when
    scale_positions.  Returns the number of rebuild steps."""
    rng = np.random.default_rng(seed)
    sw_o = O.Switch(sw)
    pairs = None
    box_prev = box_of_step(0)
    n = nA + nB
    rebuild_steps = 0
    for step in range(nsteps):
        box = box_of_step(step)
        if scale_positions:
            # fractional coordinates kept: x' = x inv(H) H'
            x = x @ np.linalg.inv(box_prev) @ box
        box_prev = box
        if step % stride == 0:
            pbc = O.Pbc(box)
            pairs = (O.nl_classic(pbc, True, 1.5, x, nA, nB) if mode == "NLIST"
                     else O.nl_cells(pbc, 1.5, x, nA, nB))
            rebuild_steps += 1
        v, d, w = act.calculate(x, box, step=step)
        ov, od, ow, npair = O.coordination(sw_o, box, x, nA, nB, pairs=pairs)
        check(v, d, w, ov, od, ow, tag=(mode, step))
        assert act.stats()["contributing_pairs"] == npair, (mode, step)
        x = x + rng.normal(0, 0.005, (n, 3))
    return rebuild_steps


@pytest.mark.parametrize("mode", ["NLIST", "NLISTCELLS"])
@pytest.mark.parametrize("groups", [1, 2])
def test_stride_random_walk_fixed_cell(mode, groups):
    """NLIST (pairs within NL_CUTOFF at the rebuild) and NLISTCELLS over
    NL_STRIDE=5, one group and two groups, against the stale-list oracle."""
    P = _P()
    n = 4000
    x, L = uniform(n, seed=11 + groups)
    box = np.diag([L] * 3)
    nA, nB = (n, 0) if groups == 1 else (n // 2, n // 2)
    grp = f"GROUPA=1-{n}" if groups == 1 else f"GROUPA=1-{nA} GROUPB={nA + 1}-{n}"
    act = P.Coordination(f"c: COORDINATION {grp} SWITCH={{{DEFN}}} {mode} NL_CUTOFF=1.5 NL_STRIDE=5")
    nreb = _stale_list_run(act, DEFN, mode, x, lambda s: box, nA, nB, 5, 17, seed=groups)
    assert act.stats()["rebuilds"] == nreb


@pytest.mark.parametrize("mode", ["NLIST", "NLISTCELLS"])
def test_cell_changes_between_rebuilds(mode):
    """A cell that changes on every step (isotropic scaling of up to 5e-4, as
    a barostat does) between rebuilds: the list of the rebuild step is kept
    (no extra rebuild) and measured in the current cell."""
    P = _P()
    n = 4000
    x, L = uniform(n, seed=21)
    act = P.Coordination(f"c: COORDINATION GROUPA=1-{n} SWITCH={{{DEFN}}} {mode} NL_CUTOFF=1.5 NL_STRIDE=6")

    def box_of_step(s):
        return np.diag([L * (1.0 + 0.0005 * np.sin(0.7 * s))] * 3)

    nreb = _stale_list_run(act, DEFN, mode, x, box_of_step, n, 0, 6, 20, seed=5)
    assert act.stats()["rebuilds"] == nreb


def test_triclinic_cell_changes_between_rebuilds():
    """Anisotropic strain and shear of a triclinic cell on every step, NLIST
    over NL_STRIDE=4, two groups."""
    P = _P()
    n = 3000
    x, L = uniform(n, seed=31)
    H0 = np.array([[L, 0, 0], [0.1 * L, L, 0], [-0.05 * L, 0.08 * L, L]])
    act = P.Coordination(f"c: COORDINATION GROUPA=1-1500 GROUPB=1501-{n} SWITCH={{{DEFN}}} "
                         f"NLIST NL_CUTOFF=1.5 NL_STRIDE=4")

    def box_of_step(s):
        e = 0.0004 * np.array([[np.sin(s), 0.3 * np.cos(s), 0], [0, -np.sin(0.5 * s), 0],
                              [0.2 * np.sin(s), 0, np.cos(0.3 * s)]])
        return H0 @ (np.eye(3) + e)

    nreb = _stale_list_run(act, DEFN, "NLIST", x @ np.linalg.inv(np.diag([L] * 3)) @ H0, box_of_step, 1500, 1500,
                           4, 13, seed=6)
    assert act.stats()["rebuilds"] == nreb


def test_cell_jump_rebuilds():
    """A cell change too large for the kept images (here 3% in one step, or
    the same lattice written with another basis) rebuilds the list on that
    step; the values stay those of the reference's list for motion within
    the stride."""
    P = _P()
    n = 4000
    x, L = uniform(n, seed=23)
    act = P.Coordination(f"c: COORDINATION GROUPA=1-{n} SWITCH={{{DEFN}}} NLIST NL_CUTOFF=1.5 NL_STRIDE=6")

    def box_of_step(s):
        if s < 3:
            return np.diag([L] * 3)
        if s < 5:
            return np.diag([1.03 * L] * 3)
        # the lattice of steps 3-4 with a sheared basis (b' = b + a)
        B = np.diag([1.03 * L] * 3)
        B[1] += B[0]
        return B

    # positions stay where they are: the cell grows around them, then changes basis
    nreb = _stale_list_run(act, DEFN, "NLIST", x, box_of_step, n, 0, 6, 8, seed=8, scale_positions=False)
    # rebuilds at steps 0 and 6 (stride), at the jump (3) and at the new basis (5)
    assert act.stats()["rebuilds"] == nreb + 2


def test_rt42_forces_full_virial():
    """regtest/basic/rt42 ff.reference (plumed driver --dump-forces
    --dump-full-virial, %10.5f): forces and full virial of the bias
    1.001 c + 0.999 d + 0.001 (e - f), every 4th frame of the triclinic
    108-atom trajectory, from the GPU derivatives and virials.  Tolerance:
    the 5e-6 of the format plus 1e-5 of the largest entry."""
    P = _P()
    frames = read_xyz(path("rt42", "trajectory.xyz"))
    ff = read_forces(path("rt42", "ff.reference"))
    lines = []
    for t, (c, a) in enumerate(RT42_BIAS_TERMS):
        if isinstance(a, dict):
            kw = f"R_0={a['r0']} NN={a['nn']}" + (f" D_0={a['d0']}" if a.get("d0") else "")
            lines.append((c, f"t{t}: COORDINATION GROUPA=1-108 GROUPB=1-108 {kw}"))
        else:
            lines.append((c, f"t{t}: COORDINATION GROUPA=1-108 GROUPB=1-108 SWITCH={{{a}}}"))
    acts = [(c, P.Coordination(line)) for c, line in lines]
    for k in range(0, len(frames), 4):
        pos, box = frames[k]
        p2 = np.concatenate([pos, pos])
        F, W = rt42_forces_virial([(c, act.calculate(p2, box, step=k)) for c, act in acts])
        vir, frc = ff[k]
        np.testing.assert_allclose(F, frc, atol=5.01e-6 + 1e-5 * np.max(np.abs(frc)), rtol=0,
                                   err_msg=f"forces frame {k}")
        np.testing.assert_allclose(W, vir, atol=5.01e-6 + 1e-5 * np.max(np.abs(vir)), rtol=0,
                                   err_msg=f"virial frame {k}")


def test_rt42_cells_nlist_equals_nlistcells():
    """regtest/basic/rt42-cells: NLIST and NLISTCELLS with D_MAX=1.5 <
    NL_CUTOFF=2.0, NL_STRIDE=4, GROUPA=GROUPB: the same value, derivatives
    and virial on every frame (check_diff = 0 and zero forces in
    ff.reference), each equal to the stale-list oracle."""
    P = _P()
    frames = read_xyz(path("rt42-cells", "trajectory.xyz"))
    ff = read_forces(path("rt42-cells", "ff.reference"))
    sw = "RATIONAL R_0=1 D_MAX=1.5"
    c = P.Coordination(f"c: COORDINATION GROUPA=1-108 GROUPB=1-108 SWITCH={{{sw}}} NLIST NL_STRIDE=4 NL_CUTOFF=2.0")
    d = P.Coordination(f"d: COORDINATION GROUPA=1-108 GROUPB=1-108 SWITCH={{{sw}}} NLISTCELLS NL_STRIDE=4 "
                       f"NL_CUTOFF=2.0")
    absidx = np.concatenate([np.arange(108), np.arange(108)])
    sw_o = O.Switch(sw)
    pc = pk = None
    for k, (pos, box) in enumerate(frames):
        p2 = np.concatenate([pos, pos])
        if k % 4 == 0:
            pbc = O.Pbc(box)
            pc = O.nl_classic(pbc, True, 2.0, p2, 108, 108)
            pk = O.nl_cells(pbc, 2.0, p2, 108, 108)
        vc, dc, wc = c.calculate(p2, box, step=k)
        vd, dd, wd = d.calculate(p2, box, step=k)
        ov, od, ow, _ = O.coordination(sw_o, box, p2, 108, 108, absidx=absidx, pairs=pc)
        check(vc, dc, wc, ov, od, ow, tag=("NLIST", k))
        ov, od, ow, _ = O.coordination(sw_o, box, p2, 108, 108, absidx=absidx, pairs=pk)
        check(vd, dd, wd, ov, od, ow, tag=("NLISTCELLS", k))
        assert abs(vd - vc) <= 5.01e-5
        F = -0.001 * ((dd[:108] + dd[108:]) - (dc[:108] + dc[108:]))
        vir, frc = ff[k]
        np.testing.assert_allclose(F, frc, atol=5.01e-6, rtol=0)
        np.testing.assert_allclose(-0.001 * (wd - wc), vir, atol=5.01e-6, rtol=0)


@pytest.mark.parametrize("form", ["NOPBC", "PAIR", "NARROW"])
def test_fp64_path_lists_over_stride(form):
    """The FP64 path (no cell, PAIR, or a cell narrower than twice NL_CUTOFF)
    keeps the reference's list over the stride too: large steps (0.08 nm per
    frame, 0.4 nm for the 27-cell stencil, which reaches further) so that the
    stale list and a fresh one differ."""
    P = _P()
    rng = np.random.default_rng(41)
    sw = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.0"
    sw_o = O.Switch(sw)
    if form == "NARROW":
        # a slab: 2.2 nm is below twice the list radius, 5 x 5 x 1 link cells
        n = 300
        x = rng.uniform(0, 1, (n, 3)) * [6.0, 6.0, 2.2]
        box = np.diag([6.0, 6.0, 2.2])
        act = P.Coordination(f"c: COORDINATION GROUPA=1-{n} SWITCH={{{sw}}} NLISTCELLS NL_CUTOFF=1.2 NL_STRIDE=3")
        nA, nB, use_pbc, dopair = n, 0, True, False
    elif form == "NOPBC":
        n = 400
        x = rng.uniform(0, 3.0, (n, 3))
        box = np.diag([3.0, 3.0, 3.0])
        act = P.Coordination(f"c: COORDINATION GROUPA=1-200 GROUPB=201-{n} SWITCH={{{sw}}} NOPBC NLIST "
                             f"NL_CUTOFF=1.2 NL_STRIDE=3")
        nA, nB, use_pbc, dopair = 200, 200, False, False
    else:
        n = 400
        x = rng.uniform(0, 3.0, (n, 3))
        box = np.diag([3.0, 3.0, 3.0])
        act = P.Coordination(f"c: COORDINATION GROUPA=1-200 GROUPB=201-{n} PAIR SWITCH={{{sw}}} NLIST "
                             f"NL_CUTOFF=1.2 NL_STRIDE=3")
        nA, nB, use_pbc, dopair = 200, 200, True, True
    pairs = None
    differs = 0
    for step in range(10):
        pbc = O.Pbc(box)
        if step % 3 == 0:
            if form == "NARROW":
                pairs = O.nl_cells(pbc, 1.2, x, nA, nB)
            else:
                pairs = O.nl_classic(pbc, use_pbc, 1.2, x, nA, nB, dopair=dopair)
        v, d, w = act.calculate(x, box, step=step)
        ov, od, ow, npair = O.coordination(sw_o, box, x, nA, nB, use_pbc=use_pbc, dopair=dopair, pairs=pairs)
        check(v, d, w, ov, od, ow, cv_tol=1e-10, f_tol=1e-10, w_tol=1e-10, tag=(form, step))
        assert act.stats()["contributing_pairs"] == npair
        fresh = O.coordination(sw_o, box, x, nA, nB, use_pbc=use_pbc, dopair=dopair)[0]
        differs += abs(fresh - ov) > 1e-8 * abs(ov)
        x = x + rng.normal(0, 0.4 if form == "NARROW" else 0.08, x.shape)
    assert differs > 0  # the case does separate stale from fresh lists
    assert act.stats()["rebuilds"] == 4
