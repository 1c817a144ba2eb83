"""COORDINATIONNUMBER values on the GPU (p2a_coordination_numbers_device)
against the oracle and the reference's symfunc goldens
(regtest/symfunc/rt-single-groupb, regtest/symfunc/rt20)."""
import numpy as np
import pytest

import oracle as O
from fixtures import path, read_colvar, read_xyz

pytestmark = pytest.mark.gpu

GOLD_TOL = 5.1e-5  # goldens printed with %8.4f


def _run(sw, pos, box, **kw):
    import torch

    import plumed2_amd as P

    t = torch.tensor(np.ascontiguousarray(pos), dtype=torch.float64, device="cuda:0")
    cn = P.coordination_numbers_torch(sw, t, box, **kw)
    torch.cuda.synchronize()
    return cn.cpu().numpy()


def test_symfunc_goldens_through_gpu():
    frames = read_xyz(path("symfunc-rt-single-groupb", "trajectory.xyz"))
    _, ref1 = read_colvar(path("symfunc-rt-single-groupb", "colvar.reference"))
    _, ref2 = read_colvar(path("symfunc-rt20", "colv2.reference"))
    more = O.Switch("EXP R_0=4.0")
    for (pos, box), r1, r2 in zip(frames, ref1, ref2):
        cn1 = _run("RATIONAL R_0=1.0", pos, box, species_a=range(5), species_b=[5])
        assert np.max(np.abs(cn1 - r1[1:])) <= GOLD_TOL
        cn2 = _run("RATIONAL R_0=0.5", pos, box, species=range(5))
        total = sum(1.0 - more.dist(float(x))[0][0] for x in cn2)
        assert abs(total - r2[1]) <= GOLD_TOL


@pytest.mark.parametrize("defn", ["RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.0", "RATIONAL R_0=0.25 D_0=0.05 NN=5 MM=9",
                                  "EXP R_0=0.2 D_MAX=1.2", "GAUSSIAN R_0=0.3 D_0=0.1"])
@pytest.mark.parametrize("box", [np.diag([3.1, 2.9, 3.3]), np.array([[3.0, 0, 0], [0.9, 2.8, 0], [-0.6, 0.7, 3.2]]),
                                 None])
def test_coordination_numbers_match_oracle(defn, box):
    rng = np.random.default_rng(11)
    pos = rng.uniform(-0.5, 3.5, size=(700, 3))
    sw = O.Switch(defn)
    obox = box if box is not None else np.diag([1.0, 1.0, 1.0])
    species = rng.permutation(700)[:400]
    got = _run(defn, pos, box, species=species)
    want = O.coordination_numbers(sw, obox, pos, species=species, use_pbc=box is not None)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-13)
    a, b = species[:150], species[100:]
    got = _run(defn, pos, box, species_a=a, species_b=b)
    want = O.coordination_numbers(sw, obox, pos, species_a=a, species_b=b, use_pbc=box is not None)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-13)


def test_coincident_atoms_and_tiny_values():
    """Two distinct atoms at one place are not in contact, and contacts whose
    switch value lies below DBL_EPSILON count nothing (ContactMatrix.h:95-121)."""
    box = np.diag([30.0, 30.0, 30.0])
    pos = np.array([[1.0, 1.0, 1.0], [1.0, 1.0, 1.0], [1.2, 1.0, 1.0], [1.0 + 0.2 * 46.0, 1.0, 1.0]])
    for defn in ("RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=0.9", "EXP R_0=0.2"):
        got = _run(defn, pos, box, species=range(4))
        want = O.coordination_numbers(O.Switch(defn), box, pos, species=range(4))
        assert np.allclose(got, want, rtol=1e-12, atol=0.0), (defn, got, want)
    s02 = O.Switch("RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=0.9").dist(0.2)[0][0]
    got = _run("RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=0.9", pos, box, species=range(4))
    np.testing.assert_allclose(got, [s02, s02, 2 * s02, 0.0], rtol=1e-12)


def test_coordination_numbers_input_errors():
    import torch

    import plumed2_amd as P

    t = torch.zeros((4, 3), dtype=torch.float64, device="cuda:0")
    with pytest.raises(ValueError):
        P.coordination_numbers_torch("RATIONAL R_0=0.3", t, np.eye(3), species=[0, 4])
    with pytest.raises(ValueError):
        P.coordination_numbers_torch("RATIONAL R_0=0.3", t, np.eye(3))
    assert P.coordination_numbers_torch("RATIONAL R_0=0.3", t, np.eye(3), species=[]).numel() == 0
