"""Differentiable COORDINATION: torch.autograd.Function over the value-only
forward kernel and the recomputing backward kernel.

Checks: forward value equals the full kernel's; gradient against the FP64
oracle derivatives (max |dF| <= 1e-5 max |F|); central finite differences
(h = 1e-5 nm, FP64) on 256 seeded atoms, taken on the local sum of the pairs
that touch the perturbed atom so that the difference does not cancel against
a CV of ~2e5; the full two-group size (10,000 x 490,000 atoms at 100 /nm^3)
with float32 positions.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
SWITCH = "RATIONAL R_0=0.35 NN=8 MM=16 D_MAX=1.0"


def system(na, nb, rho, seed):
    n = na + nb
    L = (n / rho) ** (1 / 3)
    return np.random.default_rng(seed).uniform(0, L, (n, 3)), L


def local_fd(sw, x, i, partners, L, h=1e-5):
    """d/dx_i of sum_j s(|x_j - x_i|) over the partners of i (minimum image)."""
    out = np.zeros(3)
    for k in range(3):
        vals = []
        for sgn in (1, -1):
            xi = x[i].copy()
            xi[k] += sgn * h
            d = x[partners] - xi
            d -= L * np.round(d / L)
            v, _ = sw.sqr((d * d).sum(1))
            vals.append(v.sum())
        out[k] = (vals[0] - vals[1]) / (2 * h)
    return out


def fd_check(P, torch, na, nb, rho, seed, dtype):
    x, L = system(na, nb, rho, seed)
    box = np.diag([L] * 3)
    line = f"c: COORDINATION GROUPA=1-{na} GROUPB={na + 1}-{na + nb} SWITCH={{{SWITCH}}}"
    act = P.Coordination(line)
    xt = torch.tensor(x, dtype=dtype, device="cuda", requires_grad=True)
    x64 = xt.detach().double().cpu().numpy()  # the exact positions the GPU sees
    val = P.autograd.coordination(xt, act, box)
    val.backward()
    grad = xt.grad.double().cpu().numpy()
    assert xt.grad.dtype == dtype
    # forward value equals the value of the full kernel
    v_full, d_full, _ = act.calculate(x64, box, step=1)
    assert abs(float(val.item()) - v_full) <= (1e-12 if dtype == torch.float64 else 1e-6) * abs(v_full)
    if dtype == torch.float64:
        np.testing.assert_allclose(grad, d_full, rtol=0, atol=1e-12 * np.max(np.abs(d_full)))
    # central finite differences on 128 atoms of each group
    sw = O.Switch(SWITCH)
    rng = np.random.default_rng(seed + 1)
    picks = np.concatenate([rng.choice(na, 128, replace=False), na + rng.choice(nb, 128, replace=False)])
    fmax = np.max(np.abs(grad))
    h, dmax = 1e-5, sw.dmax
    checked = 0
    for i in picks:
        other = np.arange(na, na + nb) if i < na else np.arange(na)
        d = x64[other] - x64[i]
        d -= L * np.round(d / L)
        r = np.sqrt((d * d).sum(1))
        # the switch's derivative jumps at D_MAX: a central difference that
        # straddles it for one partner is not a derivative, skip that atom
        if np.any(np.abs(r - dmax) < 3 * h):
            continue
        near = other[r <= 1.1]
        fd = local_fd(sw, x64, i, near, L, h)
        assert np.max(np.abs(fd - grad[i])) <= 1e-5 * fmax + 1e-5 * np.max(np.abs(fd)), (i, fd, grad[i])
        checked += 1
    assert checked >= 240, checked
    return act, x64, box, grad, float(val.item())


def test_autograd_small_float64():
    import torch

    import plumed2_amd as P
    import plumed2_amd.autograd  # noqa: F401
    act, x64, box, grad, val = fd_check(P, torch, 1500, 12000, 100.0, 3, torch.float64)
    ov, od, _, _ = O.coordination(O.Switch(SWITCH), box, x64, 1500, 12000)
    assert abs(val - ov) <= 1e-6 * abs(ov)
    assert np.max(np.abs(grad - od)) <= 1e-5 * np.max(np.abs(od))


def test_autograd_full_size_float32():
    """GROUPA=10,000, GROUPB=490,000 at 100 atoms/nm^3 (L = 17.0998 nm)."""
    import torch

    import plumed2_amd as P
    import plumed2_amd.autograd  # noqa: F401
    act, x64, box, grad, val = fd_check(P, torch, 10000, 490000, 100.0, 3, torch.float32)
    pairs = O.nl_cells(O.Pbc(box), 1.0, x64, 10000, 490000)
    ov, od, _, _ = O.coordination(O.Switch(SWITCH), box, x64, 10000, 490000, pairs=pairs)
    # float32 output of a CV of ~2e5 carries ~1e-7 relative rounding
    assert abs(val - ov) <= 1e-6 * abs(ov)
    assert np.max(np.abs(grad - od)) <= 1e-5 * np.max(np.abs(od))


def test_module_training_step():
    """One SGD step on the positions: the change of the CV equals the line
    integral of the gradient along the step (trapezoid rule, exact to third
    order in the step), which ties the backward to the forward values."""
    import torch

    from plumed2_amd.autograd import CoordinationModule
    x, L = system(400, 3000, 100.0, 5)
    mod = CoordinationModule(f"c: COORDINATION GROUPA=1-400 GROUPB=401-3400 SWITCH={{{SWITCH}}}", np.diag([L] * 3))
    x0 = torch.tensor(x, dtype=torch.float64, device="cuda", requires_grad=True)
    v0 = mod(x0)
    v0.backward()
    g0 = x0.grad.detach().clone()
    lr = 1e-5
    x1 = (x0.detach() - lr * g0).requires_grad_(True)
    v1 = mod(x1)
    v1.backward()
    g1 = x1.grad.detach()
    dx = (x1.detach() - x0.detach())
    dv = float(v1.detach()) - float(v0.detach())
    trap = float((0.5 * (g0 + g1) * dx).sum())
    assert dv < 0.0
    assert abs(dv - trap) <= 1e-3 * abs(dv), (dv, trap)


@pytest.mark.parametrize("skin,sigma", [(0.0, 0.05), (0.1, 0.05), (0.1, 0.005)])
def test_two_forwards_before_one_backward(skin, sigma):
    """loss = m(x1) + m(x2), then one backward.  Each forward keeps the
    derivatives at its own positions, whatever the second forward does to the
    pair structure; each gradient is checked against the oracle derivatives
    at its own positions (sigma 0.05 nm moves pairs across D_MAX between x1
    and x2)."""
    import torch

    import plumed2_amd as P
    from plumed2_amd.autograd import coordination
    na, nb = 600, 5000
    x1, L = system(na, nb, 100.0, 7)
    x2 = x1 + np.random.default_rng(8).normal(0.0, sigma, x1.shape)
    box = np.diag([L] * 3)
    act = P.Coordination(f"c: COORDINATION GROUPA=1-{na} GROUPB={na + 1}-{na + nb} SWITCH={{{SWITCH}}}")
    act.set_skin(skin)
    t1 = torch.tensor(x1, dtype=torch.float64, device="cuda", requires_grad=True)
    t2 = torch.tensor(x2, dtype=torch.float64, device="cuda", requires_grad=True)
    g_before = act.list_generation()
    loss = coordination(t1, act, box, 0) + coordination(t2, act, box, 1)
    builds = act.list_generation() - g_before
    loss.backward()
    assert builds == (1 if (skin > 0 and sigma < 0.01) else 2), builds
    sw = O.Switch(SWITCH)
    for t, xx in ((t1, x1), (t2, x2)):
        _, od, _, _ = O.coordination(sw, box, xx, na, nb)
        g = t.grad.cpu().numpy()
        assert np.max(np.abs(g - od)) <= 1e-5 * np.max(np.abs(od))


def test_forward_without_gradient_uses_value_only_pass():
    """Under torch.no_grad (and for positions that need no gradient) the
    forward takes the value-only kernels; the value equals the one of the
    forward that also forms the derivatives."""
    import torch

    import plumed2_amd as P
    from plumed2_amd.autograd import coordination
    na, nb = 500, 4000
    x, L = system(na, nb, 100.0, 11)
    box = np.diag([L] * 3)
    act = P.Coordination(f"c: COORDINATION GROUPA=1-{na} GROUPB={na + 1}-{na + nb} SWITCH={{{SWITCH}}}")
    t = torch.tensor(x, dtype=torch.float64, device="cuda", requires_grad=True)
    v_grad = coordination(t, act, box, 0)
    with torch.no_grad():
        v_nograd = coordination(t, act, box, 1)
    v_plain = coordination(t.detach(), act, box, 2)
    assert v_nograd.requires_grad is False and v_plain.requires_grad is False
    ov = O.coordination(O.Switch(SWITCH), box, x, na, nb)[0]
    for v in (v_grad, v_nograd, v_plain):
        assert abs(float(v.detach()) - ov) <= 1e-6 * abs(ov)
