"""GPU parity of the METAD grid (HIP library through the C ABI) against the FP64
oracle and the rt11/rt67/rt-backup-grid golden files.  The grid accumulator is FP64 on both
sides: values, gradients and heights are held to 1e-10 relative."""
import math
import os

import numpy as np
import pytest

import oracle as O
from fixtures import path, read_colvar, rt_backup_grid_cvs

pytestmark = pytest.mark.gpu
KB = 0.0083144621
TOL = 1e-10


def _P():
    import plumed2_amd as P
    return P


def random_walk(n, start, sigma, lo, hi, seed):
    rng = np.random.default_rng(seed)
    x = np.array(start, dtype=float)
    out = []
    for _ in range(n):
        x = np.clip(x + rng.normal(0, sigma), lo, hi)
        out.append(x.copy())
    return np.array(out)


def test_sequential_welltempered_deposition_matches_oracle():
    P = _P()
    m = P.MetaD("m: METAD ARG=c,d SIGMA=20,0.05 HEIGHT=1.2 PACE=500 BIASFACTOR=10 TEMP=300 "
                "GRID_MIN=0,0 GRID_MAX=5000,6 GRID_BIN=512,512", periodic={})
    centers = random_walk(400, [2500.0, 3.0], [15.0, 0.04], [0, 0], [4999.9, 5.999], 5)
    h = m.deposit(centers)
    g = O.Grid([0.0, 0.0], [5000.0, 6.0], [512, 512], [0, 0])
    oh = g.deposit_wt(centers, [20.0, 0.05], 1.2, 10.0, 300 * KB)
    np.testing.assert_allclose(h, oh, rtol=TOL)
    v, d = m.grid()
    ov, od = g.values(), g.derivs()
    assert np.max(np.abs(v - ov)) <= TOL * np.max(np.abs(ov))
    assert np.max(np.abs(d - od)) <= TOL * np.max(np.abs(od))
    pts = random_walk(300, [2500.0, 3.0], [40.0, 0.1], [0, 0], [4999.9, 5.999], 6)
    b, db = m.bias(pts)
    for k in range(0, 300, 7):
        ob, odb = g.spline(pts[k])
        assert abs(b[k] - ob) <= TOL * max(1.0, abs(ob))
        np.testing.assert_allclose(db[k], odb, rtol=1e-9, atol=1e-12)


def test_periodic_grid_and_restart_from_hills(tmp_path):
    P = _P()
    line = ("metad: METAD ARG=phi,psi SIGMA=0.2,0.2 HEIGHT=1.2 BIASFACTOR=10 PACE=500 TEMP=300 "
            f"GRID_MIN=-pi,-pi GRID_MAX=pi,pi GRID_BIN=150,150 FILE={tmp_path}/HILLS FMT=%20.12f")
    per = {"phi": ("-pi", "pi"), "psi": ("-pi", "pi")}
    m = P.MetaD(line, periodic=per)
    rng = np.random.default_rng(7)
    cvs = rng.uniform(-math.pi, math.pi, (60, 2))
    heights = []
    for k, c in enumerate(cvs):
        m.calculate(c)
        h = m.update(500 * k, c)
        if h is not None:
            heights.append(h)
    assert len(heights) == 59  # no hill on the first step
    g = O.Grid([-math.pi, -math.pi], [math.pi, math.pi], [150, 150], [1, 1])
    oh = g.deposit_wt(cvs[1:], [0.2, 0.2], 1.2, 10.0, 300 * KB)
    np.testing.assert_allclose(heights, oh, rtol=TOL)
    v, d = m.grid()
    np.testing.assert_allclose(v, g.values(), atol=TOL * np.max(np.abs(g.values())))
    m.close()
    # restart: a fresh action re-reads HILLS (heights rescaled back)
    m2 = P.MetaD(line + " RESTART=YES", periodic=per)
    v2, _ = m2.grid()
    # HILLS keeps 12 decimals of heights and centres
    assert np.max(np.abs(v2 - v)) < 1e-8 * np.max(np.abs(v))
    text = open(f"{tmp_path}/HILLS").read().splitlines()
    assert text[0] == "#! FIELDS time phi psi sigma_phi sigma_psi height biasf"
    assert text[1] == "#! SET multivariate false"
    assert text[2] == "#! SET kerneltype stretched-gaussian"
    assert "#! SET min_phi -pi" in text


def test_rt11_golden_through_gpu():
    P = _P()
    names, hills = read_colvar(path("rt11", "HILLS.reference"))
    _, colvar = read_colvar(path("rt11", "COLVAR.reference"))
    m = P.MetaD("md: METAD ARG=d1,vol SIGMA=0.1,0.2 HEIGHT=1.0 PACE=10 TEMP=300 GRID_MIN=0.0,120.0 "
                "GRID_MAX=2.0,130.0 BIASFACTOR=10 GRID_BIN=100,100 FMT=%14.6f", timestep=0.005)
    m.hills_path = None
    for k in range(len(hills)):
        c = hills[k, 1:3]
        b, _ = m.calculate(c)
        assert abs(b - colvar[k + 1, 3]) < 5.1e-4 + 1e-3 * abs(colvar[k + 1, 3])
        h = m.deposit([c])[0]
        assert abs(h * 10 / 9 - hills[k, names.index("height")]) < 5.1e-7 + 1e-7 * hills[k, -2]
    # HILLS record formatting matches the reference file for the same numbers
    rec = m.hill_record(hills[0, 0], hills[0, 1:3], hills[0, names.index("height")] * 0.9)
    ref_line = open(path("rt11", "HILLS.reference")).read().splitlines()[3]
    assert rec.split() == ref_line.split()


def test_rt67_golden_through_gpu():
    P = _P()
    names, col = read_colvar(path("rt67", "COLVAR.reference"))
    m = P.MetaD("metad: METAD ARG=phi,psi SIGMA=0.20,0.20 HEIGHT=1.20 BIASFACTOR=10 PACE=500 TEMP=300 "
                "GRID_MIN=-pi,-pi GRID_MAX=pi,pi GRID_BIN=150,150",
                periodic={"phi": ("-pi", "pi"), "psi": ("-pi", "pi")})
    m.hills_path = None
    ib = names.index("metad.bias")
    for k in range(len(col)):
        b, _ = m.calculate(col[k, 1:3])
        assert abs(b - col[k, ib]) < 2e-3
        m.update(500 * k, col[k, 1:3])


def test_device_step_matches_host_semantics():
    import torch

    P = _P()
    line = "m: METAD ARG=a,b SIGMA=0.3,0.2 HEIGHT=0.8 PACE=3 BIASFACTOR=6 TEMP=310 GRID_MIN=-2,-2 GRID_MAX=2,2 GRID_BIN=80,80"
    md, mh = P.MetaD(line), P.MetaD(line)
    md.hills_path = mh.hills_path = None
    cvs = random_walk(40, [0.0, 0.0], [0.2, 0.2], [-1.9, -1.9], [1.9, 1.9], 8)
    dev = torch.device("cuda:0")
    cv_t = torch.empty(2, dtype=torch.float64, device=dev)
    b_t = torch.empty(1, dtype=torch.float64, device=dev)
    f_t = torch.empty(2, dtype=torch.float64, device=dev)
    h_t = torch.empty(1, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    for step, c in enumerate(cvs, start=5):
        cv_t.copy_(torch.from_numpy(c))
        md.step_device(step, cv_t.data_ptr(), b_t.data_ptr(), f_t.data_ptr(), h_t.data_ptr(), stream=stream)
        torch.cuda.synchronize()
        hb, hf = mh.calculate(c)
        hh = mh.update(step, c)
        assert abs(b_t.item() - hb) <= TOL * max(1, abs(hb))
        np.testing.assert_allclose(f_t.cpu().numpy(), hf, rtol=1e-9, atol=1e-12)
        assert (h_t.item() == 0.0) == (hh is None)
        if hh is not None:
            assert abs(h_t.item() - hh) <= TOL * hh
    md.check()


def test_outside_grid_is_an_error():
    P = _P()
    m = P.MetaD("m: METAD ARG=a SIGMA=0.1 HEIGHT=1 PACE=1 GRID_MIN=0 GRID_MAX=1 GRID_BIN=10")
    m.hills_path = None
    with pytest.raises(P.Plumed2AmdError, match="outside the grid"):
        m.bias([[1.5]])
    with pytest.raises(P.Plumed2AmdError, match="outside the grid"):
        m.deposit([[-0.2]])
    # the context stays usable after the error
    b, _ = m.bias([[0.5]])
    assert b[0] == 0.0


def test_add_hills_equals_sequential_with_same_heights():
    P = _P()
    line = "m: METAD ARG=a,b SIGMA=0.3,0.2 HEIGHT=0.8 PACE=3 GRID_MIN=-2,-2 GRID_MAX=2,2 GRID_BIN=60,70"
    m1, m2 = P.MetaD(line), P.MetaD(line)
    c = random_walk(200, [0.0, 0.0], [0.2, 0.2], [-1.9, -1.9], [1.9, 1.9], 9)
    h = m1.deposit(c)  # not well-tempered: every height is HEIGHT
    assert np.all(h == 0.8)
    m2.add_hills(c, h)
    v1, d1 = m1.grid()
    v2, d2 = m2.grid()
    np.testing.assert_allclose(v1, v2, atol=1e-12 * np.max(np.abs(v1)))
    np.testing.assert_allclose(d1, d2, atol=1e-12 * np.max(np.abs(d1)))


def test_deposition_at_grid_edges_matches_oracle():
    """A walk that starts in a corner of the grid: hill supports are cut by the
    grid edges and the LDS tile of the deposition sits against them."""
    P = _P()
    m = P.MetaD("m: METAD ARG=c,d SIGMA=20,0.05 HEIGHT=1.2 PACE=500 BIASFACTOR=10 TEMP=300 "
                "GRID_MIN=0,0 GRID_MAX=5000,6 GRID_BIN=512,512", periodic={})
    centers = random_walk(600, [5.0, 0.02], [30.0, 0.08], [0, 0], [4999.9, 5.999], 9)
    centers = np.concatenate([centers, random_walk(300, [4995.0, 5.98], [30.0, 0.08], [0, 0], [4999.9, 5.999], 10)])
    h = m.deposit(centers)
    g = O.Grid([0.0, 0.0], [5000.0, 6.0], [512, 512], [0, 0])
    oh = g.deposit_wt(centers, [20.0, 0.05], 1.2, 10.0, 300 * KB)
    np.testing.assert_allclose(h, oh, rtol=TOL)
    v, d = m.grid()
    assert np.max(np.abs(v - g.values())) <= TOL * np.max(np.abs(g.values()))
    assert np.max(np.abs(d - g.derivs())) <= TOL * np.max(np.abs(g.derivs()))


def test_bulk_deposition_equals_hill_by_hill_bitwise():
    """Thousands of well-tempered hills in one call (runs of hills, the next
    run's heights found while the current run is added to the grid) against
    the same hills deposited one call per hill: heights and grid identical
    bit for bit, since every node and corner receives the same FP64
    additions in the same order.  The walk leaves the LDS tile now and then
    (reloads) and runs into the grid edge."""
    P = _P()
    line = ("m: METAD ARG=c,d SIGMA=20,0.05 HEIGHT=1.2 PACE=500 BIASFACTOR=10 TEMP=300 "
            "GRID_MIN=0,0 GRID_MAX=5000,6 GRID_BIN=512,512")
    centers = random_walk(1500, [2500.0, 3.0], [25.0, 0.06], [0, 0], [4999.9, 5.999], 21)
    centers = np.concatenate([centers, random_walk(500, [4990.0, 0.01], [25.0, 0.06], [0, 0], [4999.9, 5.999], 22)])
    m1, m2 = P.MetaD(line, periodic={}), P.MetaD(line, periodic={})
    h1 = m1.deposit(centers)
    h2 = np.concatenate([m2.deposit(c[None, :]) for c in centers])
    assert np.array_equal(h1, h2)
    v1, d1 = m1.grid()
    v2, d2 = m2.grid()
    assert np.array_equal(v1, v2)
    assert np.array_equal(d1, d2)


def test_rt_backup_grid_files_through_gpu(tmp_path):
    """regtest/basic/rt-backup-grid: the GRID_WFILE dumps of the GPU grid, the
    last one and the one it replaced, equal the reference files."""
    P = _P()
    g = tmp_path / "grid"
    m = P.MetaD(f"METAD ARG=d1,d2 SIGMA=0.1,0.1 PACE=2 HEIGHT=0.5 FMT=%6.2f GRID_MIN=0,0 GRID_MAX=2,2 "
                f"GRID_BIN=4,4 GRID_WFILE={g} GRID_WSTRIDE=2 LABEL=meta1 FILE={tmp_path / 'HILLS'}", timestep=0.05)
    for step, cv in enumerate(rt_backup_grid_cvs()):
        m.calculate(cv)
        m.update(step, cv)
    m.close()
    assert g.read_text() == open(path("rt-backup-grid", "grid.reference")).read()
    assert (tmp_path / "bck.last.grid").read_text() == open(path("rt-backup-grid", "bck.last.grid.reference")).read()


def test_grid_rfile_restart(tmp_path):
    """GRID_RFILE: a run restarted from a written grid holds the file's values
    (9 decimals) and its bias is the spline of that grid."""
    P = _P()
    line = ("m: METAD ARG=c,d SIGMA=20,0.05 HEIGHT=1.2 PACE=500 BIASFACTOR=10 TEMP=300 "
            "GRID_MIN=0,0 GRID_MAX=5000,6 GRID_BIN=128,96 ")
    m = P.MetaD(line + f"FILE={tmp_path / 'H1'}")
    m.deposit(random_walk(200, [2500.0, 3.0], [40.0, 0.1], [0, 0], [4999.9, 5.999], 21))
    m.write_grid(str(tmp_path / "g"))
    v0, d0 = m.grid()
    m.close()
    # the restart names the grid file, and a different GRID_BIN that the file overrides
    r = P.MetaD(line.replace("GRID_BIN=128,96", "GRID_BIN=10,10") + f"GRID_RFILE={tmp_path / 'g'} "
                f"FILE={tmp_path / 'H2'}")
    assert r.gbin == [128, 96]
    v1, d1 = r.grid()
    assert np.max(np.abs(v1 - v0)) <= 5.1e-10 and np.max(np.abs(d1 - d0)) <= 5.1e-10
    og = O.Grid([0.0, 0.0], [5000.0, 6.0], [128, 96], [0, 0])
    og.set(v1, d1)
    pts = np.column_stack([np.linspace(10, 4990, 37), np.linspace(0.01, 5.99, 37)])
    b, db = r.bias(pts)
    for k, x in enumerate(pts):
        ob, od = og.spline(x)
        assert abs(b[k] - ob) <= TOL * max(1.0, abs(ob))
        np.testing.assert_allclose(db[k], od, rtol=TOL, atol=TOL)
    # the restarted grid written again is the same text
    r.write_grid(str(tmp_path / "g2"))
    assert (tmp_path / "g2").read_text() == (tmp_path / "g").read_text()
    r.close()
