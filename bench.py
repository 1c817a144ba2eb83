"""Benchmark: COORDINATION value + derivatives + virial on one MI355X.

Workload (one step = one frame of the hot path):
  single-group COORDINATION of 50,000 atoms, uniform random at 33.4 atoms/nm^3
  (cubic periodic box, L = (50000/33.4)^(1/3) = 11.4395 nm), SWITCH={RATIONAL
  R_0=0.3 NN=6 MM=12 D_MAX=1.2}, NLISTCELLS NL_CUTOFF=1.5 NL_STRIDE=10; each
  frame moves every atom by a Gaussian step of sigma = 0.005 nm.  All frames are
  generated on the GPU before the timed region, so positions are resident in
  HBM; the neighbour-list rebuild every 10 steps is inside the timed region.

Metric: contributing pair evaluations per second, where a contributing pair is
an unordered pair within D_MAX at that frame (counted exactly on the GPU).

The CPU baseline times the reference's own NeighborList (cells) build and the
CoordinationBase pair loop compiled from the reference sources (oracle/_ref),
with OpenMP, on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_ATOMS = 50_000
DENSITY = 33.4
SIGMA = 0.005
NL_CUTOFF, NL_STRIDE = 1.5, 10
SWITCH = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2"
FLOP_PER_PAIR = 49  # algorithmic FP32 work per contributing pair, NN=6/MM=12
PROFILE_EVERY = 8   # pair-kernel HIP events on every 8th timed step
PEAK_FP32_TFLOPS = 157.3
METRIC = "COORDINATION pair-evals/s (value+forces+virial) @50k atoms; frac FP32 VALU peak"


def measured_traffic():
    """HBM bytes per launch of the pair kernel, from the committed counter runs
    (FETCH_SIZE / WRITE_SIZE scaled by a plain copy of known size)."""
    path = os.path.join(ROOT, "profiles", "k_pairs_coordination_50k_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["traffic_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def box_length(n=N_ATOMS, rho=DENSITY):
    return (n / rho) ** (1.0 / 3.0)


def initial_positions(n=N_ATOMS, seed=2):
    L = box_length(n)
    return np.random.default_rng(seed).uniform(0.0, L, (n, 3)), L


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    """Cores this process may run on (the affinity mask, not the machine)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(x0, L, nthreads, budget_s=25.0):
    """Reference NL build + OpenMP pair loop on the host (bounded sample).

    One rebuild as the reference's CoordinationBase does it at an NL_STRIDE
    step (NeighborList::update, cells branch, then getReducedAtomList, both
    compiled from the reference's sources), then the pair loop on successive
    frames of the same random walk with `nthreads` OpenMP threads, and a few
    loops on one thread.  The per-frame cost is the rebuild spread over
    NL_STRIDE frames plus one loop."""
    import oracle

    if oracle.ref() is None:
        # compiled reference subset absent: fall back to the FP64 port
        sw = oracle.Switch(SWITCH)
        t0 = time.perf_counter()
        pairs = oracle.nl_cells(oracle.Pbc(np.diag([L] * 3)), NL_CUTOFF, x0, len(x0), 0)
        t1 = time.perf_counter()
        _, _, _, npair = oracle.coordination(sw, np.diag([L] * 3), x0, len(x0), 0, pairs=pairs)
        t2 = time.perf_counter()
        per_frame = ((t1 - t0) + NL_STRIDE * (t2 - t1)) / NL_STRIDE
        return {"value": npair / per_frame, "unit": "pair-evals/s", "cores": 1, "kind": "port",
                "cpu_model": cpu_model(),
                "sample": f"1 NL build ({t1 - t0:.1f} s) + 1 evaluation ({t2 - t1:.2f} s) of frame 0, "
                          f"amortised over NL_STRIDE={NL_STRIDE}"}
    sw = oracle.RefSwitch(SWITCH)
    rc = oracle.RefCoordination(sw, np.diag([L] * 3), len(x0), 0, nlmode=2, nl_cut=NL_CUTOFF)
    t0 = time.perf_counter()
    rc.update(x0)
    t_update = time.perf_counter() - t0
    t0 = time.perf_counter()
    nreq = rc.reduce()
    t_reduce = time.perf_counter() - t0
    # every atom has a neighbour here, so the reduced indices are the full ones
    assert nreq == len(x0)
    t_build = t_update + t_reduce
    rng = np.random.default_rng(3)
    x = x0.copy()
    calc_times = []
    npair = 0
    val = None
    while True:
        t0 = time.perf_counter()
        val, _, _, npair = rc.calc(x, nthreads)
        calc_times.append(time.perf_counter() - t0)
        x = x + rng.normal(0.0, SIGMA, x.shape)
        if len(calc_times) >= NL_STRIDE or t_build + sum(calc_times) > budget_s:
            break
    one = []
    for _ in range(3):
        t0 = time.perf_counter()
        rc.calc(x, 1)
        one.append(time.perf_counter() - t0)
    t_calc = float(np.median(calc_times))
    t_one = float(np.median(one))
    per_frame = (t_build + NL_STRIDE * t_calc) / NL_STRIDE
    per_frame_1 = (t_build + NL_STRIDE * t_one) / NL_STRIDE
    return {"value": npair / per_frame, "unit": "pair-evals/s", "cores": nthreads, "kind": "reference",
            "cpu_model": cpu_model(),
            "value_1_thread": npair / per_frame_1,
            "sample": (f"frame 0 of the same workload: reference rebuild {t_build:.2f} s (serial: "
                       f"NeighborList::update cells {t_update:.2f} s + getReducedAtomList {t_reduce:.2f} s) + "
                       f"median of {len(calc_times)} pair loops {t_calc:.3f} s ({nthreads} OpenMP threads; "
                       f"median of 3 on 1 thread {t_one:.3f} s), amortised over NL_STRIDE={NL_STRIDE}; "
                       f"steady-frame rate {npair / t_calc:.3e} pair-evals/s. The pair loop follows "
                       f"CoordinationBase::calculate over the reference's NeighborList, Pbc and "
                       f"SwitchingFunction compiled from its sources"),
            "cv_frame0": val}


def _events(torch):
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def bench_autograd(P, torch, steps=20, warm=2, skin=0.1):
    """Differentiable COORDINATION: GROUPA=10,000, GROUPB=490,000 at 100 /nm^3,
    RATIONAL R_0=0.35 NN=8 MM=16 D_MAX=1.0, float32 positions on a random walk
    (sigma 0.005 nm per step, frames generated on the GPU before timing); one
    step = the forward on the next frame (value and derivatives in one pass)
    and its backward (grad x the kept derivatives).  Timed twice: with the
    pair list kept while no atom has moved more than skin/2, and rebuilt at
    every call."""
    from plumed2_amd.autograd import coordination

    na, nb = 10_000, 490_000
    n = na + nb
    L = (n / 100.0) ** (1.0 / 3.0)
    box = np.diag([L] * 3)
    line = (f"c: COORDINATION GROUPA=1-{na} GROUPB={na + 1}-{n} "
            "SWITCH={RATIONAL R_0=0.35 NN=8 MM=16 D_MAX=1.0}")
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    x = torch.rand((n, 3), generator=g, device="cuda", dtype=torch.float64) * L
    frames = []
    for _ in range(warm + steps):
        frames.append(x.to(torch.float32).requires_grad_(True))
        x = x + 0.005 * torch.randn((n, 3), generator=g, device="cuda", dtype=torch.float64)
    torch.cuda.synchronize()

    def run(sk):
        act = P.Coordination(line)
        act.set_skin(sk)
        for s in range(warm):
            coordination(frames[s], act, box, s).backward()
        torch.cuda.synchronize()
        r0 = act.stats()["rebuilds"]
        pairs = 0
        a, b = _events(torch)
        a.record()
        for s in range(warm, warm + steps):
            coordination(frames[s], act, box, s).backward()
        b.record()
        torch.cuda.synchronize()
        st = act.stats()
        act.close()
        return a.elapsed_time(b) / steps, st["contributing_pairs"], st["rebuilds"] - r0

    ms_skin, pairs, rb_skin = run(skin)
    ms_every, _, rb_every = run(0.0)
    return {"workload": f"GROUPA={na} GROUPB={nb}, rho=100/nm^3, L={L:.4f} nm, RATIONAL R_0=0.35 NN=8 MM=16 "
                        "D_MAX=1.0, float32 positions on a random walk (0.005 nm/step), torch.autograd "
                        "forward+backward",
            "ms_per_step": ms_skin, "skin_nm": skin, "rebuilds_in_timed_steps": rb_skin,
            "ms_per_step_rebuild_every_call": ms_every, "rebuilds_every_call": rb_every,
            "steps": steps, "contributing_pairs_last_step": pairs,
            "pair_evals_per_s": pairs / (ms_skin / 1e3)}


def bench_metad(P, torch, nhills=10_000, steps=2000):
    """METAD on 2 CVs, 512x512 grid over [0,5000]x[0,6], SIGMA=20,0.05,
    HEIGHT=1.2, BIASFACTOR=10, TEMP=300: 10,000 well-tempered hills at
    random-walk centres (seed 5) deposited in one launch, then the per-step
    bias+force+deposit kernel (PACE=500) on a device-resident CV."""
    import oracle

    line = ("m: METAD ARG=c,d SIGMA=20,0.05 HEIGHT=1.2 PACE=500 BIASFACTOR=10 TEMP=300 "
            "GRID_MIN=0,0 GRID_MAX=5000,6 GRID_BIN=512,512")
    rng = np.random.default_rng(5)
    c = np.array([2500.0, 3.0])
    centers = np.empty((nhills, 2))
    for k in range(nhills):
        c = np.clip(c + rng.normal(0.0, [20.0, 0.05]), [0.0, 0.0], [4999.999, 5.999999])
        centers[k] = c
    m = P.MetaD(line)
    m.hills_path = None
    cdev = torch.tensor(centers, dtype=torch.float64, device="cuda")
    hdev = torch.empty(nhills, dtype=torch.float64, device="cuda")
    import ctypes as C
    from plumed2_amd import _capi as A
    m.step_device  # noqa: B018  (binds the METAD symbols)
    A.check(A.lib().p2a_metad_set_stream(m._ctx, C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    a, b = _events(torch)
    torch.cuda.synchronize()
    a.record()
    A.check(A.lib().p2a_metad_deposit_device(m._ctx, nhills, C.c_void_p(cdev.data_ptr()), C.c_void_p(hdev.data_ptr())))
    b.record()
    torch.cuda.synchronize()
    m.check()
    dep_ms = a.elapsed_time(b)
    support = int(np.prod([2 * math.ceil(math.sqrt(12.5) * s / d) + 1 for s, d in ((20.0, 5000 / 512), (0.05, 6 / 512))]))
    # per-step bias + force (+ a hill every PACE) on a device CV
    bias = torch.empty(1, dtype=torch.float64, device="cuda")
    force = torch.empty(2, dtype=torch.float64, device="cuda")
    walk = torch.tensor(centers[np.arange(steps) % nhills], dtype=torch.float64, device="cuda")
    a.record()
    for s in range(steps):
        m.step_device(s, walk[s].data_ptr(), bias.data_ptr(), force.data_ptr())
    b.record()
    torch.cuda.synchronize()
    m.check()
    step_us = a.elapsed_time(b) * 1e3 / steps
    # FP64 host port of the same deposition, bounded sample
    g = oracle.Grid([0.0, 0.0], [5000.0, 6.0], [512, 512], [0, 0])
    t0 = time.perf_counter()
    nref = 1000
    g.deposit_wt(centers[:nref], [20.0, 0.05], 1.2, 10.0, 300 * 0.0083144621)
    cpu_s = (time.perf_counter() - t0) / nref
    m.close()
    return {"workload": "METAD 2 CVs, 512x512 grid on [0,5000]x[0,6], SIGMA=20,0.05, HEIGHT=1.2, BIASFACTOR=10, "
                        f"{nhills} well-tempered hills (random walk, seed 5)",
            "deposit_ms_total": dep_ms, "hills_per_s": nhills / (dep_ms / 1e3),
            "grid_node_updates_per_s": nhills * support / (dep_ms / 1e3), "support_nodes_per_hill": support,
            "step_us_bias_force_deposit": step_us,
            "cpu_port_hills_per_s": 1.0 / cpu_s, "cpu_port_sample": f"{nref} hills, FP64 oracle, 1 core"}


def bench_coord_100k(P, torch, steps=200, warm=20):
    """The COORDINATION CV of the METAD workload: 100,000 atoms as the 50k case."""
    n = 100_000
    L = (n / DENSITY) ** (1.0 / 3.0)
    box = np.diag([L] * 3)
    x0 = np.random.default_rng(5).uniform(0.0, L, (n, 3))
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    frames = torch.from_numpy(x0).to(dev) + torch.cumsum(
        torch.randn((warm + steps, n, 3), dtype=torch.float64, device=dev, generator=g) * SIGMA, dim=0)
    act = P.Coordination(f"c: COORDINATION GROUPA=1-{n} SWITCH={{{SWITCH}}} NLISTCELLS NL_CUTOFF={NL_CUTOFF} "
                         f"NL_STRIDE={NL_STRIDE}")
    out = (torch.empty(1, dtype=torch.float64, device=dev), torch.empty((n, 3), dtype=torch.float64, device=dev),
           torch.empty((3, 3), dtype=torch.float64, device=dev))
    for s in range(warm):
        act.calculate_torch(frames[s].contiguous(), box, s, out)
    torch.cuda.synchronize()
    act.set_profiling(True, every=PROFILE_EVERY)
    a, b = _events(torch)
    a.record()
    for s in range(warm, warm + steps):
        act.calculate_torch(frames[s].contiguous(), box, s, out)
    b.record()
    torch.cuda.synchronize()
    st = act.stats()
    ms = a.elapsed_time(b) / steps
    act.close()
    return {"workload": f"single-group COORDINATION, {n} atoms, rho={DENSITY}, L={L:.4f} nm, as the 50k case",
            "ms_per_step": ms, "pair_evals_per_s": st["total_contributing_pairs"] / (ms * steps / 1e3)}


def bench_metad_chain(P, torch, steps=200, warm=20):
    """The METAD workload as one coupled step on one stream: COORDINATION of
    100,000 atoms (as the 50k case, NLISTCELLS NL_CUTOFF=1.5 NL_STRIDE=10),
    DISTANCE between the centres of mass of two 500-atom groups, METAD on both
    (512x512 bins, SIGMA=20,0.05, HEIGHT=1.2, PACE=500, BIASFACTOR=10,
    TEMP=300), then the METAD forces on every atom and on the cell.  The grid
    spans [288000, 293000] x [0, 6]: the coordination of this system is
    ~2.9e5, outside GRID_MAX=5000.  Frames follow a random walk (sigma 0.005
    nm) generated on the GPU before timing."""
    n = 100_000
    L = (n / DENSITY) ** (1.0 / 3.0)
    box = np.diag([L] * 3)
    x0 = np.random.default_rng(5).uniform(0.0, L, (n, 3))

    def nearest(point, k, exclude=()):
        d = x0 - point
        d -= L * np.round(d / L)
        r = (d * d).sum(1)
        r[list(exclude)] = np.inf
        return np.argsort(r, kind="stable")[:k]

    g1 = nearest(np.array([3.0, 3.0, 3.0]), 500)
    g2 = nearest(np.array([5.5, 4.5, 4.0]), 500, exclude=g1)
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    frames = torch.from_numpy(x0).to(dev) + torch.cumsum(
        torch.randn((warm + steps, n, 3), dtype=torch.float64, device=dev, generator=g) * SIGMA, dim=0)
    coord = P.Coordination(f"c: COORDINATION GROUPA=1-{n} SWITCH={{{SWITCH}}} NLISTCELLS NL_CUTOFF={NL_CUTOFF} "
                           f"NL_STRIDE={NL_STRIDE}")
    cd = P.CenterDistance(n, g1, g2)
    m = P.MetaD("m: METAD ARG=c,d SIGMA=20,0.05 HEIGHT=1.2 PACE=500 BIASFACTOR=10 TEMP=300 "
                "GRID_MIN=288000,0 GRID_MAX=293000,6 GRID_BIN=512,512")
    m.hills_path = None
    chain = P.MetadStep(coord, m, cd)
    for s in range(warm):
        chain.step(s, frames[s], box)
    torch.cuda.synchronize()
    coord.set_profiling(True, every=PROFILE_EVERY)
    a, b = _events(torch)
    a.record()
    for s in range(warm, warm + steps):
        chain.step(s, frames[s], box)
    b.record()
    torch.cuda.synchronize()
    m.check()
    st = coord.stats()
    ms = a.elapsed_time(b) / steps
    coord.close()
    cd.close()
    m.close()
    return {"workload": f"METAD on COORDINATION ({n} atoms, rho={DENSITY}, L={L:.4f} nm, NLISTCELLS 1.5/10) and "
                        "DISTANCE between two 500-atom centres of mass; 512x512 grid, PACE=500, forces on atoms",
            "ms_per_step": ms, "steps": steps,
            "pair_evals_per_s": st["total_contributing_pairs"] / (ms * steps / 1e3)}


def bench_batch(P, torch, steps=10):
    """256 independent systems, 2,000-64,000 atoms each (log-uniform, seed 4),
    triclinic cells at 33.4 atoms/nm^3, SWITCH={RATIONAL R_0=0.3 NN=6 MM=12
    D_MAX=1.2}, list radius 1.2 nm rebuilt every step (NL_STRIDE=1): one step
    evaluates every system.  Autotuned decomposition, replayed as a HIP graph;
    positions take a random-walk step (sigma 0.005 nm) between steps."""
    from plumed2_amd.synthetic import batch_systems

    natoms, boxes, x = batch_systems()
    b = P.CoordinationBatch(natoms, SWITCH, 1.2)
    b.set_boxes(boxes)
    xt = torch.tensor(x, dtype=torch.float64, device="cuda")
    del x
    out = (torch.empty(len(natoms), dtype=torch.float64, device="cuda"), torch.empty_like(xt),
           torch.empty((len(natoms), 3, 3), dtype=torch.float64, device="cuda"))
    b.calculate_torch(xt, out)
    tuned = b.autotune(xt.data_ptr(), reps=2)
    b.set_graph(True)
    g = torch.Generator(device="cuda")
    g.manual_seed(4)
    for _ in range(2):
        b.calculate_torch(xt, out)
    dev_ms, wall, pairs = 0.0, 0.0, 0
    for _ in range(steps):
        xt.add_(torch.randn(xt.shape, dtype=torch.float64, device="cuda", generator=g), alpha=SIGMA)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.calculate_torch(xt, out)
        wall += time.perf_counter() - t0
        st = b.stats()
        dev_ms += st["last_step_ms"]
        pairs += st["contributing_pairs"]
    st = b.stats()
    b.close()
    return {"workload": f"{len(natoms)} systems, {int(natoms.sum())} atoms ({int(natoms.min())}-{int(natoms.max())} "
                        "per system), triclinic cells, rho=33.4/nm^3, RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2, "
                        "list radius 1.2 nm rebuilt every step, one step = every system",
            "ms_per_step": dev_ms / steps, "wall_ms_per_step": wall * 1e3 / steps,
            "contributing_pairs_per_step": pairs / steps, "pair_evals_per_s": pairs / (dev_ms / 1e3),
            "graph": bool(st["graph"]), "autotuned": {"chunk": tuned[0], "blocks_per_cu": tuned[1]},
            "clusters": st["clusters"], "list_entries": st["list_entries"], "items": st["items"]}


def run_extras(P, torch, which: str):
    runs = {"batch": ("batch_256_systems", bench_batch), "autograd": ("autograd_two_group_500k", bench_autograd),
            "coord100k": ("coordination_100k", bench_coord_100k), "metad": ("metad_grid", bench_metad),
            "chain": ("metad_coupled_step_100k", bench_metad_chain)}
    out = {}
    for key in [w.strip() for w in which.split(",") if w.strip()]:
        name, fn = runs[key]
        out[name] = fn(P, torch)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
    ap.add_argument("--no-extras", action="store_true", help="skip the autograd / METAD / 100k measurements")
    ap.add_argument("--extras", default="batch,autograd,coord100k,chain,metad",
                    help="comma list of extra measurements to run")
    ap.add_argument("--only-extras", action="store_true", help="run the extras only (no 50k main line)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 256 MiB device copy (reference size for the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "plumed2_amd runs on one GPU; runs on several GPUs are not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    if args.copy_calibration:
        n = 256 << 20
        src = torch.ones(n // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": 5}))
        return

    import plumed2_amd as P

    if args.only_extras:
        print(json.dumps(run_extras(P, torch, args.extras)))
        return

    dev = torch.device("cuda:0")
    x0, L = initial_positions()
    box = np.diag([L, L, L])
    # warmup rounded to a multiple of NL_STRIDE so the timed region starts on a rebuild step
    warm = ((args.warmup + NL_STRIDE - 1) // NL_STRIDE) * NL_STRIDE
    nframes = warm + args.steps
    g = torch.Generator(device=dev)
    g.manual_seed(2)
    frames = torch.empty((nframes, N_ATOMS, 3), dtype=torch.float64, device=dev)
    frames[0] = torch.from_numpy(x0).to(dev)
    steps = torch.randn((nframes - 1, N_ATOMS, 3), dtype=torch.float64, device=dev, generator=g) * SIGMA
    frames[1:] = frames[0] + torch.cumsum(steps, dim=0)
    del steps
    torch.cuda.synchronize()

    line = f"c: COORDINATION GROUPA=1-{N_ATOMS} SWITCH={{{SWITCH}}} NLISTCELLS NL_CUTOFF={NL_CUTOFF} NL_STRIDE={NL_STRIDE}"
    act = P.Coordination(line, device=0)
    out = (torch.empty(1, dtype=torch.float64, device=dev), torch.empty((N_ATOMS, 3), dtype=torch.float64, device=dev),
           torch.empty((3, 3), dtype=torch.float64, device=dev))
    # the timed loop calls the C entry point directly, with the pointers an
    # MD engine would pass (frame s of the device-resident trajectory)
    import ctypes as C
    from plumed2_amd import _capi as A
    act.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    fn = A.lib().p2a_coord_calculate_device
    box9 = np.ascontiguousarray(box.reshape(-1))
    bp = box9.ctypes.data
    base, stride = frames.data_ptr(), N_ATOMS * 3 * 8
    vp, dp, wp = (C.c_void_p(t.data_ptr()) for t in out)

    def step(s):
        # step s reads frame s mod nframes, always inside the frames tensor
        rc = fn(act._ctx, s, bp, C.c_void_p(base + (s % nframes) * stride), vp, dp, wp)
        if rc:
            A.check(rc)

    # autotuning inside the untimed warm-up: each decomposition of the pair
    # kernel (entries per wave x persistent grid) runs one neighbour-list
    # cycle, the fastest is kept; results do not depend on the choice
    tuned, best = (64, 0), None
    s0 = 0
    for cfg in [(64, 0), (32, 0), (64, 4), (64, 8), (32, 4), (32, 8)]:
        act.set_tuning(*cfg)
        for s in range(s0, s0 + NL_STRIDE):  # settle on the new work items
            step(s)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for s in range(s0 + NL_STRIDE, s0 + 2 * NL_STRIDE):
            step(s)
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b)
        if best is None or ms < best:
            best, tuned = ms, cfg
        s0 += 2 * NL_STRIDE
    act.set_tuning(*tuned)
    for s in range(warm):
        step(s)
    torch.cuda.synchronize()
    # the pair-kernel events are recorded on every PROFILE_EVERY-th step only:
    # each record adds a few microseconds of stream time to the step it brackets
    act.set_profiling(True, every=PROFILE_EVERY)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    for s in range(warm, warm + args.steps):
        step(s)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = start.elapsed_time(stop)
    st = act.stats()
    act.set_profiling(False)
    total_pairs = st["total_contributing_pairs"]
    elapsed = ev_ms / 1e3
    value = total_pairs / elapsed
    pairs_per_step = total_pairs / args.steps
    kern_ms = st["pair_kernel_ms"] / max(st["pair_kernel_launches"], 1)
    achieved = pairs_per_step * FLOP_PER_PAIR / (kern_ms / 1e3) / 1e12
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warm,
        "ms_per_step": ev_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"single-group COORDINATION, {N_ATOMS} atoms, rho={DENSITY}/nm^3, cubic L={L:.4f} nm, "
                        f"SWITCH={{{SWITCH}}}, NLISTCELLS NL_CUTOFF={NL_CUTOFF} NL_STRIDE={NL_STRIDE}, "
                        f"random walk sigma={SIGMA} nm/frame",
            "atoms": N_ATOMS,
            "contributing_pairs_per_step": pairs_per_step,
            "accumulation": "f32 per lane and work item, f64 across lanes, items and workgroups",
            "wall_s": wall,
            "autotuned": {"chunk": tuned[0], "blocks_per_cu": tuned[1]},
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": measured_traffic(),
            "kernel": f"k_pairs<3,false,false,false,{tuned[0]}> (FP32 pass with the FP64 cut decisions of its band pairs; value, derivatives, virial)",
            "kernel_ms": kern_ms,
            "flop_per_pair": FLOP_PER_PAIR,
        },
    }
    act.close()
    del frames
    torch.cuda.empty_cache()
    if not args.no_extras:
        result["extras"] = run_extras(P, torch, args.extras)
    if not args.no_cpu:
        nthreads = args.cpu_threads or host_cores()
        result["cpu_baseline"] = cpu_baseline(x0, L, nthreads)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
