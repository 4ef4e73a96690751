"""Benchmark: Yee cell-steps/s of the per-tick physics on one MI355X.

The timed workload is a 512 x 512 scene (544 x 544 Yee nodes with the absorbing
margin): 64 random filled conductor rectangles (sides 4-64, resistance
log-uniform 0.1-10), 16 battery cells (8 DC, 8 AC), 16 carriers per conductor
cell (about 1M electrons), seeded. A step is one full physics tick (battery
trim, electron push with P3M close-range forces, sort, deposit + smoothing,
FDTD, DST Poisson divergence cleaning); a cell-step is one Yee node advanced by
one tick. Inputs are resident in HBM before the timed region.

Also reported on the same line: the roofline of the dominant kernel (HIP
events per kernel phase on the simulation's stream), the FP64 oracle's rate on
this host (a bounded sample, one thread), and a scaled serving stream of mixed
64^2 - 1024^2 scenes run batched (persistent batched kernels + HIP graph) and
one scene per launch.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-serve]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK_TFLOPS = 157.3
BF16_MFMA_PEAK_TFLOPS = 2500.0


def algorithmic(phase: str, nx: int, ny: int, cols: int, rows: int, n: int, sub: int = 1):
    """(amount, unit, bound) of one kernel phase over one tick (`sub` FDTD substeps)."""
    nodes, cells = nx * ny, cols * rows
    if phase == "dst_gemm":
        # Two right-hand sides, four products each, every product as its two
        # folded halves: 4 nx ny (nx + ny) flops (the dense products: twice that).
        return 4.0 * nx * ny * (nx + ny), "flop", "mfma"
    if phase == "push":
        return 32.0 * n, "byte", "hbm"      # x, y, vx, vy in and out (SoA FP32)
    if phase == "fdtd":
        # Ex, Ey, Bz, Jx, Jy in and Ex, Ey, Bz out per substep; the last substep
        # also reads rho and writes the two Gauss right-hand sides.
        return 32.0 * nodes * sub + 12.0 * nodes, "byte", "hbm"
    if phase == "deposit_smooth":
        return 16.0 * n + 12.0 * nodes + 16.0 * cells, "byte", "hbm"
    if phase in ("count", "scatter"):
        return 16.0 * n, "byte", "hbm"
    if phase == "canon":
        return 36.0 * n + 8.0 * cells, "byte", "hbm"
    if phase == "scan":
        return 12.0 * cells, "byte", "hbm"
    if phase == "correct":
        return 16.0 * nodes, "byte", "hbm"
    if phase == "finish":
        return 16.0 * cells + 40.0 * cells, "byte", "hbm"   # node E, B in; display and current averages in and out
    return 0.0, "byte", "hbm"


def fp64_poisson(rhs: np.ndarray) -> np.ndarray:
    """Maxwell::poisson (crates/embloks-core/src/maxwell.rs:139-168) in FP64
    through scipy's DST-I (scipy's type-1 DST is twice the reference's
    unnormalised one): the error reference for the isolated solve."""
    import scipy.fft as F

    n, m = rhs.shape
    kx = 2.0 * np.cos(np.pi * np.arange(1, n + 1) / (n + 1)) - 2.0
    ky = 2.0 * np.cos(np.pi * np.arange(1, m + 1) / (m + 1)) - 2.0
    inv = (4.0 / ((n + 1) * (m + 1))) / (kx[:, None] + ky[None, :])
    s = F.dst(F.dst(rhs, type=1, axis=0), type=1, axis=1) / 4.0 * inv
    return F.dst(F.dst(s, type=1, axis=0), type=1, axis=1) / 4.0


def laplacian_residual(phi: np.ndarray, rhs: np.ndarray) -> float:
    """max |lap(phi) - rhs| / max |rhs| with the 5-point Laplacian and phi = 0
    outside the grid (the operator the DST-I solve inverts, maxwell.rs:56-67)."""
    lap = -4.0 * phi
    lap[1:, :] += phi[:-1, :]
    lap[:-1, :] += phi[1:, :]
    lap[:, 1:] += phi[:, :-1]
    lap[:, :-1] += phi[:, 1:]
    return float(np.abs(lap - rhs).max() / np.abs(rhs).max())


def run_single(args):
    from embloks_amd import world as W
    from embloks_amd.gpusim import GpuSim

    world = W.rectangles_circuit(512, carriers_per_cell=16, seed=2)
    sim = GpuSim(world)
    sim.run_ticks(args.warmup)
    t0 = time.perf_counter()
    sim.step(args.steps)
    sim.sync()
    elapsed = time.perf_counter() - t0
    phases = sim.kernel_times(max(10, min(args.steps, 100)))
    residual, charge, energy = sim.diagnostics()
    info = dict(nx=sim.nx, ny=sim.ny, cols=sim.cols, rows=sim.rows, n=sim.n, substeps=sim.dims().last_substeps,
                fdtd_substeps=sim.dims().fdtd_substeps, gauss_residual=residual, electron_charge=charge)
    return world, elapsed, phases, info


# Kernels that carry each tick phase (the rocprofv3 names without the namespace).
# The push phase holds the battery trim too: its blocks run in k_push_tiled's
# launch and its gain kernel before the kick.
PHASE_KERNELS = {"push": ["k_push_tiled", "k_trim_gain", "k_push_sparse<false>", "k_push_heavy<false>"], "rest": ["k_rest"],
                 "count": ["k_count"], "scan": ["k_scan"], "scatter": ["k_scatter<true>"], "canon": ["k_canon"],
                 "deposit_smooth": ["k_deposit_smooth<512>"], "fdtd": ["k_fdtd"], "correct": ["k_correct"],
                 "finish": ["k_finish"],
                 "dst_gemm": [f"k_dst_gemm<{i}>" for i in range(4)]}


def measured_traffic(phase: str):
    """HBM bytes per launch of the phase's kernel on this workload, from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-kernel calibrated)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic_512_1M.txt")
    if not os.path.exists(path):
        return None
    table = {}
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        parts = line.split()
        if len(parts) == 3:
            table[parts[0]] = int(parts[1]) + int(parts[2])
    parts = [table.get(k) for k in PHASE_KERNELS.get(phase, [])]
    return sum(parts) if parts and all(p is not None for p in parts) else None


def roofline(phases: dict, info: dict):
    dom = max(phases, key=phases.get)
    amount, unit, bound = algorithmic(dom, info["nx"], info["ny"], info["cols"], info["rows"], info["n"],
                                      info["fdtd_substeps"])
    sec = phases[dom] / 1e3
    if unit == "flop":
        achieved = amount / sec / 1e12
        peak, u = FP32_MFMA_PEAK_TFLOPS, "TFLOP/s"
    else:
        achieved = amount / sec / 1e9
        peak, u = HBM_PEAK_GBS, "GB/s"
    return dom, {"bound": bound, "achieved": round(achieved, 2), "peak": peak, "unit": u,
                 "frac": round(achieved / peak, 4), "traffic": measured_traffic(dom), "kernel": dom,
                 "avg_ms": round(phases[dom], 4)}


# FP32 arithmetic of one candidate pair in the close-range sum (close_range.rs:359-372):
# r^2 (3), the table coordinate (1), the lerp from (g, dg) (2), and the two force
# accumulations (4). The fixed-point offset, index and clamp are not counted.
FLOP_PER_PAIR = 10.0


def candidate_pairs(pos: np.ndarray, cols: int, rows: int, cutoff: float) -> float:
    """Pairs the close-range sum walks: each electron against every electron of
    its (2 reach + 1)^2 cell window (itself included)."""
    reach = int(math.ceil(cutoff))
    cell = np.clip(np.floor(pos).astype(np.int64), 0, [cols - 1, rows - 1])
    counts = np.zeros((cols, rows))
    np.add.at(counts, (cell[:, 0], cell[:, 1]), 1.0)
    pad = np.pad(counts, reach)
    window = sum(pad[reach + dc: reach + dc + cols, reach + dr: reach + dr + rows]
                 for dc in range(-reach, reach + 1) for dr in range(-reach, reach + 1))
    return float((counts * window).sum())


def push_compute(world, ms: float):
    """The push is bound by its pair arithmetic, not by HBM: FP32 flops of the
    candidate pairs per second against the FP32 peak. That peak counts packed
    FP32 (two lanes' worth per instruction); the library is built without packed
    FP32, so one VALU instruction does one FP32 operation per lane and half the
    peak is the issue limit, reported beside it."""
    pairs = candidate_pairs(world.pos, world.cols, world.rows, world.settings["electron_cutoff"])
    achieved = FLOP_PER_PAIR * pairs / (ms * 1e-3) / 1e12
    unpacked = FP32_MFMA_PEAK_TFLOPS / 2.0
    return {"bound": "valu", "achieved": round(achieved, 2), "peak": FP32_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP32_MFMA_PEAK_TFLOPS, 4),
            "peak_without_packed_fp32": unpacked, "frac_without_packed_fp32": round(achieved / unpacked, 4),
            "candidate_pairs": pairs,
            "flop_per_candidate_pair": FLOP_PER_PAIR, "candidate_pairs_per_s": round(pairs / (ms * 1e-3), 1)}


def run_poisson(n: int = 2048, repeats: int = 20):
    """The DST Poisson solve in isolation on an n x n node grid (200 Gaussian
    blobs, seed 3): FP32 and BF16x3 MFMA against FP64, with time per solve from
    HIP events over `repeats` back-to-back solves on resident data. The four
    dense products of a solve are 8 n^3 flops; FP32 runs each as its two folded
    halves through the DST-I matrix's symmetry (4 n^3 flops executed), BF16x3
    runs them whole (8 n^3, three bf16 MFMAs per product)."""
    from embloks_amd import world as W
    from embloks_amd.gpusim import Poisson

    rho = W.gaussian_blobs(n)
    want = fp64_poisson(rho)
    dense = 8.0 * n ** 3
    p = Poisson(n, n)
    p.set_rhs(rho)
    out = {"grid": f"{n}x{n}", "rhs": "200 Gaussian blobs, sigma 2-40, seed 3", "flop_per_solve_dense": dense,
           "fp64_residual": laplacian_residual(want, rho),
           "fp32_rounding_floor_residual": laplacian_residual(want.astype(np.float32).astype(np.float64), rho)}
    for prec, name, peak, executed, issued in ((0, "fp32", FP32_MFMA_PEAK_TFLOPS, 0.5 * dense, 1.0),
                                               (1, "bf16x3", BF16_MFMA_PEAK_TFLOPS, dense, 3.0)):
        p.solve(prec, 0, 2)   # warm-up
        ms = p.solve(prec, 0, repeats)
        phi = p.read()
        tflops = executed / (ms * 1e-3) / 1e12
        out[name] = {"ms_per_solve": round(ms, 4), "flop_executed": executed, "tflops": round(tflops, 2),
                     "tflops_dense_equivalent": round(dense / (ms * 1e-3) / 1e12, 2),
                     "mfma_frac": round(tflops * issued / peak, 4),
                     "max_rel_err": float(np.abs(phi - want).max() / np.abs(want).max()),
                     "residual": laplacian_residual(phi, rho)}
    p.destroy()
    return out


TABLE_SETS = ((0.25, 2, 3.0), (0.25, 2, 2.0), (0.2, 3, 4.0))


def run_tables(repeats: int = 5):
    """The close-range tables (close_range::tables, close_range.rs:120-146) built
    afresh on the device for the default parameters (softening 0.25, 2 passes,
    cutoff 3) and two more sets: milliseconds per build on the host clock
    (launches, the 64 Poisson solves and gathers of grid_force, the radial,
    square and ion tables, the copy back), best of `repeats`. The CPU build of
    the same tables is timed in cpu_baseline."""
    from embloks_amd.gpusim import close_tables

    return {f"softening {soft}, passes {passes}, cutoff {cutoff}":
            round(min(close_tables(soft, passes, cutoff, timed=True)[-1] for _ in range(repeats)), 3)
            for soft, passes, cutoff in TABLE_SETS}


def run_p3m(repeats: int = 10):
    """The close-range force alone (CloseRange::force, close_range.rs:314-376)
    on the 4096 x 4096 scene with 4M clustered electrons at cutoff 2: time per
    evaluation from HIP events over `repeats` evaluations on resident data, and
    the candidate pairs it walks (each electron against every electron of its
    (2 reach + 1)^2 cell window)."""
    from embloks_amd import world as W
    from embloks_amd.gpusim import GpuSim

    world = W.clustered_charges()
    t0 = time.perf_counter()
    sim = GpuSim(world)
    setup_s = time.perf_counter() - t0
    sim.close_force(2)
    _, ms = sim.close_force(repeats)
    reach = int(math.ceil(world.settings["electron_cutoff"]))
    counts = world.density   # electrons per cell (one carrier per cell, neutral cells)
    pad = np.pad(counts, reach)
    window = sum(pad[reach + dc: reach + dc + counts.shape[0], reach + dr: reach + dr + counts.shape[1]]
                 for dc in range(-reach, reach + 1) for dr in range(-reach, reach + 1))
    candidates = float((counts * window).sum())
    sim.destroy()
    return {"grid": f"{world.cols}x{world.rows}", "electrons": world.n_electrons, "cutoff": world.settings["electron_cutoff"],
            "placement": "70% in 1000 log-normal Gaussian clusters (spread 4-24 cells), 30% uniform, seed 4",
            "ms_per_eval": round(ms, 4), "candidate_pairs": candidates,
            "candidate_pairs_per_s": round(candidates / (ms * 1e-3), 1), "scene_setup_s": round(setup_s, 2)}


# The reference's published FP64 CPU times per tick on one core (README.md:493-501).
README_GRIDS = ((40, 32, 2.4), (100, 80, 15.0), (200, 160, 59.0))


def run_readme_grids(cpu: bool):
    """The reference bench's worlds at the README's three grids: 30% random
    1-ohm wire, 8 electrons per wire cell, seed 1 (crates/embloks-gpu/examples/
    bench.rs:16-22). GPU ms per tick over 64 ticks after 8 warm-up ticks
    (bench.rs:47-54); the FP64 oracle on one thread over 5 ticks after 1
    warm-up (bench.rs:38-46); the reference's published one-core figure."""
    from embloks_amd import world as W
    from embloks_amd.gpusim import GpuSim

    out = []
    for cols, rows, published in README_GRIDS:
        w = W.random_wire(cols, rows)
        sim = GpuSim(w)
        sim.run_ticks(8)
        t0 = time.perf_counter()
        sim.step(64)
        sim.sync()
        gpu_ms = (time.perf_counter() - t0) / 64 * 1e3
        entry = {"grid": f"{cols}x{rows}", "yee_nodes": sim.nx * sim.ny, "electrons": w.n_electrons,
                 "gpu_ms_per_tick": round(gpu_ms, 4), "reference_cpu_ms_per_tick_published": published}
        sim.destroy()
        if cpu:
            sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
            from oracle import orc
            orc.set_threads(1)
            cs = orc.Sim(w)
            cs.step(1)
            t0 = time.perf_counter()
            cs.step(5)
            entry["oracle_cpu_ms_per_tick_1_thread"] = round((time.perf_counter() - t0) / 5 * 1e3, 3)
        out.append(entry)
    return out


def run_gauss_log(ticks: int = 10000, every: int = 100):
    """The 512^2 scene's long run: `ticks` ticks with the FP64 Gauss residual
    max |div E - 4 pi k rho| and the exact deposited electron charge read every
    `every` ticks (each read waits for the ticks before it)."""
    from embloks_amd import world as W
    from embloks_amd.gpusim import GpuSim

    world = W.rectangles_circuit(512, carriers_per_cell=16, seed=2)
    sim = GpuSim(world)
    res, charge_exact, charge_dev = [], 0, 0
    t0 = time.perf_counter()
    for _ in range(ticks // every):
        sim.step(every)
        r, _q, _e = sim.diagnostics()
        res.append(r)
        # the integer deposit: -2^44 per electron, exactly; the total passes the
        # int64 range, so the cells are summed modulo 2^64 and compared there
        wrapped = int(sim.read_charge_fixed().view(np.uint64).sum(dtype=np.uint64))
        dev = (wrapped + sim.n * (1 << 44)) % (1 << 64)
        dev = dev - (1 << 64) if dev >= (1 << 63) else dev
        charge_exact += int(dev == 0)
        charge_dev = max(charge_dev, abs(dev))
    elapsed = time.perf_counter() - t0
    sim.destroy()
    return {"ticks": ticks, "checked_every": every, "checks": len(res), "residual_first": res[0],
            "residual_max": max(res), "residual_last": res[-1],
            "electron_charge_exact_checks": charge_exact, "electron_charge_max_dev_fixed": charge_dev,
            "ms_per_tick_with_checks": round(elapsed / ticks * 1e3, 4)}


def run_large(ticks: int = 5):
    """The bandwidth-bound kernels at a large grid: a 4096^2 world of 30% random
    wire at one electron per wire cell (4128^2 Yee nodes), phase times from HIP
    events over `ticks` ticks: FDTD at 32 B per node-substep plus 12 B per node
    for the Gauss right-hand sides, the push at 32 B per electron, against 8 TB/s."""
    from embloks_amd import world as W
    from embloks_amd.gpusim import GpuSim

    w = W.random_wire(4096, 4096, carriers_per_cell=1, seed=6)
    sim = GpuSim(w)
    sim.run_ticks(2)
    ph = sim.kernel_times(ticks)
    nodes = sim.nx * sim.ny
    sub = sim.dims().fdtd_substeps
    fdtd_gbs = algorithmic("fdtd", sim.nx, sim.ny, 0, 0, 0, sub)[0] / (ph["fdtd"] * 1e-3) / 1e9
    push_gbs = 32.0 * sim.n / (ph["push"] * 1e-3) / 1e9
    out = {"grid": "4096x4096", "yee_nodes": nodes, "electrons": sim.n,
           "world": "30% random 1-ohm wire, 1 electron per wire cell, seed 6",
           "phases_ms": {k: round(v, 4) for k, v in ph.items()},
           "fdtd": {"achieved": round(fdtd_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(fdtd_gbs / HBM_PEAK_GBS, 4)},
           "push": {"achieved": round(push_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(push_gbs / HBM_PEAK_GBS, 4)}}
    sim.destroy()
    return out


def cpu_baseline(world, budget_s: float = 20.0):
    """The FP64 oracle (the CPU restatement of embloks-core's tick) on this
    host over a bounded sample of the same scene: one thread (the reference's
    published CPU figures are one core), then every core of the GPU's share of
    the host (OpenMP over electrons and DST lanes). Returns the all-core figure
    with the one-thread one beside it."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import orc

    m = world.settings["absorbing_margin"]
    nodes = (world.cols + 2 * m) * (world.rows + 2 * m)

    def timed(threads, seconds):
        orc.set_threads(threads)
        sim = orc.Sim(world)
        sim.step(1)
        ticks, t0 = 0, time.perf_counter()
        while ticks < 2 or (time.perf_counter() - t0 < seconds and ticks < 10):
            sim.step(1)
            ticks += 1
        return (time.perf_counter() - t0) / ticks, ticks

    one, n1 = timed(1, budget_s * 0.5)
    # Every core of this process's share: the affinity mask and OMP_NUM_THREADS
    # (a GPU host's share is set by its scheduler), not the machine's count.
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    cores = max(1, min(share, int(omp)) if omp and omp.isdigit() else share)
    many, nm = timed(cores, budget_s * 0.25)
    orc.set_threads(1)
    tables_ms = {}
    for soft, passes, cutoff in TABLE_SETS:   # the close-range table build, one thread
        t0 = time.perf_counter()
        orc.tables(soft, passes, cutoff * (1.0 + 1e-12))   # a parameter set the oracle has not cached
        tables_ms[f"softening {soft}, passes {passes}, cutoff {cutoff}"] = round((time.perf_counter() - t0) * 1e3, 1)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": round(nodes / many, 1), "unit": "cell-steps/s", "cores": cores, "kind": "port",
            "cpu_model": model, "host_cpus": os.cpu_count(), "cpu_share": share,
            "single_thread_value": round(nodes / one, 1), "close_range_tables_ms": tables_ms,
            "sample": (f"FP64 oracle, same 512x512 scene: {cores} threads {nm} ticks after 1 warm-up, "
                       f"{many * 1e3:.0f} ms/tick; 1 thread {n1} ticks, {one * 1e3:.0f} ms/tick")}


def run_serving(stream, batched: bool):
    """Feed `stream` (arrival s, steps, world) to a server at its arrival times
    (wall clock from the first submit) and poll until every request is done."""
    from embloks_amd import abi
    from embloks_amd.gpusim import Server

    srv = Server(abi.settings(carriers_per_cell=1), hbm_budget_bytes=64 << 30, max_slots=256, batched=batched,
                 ticks_per_round=16, use_graph=batched)
    t0 = time.perf_counter()
    pending = list(stream)
    active = 0
    shown = 0.0
    while pending or active:
        now = time.perf_counter() - t0
        while pending and pending[0][0] <= now:
            _, steps, w = pending.pop(0)
            srv.submit(w, steps)
        active = srv.poll()
        if now - shown >= 30.0:   # a progress line on stderr for long streams
            shown = now
            print(f"serving ({'batched' if batched else 'plain'}): {now:.0f} s, {active} active, {len(pending)} to arrive",
                  file=sys.stderr, flush=True)
        if not active and pending:
            time.sleep(max(0.0, min(0.001, pending[0][0] - (time.perf_counter() - t0))))
    st = srv.stats()
    srv.destroy()
    return st


def copy_calibration(nbytes: int = 1 << 30, reps: int = 5):
    """A copy kernel of known size: each launch reads and writes `nbytes`, so the
    HBM byte counters it shows give the scale of the same counters elsewhere."""
    import torch

    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copy calibration: {reps} copies of {nbytes} bytes", file=sys.stderr)
    del src, dst


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=256)
    ap.add_argument("--serve-scale", type=float, default=0.1, help="fraction of each request's 1000-20000 steps")
    ap.add_argument("--no-poisson", action="store_true", help="skip the isolated 2048^2 DST solve")
    ap.add_argument("--no-p3m", action="store_true", help="skip the 4096^2 clustered close-range force")
    ap.add_argument("--no-tables", action="store_true", help="skip the close-range table build")
    ap.add_argument("--no-readme", action="store_true", help="skip the README grids (40x32, 100x80, 200x160)")
    ap.add_argument("--no-gauss-log", action="store_true", help="skip the 10^4-tick Gauss residual log")
    ap.add_argument("--no-large", action="store_true", help="skip the 4096^2 FDTD / push bandwidth leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 1 GiB device-to-device copy 5 times (scale for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU design is not built.",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration()
    world, elapsed, phases, info = run_single(args)
    nodes = info["nx"] * info["ny"]
    value = nodes * args.steps / elapsed
    dom, roof = roofline(phases, info)
    if dom == "push":
        roof["compute"] = push_compute(world, phases[dom])
    out = {
        "metric": "Yee cell-steps/s + p99 step latency, batched vs plain, <=1e-5 rel err vs FP64",
        "value": round(value, 1),
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random-rectangle circuit)",
        "config": {"workload": "512x512 scene, 64 conductor rectangles, 16 batteries, 16 carriers/cell",
                   "yee_nodes": nodes, "visible_cells": info["cols"] * info["rows"], "electrons": info["n"],
                   "parallelism": "single GPU"},
        "phases_ms": {k: round(v, 4) for k, v in phases.items()},
        "roofline": roof,
        "gauss_residual_fp64": info["gauss_residual"],
    }
    if not args.no_poisson:
        out["poisson"] = run_poisson()
    if not args.no_p3m:
        out["p3m"] = run_p3m()
    if not args.no_tables:
        out["close_range_tables_ms"] = run_tables()
    if not args.no_large:
        out["large"] = run_large()
    if not args.no_gauss_log:
        out["gauss_log"] = run_gauss_log()
    if not args.no_readme:
        out["readme_grids"] = run_readme_grids(cpu=not args.no_cpu)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(world)
    if not args.no_serve:
        from embloks_amd import world as W

        serve = {}
        stream = W.serving_stream(n_requests=args.serve_requests, seed=5,
                                  min_steps=max(1, int(1000 * args.serve_scale)),
                                  max_steps=max(2, int(20000 * args.serve_scale)))
        print(f"serving stream: {len(stream)} requests generated", file=sys.stderr, flush=True)
        for mode in ("batched", "plain"):
            st = run_serving(stream, mode == "batched")
            print(f"serving {mode}: {st['cell_steps_per_s'] / 1e9:.3f} G cell-steps/s", file=sys.stderr, flush=True)
            serve[mode] = {"cell_steps_per_s": round(st["cell_steps_per_s"], 1),
                           "p50_ms": round(st["step_latency_p50_ms"], 4), "p99_ms": round(st["step_latency_p99_ms"], 4),
                           "request_p50_ms": round(st["request_latency_p50_ms"], 1),
                           "request_p99_ms": round(st["request_latency_p99_ms"], 1),
                           "queue_wait_p50_ms": round(st["queue_wait_p50_ms"], 1),
                           "queue_wait_p99_ms": round(st["queue_wait_p99_ms"], 1),
                           "completed": st["completed"], "peak_resident": st["peak_resident"],
                           "elapsed_s": round(st["elapsed_s"], 3)}
        b, p = serve["batched"]["cell_steps_per_s"], serve["plain"]["cell_steps_per_s"]
        serve["speedup"] = round(b / p, 2) if p > 0 else None
        serve["speedup_target"] = 5.0
        serve["stream"] = (f"{args.serve_requests} requests, Poisson arrivals mean 2 ms, sides log-uniform 64-1024, "
                           f"steps uniform {max(1, int(1000 * args.serve_scale))}-{max(2, int(20000 * args.serve_scale))}")
        out["serving"] = serve
    print(json.dumps(out))


if __name__ == "__main__":
    main()
