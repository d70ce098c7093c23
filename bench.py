"""Benchmark: FP64 ideal SRMHD 3-D Kelvin-Helmholtz, 256^3, WENO5 FVS, RK2.

One step = one RK2 step (two stages, each with the flux sweeps, the update and
the per-cell Newton cons2prim) of the whole grid, state resident in HBM.
Prints one JSON line. The HBM roofline counts 784 algorithmic bytes per zone
update (BYTES_PER_ZONE_UPDATE below, the unit's definition; this engine's step
needs 672, ENGINE_BYTES_PER_ZONE_UPDATE); the FP64 compute roofline takes its
FLOPs and VALU instructions per zone update from the committed rocprof counter
summary of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "zone-updates/s FP64 SRMHD 256^3; cons2prim cells/s learned vs plain guess"
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes of one SRMHD RK2 zone update: 98 FP64 values
# (stage 1 reads U^n, V^n and writes U1, V1; stage 2 reads U^n, U1, V1 and the
# V^n guess, writes U^{n+1}, V^{n+1}, A^{n+1}).
BYTES_PER_ZONE_UPDATE = 784
# The same count for this engine's step, which writes only rho .. p of V1 (the
# next stage reads no more) and the cons2prim root (v^2, Z) in place of
# A^{n+1} (the aux is formed from it when read): 84 values.
ENGINE_BYTES_PER_ZONE_UPDATE = 672
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "kh3d_256_weno5_rk2_traffic.json")
# FP64 FLOPs and VALU instructions per zone update from the rocprof
# instruction counters of the same workload (the file names the command).
FLOPS_FILE = os.path.join(ROOT, "profiles", "kh3d_256_weno5_rk2_flops.json")
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector / matrix peak
# VALU issue: one wave64 instruction per 4 cycles per SIMD, 256 CUs x 4 SIMDs at 2.4 GHz
VALU_ISSUE_PER_S = 256 * 4 * 2.4e9 / 4 * 64


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256, help="cells per side")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=96, help="cells per side of the CPU baseline sample")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-c2p", action="store_true", help="skip the cons2prim guess comparison")
    ap.add_argument("--no-srrmhd", action="store_true", help="skip the SRRMHD SSP2 1024^2 leg")
    ap.add_argument("--srrmhd-steps", type=int, default=10)
    ap.add_argument("--srrmhd-only", action="store_true",
                    help="run only the SRRMHD SSP2 leg and print its JSON (for profiling its kernels)")
    ap.add_argument("--mlp-steps", type=int, default=12000)
    ap.add_argument("--mlp-params", default=None,
                    help="npz of trained guess-network parameters: loaded if it exists, else written after training")
    ap.add_argument("--c2p-only", action="store_true",
                    help="run only the cons2prim guess comparison and print its JSON (for profiling its kernels)")
    ap.add_argument("--no-regime", action="store_true", help="skip the SRMHD + REGIME RKSplit leg")
    ap.add_argument("--regime-n", type=int, default=256)
    ap.add_argument("--regime-steps", type=int, default=10)
    ap.add_argument("--regime-only", action="store_true",
                    help="run only the REGIME leg and print its JSON (for profiling its kernels)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--batch-steps", type=int, default=50)
    ap.add_argument("--batch-members", type=int, default=64)
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched-ensemble leg and print its JSON (for profiling its kernels)")
    ap.add_argument("--sweep-plan", type=int, default=None,
                    help="sweep plan of the headline engine and the SRRMHD leg (mgpu_set_sweep_plan); "
                         "default: the engine's")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (HBM counter calibration) and exit")
    return ap.parse_args()


def calibrate_copy():
    """A plain copy kernel of known size: 2^27 doubles read and written per launch.
    Run under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to scale those counters;
    its timed copies (HIP events on the copy's stream) give the practical HBM
    rate of a pure streaming kernel on this GPU."""
    import torch

    n = 1 << 27
    a = torch.empty(n, dtype=torch.float64, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    reps = 20
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        b.copy_(a)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / reps
    print(json.dumps({"calibration": "copy", "bytes_read_per_launch": 8 * n, "bytes_written_per_launch": 8 * n,
                      "launches": 3 + reps, "ms_per_copy": ms, "copy_GBs": 16 * n / (ms * 1e-3) / 1e9}))


REF_BIN = os.path.join(ROOT, "oracle", "_ref", "method_ref")
REF_OMP_BIN = os.path.join(ROOT, "oracle", "_ref", "method_ref_omp")


def host_cpu():
    """(model name, threads this process may run on) of the host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    # the box grants this process a share of its cores through OMP_NUM_THREADS
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        threads = min(threads, int(env))
    return model, max(threads, 1)


def cpu_baseline(n: int, steps: int):
    """The reference's own CPU builds (oracle/_ref, compiled from the reference
    sources) when they are present, else the C oracle, over a bounded sample of
    the workload: KH 3-D n^3, WENO5, RK2. The reported figure is the OpenMP
    build on the host threads this process may use; the serial build's
    figure sits beside it."""
    model, threads = host_cpu()
    if os.path.exists(REF_BIN):
        serial = reference_baseline(REF_BIN, n, steps, 1, model)
        if os.path.exists(REF_OMP_BIN) and threads > 1:
            out = reference_baseline(REF_OMP_BIN, n, steps, threads, model)
            out["serial"] = serial
            return out
        return serial
    out = port_baseline(n, steps)
    out["cpu_model"] = model
    return out


def reference_baseline(binary: str, n: int, steps: int, threads: int, model: str):
    import subprocess
    import tempfile

    from method_amd import initdata as idt

    g = idt.kh3d_grid(n)
    prims0 = idt.kh_prims(g, seed=1234)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    with tempfile.TemporaryDirectory() as td:
        pin = os.path.join(td, "prims.bin")
        np.ascontiguousarray(prims0).tofile(pin)
        args = ["model=srmhd", f"nx={n}", f"ny={n}", f"nz={n}", "xmin=-0.5", "xmax=0.5", "ymin=-1.0", "ymax=1.0",
                "zmin=0.0", "zmax=1.0", "Ng=4", f"gamma={4.0 / 3.0!r}", "cfl=0.4", "weno=5", "integrator=rk2",
                "bcx=periodic", "bcy=outflow", "bcz=periodic", "mode=run", f"steps={steps}", f"prims_in={pin}",
                f"out={os.path.join(td, 'out')}", "dump=-1"]
        r = subprocess.run([binary] + args, capture_output=True, text=True, cwd=td, env=env)
    if r.returncode != 0:
        raise RuntimeError("reference CPU build failed: " + r.stderr[-2000:])
    secs = float(r.stdout.split("step_seconds=")[1].split()[0])
    build = "OpenMP (-fopenmp; FVS and WENO loops parallel, cons2prim serial)" if threads > 1 else "serial"
    return {"value": n ** 3 * steps / secs, "unit": "zone-updates/s", "cores": threads, "kind": "reference",
            "cpu_model": model,
            "sample": f"reference METHOD {build} CPU build (Simulation::updateTime with RK2, FVS, WENO5, SRMHD "
                      f"with cminpack hybrd C2P), KH 3-D {n}^3, {steps} steps, {secs:.1f} s inside updateTime "
                      f"on {threads} host thread{'s' if threads > 1 else ''} of {model}"}


def port_baseline(n: int, steps: int):
    from method_amd import initdata as idt
    from oracle import oracle as orc

    g = idt.kh3d_grid(n)
    prims0 = idt.kh_prims(g, seed=1234)
    o = orc.Oracle(orc.Case(nx=n, ny=n, nz=n, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, zmin=0.0, zmax=1.0,
                            weno_order=5, periodic=(1, 0, 1), cfl=0.4))
    c, p, a = o.set(prims0)
    t0 = time.perf_counter()
    t = 0.0
    for it in range(steps):
        dt = o.timestep(t, it)
        o.step(c, p, a, dt)
        t += dt
    el = time.perf_counter() - t0
    return {"value": n ** 3 * steps / el, "unit": "zone-updates/s", "cores": 1, "kind": "port",
            "sample": f"C oracle (serial restatement of the reference RK2 step), KH 3-D {n}^3 WENO5 RK2, "
                      f"{steps} steps, {el:.1f} s on 1 host core"}


def c2p_comparison(e, dt, args):
    """SRMHD cons2prim of the 256^3 state with the plain, learned and cold
    guesses: kernel time (HIP events), Newton steps, fallbacks, and the
    learned/cold results against the plain one."""
    from method_amd import c2p_guess
    from method_amd import engine as eng

    t0 = time.perf_counter()
    if args.mlp_params and os.path.exists(args.mlp_params):
        with np.load(args.mlp_params, allow_pickle=False) as z:
            params, scales, loss = z["params"], z["scales"], float(z["loss"])
    else:
        print("bench: training the cons2prim guess network", file=sys.stderr, flush=True)
        params, scales, loss = c2p_guess.train_guess(steps=args.mlp_steps)
        if args.mlp_params:
            np.savez(args.mlp_params, params=params, scales=scales, loss=np.float64(loss))
    train_s = time.perf_counter() - t0
    _, p_prev, _ = e.download()
    _, st_step = e.step(dt, stats=True)
    out = {"workload": "SRMHD cons2prim of every cell of the 256^3 KH state; the guess state is one step older",
           "in_step_newton_steps_per_cell": st_step["iterations"] / max(st_step["cells"], 1),
           "training": {"steps": args.mlp_steps, "final_loss": loss, "seconds": train_s,
                        "network": "8-64-64-2 GELU, BF16 weights on MFMA, FP32 accumulate"}}
    prims = {}
    for name, mode in (("plain", eng.GUESS_PLAIN), ("learned", eng.GUESS_LEARNED), ("cold", eng.GUESS_COLD)):
        e.upload(prims=p_prev)
        e.set_c2p_guess(mode, params if mode == eng.GUESS_LEARNED else None,
                        scales if mode == eng.GUESS_LEARNED else None)
        ms, st, rc = e.timed_c2p(reps=5)
        _, prims[name], _ = e.download()
        cells = st["cells"]
        out[name] = {"cells_per_s": cells / (ms * 1e-3), "kernel_ms": ms,
                     "newton_steps_per_cell": st["iterations"] / max(cells, 1), "max_newton_steps": st["max_iterations"],
                     "fallback_rate": st["fallbacks"] / max(cells, 1), "failures": st["failures"]}
        # Newton steps per cell over every cell (a separate, untimed call)
        e.upload(prims=p_prev)
        e.set_c2p_histogram(True)
        e.c2p()
        hist = e.c2p_histogram()
        e.set_c2p_histogram(False)
        c = np.cumsum(hist)
        out[name]["newton_steps_p50"] = int(np.searchsorted(c, 0.50 * c[-1]))
        out[name]["newton_steps_p99"] = int(np.searchsorted(c, 0.99 * c[-1]))
    e.set_c2p_guess(eng.GUESS_PLAIN)
    out["learned"]["guess_error"] = learned_guess_error(e, params, scales, prims["plain"])
    ref = prims["plain"]
    scale = np.abs(ref).reshape(ref.shape[0], -1).max(axis=1)
    for name in ("learned", "cold"):
        d = np.abs(prims[name] - ref).reshape(ref.shape[0], -1).max(axis=1)
        out[name]["max_rel_diff_vs_plain"] = float(np.max(d / np.maximum(scale, 1e-300)))
    out["learned_vs_plain_speedup"] = out["learned"]["cells_per_s"] / out["plain"]["cells_per_s"]
    out["learned_vs_cold_speedup"] = out["learned"]["cells_per_s"] / out["cold"]["cells_per_s"]
    return out


def learned_guess_error(e, params, scales, prims):
    """Relative error of the network's decoded guess (W, p) against the
    converged primitives on every 16th interior cell of the evaluation state:
    median and 90th percentile."""
    from method_amd import engine as eng

    cons, _, _ = e.download()
    ng = 4
    c8 = cons[:8, ng:-ng, ng:-ng, ng:-ng].reshape(8, -1)[:, ::16].T.copy()
    pr = prims[:, ng:-ng, ng:-ng, ng:-ng].reshape(prims.shape[0], -1)[:, ::16]
    m = eng.Mlp(seed=0)
    try:
        m.set_params(params, scales)
        y = m.predict(c8).astype(np.float64)
    finally:
        m.close()
    ym, ys = scales[:2].astype(np.float64), scales[2:].astype(np.float64)
    W_g = np.exp(np.clip(y[:, 0] * ys[0] + ym[0], 0.0, 6.0))
    p_g = (c8[:, 4] + c8[:, 0]) * np.exp(y[:, 1] * ys[1] + ym[1])
    W = 1.0 / np.sqrt(1.0 - np.sum(pr[1:4] ** 2, axis=0))
    eW, ep = np.abs(W_g / W - 1.0), np.abs(p_g / pr[4] - 1.0)
    return {"cells": int(c8.shape[0]), "W_rel_err_p50": float(np.median(eW)),
            "W_rel_err_p90": float(np.quantile(eW, 0.9)), "p_rel_err_p50": float(np.median(ep)),
            "p_rel_err_p90": float(np.quantile(ep, 0.9))}


def srrmhd_leg(steps: int, plan=None):
    """Resistive SRMHD Orszag-Tang 1024^2, sigma 1e3, SSP2 IMEX, WENO3 (sweep
    plan `plan` if given, else the engine's default)."""
    from method_amd import engine as eng
    from method_amd import initdata as idt

    g = idt.ot2d_grid(1024)
    cfg = eng.Config(nx=1024, ny=1024, nz=0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, model=eng.SRRMHD,
                     integrator=eng.SSP2, weno_order=3, bc=(eng.PERIODIC,) * 3, sigma=1e3, cp=0.1, cfl=0.5)
    with eng.Engine(cfg) as e:
        if plan is not None:
            e.set_sweep_plan(plan)
        e.set_prims(idt.ot_prims_resistive(g))
        for _ in range(2):
            e.update_time()
        base = 0.5 * (1.0 / 1024) / np.sqrt(2.0)
        dts = [base * 0.1] * steps
        e.sync()
        w0 = time.perf_counter()
        ev_ms, slots = e.timed_steps(dts)
        wall = time.perf_counter() - w0
        _, st = e.step(dts[-1], stats=True)
    cells = 1024 * 1024
    return {"workload": "resistive SRMHD Orszag-Tang 1024^2, sigma 1e3, SSP2 IMEX, WENO3, FP64",
            "zone_updates_per_s": cells * steps / wall, "ms_per_step": 1e3 * wall / steps,
            "event_ms_per_step": ev_ms / steps, "launch_ms": slots,
            "secant_steps_per_cell_solve": st["iterations"] / max(st["cells"], 1), "failures": st["failures"]}


# Minimal HBM bytes of REGIME's source per cell: the eight primitives read and
# the eight source values written once (every intermediate field kept on chip).
REGIME_SOURCE_BYTES = 16 * 8


def regime_leg(n: int, steps: int):
    """Ideal SRMHD + the REGIME resistive extension, 3-D Kelvin-Helmholtz n^3,
    WENO5, RKSplit (RK2, then the split source with REGIME's added), sigma 100."""
    from method_amd import engine as eng
    from method_amd import initdata as idt

    g = idt.kh3d_grid(n)
    cfg = eng.Config(nx=n, ny=n, nz=n, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, zmin=0.0, zmax=1.0,
                     weno_order=5, bc=(eng.PERIODIC, eng.OUTFLOW, eng.PERIODIC), cfl=0.4, sigma=100.0,
                     integrator=eng.RKSPLIT)
    cells = cfg.interior_cells
    with eng.Engine(cfg) as e:
        e.set_prims(idt.kh_prims(g, seed=1234, bz=0.1))
        e.set_model_extension(eng.EXT_REGIME)
        for _ in range(2):
            e.update_time()
        base = 0.4 * (1.0 / n) / np.sqrt(3.0)
        dts = [base * 0.1] * steps
        e.sync()
        w0 = time.perf_counter()
        ev_ms, slots = e.timed_steps(dts)
        wall = time.perf_counter() - w0
        _, st = e.step(dts[-1], stats=True)
    src_ms = slots.get("regime_source", 0.0)
    padded = (n + 8) ** 3
    achieved = REGIME_SOURCE_BYTES * padded / (src_ms * 1e-3) / 1e9 if src_ms > 0 else None
    return {"workload": f"ideal SRMHD + REGIME 3-D Kelvin-Helmholtz {n}^3, WENO5, RKSplit, sigma 100, FP64",
            "zone_updates_per_s": cells * steps / wall, "ms_per_step": 1e3 * wall / steps,
            "event_ms_per_step": ev_ms / steps, "launch_ms": slots,
            "regime_source_roofline": {"bound": "hbm", "bytes_per_padded_cell": REGIME_SOURCE_BYTES,
                                       "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                       "frac": achieved / HBM_PEAK_GBS if achieved else None},
            "newton_steps_per_cell": st["iterations"] / max(st["cells"], 1), "failures": st["failures"]}


def batch_leg(steps: int, members: int):
    """64 independent 2-D ideal SRMHD KH runs of uneven size, WENO5, RK2, per-member
    dt: the batched engine (autotuned persistent stage kernels under a hipGraph)
    against the same members launched one at a time on the single-instance
    engine, and one at a time as one-member batches (the fused stage kernels
    without the batching), which separates the two effects."""
    from method_amd import engine as eng
    from method_amd import initdata as idt

    ms_list = idt.kh_ensemble(members)
    cells = sum(m[0] * m[1] for m in ms_list)
    inst = [eng.Instance(m[0], m[1]) for m in ms_list]
    with eng.Batch(inst, weno_order=5, bc=(eng.PERIODIC, eng.OUTFLOW)) as b:
        for k, m in enumerate(ms_list):
            b.set_prims(k, idt.kh_prims(idt.Grid(m[0], m[1], 0, -0.5, 0.5, -1.0, 1.0), seed=m[4], shear=m[2], bz=m[3]))
        best, cands = b.autotune(trial_steps=3)
        b.step(2)
        b.sync()
        w0 = time.perf_counter()
        ev_ms = b.timed_steps(steps)
        wall = time.perf_counter() - w0
        _, st = b.step(1, stats=True)
    batched = cells * steps / wall

    # one at a time: each member on its own engine, its steps queued back to
    # back and timed with events on that engine's stream
    single_ms = 0.0
    for m in ms_list:
        nx, ny = m[0], m[1]
        cfg = eng.Config(nx=nx, ny=ny, nz=0, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, weno_order=5,
                         bc=(eng.PERIODIC, eng.OUTFLOW, eng.PERIODIC), cfl=0.4)
        with eng.Engine(cfg) as e:
            e.set_prims(idt.kh_prims(idt.Grid(nx, ny, 0, -0.5, 0.5, -1.0, 1.0), seed=m[4], shear=m[2], bz=m[3]))
            for _ in range(2):
                e.update_time()
            base = 0.4 * min(1.0 / nx, 2.0 / ny) / np.sqrt(2.0)
            dts = [base * (0.1 if it < 15 else (0.25 if it < 30 else 1.0)) for it in range(2, 2 + steps)]
            ms, _ = e.timed_steps(dts)
            single_ms += ms
    one_at_a_time = cells * steps / (single_ms * 1e-3)

    # fusion without batching: each member alone as a one-member batch (the
    # same fused stage kernels and hipGraph, autotuned for that member),
    # launched one member after another
    fused_ms = 0.0
    for m in ms_list:
        with eng.Batch([eng.Instance(m[0], m[1])], weno_order=5, bc=(eng.PERIODIC, eng.OUTFLOW)) as b1:
            b1.set_prims(0, idt.kh_prims(idt.Grid(m[0], m[1], 0, -0.5, 0.5, -1.0, 1.0), seed=m[4], shear=m[2],
                                         bz=m[3]))
            b1.autotune(trial_steps=3)
            b1.step(2)
            b1.sync()
            fused_ms += b1.timed_steps(steps)
    fused_one_at_a_time = cells * steps / (fused_ms * 1e-3)
    return {"workload": f"{members} independent ideal SRMHD 2-D KH runs, nx, ny from {{128..768}}, WENO5, RK2, "
                        "per-member dt, FP64",
            "cells": cells, "steps": steps,
            "zone_updates_per_s": batched, "ms_per_step": 1e3 * wall / steps, "event_ms_per_step": ev_ms / steps,
            "hbm_frac_784B": batched * BYTES_PER_ZONE_UPDATE / 1e9 / HBM_PEAK_GBS,
            "one_at_a_time_zone_updates_per_s": one_at_a_time, "one_at_a_time_ms_per_step": single_ms / steps,
            "speedup_vs_one_at_a_time": batched / one_at_a_time,
            "fused_one_member_per_launch_zone_updates_per_s": fused_one_at_a_time,
            "fused_one_member_per_launch_ms_per_step": fused_ms / steps,
            "speedup_from_fusion": fused_one_at_a_time / one_at_a_time,
            "speedup_from_batching": batched / fused_one_at_a_time,
            "tuning": best, "candidates": len(cands),
            "newton_steps_per_cell": st["iterations"] / max(st["cells"], 1), "failures": st["failures"]}


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1 (multi-GPU halo exchange "
                                   "is a design note only)"}))
        sys.exit(2)
    if args.batch_only:
        print(json.dumps({"batched_ensemble": batch_leg(args.batch_steps, args.batch_members)}))
        return
    if args.srrmhd_only:
        print(json.dumps({"srrmhd_ssp2": srrmhd_leg(args.srrmhd_steps, args.sweep_plan)}))
        return
    if args.regime_only:
        print(json.dumps({"regime_rksplit": regime_leg(args.regime_n, args.regime_steps)}))
        return
    from method_amd import engine as eng
    from method_amd import initdata as idt

    n = args.n
    g = idt.kh3d_grid(n)
    prims0 = idt.kh_prims(g, seed=1234)
    cfg = eng.Config(nx=n, ny=n, nz=n, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, zmin=0.0, zmax=1.0,
                     weno_order=5, bc=(eng.PERIODIC, eng.OUTFLOW, eng.PERIODIC), cfl=0.4)
    cells = cfg.interior_cells
    e = eng.Engine(cfg)
    if args.sweep_plan is not None:
        e.set_sweep_plan(args.sweep_plan)
    e.set_prims(prims0)
    del prims0
    print(f"bench: {n}^3 set up; warmup {args.warmup} steps", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        e.update_time()
    e.sync()
    if args.c2p_only:
        t, dt, _ = e.time()
        print(json.dumps({"c2p": c2p_comparison(e, dt, args)}))
        e.close()
        return

    # dt sequence of the next K updateTime calls (slow start changes dt only)
    t, _, iters = e.time()
    dts = []
    for k in range(args.steps):
        it = iters + k
        base = 0.4 * (1.0 / n) / np.sqrt(3.0)
        dts.append(base * (0.1 if it < 15 else (0.25 if it < 30 else 1.0)))

    print(f"bench: timing {args.steps} steps", file=sys.stderr, flush=True)
    e.sync()
    w0 = time.perf_counter()
    event_ms, slots = e.timed_steps(dts)
    e.sync()
    wall = time.perf_counter() - w0
    e.set_time(t + float(np.sum(dts)), iters + args.steps)

    # cons2prim of the same state with three starting guesses (plain = the
    # previous step's primitives, as the reference; learned; cold)
    c2p = c2p_comparison(e, dts[-1], args) if not args.no_c2p else None

    ms_per_step = 1e3 * wall / args.steps
    value = cells * args.steps / wall
    step_kernel_ms = sum(slots.values())
    dom = max(slots, key=slots.get)
    achieved_step = BYTES_PER_ZONE_UPDATE * cells / (event_ms / args.steps * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("bytes_per_step")
    compute = None
    if os.path.exists(FLOPS_FILE) and n == 256:
        with open(FLOPS_FILE) as f:
            fl = json.load(f)
        step_s = event_ms / args.steps * 1e-3
        tflops = fl["fp64_flops_per_zone_update"] * cells / step_s / 1e12
        compute = {"bound": "fp64 valu", "fp64_flops_per_zone_update": fl["fp64_flops_per_zone_update"],
                   "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tflops / FP64_PEAK_TFLOPS,
                   "valu_instructions_per_zone_update": fl["valu_instructions_per_zone_update"],
                   "valu_issue_frac": fl["valu_instructions_per_zone_update"] * cells / step_s / VALU_ISSUE_PER_S}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Kelvin-Helmholtz shear layers, splitmix64 perturbation seed 1234)",
        "config": {"workload": f"ideal SRMHD 3-D Kelvin-Helmholtz {n}^3, WENO5 FVS, RK2, cfl 0.4, "
                               "periodic x,z / outflow y, FP64", "cells": cells, "parallelism": "single GPU"},
        "roofline": {
            "bound": "hbm",
            "kernel": "one RK2 step (the fused sweep launches of both stages), 784 B per zone update",
            "achieved": achieved_step,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved_step / HBM_PEAK_GBS,
            "engine_algorithmic_B_per_zone_update": ENGINE_BYTES_PER_ZONE_UPDATE,
            "engine_frac": achieved_step * ENGINE_BYTES_PER_ZONE_UPDATE / BYTES_PER_ZONE_UPDATE / HBM_PEAK_GBS,
            "traffic": traffic,
            "event_ms_per_step": event_ms / args.steps,
            "dominant_launch": dom,
            "dominant_launch_ms": slots[dom],
            "launch_ms": slots,
            "launch_ms_sum": step_kernel_ms,
        },
        "compute_roofline": compute,
        "c2p": c2p,
    }
    e.close()
    if not args.no_srrmhd:
        print("bench: SRRMHD SSP2 1024^2", file=sys.stderr, flush=True)
        out["srrmhd_ssp2"] = srrmhd_leg(args.srrmhd_steps, args.sweep_plan)
    if not args.no_regime:
        print(f"bench: SRMHD + REGIME RKSplit {args.regime_n}^3", file=sys.stderr, flush=True)
        out["regime_rksplit"] = regime_leg(args.regime_n, args.regime_steps)
    if not args.no_batch:
        print("bench: batched ensemble", file=sys.stderr, flush=True)
        out["batched_ensemble"] = batch_leg(args.batch_steps, args.batch_members)
    if not args.no_cpu_baseline:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
