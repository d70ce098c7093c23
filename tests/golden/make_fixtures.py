"""Write the golden fixtures under tests/golden/ from the reference build.

Runs ``oracle/_ref/method_ref`` -- the reference METHOD serial CPU sources,
compiled unmodified from the reference tree by ``oracle/ref/Makefile`` -- on
initial data from ``method_amd.initdata`` and stores inputs and outputs as
compressed .npz files. Only inputs and outputs are stored (no reference
source). Re-run with ``python tests/golden/make_fixtures.py`` in a container
that has the reference tree; the GPU box only reads the .npz files.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from method_amd import initdata as idt  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "method_ref")
NV = {"srmhd": (9, 8, 13), "srrmhd": (14, 11, 17)}


def _run(args, cwd):
    r = subprocess.run([REF_BIN] + args, capture_output=True, text=True, cwd=cwd)
    if r.returncode != 0:
        raise RuntimeError(f"reference run failed: {args}\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    return r.stdout


def _load_state(path, model, shape):
    nc, np_, na = NV[model]
    n = shape[0] * shape[1] * shape[2]
    a = np.fromfile(path)
    assert a.size == (nc + np_ + na) * n, path
    return (a[: nc * n].reshape((nc,) + shape), a[nc * n:(nc + np_) * n].reshape((np_,) + shape),
            a[(nc + np_) * n:].reshape((na,) + shape))


def _grid_args(g: idt.Grid, gamma, cfl, weno, integrator, bc, model, sigma=0.0, cp=0.1):
    return [
        f"model={model}", f"nx={g.nx}", f"ny={g.ny}", f"nz={g.nz}",
        f"xmin={g.xmin!r}", f"xmax={g.xmax!r}", f"ymin={g.ymin!r}", f"ymax={g.ymax!r}",
        f"zmin={g.zmin!r}", f"zmax={g.zmax!r}", f"Ng={g.Ng}", f"gamma={gamma!r}", f"cfl={cfl!r}",
        f"sigma={sigma!r}", f"cp={cp!r}", f"weno={weno}", f"integrator={integrator}",
        f"bcx={bc[0]}", f"bcy={bc[1]}", f"bcz={bc[2]}",
    ]


def run_case(name, g, prims0, model, weno, integrator, bc, steps, dumps, gamma=4.0 / 3.0, cfl=0.4,
             sigma=0.0, cp=0.1, interior_only=False, with_fnet=False, extra=()):
    shape = g.shape
    with tempfile.TemporaryDirectory() as td:
        pin = os.path.join(td, "prims.bin")
        np.ascontiguousarray(prims0).tofile(pin)
        base = _grid_args(g, gamma, cfl, weno, integrator, bc, model, sigma, cp)
        out = os.path.join(td, "out")
        _run(base + ["mode=run", f"steps={steps}", f"prims_in={pin}", f"out={out}",
                     "dump=" + ",".join(str(s) for s in dumps)] + list(extra), td)
        log = open(out + "_log.txt").read().split("\n")
        dts = [float(l.split()[5]) for l in log if l.startswith("step ")]
        fail_line = [l for l in log if l.startswith("c2p_calls")]
        rec = {
            "prims0": prims0, "dts": np.array(dts),
            "meta": np.array([g.nx, g.ny, g.nz, g.Ng, weno, steps]),
            "bounds": np.array([g.xmin, g.xmax, g.ymin, g.ymax, g.zmin, g.zmax]),
            "params": np.array([gamma, cfl, sigma, cp]),
            "periodic": np.array([1 if b == "periodic" else 0 for b in bc]),
            "dumps": np.array(dumps),
        }
        if fail_line:
            rec["c2p_fail"] = np.array([int(fail_line[0].split()[-1])])
        sl = (slice(None), slice(g.Ng, shape[0] - g.Ng),
              slice(g.Ng, shape[1] - g.Ng) if g.ny > 0 else slice(None),
              slice(g.Ng, shape[2] - g.Ng) if g.nz > 0 else slice(None))
        for s in dumps:
            c, p, a = _load_state(f"{out}_step{s}.bin", model, shape)
            if interior_only:
                c, p, a = c[sl], p[sl], a[sl]
            rec[f"cons{s}"], rec[f"prims{s}"], rec[f"aux{s}"] = c.copy(), p.copy(), a.copy()
        if with_fnet:
            _run(base + ["mode=fnet", f"prims_in={pin}", f"out={out}"], td)
            nc = NV[model][0]
            fnet = np.fromfile(out + "_fnet.bin").reshape((nc,) + shape)
            rec["fnet0"] = fnet[sl].copy()
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **rec)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB)")


def c2p_vectors(name, n=2000, seed=7):
    """Reference SRMHD::getPrimitiveVars on a 1-D row of synthetic states.

    Primitive states are drawn over rho in [1e-2, 10], p/rho in [1e-2, 10],
    |v| < 0.9 and |B| < 2 in random directions; cons/aux come from the
    reference primsToAll (step 0 of a run with steps=0), and the C2P guess is
    the true state perturbed by up to 1e-3 relative, as after one small step.
    """
    rng = np.random.default_rng(seed)
    g = idt.Grid(n, 0, 0, 0.0, 1.0, 0.0, 1.0)
    shape = g.shape
    prims = np.zeros((8,) + shape)
    rho = 10 ** rng.uniform(-2, 1, size=shape)
    prims[0] = rho
    prims[4] = rho * 10 ** rng.uniform(-2, 1, size=shape)
    vdir = rng.normal(size=(3,) + shape)
    vdir /= np.linalg.norm(vdir, axis=0)
    prims[1:4] = vdir * rng.uniform(0, 0.9, size=shape)
    bdir = rng.normal(size=(3,) + shape)
    bdir /= np.linalg.norm(bdir, axis=0)
    prims[5:8] = bdir * rng.uniform(0, 2.0, size=shape)
    with tempfile.TemporaryDirectory() as td:
        pin = os.path.join(td, "prims.bin")
        prims.tofile(pin)
        base = _grid_args(g, 4.0 / 3.0, 0.4, 3, "rk2", ("outflow", "outflow", "outflow"), "srmhd")
        out = os.path.join(td, "out")
        _run(base + ["mode=run", "steps=0", f"prims_in={pin}", f"out={out}", "dump=0"], td)
        cons, ptrue, atrue = _load_state(out + "_step0.bin", "srmhd", shape)
        guess_p = ptrue * (1.0 + 1e-3 * rng.uniform(-1, 1, size=ptrue.shape))
        guess_a = atrue * (1.0 + 1e-3 * rng.uniform(-1, 1, size=atrue.shape))
        for arr, nm in ((cons, "cons"), (guess_p, "gp"), (guess_a, "ga")):
            np.ascontiguousarray(arr).tofile(os.path.join(td, nm + ".bin"))
        txt = _run(base + ["mode=c2p", f"cons_in={td}/cons.bin", f"prims_in={td}/gp.bin", f"aux_in={td}/ga.bin",
                           f"out={out}"], td)
        res = np.fromfile(out + "_c2p.bin")
        m = 8 * shape[0]
        p_ref = res[:m].reshape((8,) + shape)
        a_ref = res[m:].reshape((13,) + shape)
    fail = int(txt.split("fail=")[1].split()[0])
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, prims_true=ptrue, aux_true=atrue, cons=cons, guess_prims=guess_p, guess_aux=guess_a,
                        prims_ref=p_ref, aux_ref=a_ref, n=np.array([n]), c2p_fail=np.array([fail]))
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB), reference hybrd failures: {fail}")


def kh_divb_prims(g: idt.Grid):
    """KH shear layers plus Bx = 0.05 sin(2 pi x), whose discrete divergence
    feeds phi, so RKSplit's phi source term is exercised."""
    p = idt.kh_prims(g, seed=77, shear=0.5, bz=0.1)
    x, _, _ = g.centres()
    p[5] = 0.05 * np.sin(2.0 * np.pi * x)[:, None, None]
    return p


def kh_example_save(name, n=32, steps=5):
    """The reference's own Kelvin-Helmholtz example set-up
    (Examples/KelvinHelmholtz/SingleFluidIdealSerial/main.cc: KHInstabilitySingleFluid
    with mag = 1, SRMHD, WENO3, FVS, Periodic, RKSplit, cfl 0.6, sigma 10, cp 1,
    gamma 4/3) at n x n for `steps` updateTime calls, then SerialSaveData::saveAll.
    Stores the state after set-up and the five text files saveAll writes
    after set-up (text0_*) and after the steps (text_*)."""
    g = idt.Grid(n, n, 0, -0.5, 0.5, -1.0, 1.0, 0.0, 1.0)
    files = {}
    base = _grid_args(g, 4.0 / 3.0, 0.6, 3, "rksplit", ("periodic", "periodic", "periodic"), "srmhd", sigma=10.0,
                      cp=1.0)
    for tag, nsteps in (("text0_", 0), ("text_", steps)):
        with tempfile.TemporaryDirectory() as td:
            out = os.path.join(td, "out")
            _run(base + ["mode=run", "init=kh", "mag=1", "endTime=3.0", f"steps={nsteps}", f"out={out}", "dump=0",
                         "save=1"], td)
            if nsteps == 0:
                c0, p0, a0 = _load_state(out + "_step0.bin", "srmhd", g.shape)
            for key, rel in (("cons", "Conserved/cons.dat"), ("prims", "Primitive/prims.dat"),
                             ("aux", "Auxiliary/aux.dat"), ("domain", "Domain/domain.dat"),
                             ("constants", "Constants/constants.dat")):
                with open(os.path.join(td, "Data", "Final", rel), "rb") as f:
                    files[tag + key] = np.frombuffer(f.read(), dtype=np.uint8)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, prims0=p0, cons0=c0, aux0=a0, meta=np.array([n, steps]), **files)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB)")


def init_case(name, g, init_args, model, integrator, bc, steps, gamma, cfl, sigma=0.0, cp=0.1, weno=3):
    """A reference example's set-up: one of the reference's own initial states
    (init=...), run for `steps` updateTime calls. Stores the state after
    Simulation::set (step 0) and after the steps."""
    shape = g.shape
    with tempfile.TemporaryDirectory() as td:
        base = _grid_args(g, gamma, cfl, weno, integrator, bc, model, sigma, cp)
        out = os.path.join(td, "out")
        _run(base + ["mode=run", f"steps={steps}", f"out={out}", f"dump=0,{steps}"] + init_args, td)
        log = open(out + "_log.txt").read().split("\n")
        dts = [float(l.split()[5]) for l in log if l.startswith("step ")]
        fail = [int(l.split()[-1]) for l in log if l.startswith("c2p_calls")]
        rec = {"dts": np.array(dts), "meta": np.array([g.nx, g.ny, g.nz, g.Ng, weno, steps]),
               "bounds": np.array([g.xmin, g.xmax, g.ymin, g.ymax, g.zmin, g.zmax]),
               "params": np.array([gamma, cfl, sigma, cp]),
               "periodic": np.array([1 if b == "periodic" else 0 for b in bc]), "c2p_fail": np.array(fail)}
        for st in (0, steps):
            c, p, a = _load_state(f"{out}_step{st}.bin", model, shape)
            rec[f"cons{st}"], rec[f"prims{st}"], rec[f"aux{st}"] = c, p, a
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **rec)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB), reference C2P failures: {fail}")


def reference_examples():
    """The reference's serial examples (Examples/*/main.cc) at their own
    parameters, for a fixed number of steps."""
    g1 = idt.Grid(100, 0, 0, 0.0, 1.0, -1.0, 1.0, 0.0, 1.0)
    out3 = ("outflow", "outflow", "outflow")
    per3 = ("periodic", "periodic", "periodic")
    # BrioWuShockTube/SingleFluidIdealSerial: SRMHD, gamma 2, cfl 0.4, Outflow, RKSplit
    init_case("briowu_100_rksplit", g1, ["init=briowu", "endTime=0.4"], "srmhd", "rksplit", out3, 60, 2.0, 0.4)
    # BrioWuShockTube/SingleFluidResistiveSerial: SRRMHD, gamma 2, sigma 10, Outflow, SSP2
    init_case("briowu_100_ssp2", g1, ["init=briowu", "endTime=0.4"], "srrmhd", "ssp2", out3, 60, 2.0, 0.4,
              sigma=10.0)
    # OrszagTangVortex/SingleFluidIdealSerial at 64^2: gamma 5/3, cfl 0.6, sigma 10, cp 1, Periodic, RKSplit
    g2 = idt.Grid(64, 64, 0, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    init_case("ot_64_rksplit", g2, ["init=ot", "endTime=1.0"], "srmhd", "rksplit", per3, 10, 5.0 / 3.0, 0.6,
              sigma=10.0, cp=1.0)
    # KelvinHelmholtz/SingleFluidIdealRandomSerial at 64^2, seed 10
    init_case("khrandom_64_rksplit", g2, ["init=khrandom", "mag=1", "seed=10", "endTime=3.0"], "srmhd", "rksplit",
              per3, 10, 4.0 / 3.0, 0.6, sigma=10.0, cp=1.0)


def rkplus_cases():
    """The RKPlus integrators (RKPlus.cc): RK2B, RK3, RK4 and RK4_10 on the
    divergent-Bx KH state (so phi's source acts) and two resistive set-ups
    with a conductivity the explicit source step resolves (sigma dt < 0.1)."""
    gs = idt.kh2d_grid(32)
    for integ in ("rk2b", "rk3", "rk4", "rk4_10"):
        run_case(f"kh2d_32_weno3_{integ}", gs, kh_divb_prims(gs), "srmhd", 3, integ,
                 ("periodic", "outflow", "periodic"), 4, [1, 4], cp=0.1, interior_only=True)
    g1 = idt.Grid(100, 0, 0, 0.0, 1.0, -1.0, 1.0, 0.0, 1.0)
    init_case("briowu_100_srrmhd_rk3", g1, ["init=briowu", "endTime=0.4"], "srrmhd", "rk3",
              ("outflow", "outflow", "outflow"), 20, 2.0, 0.4, sigma=10.0)
    go = idt.ot2d_grid(32)
    run_case("ot2d_32_weno3_rk4", go, idt.ot_prims_resistive(go), "srrmhd", 3, "rk4",
             ("periodic", "periodic", "periodic"), 3, [1, 3], cfl=0.5, sigma=10.0, cp=0.1, interior_only=True)


def regime_source_case(name, g, init_args, weno, bc, gamma, sigma, prims0=None):
    """REGIME::sourceExtension (REGIME.cc:83-138) of a set-up state: the
    reference's post-set prims and the source over the interior cells."""
    shape = g.shape
    with tempfile.TemporaryDirectory() as td:
        base = _grid_args(g, gamma, 0.4, weno, "rksplit", bc, "srmhd", sigma, 0.1)
        out = os.path.join(td, "out")
        extra = list(init_args)
        if prims0 is not None:
            pin = os.path.join(td, "prims.bin")
            np.ascontiguousarray(prims0).tofile(pin)
            extra.append(f"prims_in={pin}")
        _run(base + ["mode=regime", f"out={out}", "regime=1"] + extra, td)
        src = np.fromfile(out + "_regime.bin").reshape((9,) + shape)
        c, p, a = _load_state(out + "_step0.bin", "srmhd", shape)
    sl = (slice(None), slice(g.Ng, shape[0] - g.Ng), slice(g.Ng, shape[1] - g.Ng) if g.ny > 0 else slice(None),
          slice(g.Ng, shape[2] - g.Ng) if g.nz > 0 else slice(None))
    rec = {"prims0": p, "src": src[sl].copy(), "meta": np.array([g.nx, g.ny, g.nz, g.Ng, weno, 0]),
           "bounds": np.array([g.xmin, g.xmax, g.ymin, g.ymax, g.zmin, g.zmax]),
           "params": np.array([gamma, 0.4, sigma, 0.1]),
           "periodic": np.array([1 if b == "periodic" else 0 for b in bc])}
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **rec)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB)")


def regime_cases():
    """The REGIME extension (REGIME.cc): its source on four set-ups, and runs
    of the reference's two REGIME examples (Examples/REGIME/*Serial/main.cc,
    RKSplit) plus 2-D WENO5 with RK2B and 3-D with RKSplit."""
    out3 = ("outflow", "outflow", "outflow")
    mixed = ("periodic", "outflow", "periodic")
    gcs = idt.Grid(128, 0, 0, -3.0, 3.0, -1.0, 1.0, -1.5, 1.5)
    gbw = idt.Grid(128, 0, 0, 0.0, 1.0, -1.0, 1.0, -1.5, 1.5)
    g2 = idt.kh2d_grid(32)
    g3 = idt.kh3d_grid(12)
    p2 = idt.kh_prims(g2, seed=3, shear=0.5, bz=0.3)
    p3 = idt.kh_prims(g3, seed=4, shear=0.5, bz=0.3)
    regime_source_case("regime_src_currentsheet_128", gcs, ["init=currentsheet"], 3, out3, 2.0, 1e4)
    regime_source_case("regime_src_briowu_128", gbw, ["init=briowu"], 3, out3, 2.0, 50.0)
    regime_source_case("regime_src_kh2d_32_weno5", g2, [], 5, mixed, 4.0 / 3.0, 100.0, p2)
    regime_source_case("regime_src_kh3d_12_weno3", g3, [], 3, mixed, 4.0 / 3.0, 100.0, p3)
    # Examples/REGIME/CurrentSheetSerial: gamma 2, sigma 1e4, cfl 0.2, WENO3, Outflow, RKSplit + REGIME
    init_case("currentsheet_128_regime_rksplit", gcs, ["init=currentsheet", "regime=1", "endTime=7.0"], "srmhd",
              "rksplit", out3, 20, 2.0, 0.2, sigma=1e4)
    # Examples/REGIME/BrioWuSerial: gamma 2, sigma 50, cfl 0.15, WENO3, Outflow, RKSplit + REGIME
    init_case("briowu_128_regime_rksplit", gbw, ["init=briowu", "regime=1", "endTime=0.4"], "srmhd", "rksplit",
              out3, 20, 2.0, 0.15, sigma=50.0)
    run_case("kh2d_32_weno5_regime_rk2b", g2, p2, "srmhd", 5, "rk2b", mixed, 3, [1, 3], sigma=100.0,
             interior_only=True, extra=["regime=1"])
    run_case("kh3d_12_weno3_regime_rksplit", g3, p3, "srmhd", 3, "rksplit", mixed, 2, [2], sigma=100.0,
             interior_only=True, extra=["regime=1"])


def regime_rkplus_cases():
    """REGIME with the remaining RKPlus schemes (RK3, RK4 with its stored rhs,
    RK4_10 whose stage after the second preparation forms the extension
    source from the step's own primitives), 2-D KH WENO5, 3 steps."""
    mixed = ("periodic", "outflow", "periodic")
    g2 = idt.kh2d_grid(32)
    p2 = idt.kh_prims(g2, seed=3, shear=0.5, bz=0.3)
    for integ in ("rk3", "rk4", "rk4_10"):
        run_case(f"kh2d_32_weno5_regime_{integ}", g2, p2, "srmhd", 5, integ, mixed, 3, [1, 3], sigma=100.0,
                 interior_only=True, extra=["regime=1"])


def evolve_steps(g, cfl, end_time):
    """The number of updateTime calls Simulation::evolve makes to reach
    end_time on a 2-D grid (simulation.cc:104-133: light-speed dt, slow start)."""
    base = cfl * min((g.xmax - g.xmin) / g.nx, (g.ymax - g.ymin) / g.ny) / np.sqrt(2.0)
    t, it = 0.0, 0
    while t < end_time:
        dt = base * (0.1 if it < 15 else (0.25 if it < 30 else 1.0))
        if t + dt > end_time:
            dt = end_time - t
        t += dt
        it += 1
    return it


def reference_test_cases():
    """The set-ups of the reference's RK2 and RKSplit serial-output tests
    (Tests/CPU/Src/test_rk2.cc, test_rkSplit.cc): 40x40 on [0,1]^2, WENO3,
    Orszag-Tang (cfl 0.6, gamma 2, one updateTime), random KH (seed 10, cfl 0.6,
    gamma 4/3, sigma 10, cp 1, evolve to 0.004) and Brio-Wu (the Data defaults:
    cfl 0.5, gamma 5/3, sigma 1e3, cp 0.1, evolve to 0.004), under periodic,
    outflow and Flow boundaries."""
    g = idt.Grid(40, 40, 0, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    per = ("periodic", "periodic", "periodic")
    out = ("outflow", "outflow", "outflow")
    flow = ("periodic", "outflow", "outflow")
    kh = ["init=khrandom", "mag=1", "seed=10", "endTime=0.004"]
    bw = ["init=briowu", "endTime=0.004"]
    ot = ["init=ot", "endTime=0.004"]
    nkh, nbw = evolve_steps(g, 0.6, 0.004), evolve_steps(g, 0.5, 0.004)
    init_case("reftest_rk2_ot40_periodic", g, ot, "srmhd", "rk2", per, 1, 2.0, 0.6)
    init_case("reftest_rk2_ot40_outflow", g, ot, "srmhd", "rk2", out, 1, 2.0, 0.6)
    init_case("reftest_rk2_khrandom40_flow", g, kh, "srmhd", "rk2", flow, nkh, 4.0 / 3.0, 0.6, sigma=10.0, cp=1.0)
    init_case("reftest_rk2_briowu40_periodic", g, bw, "srmhd", "rk2", per, nbw, 5.0 / 3.0, 0.5, sigma=1e3, cp=0.1)
    # test_rkSplit.cc's Orszag-Tang cases run RK2 with the KH parameters to 0.004
    init_case("reftest_rk2_ot40_outflow_evolve", g, ot, "srmhd", "rk2", out, nkh, 4.0 / 3.0, 0.6, sigma=10.0, cp=1.0)
    init_case("reftest_rksplit_khrandom40_periodic", g, kh, "srmhd", "rksplit", per, nkh, 4.0 / 3.0, 0.6, sigma=10.0,
              cp=1.0)
    init_case("reftest_rksplit_briowu40_flow", g, bw, "srmhd", "rksplit", flow, nbw, 5.0 / 3.0, 0.5, sigma=1e3, cp=0.1)
    # Tests/CPU/Src/test_imex.cc SSP2FlowKHSingleFluid: resistive KH 64x16 on [-0.5,0.5]x[-1,1], sigma 0,
    # cfl 0.5, gamma 4/3, Flow, SSP2, evolve to 0.05
    gk = idt.Grid(64, 16, 0, -0.5, 0.5, -1.0, 1.0, 0.0, 1.0)
    init_case("reftest_ssp2_kh64x16_flow", gk, ["init=kh", "endTime=0.05"], "srrmhd", "ssp2", flow,
              evolve_steps(gk, 0.5, 0.05), 4.0 / 3.0, 0.5, sigma=0.0, cp=0.1)


FULL_PLANE_VARS = ("D", "Sx", "tau", "By")  # cons fields kept on the REGIME workload's plane
FULL_STEPS = 10  # reference steps of the full-size fixtures


def volume_stats(a):
    """Per-field min, max, mean and L2 norm over the interior of a [var][i][j][k] array."""
    out = np.zeros((a.shape[0], 4))
    for v in range(a.shape[0]):
        x = a[v]
        out[v] = (x.min(), x.max(), x.mean(dtype=np.float64), np.sqrt(np.sum(x * x, dtype=np.float64)))
    return out


def full_size_case(steps=FULL_STEPS):
    """KH 3-D at its benchmark size (256^3, WENO5, RK2, cfl 0.4, periodic x,z /
    outflow y, seed 1234: the state bench.py times), `steps` steps of the
    reference build. The 4.4 GB result is kept as the dts, the x-plane through
    the middle of every cons and prim field, and every cons and prim field's
    min, max, mean and L2 norm over the interior."""
    g = idt.kh3d_grid(256)
    prims0 = idt.kh_prims(g, seed=1234)
    bc = ("periodic", "outflow", "periodic")
    with tempfile.TemporaryDirectory(dir=HERE) as td:
        pin = os.path.join(td, "prims.bin")
        prims0.tofile(pin)
        del prims0
        base = _grid_args(g, 4.0 / 3.0, 0.4, 5, "rk2", bc, "srmhd")
        out = os.path.join(td, "out")
        _run(base + ["mode=run", f"steps={steps}", f"prims_in={pin}", f"out={out}", f"dump={steps}"], td)
        log = open(out + "_log.txt").read().split("\n")
        dts = [float(l.split()[5]) for l in log if l.startswith("step ")]
        shape = g.shape
        n = shape[0] * shape[1] * shape[2]
        raw = np.memmap(f"{out}_step{steps}.bin", dtype=np.float64, mode="r")
        nc, np_, _ = NV["srmhd"]
        inner = (slice(g.Ng, shape[0] - g.Ng), slice(g.Ng, shape[1] - g.Ng), slice(g.Ng, shape[2] - g.Ng))
        cons = np.array(raw[: nc * n].reshape((nc,) + shape)[(slice(None),) + inner])
        cs = volume_stats(cons)
        cplane = cons[:, 128].copy()
        del cons
        prims = np.array(raw[nc * n:(nc + np_) * n].reshape((np_,) + shape)[(slice(None),) + inner])
        ps = volume_stats(prims)
        pplane = prims[:, 128].copy()
        del prims, raw
    path = os.path.join(HERE, "kh3d_256_weno5_rk2_full.npz")
    np.savez_compressed(path, steps=np.array(steps), dts=np.array(dts), cons_stats=cs, prims_stats=ps,
                        cons_plane_x128=cplane, prims_plane_x128=pplane)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB), dts {dts}")


def full_size_resistive_case(steps=FULL_STEPS):
    """The resistive workload at its own size (Orszag-Tang 1024^2, sigma 1e3,
    SSP2, WENO3, cfl 0.5, periodic: the state bench.py times), `steps` steps of
    the reference build: the dts, the middle row along each axis of every cons
    and prim field, and every cons and prim field's min, max, mean and L2 norm."""
    g = idt.ot2d_grid(1024)
    prims0 = idt.ot_prims_resistive(g)
    per = ("periodic", "periodic", "periodic")
    with tempfile.TemporaryDirectory(dir=HERE) as td:
        pin = os.path.join(td, "prims.bin")
        prims0.tofile(pin)
        base = _grid_args(g, 4.0 / 3.0, 0.5, 3, "ssp2", per, "srrmhd", sigma=1e3, cp=0.1)
        out = os.path.join(td, "out")
        _run(base + ["mode=run", f"steps={steps}", f"prims_in={pin}", f"out={out}", f"dump={steps}"], td)
        log = open(out + "_log.txt").read().split("\n")
        dts = [float(l.split()[5]) for l in log if l.startswith("step ")]
        c, p, _ = _load_state(f"{out}_step{steps}.bin", "srrmhd", g.shape)
        inner = (slice(None), slice(g.Ng, -g.Ng), slice(g.Ng, -g.Ng), slice(None))
        c, p = c[inner], p[inner]
        rec = {"steps": np.array(steps), "dts": np.array(dts), "cons_stats": volume_stats(c),
               "prims_stats": volume_stats(p),
               "row_x512": c[:, 512, :, 0].copy(), "row_y512": c[:, :, 512, 0].copy(),
               "prow_x512": p[:, 512, :, 0].copy(), "prow_y512": p[:, :, 512, 0].copy()}
    path = os.path.join(HERE, "ot2d_1024_srrmhd_ssp2_full.npz")
    np.savez_compressed(path, **rec)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB), dts {dts}")


def large_regime_case(n=128):
    """The REGIME workload's set-up (KH 3-D, bz 0.1, WENO5, RKSplit with
    REGIME, sigma 100) at 128^3, one step of the reference build: the dt, the
    middle x-plane of four cons fields and every cons and prim field's min,
    max, mean and L2 norm. (At 256^3 the reference's REGIME arrays alone take
    about 60 GB, more than this machine has.)"""
    g = idt.kh3d_grid(n)
    prims0 = idt.kh_prims(g, seed=1234, bz=0.1)
    bc = ("periodic", "outflow", "periodic")
    with tempfile.TemporaryDirectory(dir=HERE) as td:
        pin = os.path.join(td, "prims.bin")
        prims0.tofile(pin)
        del prims0
        base = _grid_args(g, 4.0 / 3.0, 0.4, 5, "rksplit", bc, "srmhd", sigma=100.0)
        out = os.path.join(td, "out")
        _run(base + ["mode=run", "steps=1", f"prims_in={pin}", f"out={out}", "dump=1", "regime=1"], td)
        log = open(out + "_log.txt").read().split("\n")
        dts = [float(l.split()[5]) for l in log if l.startswith("step ")]
        shape = g.shape
        n = shape[0] * shape[1] * shape[2]
        raw = np.memmap(f"{out}_step1.bin", dtype=np.float64, mode="r")
        nc, np_, _ = NV["srmhd"]
        inner = (slice(g.Ng, shape[0] - g.Ng), slice(g.Ng, shape[1] - g.Ng), slice(g.Ng, shape[2] - g.Ng))
        cons = np.array(raw[: nc * n].reshape((nc,) + shape)[(slice(None),) + inner])
        cs = volume_stats(cons)
        names = ["D", "Sx", "Sy", "Sz", "tau", "Bx", "By", "Bz", "phi"]
        plane = np.stack([cons[names.index(v), g.nx // 2] for v in FULL_PLANE_VARS])
        del cons
        prims = np.array(raw[nc * n:(nc + np_) * n].reshape((np_,) + shape)[(slice(None),) + inner])
        ps = volume_stats(prims)
        del prims, raw
    path = os.path.join(HERE, f"kh3d_{g.nx}_weno5_regime_rksplit_large.npz")
    np.savez_compressed(path, dts=np.array(dts), cons_stats=cs, prims_stats=ps, plane_mid_x=plane,
                        plane_vars=np.array(FULL_PLANE_VARS))
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB), dts {dts}")


# (nx, ny, shear, bz, seed) of the batch member checked against the reference at its real size
BATCH_MEMBER = (768, 384, 0.6, 0.2, 77)


def batch_member_case():
    """One member of the batched-ensemble workload at a real ensemble size
    (KH 2-D 768 x 384, WENO5, RK2, cfl 0.4, periodic x / outflow y), 3 steps
    of the reference build: the dts, the middle row along each axis of every
    cons field, and every cons and prim field's min, max, mean and L2 norm."""
    nx, ny, shear, bz, seed = BATCH_MEMBER
    g = idt.Grid(nx, ny, 0, -0.5, 0.5, -1.0, 1.0)
    prims0 = idt.kh_prims(g, seed=seed, shear=shear, bz=bz)
    with tempfile.TemporaryDirectory(dir=HERE) as td:
        pin = os.path.join(td, "prims.bin")
        prims0.tofile(pin)
        base = _grid_args(g, 4.0 / 3.0, 0.4, 5, "rk2", ("periodic", "outflow", "periodic"), "srmhd")
        out = os.path.join(td, "out")
        _run(base + ["mode=run", "steps=3", f"prims_in={pin}", f"out={out}", "dump=3"], td)
        log = open(out + "_log.txt").read().split("\n")
        dts = [float(l.split()[5]) for l in log if l.startswith("step ")]
        c, p, _ = _load_state(f"{out}_step3.bin", "srmhd", g.shape)
        inner = (slice(None), slice(g.Ng, -g.Ng), slice(g.Ng, -g.Ng), slice(None))
        c, p = c[inner], p[inner]
        rec = {"member": np.array(BATCH_MEMBER, dtype=np.float64), "dts": np.array(dts),
               "cons_stats": volume_stats(c), "prims_stats": volume_stats(p),
               "row_x": c[:, nx // 2, :, 0].copy(), "row_y": c[:, :, ny // 2, 0].copy()}
    path = os.path.join(HERE, f"batch_member_{nx}x{ny}_weno5_rk2.npz")
    np.savez_compressed(path, **rec)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB), dts {dts}")


def frecon_case(name, g, init_args, weno, bc, gamma, cfl):
    """FVS::fluxReconstruction of a reference set-up state along every axis,
    as Tests/CPU/Src/test_fvs.cc:56-170 calls it after Model::fluxVector: the
    post-set prims and cons, and per axis the flux vector f and frecon, over
    the whole padded grid."""
    with tempfile.TemporaryDirectory() as td:
        base = _grid_args(g, gamma, cfl, weno, "rk2", bc, "srmhd")
        out = os.path.join(td, "out")
        _run(base + ["mode=frecon", f"out={out}"] + init_args, td)
        c, p, _ = _load_state(f"{out}_step0.bin", "srmhd", g.shape)
        n = c.size
        rec = {"prims0": p, "cons0": c}
        for d in range(3):
            raw = np.fromfile(f"{out}_frecon{d}.bin")
            rec[f"f{d}"] = raw[:n].reshape(c.shape)
            rec[f"frecon{d}"] = raw[n:].reshape(c.shape)
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **rec)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB)")


def frecon_cases():
    # test_fvs.cc's set-up (OTVortexSingleFluid, SRMHD, Weno3, Outflow, the
    # Data defaults gamma 5/3 and cfl 0.5) on a small ragged 3-D grid with Ng 3
    g = idt.Grid(10, 12, 8, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0, Ng=3)
    frecon_case("frecon_ot3d_weno3", g, ["init=ot"], 3, ("outflow", "outflow", "outflow"), 5.0 / 3.0, 0.5)
    # WENO5 on the reference's KH example state, 2-D, Flow
    g2 = idt.Grid(16, 12, 0, -0.5, 0.5, -1.0, 1.0)
    frecon_case("frecon_kh2d_weno5", g2, ["init=kh", "mag=1"], 5, ("periodic", "outflow", "periodic"), 4.0 / 3.0,
                0.4)


def srrmhd3d_prims(g: idt.Grid, bc):
    """The resistive Orszag-Tang profile in x-y with vz = 0.05 sin 2 pi z and
    Bz = 0.1 added and E = -v x B of the whole 3-D state (the ideal field, as
    the 2-D profile has), ghosts holding the images the boundaries give."""
    from oracle import oracle as orc

    p = idt.ot_prims_resistive(g)
    _, _, z = g.centres()
    p[3] = 0.05 * np.sin(2 * np.pi * z)[None, None, :]
    p[7] = 0.1
    v, B = p[1:4], p[5:8]
    p[8:11] = -np.cross(v, B, axis=0)
    periodic = tuple(1 if b == "periodic" else 0 for b in bc)
    o = orc.Oracle(orc.Case(nx=g.nx, ny=g.ny, nz=g.nz, xmin=g.xmin, xmax=g.xmax, ymin=g.ymin, ymax=g.ymax,
                            zmin=g.zmin, zmax=g.zmax, Ng=g.Ng, model=orc.SRRMHD, periodic=periodic))
    p = np.ascontiguousarray(p)
    o.apply_bcs(p)
    return p


def srrmhd3d_cases():
    """3-D resistive runs, so the z-direction flux of srrmhd.cc:163-171 (Sy
    written twice, Sz left from the y call) is checked against the reference
    itself: SSP2 (sigma 50) and RKSplit (sigma 10), periodic and Flow, on a
    ragged 12 x 14 x 10 grid, 3 steps. Interior cells are stored."""
    g = idt.Grid(12, 14, 10, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    per, flow = ("periodic",) * 3, ("periodic", "outflow", "outflow")
    for integ, sigma, weno in (("ssp2", 50.0, 3), ("rksplit", 10.0, 5)):
        for bc, bname in ((per, "periodic"), (flow, "flow")):
            prims0 = srrmhd3d_prims(g, bc)
            name = f"srrmhd3d_12x14x10_{integ}_{bname}"
            with tempfile.TemporaryDirectory() as td:
                pin = os.path.join(td, "prims.bin")
                prims0.tofile(pin)
                base = _grid_args(g, 4.0 / 3.0, 0.5, weno, integ, bc, "srrmhd", sigma=sigma, cp=0.1)
                out = os.path.join(td, "out")
                _run(base + ["mode=run", "steps=3", f"prims_in={pin}", f"out={out}", "dump=3"], td)
                log = open(out + "_log.txt").read().split("\n")
                dts = [float(l.split()[5]) for l in log if l.startswith("step ")]
                c, p, _ = _load_state(f"{out}_step3.bin", "srrmhd", g.shape)
            inner = (slice(None), slice(g.Ng, -g.Ng), slice(g.Ng, -g.Ng), slice(g.Ng, -g.Ng))
            rec = {"prims0": prims0[inner], "cons3": c[inner], "prims3": p[inner], "dts": np.array(dts),
                   "sigma": np.array(sigma), "weno": np.array(weno), "periodic": np.array([b == "periodic" for b in bc])}
            path = os.path.join(HERE, f"{name}.npz")
            np.savez_compressed(path, **rec)
            print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB), dts {dts}")


GROUPS = {"frecon": frecon_cases, "srrmhd3d": srrmhd3d_cases, "regime_rkplus": regime_rkplus_cases, "reference_tests": reference_test_cases, "batch_member": batch_member_case,
          "full_size": full_size_case, "full_size_resistive": full_size_resistive_case,
          "large_regime": large_regime_case}


def main():
    if not os.path.exists(REF_BIN):
        raise SystemExit(f"{REF_BIN} missing: build it with `make -C oracle/ref`")
    if len(sys.argv) > 1:
        # only the named groups
        for name in sys.argv[1:]:
            GROUPS[name]()
        return
    # Ideal SRMHD 2-D Kelvin-Helmholtz, 64x64, WENO3, RK2, cfl 0.4, 10 steps,
    # periodic x / outflow y (the reference's Flow), seed 1234.
    g = idt.kh2d_grid(64)
    run_case("kh2d_64_weno3_rk2", g, idt.kh_prims(g, seed=1234), "srmhd", 3, "rk2",
             ("periodic", "outflow", "periodic"), 10, [1, 10], with_fnet=True)
    # Ideal SRMHD 3-D Kelvin-Helmholtz, 16^3, WENO5, RK2, periodic x,z / outflow y.
    g3 = idt.kh3d_grid(16)
    run_case("kh3d_16_weno5_rk2", g3, idt.kh_prims(g3, seed=99), "srmhd", 5, "rk2",
             ("periodic", "outflow", "periodic"), 3, [1, 3], interior_only=True, with_fnet=True)
    # Resistive SRRMHD 2-D Orszag-Tang, 32x32, WENO3, SSP2, sigma=1e3, cp 0.1, cfl 0.5.
    go = idt.ot2d_grid(32)
    run_case("ot2d_32_weno3_ssp2", go, idt.ot_prims_resistive(go), "srrmhd", 3, "ssp2",
             ("periodic", "periodic", "periodic"), 3, [1, 3], cfl=0.5, sigma=1e3, cp=0.1, with_fnet=True)
    c2p_vectors("srmhd_c2p_vectors")
    # RKSplit (rkSplit.cc:31-53): KH 48x48 with a divergent Bx so phi's source
    # acts, cp 0.1, WENO3, periodic x / outflow y, 8 steps.
    gs = idt.kh2d_grid(48)
    run_case("kh2d_48_weno3_rksplit", gs, kh_divb_prims(gs), "srmhd", 3, "rksplit",
             ("periodic", "outflow", "periodic"), 8, [8], cp=0.1)
    kh_example_save("kh_example_32_rksplit_save")
    reference_examples()
    rkplus_cases()
    regime_cases()
    regime_rkplus_cases()
    reference_test_cases()
    full_size_case()
    full_size_resistive_case()
    large_regime_case()
    batch_member_case()
    frecon_cases()
    srrmhd3d_cases()


if __name__ == "__main__":
    main()
