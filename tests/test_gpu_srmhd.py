"""Parity of the HIP SRMHD time step with the reference (golden fixtures) and
with the CPU oracle, through the C ABI. Needs an MI355X."""
import sys

import numpy as np
import pytest

from method_amd import engine as eng
from method_amd import initdata as idt
from oracle import oracle as orc
from tests import parity

pytestmark = pytest.mark.gpu

KH2D = dict(nx=64, ny=64, nz=0, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0)
KH3D16 = dict(nx=16, ny=16, nz=16, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, zmin=0.0, zmax=1.0)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def engine(weno=3, bc=(eng.PERIODIC, eng.OUTFLOW, eng.PERIODIC), quirks=1, **grid):
    return eng.Engine(eng.Config(weno_order=weno, bc=bc, quirks=quirks, **grid))


def oracle_for(weno=3, periodic=(1, 0, 1), quirks=1, **grid):
    return orc.Oracle(orc.Case(weno_order=weno, periodic=periodic, quirks=quirks, **grid))


def check_rows(rows, tol, use_group=False):
    for n, dm, fm, rf, rg, pw in rows:
        err = rg if use_group else rf
        assert err <= tol, f"{n}: {err:.3e} > {tol:.1e}"
        if fm == 0.0:
            assert dm == 0.0, f"{n} must stay exactly zero"


def test_set_matches_oracle_including_ghosts(golden):
    z = golden("kh2d_64_weno3_rk2")
    say("gpu: set_prims KH 2-D 64^2")
    with engine(**KH2D) as e:
        e.set_prims(z["prims0"])
        c, p, a = e.download()
    oc, op, oa = oracle_for(**KH2D).set(z["prims0"])
    for got, ref, names in ((c, oc, parity.SRMHD_CONS), (p, op, parity.SRMHD_PRIMS), (a, oa, parity.SRMHD_AUX)):
        rows = parity.field_errors(got, ref, names)  # whole padded grid, ghosts included
        say(parity.report("gpu vs oracle after Simulation::set (padded grid)", rows))
        check_rows(rows, 1e-14, use_group=True)


@pytest.mark.parametrize("dir", [0, 1])
def test_flux_vector_matches_oracle_2d(golden, dir):
    z = golden("kh2d_64_weno3_rk2")
    say(f"gpu: flux_vector dir={dir}")
    with engine(**KH2D) as e:
        e.set_prims(z["prims0"])
        f = e.flux_vector(dir)
    o = oracle_for(**KH2D)
    c, p, a = o.set(z["prims0"])
    ref = o.flux_vector(c, p, a, dir)
    rows = parity.field_errors(f, ref, parity.SRMHD_CONS)
    check_rows(rows, 1e-13, use_group=True)


@pytest.mark.parametrize("quirks", [1, 0])
def test_flux_vector_z_phi_quirk(quirks):
    g = idt.kh3d_grid(16)
    prims0 = idt.kh_prims(g, seed=5)
    prims0[5] = 0.03  # Bx != 0 so B has all components
    prims0[6] = -0.02
    say(f"gpu: flux_vector z quirks={quirks}")
    with engine(weno=5, quirks=quirks, **KH3D16) as e:
        e.set_prims(prims0)
        f = e.flux_vector(2)
    o = oracle_for(weno=5, quirks=quirks, **KH3D16)
    c, p, a = o.set(prims0)
    ref = o.flux_vector(c, p, a, 2)
    check_rows(parity.field_errors(f, ref, parity.SRMHD_CONS), 1e-13, use_group=True)
    assert np.allclose(f[8], p[6] if quirks else p[7], rtol=0, atol=0)


# FVS::F is compared at 1e-9 of the field scale, not at rounding level: the
# reference forms the WENO smoothness indicators as expanded monomials
# (wenoUpwinds.cc:48-67), whose cancellation leaves ~1e-15 absolute noise in
# beta; where beta^2 is near eps = 1e-16 (variations ~1e-4, as in these
# states) that moves its weights by ~1e-7 relative, i.e. face fluxes by ~1e-11.
# The engine uses the factored Jiang-Shu form, accurate to rounding.
F_TOL = 1e-9


def test_fnet_matches_reference_2d_weno3(golden):
    z = golden("kh2d_64_weno3_rk2")
    say("gpu: fnet KH 2-D WENO3")
    with engine(**KH2D) as e:
        e.set_prims(z["prims0"])
        f = e.fnet()
    rows = parity.field_errors(parity.interior(f, 4, 2), z["fnet0"], parity.SRMHD_CONS)
    say(parity.report("gpu vs reference FVS::F, KH 2-D WENO3", rows))
    check_rows(rows, F_TOL, use_group=True)
    # ghost cells of fnet stay zero, as FVS::F writes the interior only
    assert np.all(f[:, :4] == 0) and np.all(f[:, -4:] == 0)


def test_fnet_matches_reference_3d_weno5(golden):
    z = golden("kh3d_16_weno5_rk2")
    say("gpu: fnet KH 3-D WENO5")
    with engine(weno=5, **KH3D16) as e:
        e.set_prims(z["prims0"])
        f = e.fnet()
    rows = parity.field_errors(parity.interior(f, 4, 3), z["fnet0"], parity.SRMHD_CONS)
    say(parity.report("gpu vs reference FVS::F, KH 3-D WENO5", rows))
    check_rows(rows, F_TOL, use_group=True)


def test_c2p_vectors_round_trip_and_reference(golden):
    z = golden("srmhd_c2p_vectors")
    n = int(z["n"][0])
    # a 2-D strip n x 4 holding the 1-D vectors in its first row is not the
    # same grid; use an n x Ng grid whose every row repeats the vectors
    ny = 4
    shape = (n + 8, ny + 8, 1)

    def widen(a):
        out = np.zeros((a.shape[0],) + shape)
        out[:, :, :, 0] = a[:, :, 0, 0][:, :, None]
        return out

    say("gpu: c2p vectors")
    with eng.Engine(eng.Config(nx=n, ny=ny, nz=0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0)) as e:
        e.upload(cons=widen(z["cons"]), prims=widen(z["guess_prims"]), aux=widen(z["guess_aux"]))
        rc, st = e.c2p()
        _, p, _ = e.download()
    assert rc == 0 and st["failures"] == 0 and st["cells"] == n * ny
    got = p[:, 4:-4, 4, 0]
    pt, pr = z["prims_true"][:, 4:-4, 0, 0], z["prims_ref"][:, 4:-4, 0, 0]
    scale = np.abs(pt).max(axis=1, keepdims=True)
    say(f"c2p round trip max rel {np.max(np.abs(got - pt) / scale):.3e}, vs reference hybrd "
        f"{np.max(np.abs(got - pr) / scale):.3e}, mean Newton steps {st['iterations'] / st['cells']:.3f}")
    assert (np.abs(got - pt) / scale).max() < 1e-11
    assert (np.abs(got - pr) / scale).max() < 1e-7


def test_kh2d_ten_steps_match_reference(golden):
    """The parity gate: KH 2-D 64^2 WENO3 RK2 cfl 0.4, 10 updateTime calls."""
    z = golden("kh2d_64_weno3_rk2")
    say("gpu: KH 2-D 64^2 10 steps vs reference")
    with engine(**KH2D) as e:
        e.set_prims(z["prims0"])
        total = {"cells": 0, "iterations": 0, "failures": 0}
        for s in range(10):
            rc, st = e.update_time(stats=True)
            assert rc == 0
            for k in total:
                total[k] += st[k]
            t, dt, it = e.time()
            assert dt == z["dts"][s]
        c, p, a = e.download()
    say(f"C2P: {total['cells']} cells, {total['iterations'] / total['cells']:.3f} Newton steps/cell, "
        f"{total['failures']} failures")
    for got, ref, names in ((c, z["cons10"], parity.SRMHD_CONS), (p, z["prims10"], parity.SRMHD_PRIMS),
                            (a, z["aux10"], parity.SRMHD_AUX)):
        rows = parity.field_errors(parity.interior(got, 4, 2), parity.interior(ref, 4, 2), names)
        say(parity.report("gpu vs reference, KH 2-D 64^2 WENO3 RK2, 10 steps", rows))
        check_rows(rows, 1e-9)
    # ghost cells hold what the reference's Flow boundaries leave there
    for got, ref in ((c, z["cons10"]), (p, z["prims10"])):
        assert np.max(np.abs(got - ref)) <= 1e-9 * max(1.0, np.abs(ref).max())


def test_kh3d_weno5_three_steps_match_reference(golden):
    z = golden("kh3d_16_weno5_rk2")
    say("gpu: KH 3-D 16^3 WENO5 3 steps vs reference")
    with engine(weno=5, **KH3D16) as e:
        e.set_prims(z["prims0"])
        for s in range(3):
            assert e.update_time() == 0
        c, p, a = e.download()
    for got, ref, names in ((c, z["cons3"], parity.SRMHD_CONS), (p, z["prims3"], parity.SRMHD_PRIMS),
                            (a, z["aux3"], parity.SRMHD_AUX)):
        rows = parity.field_errors(parity.interior(got, 4, 3), ref, names)
        say(parity.report("gpu vs reference, KH 3-D 16^3 WENO5 RK2, 3 steps", rows))
        check_rows(rows, 1e-9, use_group=True)


@pytest.mark.parametrize("n, weno, bc", [(40, 5, (1, 0, 1)), (37, 3, (0, 1, 0))])
def test_3d_multi_tile_matches_oracle(n, weno, bc):
    """Grids spanning several tiles (and ragged ones) against the oracle, 2 steps."""
    g = idt.Grid(n, n + 3, n - 2, -0.5, 0.5, -1.0, 1.0, 0.0, 1.0)
    prims0 = idt.kh_prims(g, seed=11)
    grid = dict(nx=g.nx, ny=g.ny, nz=g.nz, xmin=g.xmin, xmax=g.xmax, ymin=g.ymin, ymax=g.ymax, zmin=0.0, zmax=1.0)
    say(f"gpu: 3-D {g.nx}x{g.ny}x{g.nz} WENO{weno} bc={bc} vs oracle")
    o = oracle_for(weno=weno, periodic=bc, **grid)
    oc, op, oa, _, _ = o.run(prims0, 2)
    with engine(weno=weno, bc=tuple(bc), **grid) as e:
        e.set_prims(prims0)
        for _ in range(2):
            assert e.update_time() == 0
        c, p, a = e.download()
    for got, ref, names in ((c, oc, parity.SRMHD_CONS), (p, op, parity.SRMHD_PRIMS)):
        rows = parity.field_errors(got, ref, names)  # padded, ghosts included
        say(parity.report(f"gpu vs oracle {g.nx}x{g.ny}x{g.nz}", rows))
        check_rows(rows, 1e-11, use_group=True)


def test_uniform_state_is_stationary():
    # test_simulation.cc:78-125: a homogeneous state is unchanged by evolve
    g = idt.kh3d_grid(32)
    prims0 = np.zeros((8,) + g.shape)
    prims0[0], prims0[1], prims0[2], prims0[3], prims0[4] = 0.8, 0.3, -0.2, 0.1, 1.3
    prims0[5], prims0[6], prims0[7] = 0.2, -0.1, 0.4
    say("gpu: uniform state 32^3")
    with engine(weno=5, nx=32, ny=32, nz=32, xmin=-0.5, xmax=0.5, ymin=-1, ymax=1, zmin=0, zmax=1) as e:
        e.set_prims(prims0)
        c0, p0, _ = e.download()
        for _ in range(3):
            e.update_time()
        c, p, _ = e.download()
    assert np.max(np.abs(c - c0)) <= 1e-14 * np.abs(c0).max()
    assert np.max(np.abs(p - p0)) <= 1e-13 * np.abs(p0).max()


def test_full_size_periodic_shift_invariance():
    """256^3 (the benchmark grid): shifting the initial data periodically in x
    and z shifts the result exactly; no cell fails its Newton solve."""
    g = idt.kh3d_grid(256)
    prims0 = idt.kh_prims(g, seed=1234)
    grid = dict(nx=256, ny=256, nz=256, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, zmin=0.0, zmax=1.0)
    sx, sz = 37, 101
    inner = prims0[:, 4:-4, :, 4:-4]
    shifted = prims0.copy()
    shifted[:, 4:-4, :, 4:-4] = np.roll(np.roll(inner, sx, axis=1), sz, axis=3)
    outs = []
    for k, p0 in enumerate((prims0, shifted)):
        say(f"gpu: 256^3 WENO5 run {k}")
        with engine(weno=5, **grid) as e:
            e.set_prims(p0)
            for _ in range(2):
                rc, st = e.update_time(stats=True)
                assert rc == 0 and st["failures"] == 0
            c, _, _ = e.download()
            outs.append(c[:, 4:-4, 4:-4, 4:-4].copy())
            del c
    a, b = outs
    assert np.isfinite(a).all()
    assert np.array_equal(np.roll(np.roll(a, sx, axis=1), sz, axis=3), b)


@pytest.mark.parametrize("grid", ["2d", "3d"])
def test_max_wavespeed_matches_oracle(grid):
    """CFL max-wavespeed reduction on the engine's own state after two steps,
    against the oracle's reduction of the same downloaded prims and aux."""
    if grid == "2d":
        g, cfg = idt.kh2d_grid(64), KH2D
    else:
        g = idt.Grid(24, 20, 28, -0.5, 0.5, -1.0, 1.0, 0.0, 1.0)
        cfg = dict(nx=24, ny=20, nz=28, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, zmin=0.0, zmax=1.0)
    say(f"gpu: max wavespeed {grid}")
    with engine(weno=5, **cfg) as e:
        e.set_prims(idt.kh_prims(g, shear=0.7, bz=0.3))
        for _ in range(2):
            e.update_time()
        got = e.max_wavespeed()
        c, p, a = e.download()
    want = oracle_for(weno=5, **cfg).max_wavespeed(p, a)
    say(f"  gpu {got}  oracle {want}")
    dims = 2 if grid == "2d" else 3
    assert np.all(got[:dims] > 0.0) and np.all(got < 1.0)
    assert np.max(np.abs(got - want)) <= 1e-14


@pytest.mark.parametrize("n, weno, bc", [((20, 37, 27), 5, (1, 0, 1)), ((33, 18, 41), 3, (0, 1, 0)),
                                         ((128, 128, 128), 5, (1, 0, 1))])
def test_fused_yz_sweep_equals_separate_sweeps_bitwise(n, weno, bc):
    """The fused z+y sweep launch must give exactly the bits of the separate
    z and y sweeps (same faces, same accumulation order)."""
    nx, ny, nz = n
    g = idt.Grid(nx, ny, nz, -0.5, 0.5, -1.0, 1.0, 0.0, 1.0)
    cfg = dict(nx=nx, ny=ny, nz=nz, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, zmin=0.0, zmax=1.0)
    p0 = idt.kh_prims(g, seed=7, shear=0.6, bz=0.2)
    say(f"gpu: fused yz vs separate {n} WENO{weno} bc={bc}")
    out = []
    for fused in (True, False):
        with engine(weno=weno, bc=tuple(bc), **cfg) as e:
            e.set_sweep_plan(fused)
            e.set_prims(p0)
            f0 = e.fnet()
            for _ in range(2):
                assert e.update_time() == 0
            out.append((f0,) + e.download())
    for a, b, name in zip(out[0], out[1], ("fnet", "cons", "prims", "aux")):
        assert np.array_equal(a, b), f"{name}: max diff {np.max(np.abs(a - b)):.3e}"


@pytest.mark.parametrize("n, weno, bc, integrator", [((20, 37, 27), 5, (1, 0, 1), eng.RK2),
                                                    ((33, 18, 41), 3, (0, 1, 0), eng.RK2),
                                                    ((24, 20, 17), 5, (1, 1, 1), eng.RKSPLIT),
                                                    ((16, 21, 18), 3, (1, 0, 1), eng.RK3)])
def test_fused_yz_finishing_sweep_equals_separate_sweeps_bitwise(n, weno, bc, integrator):
    """Sweep plan 3 (an x sweep, then the y and z sweeps in one launch that
    finishes the stage, or leaves the net flux for an RKPlus stage) must give
    exactly the bits of plan 2's separate x, y, z sweeps: the same faces and
    the same sums in the same order, on ragged grids with partial tiles."""
    nx, ny, nz = n
    g = idt.Grid(nx, ny, nz, -0.5, 0.5, -1.0, 1.0, 0.0, 1.0)
    cfg = dict(nx=nx, ny=ny, nz=nz, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, zmin=0.0, zmax=1.0)
    p0 = idt.kh_prims(g, seed=9, shear=0.6, bz=0.2)
    say(f"gpu: fused y+z finishing sweep (plan 3) vs plan 2, {n} WENO{weno} bc={bc} integrator {integrator}")
    out = []
    for plan in (3, 2):
        with eng.Engine(eng.Config(weno_order=weno, bc=tuple(bc), quirks=1, integrator=integrator, cp=0.1,
                                   **cfg)) as e:
            e.set_sweep_plan(plan)
            e.set_prims(p0)
            st = None
            for _ in range(2):
                rc, st = e.update_time(stats=True)
                assert rc == 0
            out.append(e.download() + (st["cells"], st["iterations"]))
    for a, b, name in zip(out[0], out[1], ("cons", "prims", "aux", "cells", "iterations")):
        assert np.array_equal(a, b), f"{name} differs"


@pytest.mark.parametrize("n, weno, bc, integrator", [((37, 20), 5, (1, 0, 1), eng.RK2),
                                                    ((18, 33), 3, (0, 1, 0), eng.RK2),
                                                    ((24, 21), 5, (1, 1, 1), eng.RKSPLIT),
                                                    ((21, 16), 3, (1, 0, 1), eng.RK3),
                                                    ((64, 64), 3, (1, 0, 1), eng.RK2)])
def test_fused_yx_finishing_sweep_2d_equals_separate_sweeps_bitwise(n, weno, bc, integrator):
    """Sweep plan 3 on a 2-D grid (the y and x sweeps in one launch that
    finishes the stage, or leaves the net flux for an RKPlus stage) must give
    exactly the bits of plan 0's separate y and x sweeps, on ragged grids with
    partial tiles and on the 64^2 KH case."""
    nx, ny = n
    g = idt.Grid(nx, ny, 0, -0.5, 0.5, -1.0, 1.0)
    cfg = dict(nx=nx, ny=ny, nz=0, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0)
    p0 = idt.kh_prims(g, seed=11, shear=0.6, bz=0.2)
    say(f"gpu: 2-D fused y+x finishing sweep (plan 3) vs plan 0, {n} WENO{weno} bc={bc} integrator {integrator}")
    out = []
    for plan in (3, 0):
        with eng.Engine(eng.Config(weno_order=weno, bc=tuple(bc), quirks=1, integrator=integrator, cp=0.1,
                                   **cfg)) as e:
            e.set_sweep_plan(plan)
            e.set_prims(p0)
            st = None
            for _ in range(3):
                rc, st = e.update_time(stats=True)
                assert rc == 0
            out.append(e.download() + (e.fnet(), st["cells"], st["iterations"]))
    for a, b, name in zip(out[0], out[1], ("cons", "prims", "aux", "fnet", "cells", "iterations")):
        assert np.array_equal(a, b), f"{name} differs"


def test_rksplit_steps_match_reference(golden):
    """RKSplit on the GPU (RK2, split phi source, cons2prim) against the
    reference build: KH 48^2 with a divergent Bx, cp 0.1, 8 steps."""
    z = golden("kh2d_48_weno3_rksplit")
    grid = dict(nx=48, ny=48, nz=0, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0)
    say("gpu: RKSplit KH 48^2 8 steps vs reference")
    with eng.Engine(eng.Config(weno_order=3, integrator=eng.RKSPLIT, cp=0.1, **grid)) as e:
        e.set_prims(z["prims0"])
        for s in range(8):
            rc, st = e.update_time(stats=True)
            assert rc == 0 and st["failures"] == 0
            assert e.time()[1] == z["dts"][s]
        c, p, a = e.download()
    for got, ref, names in ((c, z["cons8"], parity.SRMHD_CONS), (p, z["prims8"], parity.SRMHD_PRIMS)):
        rows = parity.field_errors(parity.interior(got, 4, 2), parity.interior(ref, 4, 2), names)
        say(parity.report("gpu vs reference, RKSplit KH 48^2, 8 steps", rows))
        check_rows(rows, 1e-9, use_group=True)


@pytest.mark.parametrize("quirks", [0, 1])
def test_c2p_retry_matches_oracle(quirks):
    """Three cells whose Newton solve fails take the neighbour-mean retry
    (srmhd.cc:465-507): the failure count, and with the quirk on the reference's
    use of the last cell's arguments, must match the oracle. The retry starts
    from the neighbours' roots summed in another order than the oracle's, so the
    agreement is to the Newton tolerance, not bitwise."""
    say(f"gpu: c2p retry quirks={quirks}")
    o = oracle_for(weno=3, quirks=quirks, **parity.C2P_RETRY_GRID)
    cons, prims, aux, gp, ga = parity.c2p_retry_case(o)
    op, oa = gp.copy(), ga.copy()
    fails, _ = o.c2p(cons, op, oa)
    with engine(weno=3, quirks=quirks, **parity.C2P_RETRY_GRID) as e:
        e.upload(cons=cons, prims=gp, aux=ga)
        rc, st = e.c2p()
        _, p, a = e.download()
    say(f"  failures {st['failures']} retry failures {st['fallbacks']}")
    assert fails == len(parity.C2P_RETRY_CELLS)
    assert rc == fails and st["failures"] == fails and st["fallbacks"] == 0
    for got, ref, names in ((p, op, parity.SRMHD_PRIMS), (a, oa, parity.SRMHD_AUX)):
        rows = parity.field_errors(parity.interior(got, 4, 2), parity.interior(ref, 4, 2), names)
        say(parity.report(f"c2p retry quirks={quirks}", rows))
        check_rows(rows, 1e-12)
    if not quirks:
        sl = (slice(None), slice(4, -4), slice(4, -4), 0)
        assert np.allclose(p[sl], prims[sl], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("quirks", [1, 0])
def test_c2p_retry_3d_matches_oracle(quirks):
    """The retry on a 3-D grid, whose neighbour search runs over a runtime
    axis index: four failing cells (a pair that are z-neighbours, one next to
    the z edge), against the oracle to the Newton tolerance."""
    say(f"gpu: c2p retry 3-D quirks={quirks}")
    o = oracle_for(weno=3, quirks=quirks, **parity.C2P_RETRY_GRID_3D)
    cons, prims, aux, gp, ga = parity.c2p_retry_case_3d(o)
    op, oa = gp.copy(), ga.copy()
    fails, _ = o.c2p(cons, op, oa)
    with engine(weno=3, quirks=quirks, **parity.C2P_RETRY_GRID_3D) as e:
        e.upload(cons=cons, prims=gp, aux=ga)
        rc, st = e.c2p()
        _, p, a = e.download()
    assert fails == len(parity.C2P_RETRY_CELLS_3D)
    assert rc == fails and st["failures"] == fails and st["fallbacks"] == 0
    for got, ref, names in ((p, op, parity.SRMHD_PRIMS), (a, oa, parity.SRMHD_AUX)):
        rows = parity.field_errors(parity.interior(got, 4, 3), parity.interior(ref, 4, 3), names)
        say(parity.report(f"c2p retry 3-D quirks={quirks}", rows))
        check_rows(rows, 1e-12)
    if not quirks:
        sl = (slice(None), slice(4, -4), slice(4, -4), slice(4, -4))
        assert np.allclose(p[sl], prims[sl], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("integrator", ["rk2", "rksplit"])
def test_step_retry_keeps_root_for_the_aux(integrator):
    """A step whose cons2prim fails in both stages at three cells
    (parity.c2p_step_retry_case) takes the retry branch that keeps the root in
    place of the aux; the aux formed from it on download, and the cons and
    prims, must match the oracle's step, with the oracle's failure count.
    (Without the last-cell quirk: with it the retried cells' prims are
    recovered from another cell's root and disagree with their own aux, and
    the reference's flux reads W and b from that aux where the engine's forms
    them from the prims, so the two steps part there by design.)"""
    quirks = 0
    say(f"gpu: {integrator} step with failing cells, quirks={quirks}")
    o = oracle_for(weno=3, quirks=quirks, **parity.C2P_RETRY_GRID)
    cons, prims, aux, gp, ga = parity.c2p_step_retry_case(o)
    dt = o.timestep(0.0, 0)
    oc, op, oa = cons.copy(), gp.copy(), ga.copy()
    ofails, ost = o.step(oc, op, oa, dt, integrator)
    integ = eng.RK2 if integrator == "rk2" else eng.RKSPLIT
    with engine(weno=3, quirks=quirks, integrator=integ, **parity.C2P_RETRY_GRID) as e:
        e.upload(cons=cons, prims=gp, aux=ga)
        rc, st = e.step(dt, stats=True)
        c, p, a = e.download()
    say(f"  gpu failures {st['failures']} retry failures {st['fallbacks']}, oracle failures {ost.failures}")
    assert st["failures"] == ost.failures == 2 * len(parity.C2P_RETRY_CELLS)
    assert st["fallbacks"] == 0
    for got, ref, names in ((c, oc, parity.SRMHD_CONS), (p, op, parity.SRMHD_PRIMS), (a, oa, parity.SRMHD_AUX)):
        rows = parity.field_errors(parity.interior(got, 4, 2), parity.interior(ref, 4, 2), names)
        say(parity.report(f"{integrator} step with retries, quirks={quirks}", rows))
        check_rows(rows, 1e-10)


@pytest.mark.parametrize("model", [eng.SRMHD, eng.SRRMHD])
def test_flux_vector_splitting_stationary(model):
    """FluxVectorSplittingStationary (test_srmhd.cc:33-110, test_srrmhd.cc:
    31-113): a state at rest with rho = 0.5 and every other primitive zero on a
    periodic 10^3 grid (gamma 5/3, sigma 1000, cp 0.5) has zero net flux, to
    the tests' 1e-15."""
    cfg = eng.Config(nx=10, ny=10, nz=10, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0,
                     gamma=5.0 / 3.0, sigma=1000.0, cp=0.5, model=model, weno_order=3,
                     bc=(eng.PERIODIC,) * 3, integrator=eng.SSP2 if model == eng.SRRMHD else eng.RK2)
    say(f"gpu: stationary state, model {model}")
    with eng.Engine(cfg) as e:
        prims = e.zeros(e.nprims)
        prims[0] = 0.5
        e.set_prims(prims)
        fnet = e.fnet()
    inner = fnet[:, 4:-4, 4:-4, 4:-4]
    assert np.abs(inner).max() <= 1e-15


def _volume_stats(a):
    # the same per-field reduction as the fixture's generator (tests/golden/make_fixtures.py)
    out = np.zeros((a.shape[0], 4))
    for v in range(a.shape[0]):
        x = a[v]
        out[v] = (x.min(), x.max(), x.mean(dtype=np.float64), np.sqrt(np.sum(x * x, dtype=np.float64)))
    return out


def test_full_size_ten_steps_match_reference(golden):
    """The benchmark workload at its own size (KH 3-D 256^3, WENO5, RK2, the
    seeded state bench.py times), ten steps against the reference build: dt
    bit for bit; on the middle x-plane every cons and prim field to 1e-9 of
    its group scale, and pointwise to 1e-8 relative wherever |value| is at
    least 1e-3 of that scale (parity.pointwise_floor); and every cons and prim
    field's min, max, mean and L2 norm over the 16.8 M interior cells to 1e-9
    of that field's scale."""
    z = golden("kh3d_256_weno5_rk2_full")
    steps = int(z["steps"])
    g = idt.kh3d_grid(256)
    prims0 = idt.kh_prims(g, seed=1234)
    say(f"gpu: 256^3 WENO5 RK2, {steps} steps against the reference build")
    with engine(weno=5, nx=256, ny=256, nz=256, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, zmin=0.0,
                zmax=1.0) as e:
        e.set_prims(prims0)
        del prims0
        for s in range(steps):
            rc, st = e.update_time(stats=True)
            assert rc == 0 and st["failures"] == 0
            assert e.time()[1] == z["dts"][s]
        c, p, _ = e.download()
    inner = (slice(None), slice(4, -4), slice(4, -4), slice(4, -4))
    c, p = c[inner], p[inner]
    for got, ref, names in ((c[:, 128], z["cons_plane_x128"], parity.SRMHD_CONS),
                            (p[:, 128], z["prims_plane_x128"], parity.SRMHD_PRIMS)):
        rows = parity.field_errors(got, ref, names)
        say(parity.report(f"gpu vs reference, 256^3 x-plane 128 after {steps} steps", rows))
        pw = parity.pointwise_floor(got, ref, names, 1e-3)
        for (n, dm, fm, rf, rg, _), e_pw in zip(rows, pw):
            say(f"  {n}: pointwise (floor 1e-3 of scale) {e_pw:.3e}")
            assert rg <= 1e-9, f"{n}: {rg:.3e}"
            assert e_pw <= 1e-8, f"{n}: pointwise {e_pw:.3e}"
            if fm == 0.0:
                assert dm == 0.0, n
    for got, ref, fields in ((_volume_stats(c), z["cons_stats"], parity.SRMHD_CONS),
                             (_volume_stats(p), z["prims_stats"], parity.SRMHD_PRIMS)):
        # a field's scale is its largest |value|, taken over its vector group (S, v, B + phi) as the
        # other parity checks do, so a near-zero component such as Bx is measured on the field's scale
        scale = {n: max(abs(ref[v, 0]), abs(ref[v, 1])) for v, n in enumerate(fields)}
        norm = {n: ref[v, 3] for v, n in enumerate(fields)}
        for grp in parity._GROUPS:
            members = [n for n in grp if n in scale]
            for n in members:
                scale[n] = max(scale[m] for m in members)
                norm[n] = max(norm[m] for m in members)
        for v, name in enumerate(fields):
            diff = np.abs(got[v] - ref[v])
            say(f"  {name}: min/max/mean/L2 differences {diff} (scale {scale[name]:.3e})")
            assert np.all(diff[:3] <= 1e-9 * max(scale[name], 1e-300)), name
            assert diff[3] <= 1e-9 * max(norm[name], 1e-300), name


@pytest.mark.parametrize("integrator", [eng.RK2, eng.RKSPLIT], ids=["rk2", "rksplit"])
def test_aux_formed_from_kept_root_is_consistent(golden, integrator):
    """The SRMHD RK2 and RKSplit steps keep each cell's cons2prim root (v^2, Z)
    in place of its 13 aux values, and the aux is formed from it when read. It
    must agree with primsToAll of the downloaded prims (srmhd.cc:573-673) to the
    Newton tolerance, carry S.B and S^2 of the downloaded cons, and be the same
    bits on every later read (a wavespeed reduction in between included)."""
    z = golden("kh3d_16_weno5_rk2")
    say(f"gpu: aux formed after {'RK2' if integrator == eng.RK2 else 'RKSplit'} steps, KH 3-D 16^3")
    cfg = eng.Config(weno_order=5, bc=(eng.PERIODIC, eng.OUTFLOW, eng.PERIODIC), quirks=1, integrator=integrator,
                     **KH3D16)
    with eng.Engine(cfg) as e:
        e.set_prims(z["prims0"])
        for _ in range(3):
            rc, st = e.update_time(stats=True)
            assert rc == 0 and st["failures"] == 0
        c1, p1, a1 = e.download()
        speeds = e.max_wavespeed()
        c2, p2, a2 = e.download()
    assert np.all(speeds[:3] > 0.0) and np.all(speeds < 1.0)
    for x, y in ((c1, c2), (p1, p2), (a1, a2)):
        assert np.array_equal(x, y)
    _, _, oa = oracle_for(weno=5, **KH3D16).prims_to_all(p1)
    a, c, p = (parity.interior(x, 4, 3) for x in (a1, c1, p1))
    rows = parity.field_errors(a, parity.interior(oa, 4, 3), parity.SRMHD_AUX)
    say(parity.report("aux formed from the root vs primsToAll of the prims", rows))
    check_rows(rows, 1e-10, use_group=True)
    S, B = c[1:4], p[5:8]
    assert np.max(np.abs(a[10] - (B * S).sum(axis=0))) <= 1e-14 * np.abs(a[10]).max()
    assert np.max(np.abs(a[12] - (S * S).sum(axis=0))) <= 1e-14 * np.abs(a[12]).max()


def test_partial_upload_keeps_the_aux_it_does_not_replace(golden):
    """mgpu_upload_state with cons and prims but no aux leaves the device aux
    as it was. After RK2 steps, which keep the cons2prim root in place of the
    aux, that is the stepped state's aux, formed before the upload."""
    z = golden("kh3d_16_weno5_rk2")
    say("gpu: partial upload after RK2 steps, KH 3-D 16^3")
    with engine(weno=5, **KH3D16) as e:
        e.set_prims(z["prims0"])
        for _ in range(2):
            assert e.update_time() == 0
        _, _, want = e.download()
    with engine(weno=5, **KH3D16) as e:
        e.set_prims(z["prims0"])
        c0, p0, _ = e.download()
        for _ in range(2):
            assert e.update_time() == 0
        e.upload(cons=c0, prims=p0)
        c, p, a = e.download()
    assert np.array_equal(c, c0) and np.array_equal(p, p0)
    assert np.array_equal(a, want)
