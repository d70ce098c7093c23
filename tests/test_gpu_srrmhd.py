"""Parity of the HIP resistive SRMHD model and the SSP2 IMEX step with the
reference (golden fixture) and the CPU oracle, through the C ABI. Needs an MI355X."""
import sys

import numpy as np
import pytest

from method_amd import engine as eng
from method_amd import initdata as idt
from oracle import oracle as orc
from tests import parity

pytestmark = pytest.mark.gpu

OT32 = dict(nx=32, ny=32, nz=0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0)
RES = dict(sigma=1e3, cp=0.1, cfl=0.5)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def engine(integrator=eng.SSP2, weno=3, bc=(eng.PERIODIC,) * 3, quirks=1, **grid):
    return eng.Engine(eng.Config(model=eng.SRRMHD, integrator=integrator, weno_order=weno, bc=bc, quirks=quirks,
                                 **RES, **grid))


def oracle_for(weno=3, periodic=(1, 1, 1), quirks=1, **grid):
    return orc.Oracle(orc.Case(model=orc.SRRMHD, weno_order=weno, periodic=periodic, quirks=quirks, **RES, **grid))


def check_rows(rows, tol):
    for n, dm, fm, rf, rg, pw in rows:
        assert rg <= tol, f"{n}: {rg:.3e} > {tol:.1e}"


def ot3d_prims(n):
    """Orszag-Tang in x-y with a z-dependent vz, on a 3-D periodic box."""
    g = idt.Grid(n, n + 2, n - 2, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    p = idt.ot_prims_resistive(g)
    _, _, z = g.centres()
    p[3] = 0.05 * np.sin(2 * np.pi * z)[None, None, :]
    p[7] = 0.1
    return g, p


def test_set_matches_oracle(golden):
    z = golden("ot2d_32_weno3_ssp2")
    say("gpu: SRRMHD set OT 32^2")
    with engine(**OT32) as e:
        e.set_prims(z["prims0"])
        c, p, a = e.download()
    oc, op, oa = oracle_for(**OT32).set(z["prims0"])
    rows = parity.field_errors(parity.interior(c, 4, 2), parity.interior(oc, 4, 2), parity.SRRMHD_CONS)
    say(parity.report("gpu vs oracle, SRRMHD primsToAll", rows))
    check_rows(rows, 1e-13)
    assert np.max(np.abs(a - oa)) <= 1e-11 * np.abs(oa).max()


def test_fnet_and_source_match(golden):
    z = golden("ot2d_32_weno3_ssp2")
    say("gpu: SRRMHD fnet/source OT 32^2")
    with engine(**OT32) as e:
        e.set_prims(z["prims0"])
        f = e.fnet()
        s = e.source()
    rows = parity.field_errors(parity.interior(f, 4, 2), z["fnet0"], parity.SRRMHD_CONS)
    say(parity.report("gpu vs reference FVS::F, SRRMHD OT", rows))
    check_rows(rows, 1e-9)
    o = oracle_for(**OT32)
    oc, op, oa = o.set(z["prims0"])
    os_ = o.source(oc, op, oa)
    # for this state E = -v x B, so J and the E source are rounding noise of
    # size sigma |E| eps in both codes: compare on that scale
    scale = RES["sigma"] * np.abs(oc[8:11]).max()
    assert np.max(np.abs(parity.interior(s - os_, 4, 2))) <= 1e-12 * scale


@pytest.mark.parametrize("quirks", [1, 0])
def test_flux_vectors_3d_match_oracle(quirks):
    g, p0 = ot3d_prims(12)
    grid = dict(nx=g.nx, ny=g.ny, nz=g.nz, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0)
    say(f"gpu: SRRMHD flux vectors 3-D quirks={quirks}")
    o = oracle_for(quirks=quirks, **grid)
    oc, op, oa = o.set(p0)
    with engine(quirks=quirks, **grid) as e:
        e.set_prims(p0)
        buf = None
        for d in range(3):
            f = e.flux_vector(d)
            # the oracle's quirky z flux keeps the y flux in Sz, as the reference's shared buffer does
            buf = o.flux_vector(oc, op, oa, d, f=buf.copy() if buf is not None else None)
            rows = parity.field_errors(parity.interior(f, 4, 3), parity.interior(buf, 4, 3), parity.SRRMHD_CONS)
            check_rows(rows, 1e-11)


def test_ssp2_three_steps_match_reference(golden):
    """OT 2-D 32^2 SRRMHD, sigma 1e3, SSP2 IMEX, WENO3: 3 updateTime calls."""
    z = golden("ot2d_32_weno3_ssp2")
    say("gpu: SRRMHD SSP2 OT 32^2 3 steps vs reference")
    with engine(**OT32) as e:
        e.set_prims(z["prims0"])
        for s in range(3):
            rc, st = e.update_time(stats=True)
            assert rc == 0 and st["failures"] == 0
            assert e.time()[1] == z["dts"][s]
        c, p, a = e.download()
    for got, ref, names in ((c, z["cons3"], parity.SRRMHD_CONS), (p, z["prims3"], parity.SRRMHD_PRIMS)):
        rows = parity.field_errors(parity.interior(got, 4, 2), parity.interior(ref, 4, 2), names)
        say(parity.report("gpu vs reference, OT 2-D 32^2 SRRMHD SSP2, 3 steps", rows))
        check_rows(rows, 1e-9)


@pytest.mark.parametrize("weno, bc, integrator", [(5, (1, 1, 1), eng.SSP2), (3, (1, 0, 1), eng.SSP2),
                                                  (3, (1, 1, 1), eng.RK2)])
def test_3d_matches_oracle(weno, bc, integrator):
    g, p0 = ot3d_prims(20)
    grid = dict(nx=g.nx, ny=g.ny, nz=g.nz, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0)
    name = "ssp2" if integrator == eng.SSP2 else "rk2"
    say(f"gpu: SRRMHD 3-D {g.nx}x{g.ny}x{g.nz} WENO{weno} bc={bc} {name} vs oracle")
    o = oracle_for(weno=weno, periodic=bc, **grid)
    oc, op, oa, _, _ = o.run(p0, 2, integrator=name)
    with engine(integrator=integrator, weno=weno, bc=tuple(bc), **grid) as e:
        e.set_prims(p0)
        for _ in range(2):
            rc, st = e.update_time(stats=True)
            assert rc == 0 and st["failures"] == 0
        c, p, a = e.download()
    for got, ref, names in ((c, oc, parity.SRRMHD_CONS), (p, op, parity.SRRMHD_PRIMS)):
        rows = parity.field_errors(parity.interior(got, 4, 3), parity.interior(ref, 4, 3), names)
        say(parity.report("gpu vs oracle SRRMHD 3-D", rows))
        check_rows(rows, 1e-9)


def test_ot_1024_two_steps_stay_finite_and_symmetric():
    """The config-2 grid: 1024^2, two SSP2 steps, every cell converges, and the
    Orszag-Tang point symmetry (x, y) -> (1-x, 1-y) is kept."""
    g = idt.ot2d_grid(1024)
    p0 = idt.ot_prims_resistive(g)
    say("gpu: SRRMHD OT 1024^2 2 steps")
    with engine(nx=1024, ny=1024, nz=0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0) as e:
        e.set_prims(p0)
        for _ in range(2):
            rc, st = e.update_time(stats=True)
            assert rc == 0 and st["failures"] == 0
        c, p, a = e.download()
    ci = parity.interior(c, 4, 2)[:, :, :, 0]
    assert np.isfinite(ci).all()
    rho = ci[0]
    assert np.max(np.abs(rho - rho[::-1, ::-1])) <= 1e-10 * rho.max()


def test_fused_yz_sweep_equals_separate_sweeps_bitwise():
    g, p0 = ot3d_prims(20)
    grid = dict(nx=g.nx, ny=g.ny, nz=g.nz, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0)
    say("gpu: SRRMHD fused yz vs separate sweeps")
    out = []
    for fused in (True, False):
        with engine(weno=5, **grid) as e:
            e.set_sweep_plan(fused)
            e.set_prims(p0)
            for _ in range(2):
                assert e.update_time() == 0
            out.append(e.download())
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("weno, integrator", [(5, eng.SSP2), (3, eng.SSP2), (3, eng.RK2), (5, eng.RK2)])
def test_fused_yz_finishing_sweep_equals_separate_sweeps_bitwise(weno, integrator):
    """Sweep plan 3 with the resistive model (its fused y+z launch runs each
    phase in two groups of 8 pencils) under SSP2 (the implicit stage-two and
    final modes) and RK2 gives plan 2's bits."""
    g, p0 = ot3d_prims(20)
    grid = dict(nx=g.nx, ny=g.ny, nz=g.nz, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0)
    say(f"gpu: SRRMHD fused y+z finishing sweep vs plan 2, WENO{weno} integrator {integrator}")
    out = []
    for plan in (3, 2):
        with engine(integrator=integrator, weno=weno, **grid) as e:
            e.set_sweep_plan(plan)
            e.set_prims(p0)
            for _ in range(2):
                assert e.update_time() == 0
            out.append(e.download())
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("n, weno, integrator, bc", [((23, 20), 3, eng.SSP2, (eng.PERIODIC,) * 3),
                                                    ((32, 32), 3, eng.SSP2, (eng.PERIODIC,) * 3),
                                                    ((20, 37), 5, eng.SSP2, (eng.OUTFLOW,) * 3),
                                                    ((19, 26), 5, eng.RK2, (eng.PERIODIC,) * 3)])
def test_fused_yx_finishing_sweep_2d_equals_separate_sweeps_bitwise(n, weno, integrator, bc):
    """On 2-D grids plan 3 runs the y and x sweeps as one launch (in two groups
    of 8 pencils for the resistive model); it must give plan 0's bits under
    SSP2's implicit modes and RK2, on ragged grids."""
    nx, ny = n
    g = idt.Grid(nx, ny, 0, 0.0, 1.0, 0.0, 1.0)
    p0 = idt.ot_prims_resistive(g)
    grid = dict(nx=nx, ny=ny, nz=0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0)
    say(f"gpu: SRRMHD 2-D fused y+x finishing sweep vs plan 0, {n} WENO{weno} integrator {integrator} bc={bc}")
    out = []
    for plan in (3, 0):
        with engine(integrator=integrator, weno=weno, bc=bc, **grid) as e:
            e.set_sweep_plan(plan)
            e.set_prims(p0)
            for _ in range(3):
                assert e.update_time() == 0
            out.append(e.download())
    for a, b, name in zip(out[0], out[1], ("cons", "prims", "aux")):
        assert np.array_equal(a, b), f"{name}: max diff {np.max(np.abs(a - b)):.3e}"


def test_ssp2_aux_formed_from_kept_root_is_consistent(golden):
    """SSP2 steps keep each cell's cons2prim root Z = rho h W^2 in place of its
    17 aux values, and the aux is formed from it when the state is read. It
    must be the same bits on every later read, hold Z itself (aux 10) and B^2,
    E^2 of the downloaded cons, and agree with the aux a fresh cons2prim of the
    same cons writes (srrmhd.cc:256-312) to the secant tolerance."""
    z = golden("ot2d_32_weno3_ssp2")
    say("gpu: SRRMHD aux formed from the kept root after SSP2 steps, OT 32^2")
    with engine(**OT32) as e:
        e.set_prims(z["prims0"])
        for _ in range(3):
            assert e.update_time() == 0
        c1, p1, a1 = e.download()
        c2, p2, a2 = e.download()
        rc, st = e.c2p()
        assert rc == 0 and st["failures"] == 0
        c3, p3, a3 = e.download()
    for x, y in ((c1, c2), (p1, p2), (a1, a2)):
        assert np.array_equal(x, y)
    a, c, a_fresh = (parity.interior(x, 4, 2) for x in (a1, c1, a3))
    B2, E2 = (c[5:8] ** 2).sum(axis=0), (c[8:11] ** 2).sum(axis=0)
    assert np.max(np.abs(a[7] - B2)) <= 1e-14 * B2.max()
    assert np.max(np.abs(a[8] - E2)) <= 1e-14 * E2.max()
    for k in range(a.shape[0]):
        scale = max(np.abs(a_fresh[k]).max(), 1e-300)
        assert np.max(np.abs(a[k] - a_fresh[k])) <= 1e-10 * scale, f"aux {k}"


def test_rksplit_matches_oracle():
    """RKSplit with the resistive model (explicit split source, sigma 10)."""
    g = idt.ot2d_grid(24)
    p0 = idt.ot_prims_resistive(g)
    grid = dict(nx=24, ny=24, nz=0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0)
    say("gpu: SRRMHD RKSplit OT 24^2 vs oracle")
    o = orc.Oracle(orc.Case(model=orc.SRRMHD, weno_order=3, periodic=(1, 1, 1), quirks=1, sigma=10.0, cp=0.1,
                            cfl=0.5, **grid))
    oc, op, oa, _, _ = o.run(p0, 3, integrator="rksplit")
    with eng.Engine(eng.Config(model=eng.SRRMHD, integrator=eng.RKSPLIT, weno_order=3, bc=(eng.PERIODIC,) * 3,
                               quirks=1, sigma=10.0, cp=0.1, cfl=0.5, **grid)) as e:
        e.set_prims(p0)
        for _ in range(3):
            rc, st = e.update_time(stats=True)
            assert rc == 0 and st["failures"] == 0
        c, p, a = e.download()
    for got, ref, names in ((c, oc, parity.SRRMHD_CONS), (p, op, parity.SRRMHD_PRIMS)):
        rows = parity.field_errors(parity.interior(got, 4, 2), parity.interior(ref, 4, 2), names)
        say(parity.report("gpu vs oracle SRRMHD RKSplit", rows))
        check_rows(rows, 1e-9)


def _volume_stats(a):
    # the same per-field reduction as the fixture's generator (tests/golden/make_fixtures.py)
    out = np.zeros((a.shape[0], 4))
    for v in range(a.shape[0]):
        x = a[v]
        out[v] = (x.min(), x.max(), x.mean(dtype=np.float64), np.sqrt(np.sum(x * x, dtype=np.float64)))
    return out


def _group_scales(ref, fields):
    # a field's scale is its largest |value| over its vector group, as in the other parity checks
    scale = {n: max(abs(ref[v, 0]), abs(ref[v, 1])) for v, n in enumerate(fields)}
    norm = {n: ref[v, 3] for v, n in enumerate(fields)}
    for grp in parity._GROUPS:
        members = [n for n in grp if n in scale]
        for n in members:
            scale[n] = max(scale[m] for m in members)
            norm[n] = max(norm[m] for m in members)
    return scale, norm


def test_full_size_resistive_ten_steps_match_reference(golden):
    """The resistive workload at its own size (Orszag-Tang 1024^2, sigma 1e3,
    SSP2, WENO3, the state bench.py times), ten steps against the reference
    build: dt bit for bit; the middle row along each axis of every cons and
    prim field to 1e-9 of the field's group scale, and pointwise to 1e-8
    relative wherever |value| is at least 1e-3 of that scale; and every cons
    and prim field's min, max, mean and L2 norm over the 1 M interior cells to
    1e-9."""
    z = golden("ot2d_1024_srrmhd_ssp2_full")
    steps = int(z["steps"])
    g = idt.ot2d_grid(1024)
    say(f"gpu: resistive OT 1024^2 SSP2, {steps} steps against the reference build")
    with engine(nx=1024, ny=1024, nz=0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0) as e:  # RES: sigma 1e3, cp 0.1, cfl 0.5
        e.set_prims(idt.ot_prims_resistive(g))
        for s in range(steps):
            assert e.update_time() == 0
            assert e.time()[1] == z["dts"][s]
        c, p, _ = e.download()
    c, p = c[:, 4:-4, 4:-4, :], p[:, 4:-4, 4:-4, :]
    for got, ref, names, what in ((c[:, 512, :, 0], z["row_x512"], parity.SRRMHD_CONS, "cons row x=512"),
                                  (c[:, :, 512, 0], z["row_y512"], parity.SRRMHD_CONS, "cons row y=512"),
                                  (p[:, 512, :, 0], z["prow_x512"], parity.SRRMHD_PRIMS, "prims row x=512"),
                                  (p[:, :, 512, 0], z["prow_y512"], parity.SRRMHD_PRIMS, "prims row y=512")):
        rows = parity.field_errors(got, ref, names)
        say(parity.report(f"gpu vs reference, {what}", rows))
        pw = parity.pointwise_floor(got, ref, names, 1e-3)
        for (n, dm, fm, rf, rg, _), e_pw in zip(rows, pw):
            assert rg <= 1e-9, f"{what} {n}: {rg:.3e}"
            assert e_pw <= 1e-8, f"{what} {n}: pointwise {e_pw:.3e}"
    for got, ref, fields in ((_volume_stats(c), z["cons_stats"], parity.SRRMHD_CONS),
                             (_volume_stats(p), z["prims_stats"], parity.SRRMHD_PRIMS)):
        scale, norm = _group_scales(ref, fields)
        for v, name in enumerate(fields):
            diff = np.abs(got[v] - ref[v])
            say(f"  {name}: min/max/mean/L2 differences {diff} (scale {scale[name]:.3e})")
            assert np.all(diff[:3] <= 1e-9 * max(scale[name], 1e-300)), name
            assert diff[3] <= 1e-9 * max(norm[name], 1e-300), name


@pytest.mark.parametrize("name", ["srrmhd3d_12x14x10_ssp2_periodic", "srrmhd3d_12x14x10_ssp2_flow",
                                  "srrmhd3d_12x14x10_rksplit_periodic", "srrmhd3d_12x14x10_rksplit_flow"])
def test_3d_matches_reference(golden, name):
    """3-D resistive runs against the reference build itself (SSP2 sigma 50,
    RKSplit sigma 10; periodic and Flow; ragged 12 x 14 x 10; 3 steps): the
    z-direction flux quirk of srrmhd.cc:163-171 and the 3-D stencils to 1e-9
    of each field's group scale, dt bit for bit."""
    z = golden(name)
    grid = dict(nx=12, ny=14, nz=10, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0)
    integ = eng.SSP2 if "ssp2" in name else eng.RKSPLIT
    bc = tuple(eng.PERIODIC if b else eng.OUTFLOW for b in z["periodic"])
    say(f"gpu: SRRMHD 3-D {name} vs reference")
    p0 = np.zeros((11, 20, 22, 18))
    p0[:, 4:-4, 4:-4, 4:-4] = z["prims0"]
    cfg = eng.Config(model=eng.SRRMHD, integrator=integ, weno_order=int(z["weno"]), bc=bc, quirks=1,
                     sigma=float(z["sigma"]), cp=0.1, cfl=0.5, **grid)
    with eng.Engine(cfg) as e:
        e.set_prims(p0)
        for s in range(3):
            rc, st = e.update_time(stats=True)
            assert rc == 0 and st["failures"] == 0
            assert e.time()[1] == z["dts"][s]
        c, p, _ = e.download()
    for got, ref, names in ((c, z["cons3"], parity.SRRMHD_CONS), (p, z["prims3"], parity.SRRMHD_PRIMS)):
        rows = parity.field_errors(parity.interior(got, 4, 3), ref, names)
        say(parity.report(f"gpu vs reference, {name}", rows))
        check_rows(rows, 1e-9)
