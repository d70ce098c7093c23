"""The REGIME extension of ideal SRMHD (REGIME.cc) on the GPU through the C
ABI: its source against the reference build's REGIME::sourceExtension, and
RKSplit / RK2B runs with it against the reference's trajectories and the
oracle. Needs an MI355X."""
import ctypes
import sys

import numpy as np
import pytest

from method_amd import engine as eng
from oracle import oracle as orc
from tests import parity

pytestmark = pytest.mark.gpu

INTEG = {"rksplit": eng.RKSPLIT, "rk2b": eng.RK2B, "rk3": eng.RK3, "rk4": eng.RK4, "rk4_10": eng.RK4_10}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def config_of(z, integrator="rksplit"):
    nx, ny, nz, Ng, weno, _ = (int(v) for v in z["meta"])
    b = z["bounds"]
    gam, cfl, sig, cp = (float(v) for v in z["params"])
    bc = tuple(eng.PERIODIC if int(v) else eng.OUTFLOW for v in z["periodic"])
    return eng.Config(nx=nx, ny=ny, nz=nz, xmin=b[0], xmax=b[1], ymin=b[2], ymax=b[3], zmin=b[4], zmax=b[5],
                      gamma=gam, sigma=sig, cp=cp, cfl=cfl, weno_order=weno, integrator=INTEG[integrator], bc=bc)


def dims_of(cfg):
    return 1 if cfg.ny == 0 else (2 if cfg.nz == 0 else 3)


@pytest.mark.parametrize("name", ["regime_src_currentsheet_128", "regime_src_briowu_128",
                                  "regime_src_kh2d_32_weno5", "regime_src_kh3d_12_weno3"])
def test_regime_source_matches_reference(golden, name):
    """The source on the padded grid with the reference's loop ranges; the
    GPU's WENO arithmetic differs from the reference's in the last bits, which
    the two centred differences amplify, hence 1e-9 of each field's scale."""
    z = golden(name)
    cfg = config_of(z)
    say(f"gpu: REGIME source {name}")
    with eng.Engine(cfg) as e:
        e.set_prims(z["prims0"])
        e.set_model_extension(eng.EXT_REGIME)
        src = e.source_extension()
    got, ref = parity.interior(src, 4, dims_of(cfg)), z["src"]
    rows = parity.field_errors(got, ref, parity.SRMHD_CONS)
    say(parity.report(f"REGIME source vs reference, {name}", rows))
    for n, dm, fm, rf, rg, pw in rows:
        assert rf <= 1e-9, f"{n}: {rf:.3e}"
        if fm == 0.0:
            assert dm == 0.0, f"{n} must stay exactly zero"


# the current sheet's bar is the reference cons2prim's tolerance amplified by
# the stiff source at sigma 1e4 (the oracle's test gives the reasoning)
RUNS = [("currentsheet_128_regime_rksplit", "rksplit", 5e-9), ("briowu_128_regime_rksplit", "rksplit", 1e-9),
        ("kh2d_32_weno5_regime_rk2b", "rk2b", 1e-9), ("kh3d_12_weno3_regime_rksplit", "rksplit", 1e-9),
        ("kh2d_32_weno5_regime_rk3", "rk3", 1e-9), ("kh2d_32_weno5_regime_rk4", "rk4", 1e-9),
        ("kh2d_32_weno5_regime_rk4_10", "rk4_10", 1e-9)]


@pytest.mark.parametrize("name, integrator, tol", RUNS)
def test_regime_runs_match_reference(golden, name, integrator, tol):
    z = golden(name)
    cfg = config_of(z, integrator)
    steps = int(z["meta"][5])
    dims = dims_of(cfg)
    say(f"gpu: {integrator} + REGIME {name}, {steps} steps vs reference")
    with eng.Engine(cfg) as e:
        e.set_prims(z["prims0"])
        e.set_model_extension(eng.EXT_REGIME)
        for s in range(steps):
            rc, st = e.update_time(stats=True)
            assert rc == 0 and st["failures"] == 0
            assert e.time()[1] == z["dts"][s]
        c, p, a = e.download()
    for got, ref, names in ((c, z[f"cons{steps}"], parity.SRMHD_CONS), (p, z[f"prims{steps}"], parity.SRMHD_PRIMS)):
        got = parity.interior(got, 4, dims)
        if ref.shape != got.shape:
            ref = parity.interior(ref, 4, dims)
        rows = parity.field_errors(got, ref, names)
        say(parity.report(f"gpu vs reference, {name}", rows))
        for n, dm, fm, rf, rg, pw in rows:
            assert rg <= tol, f"{n}: {rg:.3e}"


def test_regime_current_sheet_matches_oracle(golden):
    """Against the oracle, whose cons2prim is the same Newton solve, the current
    sheet agrees to the 1e-9 bar after 20 steps without the reference's
    cons2prim tolerance in the way; what remains is last-bit WENO rounding
    amplified by the stiff source (sigma 1e4) into Sx, 2000 times smaller than
    tau, which itself agrees to ~1e-13."""
    z = golden("currentsheet_128_regime_rksplit")
    cfg = config_of(z)
    o = orc.Oracle(orc.Case(nx=128, ny=0, nz=0, xmin=-3.0, xmax=3.0, ymin=-1.0, ymax=1.0, zmin=-1.5, zmax=1.5,
                            gamma=2.0, cfl=0.2, sigma=1e4, periodic=(0, 0, 0), regime=1))
    oc, op, oa, dts, _ = o.run(z["prims0"], 20, integrator="rksplit")
    say("gpu: current sheet + REGIME vs oracle, 20 steps")
    with eng.Engine(cfg) as e:
        e.set_prims(z["prims0"])
        e.set_model_extension(eng.EXT_REGIME)
        for _ in range(20):
            assert e.update_time() == 0
        c, p, a = e.download()
    rows = parity.field_errors(parity.interior(c, 4, 1), parity.interior(oc, 4, 1), parity.SRMHD_CONS)
    say(parity.report("current sheet + REGIME vs oracle", rows))
    for n, dm, fm, rf, rg, pw in rows:
        assert rg <= 1e-9, f"{n}: {rg:.3e}"
    assert max(r[4] for r in rows if r[0] in ("D", "tau", "By")) <= 1e-12


def test_regime_needs_ideal_model_and_is_required_for_its_source():
    grid = dict(nx=16, ny=16, nz=0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0)
    with eng.Engine(eng.Config(model=eng.SRRMHD, integrator=eng.SSP2, sigma=10.0, **grid)) as e:
        with pytest.raises(eng.EngineError, match="ideal SRMHD"):
            e.set_model_extension(eng.EXT_REGIME)
    with eng.Engine(eng.Config(**grid)) as e:
        with pytest.raises(eng.EngineError, match="no model extension"):
            e.source_extension()


def _gpu_sheet(grid, direction, end_time=0.1):
    """The reference's REGIME symmetry set-ups (test_REGIME.cc:943-1088) on the
    engine: current sheet, gamma 2, sigma 100, cfl 0.4, WENO3, Outflow,
    RKSplit + REGIME, evolved to t = 0.1."""
    from method_amd import initdata as idt

    cfg = eng.Config(gamma=2.0, sigma=100.0, cfl=0.4, cp=0.1, weno_order=3, integrator=eng.RKSPLIT,
                     bc=(eng.OUTFLOW,) * 3, **grid)
    with eng.Engine(cfg) as e:
        e.set_prims(idt.current_sheet_prims(idt.Grid(**grid), 100.0, direction))
        e.set_model_extension(eng.EXT_REGIME)
        while e.time()[0] < end_time:
            assert e.update_time(end_time) == 0
        return e.download()


def test_regime_sheet_symmetries():
    """YAxis / ZAxis / RotSymmetries of the reference's REGIME tests, to their
    1e-15, and the 3-D sheet against the oracle."""
    say("gpu: REGIME current-sheet symmetries")
    c2, p2, _ = _gpu_sheet(dict(nx=10, ny=10, nz=0, xmin=-3.0, xmax=3.0, ymin=-1.0, ymax=1.0), 0)
    assert np.abs(np.diff(p2[:, 4:-4, 4:-4, 0], axis=2)).max() <= 1e-15
    g3 = dict(nx=10, ny=10, nz=10, xmin=-3.0, xmax=3.0, ymin=-1.0, ymax=1.0, zmin=-1.0, zmax=1.0)
    c3, p3, _ = _gpu_sheet(g3, 0)
    inner = p3[:, 4:-4, 4:-4, 4:-4]
    assert np.abs(np.diff(inner, axis=2)).max() <= 1e-15 and np.abs(np.diff(inner, axis=3)).max() <= 1e-15
    grid = dict(nx=10, ny=10, nz=10, xmin=-3.0, xmax=3.0, ymin=-3.0, ymax=3.0, zmin=-3.0, zmax=3.0)
    ca, _, _ = _gpu_sheet(grid, 0)
    cb, _, _ = _gpu_sheet(grid, 1)
    cc, _, _ = _gpu_sheet(grid, 2)
    a = ca[6, 4:-4, 4:-4, 4:-4]
    b = np.transpose(cb[7, 4:-4, 4:-4, 4:-4], (1, 2, 0))
    c = np.transpose(cc[5, 4:-4, 4:-4, 4:-4], (2, 0, 1))
    say(f"  rotations: max |A - B| {np.abs(a - b).max():.3e}, max |A - C| {np.abs(a - c).max():.3e}")
    assert np.abs(a - b).max() <= 1e-15 and np.abs(a - c).max() <= 1e-15
    # the 3-D sheet against the oracle
    from method_amd import initdata as idt
    o = orc.Oracle(orc.Case(gamma=2.0, sigma=100.0, cfl=0.4, cp=0.1, periodic=(0, 0, 0), regime=1, **g3))
    oc, op, oa = o.set(idt.current_sheet_prims(idt.Grid(**g3), 100.0, 0))
    t, it = 0.0, 0
    while t < 0.1:
        dt = o.timestep(t, it, 0.1)
        assert o.step(oc, op, oa, dt, "rksplit")[0] == 0
        t += dt
        it += 1
    rows = parity.field_errors(parity.interior(c3, 4, 3), parity.interior(oc, 4, 3), parity.SRMHD_CONS)
    say(parity.report("3-D current sheet + REGIME vs oracle", rows))
    for n, dm, fm, rf, rg, pw in rows:
        assert rg <= 1e-9, f"{n}: {rg:.3e}"


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_regime_fields_match_reference_known_answers(axis):
    """The reference's DataAssignment1D/2D/3D known answers
    (test_REGIME.cc:36-353) on the GPU's REGIME fields: E, q = dE/ds, alpha
    and dw/ds_b at the cell where the linear coordinate is 1 (its stencils lie
    inside the interior, so the engine's boundary images do not enter)."""
    from method_amd import initdata as idt
    from tests.test_oracle import REGIME_KAT

    name = ("1D", "2D", "3D")[axis]
    n = (100, 10, 10)[axis]
    lo, hi = ((0.01, 2.01), (0.1, 2.1), (0.1, 2.1))[axis]
    grid = dict(nx=n if axis == 0 else 10, ny=0 if axis == 0 else 10, nz=10 if axis == 2 else 0,
                xmin=lo if axis == 0 else 0.1, xmax=hi if axis == 0 else 2.1, ymin=0.1 if axis else 0.0,
                ymax=2.1 if axis else 1.0, zmin=0.1 if axis == 2 else 0.0, zmax=2.1 if axis == 2 else 1.0)
    g = idt.Grid(**grid)
    s = g.centres()[axis]
    mid = len(s) // 2 - 1
    sb = s.reshape([-1 if a == axis else 1 for a in range(3)])
    p = np.zeros((8,) + g.shape)
    lin_v, const_v, lin_b, const_b = [(2, 3, 6, 7), (3, 1, 7, 5), (1, 2, 5, 6)][axis]
    p[0], p[4] = 0.1, 0.66
    p[lin_v] = 0.41 * sb
    p[const_v] = 0.51
    p[lin_b] = 0.22 * sb
    p[const_b] = 0.33
    say(f"gpu: REGIME known answers {name}")
    cfg = eng.Config(gamma=2.0, sigma=50.0, bc=(eng.PERIODIC,) * 3, **grid)
    with eng.Engine(cfg) as e:
        e.upload(prims=p)
        e.set_model_extension(eng.EXT_REGIME)
        E, q, alpha, dw, K = e.regime_fields()
    cell = tuple(mid if a <= axis else 0 for a in range(3))
    e_exp = -(0.41 * 0.33 - 0.51 * 0.22)
    assert abs(E[axis][cell] - e_exp) <= 1e-15
    assert abs(q[cell] - e_exp) <= 1e-15
    assert abs(alpha[cell] - 1.3825277485504367e-07) <= 1e-13
    for l, m, want, tol in REGIME_KAT[name]:
        assert abs(dw[l, m][cell] - want) <= tol, (name, l, m)


@pytest.mark.parametrize("name", ["regime_src_currentsheet_128", "regime_src_briowu_128",
                                  "regime_src_kh2d_32_weno5", "regime_src_kh3d_12_weno3"])
def test_regime_fused_source_equals_two_launch_source_bitwise(golden, name):
    """D formed on chip in the differencing launch (the default plan) gives
    the source of the two-launch plan, which stores D, bit for bit."""
    z = golden(name)
    cfg = config_of(z)
    say(f"gpu: REGIME fused vs two-launch source {name}")
    out = []
    for fused in (True, False):
        with eng.Engine(cfg) as e:
            e.set_prims(z["prims0"])
            e.set_model_extension(eng.EXT_REGIME)
            e.set_regime_plan(fused)
            out.append(e.source_extension())
    assert np.array_equal(out[0], out[1])
    assert np.abs(out[0]).max() > 0


def test_regime_fused_plan_run_equals_two_launch_run_bitwise(golden):
    """Two RKSplit steps of 3-D KH with REGIME under both plans: the same
    state bit for bit."""
    z = golden("kh3d_12_weno3_regime_rksplit")
    cfg = config_of(z)
    say("gpu: REGIME fused vs two-launch, kh3d 12^3 RKSplit 2 steps")
    states = []
    for fused in (True, False):
        with eng.Engine(cfg) as e:
            e.set_prims(z["prims0"])
            e.set_model_extension(eng.EXT_REGIME)
            e.set_regime_plan(fused)
            for _ in range(2):
                assert e.update_time() == 0
            states.append(e.download())
    for a, b in zip(*states):
        assert np.array_equal(a, b)


# ragged grids for the source's tiles (4x8x8 in 3-D, 16x16 in 2-D), mixed boundaries, both orders
RAGGED = [(13, 10, 9, 5, (1, 0, 1)), (9, 17, 11, 3, (0, 1, 0)), (37, 21, 0, 5, (1, 0, 1)), (19, 35, 0, 3, (0, 0, 0))]


@pytest.mark.parametrize("nx, ny, nz, order, per", RAGGED)
def test_regime_source_on_ragged_grids_matches_oracle(nx, ny, nz, order, per):
    """REGIME's source on grids that leave partial tiles at every face, with
    periodic and outflow axes, against the oracle's restatement of
    REGIME::sourceExtension; both source plans."""
    from method_amd import initdata as idt

    bounds = dict(xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, zmin=0.0, zmax=1.0)
    g = idt.Grid(nx, ny, nz, **bounds)
    prims0 = idt.kh_prims(g, seed=nx * 100 + ny, shear=0.5, bz=0.3)
    case = orc.Case(nx=nx, ny=ny, nz=nz, weno_order=order, periodic=per, sigma=100.0, regime=1, **bounds)
    o = orc.Oracle(case)
    cons, prims, aux = o.set(prims0)
    want = np.zeros_like(cons)
    o.L.oc_regime_source(ctypes.byref(o.c), orc._p(cons), orc._p(prims), orc._p(aux), orc._p(want))
    dims = 2 if nz == 0 else 3
    bc = tuple(eng.PERIODIC if p else eng.OUTFLOW for p in per)
    cfg = eng.Config(nx=nx, ny=ny, nz=nz, weno_order=order, bc=bc, sigma=100.0, integrator=eng.RKSPLIT, **bounds)
    for fused in (True, False):
        say(f"gpu: REGIME source {nx}x{ny}x{nz} WENO{order} periodic {per}, fused={fused}")
        with eng.Engine(cfg) as e:
            e.set_prims(prims0)
            e.set_model_extension(eng.EXT_REGIME)
            e.set_regime_plan(fused)
            got = e.source_extension()
        rows = parity.field_errors(parity.interior(got, 4, dims), parity.interior(want, 4, dims), parity.SRMHD_CONS)
        say(parity.report(f"REGIME source vs oracle, {nx}x{ny}x{nz}", rows))
        for n, dm, fm, rf, rg, pw in rows:
            assert rf <= 1e-9, f"{n}: {rf:.3e}"
        assert np.all(got[8] == 0.0)


def _volume_stats(a):
    # the same per-field reduction as the fixture's generator (tests/golden/make_fixtures.py)
    out = np.zeros((a.shape[0], 4))
    for v in range(a.shape[0]):
        x = a[v]
        out[v] = (x.min(), x.max(), x.mean(dtype=np.float64), np.sqrt(np.sum(x * x, dtype=np.float64)))
    return out


def test_regime_large_grid_step_matches_reference(golden):
    """The REGIME workload's set-up (KH 3-D, WENO5, RKSplit with REGIME, sigma
    100) at 128^3, one step against the reference build: dt bit for bit, the
    middle x-plane of D, Sx, tau and By to 1e-9 of each field's scale, and
    every cons and prim field's min, max, mean and L2 norm over the 2.1 M
    interior cells to 1e-9 of the field's group scale."""
    from method_amd import initdata as idt

    z = golden("kh3d_128_weno5_regime_rksplit_large")
    g = idt.kh3d_grid(128)
    cfg = eng.Config(nx=128, ny=128, nz=128, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, zmin=0.0, zmax=1.0,
                     weno_order=5, bc=(eng.PERIODIC, eng.OUTFLOW, eng.PERIODIC), cfl=0.4, sigma=100.0,
                     integrator=eng.RKSPLIT)
    say("gpu: REGIME 128^3 WENO5 RKSplit, 1 step against the reference build")
    with eng.Engine(cfg) as e:
        e.set_prims(idt.kh_prims(g, seed=1234, bz=0.1))
        e.set_model_extension(eng.EXT_REGIME)
        rc, st = e.update_time(stats=True)
        assert rc == 0 and st["failures"] == 0
        assert e.time()[1] == z["dts"][0]
        c, p, _ = e.download()
    inner = (slice(None), slice(4, -4), slice(4, -4), slice(4, -4))
    c, p = c[inner], p[inner]
    names = list(parity.SRMHD_CONS)
    plane = np.stack([c[names.index(str(v)), 64] for v in z["plane_vars"]])
    rows = parity.field_errors(plane, z["plane_mid_x"], [str(v) for v in z["plane_vars"]])
    say(parity.report("gpu vs reference, REGIME 128^3 middle x-plane after 1 step", rows))
    for n, dm, fm, rf, rg, pw in rows:
        assert rf <= 1e-9, f"{n}: {rf:.3e}"
    for got, ref, fields in ((_volume_stats(c), z["cons_stats"], parity.SRMHD_CONS),
                             (_volume_stats(p), z["prims_stats"], parity.SRMHD_PRIMS)):
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
