"""The CPU oracle, pinned before it is trusted as the checker.

Pins: the reference's own known-answer tests (Tests/CPU/Src/test_weno.cc,
test_srmhd.cc, test_simulation.cc, test_boundaryConds.cc), the reference build's
outputs stored under tests/golden/ (make_fixtures.py), and closed forms.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle import oracle as orc
from tests import parity

KH2D = dict(nx=64, ny=64, nz=0, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0)
KH3D = dict(nx=16, ny=16, nz=16, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, zmin=0.0, zmax=1.0)


# --------------------------------------------------------------------- WENO
def test_weno_linear_kats():
    # test_weno.cc:37-82 (WENO3 up 1.5, down 0.5) and :161-210 (WENO5 up 2.5, down 1.5)
    assert abs(orc.weno3(0.0, 1.0, 2.0) - 1.5) < 1e-15
    assert abs(orc.weno3(2.0, 1.0, 0.0) - 0.5) < 1e-15
    assert abs(orc.weno5(0.0, 1.0, 2.0, 3.0, 4.0) - 2.5) < 1e-15
    assert abs(orc.weno5(4.0, 3.0, 2.0, 1.0, 0.0) - 1.5) < 1e-15


def _jiang_shu5(v):
    """Jiang-Shu WENO5 from its textbook definition, in exact rationals then float."""
    from fractions import Fraction as Fr

    v = [Fr(x) for x in v]
    b = [Fr(13, 12) * (v[0] - 2 * v[1] + v[2]) ** 2 + Fr(1, 4) * (v[0] - 4 * v[1] + 3 * v[2]) ** 2,
         Fr(13, 12) * (v[1] - 2 * v[2] + v[3]) ** 2 + Fr(1, 4) * (v[1] - v[3]) ** 2,
         Fr(13, 12) * (v[2] - 2 * v[3] + v[4]) ** 2 + Fr(1, 4) * (3 * v[2] - 4 * v[3] + v[4]) ** 2]
    q = [(2 * v[0] - 7 * v[1] + 11 * v[2]) / 6, (-v[1] + 5 * v[2] + 2 * v[3]) / 6, (2 * v[2] + 5 * v[3] - v[4]) / 6]
    d = [Fr(1, 10), Fr(6, 10), Fr(3, 10)]
    eps = Fr(1, 10 ** 16)
    a = [d[k] / (eps + b[k] ** 2) for k in range(3)]
    return float(sum(a[k] * q[k] for k in range(3)) / sum(a))


def _jiang_shu3(v):
    from fractions import Fraction as Fr

    v = [Fr(x) for x in v]
    b = [(v[1] - v[0]) ** 2, (v[2] - v[1]) ** 2]
    q = [(-v[0] + 3 * v[1]) / 2, (v[1] + v[2]) / 2]
    d = [Fr(1, 3), Fr(2, 3)]
    eps = Fr(1, 10 ** 16)
    a = [d[k] / (eps + b[k] ** 2) for k in range(2)]
    return float(sum(a[k] * q[k] for k in range(2)) / sum(a))


def test_weno_against_exact_rational_definition():
    rng = np.random.default_rng(3)
    for scale in (1e-3, 1.0, 10.0):
        for _ in range(50):
            v = (rng.normal(size=5) * scale).tolist()
            ref5 = _jiang_shu5(v)
            assert abs(orc.weno5(*v) - ref5) <= 1e-13 * max(1.0, abs(ref5)) * max(1.0, scale)
            ref3 = _jiang_shu3(v[:3])
            assert abs(orc.weno3(*v[:3]) - ref3) <= 1e-13 * max(1.0, abs(ref3)) * max(1.0, scale)
    # a discontinuity picks the smooth side (essentially non-oscillatory)
    assert abs(orc.weno5(0, 0, 0, 1, 1) - 0.0) < 1e-6


# ------------------------------------------------------------ primsToAll
def test_prims_to_all_closed_form():
    # test_srmhd.cc:228-282: prims = 1e-16, gamma = 5/3 -> h 3.5, W 1, e 1.5, c sqrt(10/21)
    case = orc.Case(10, 10, 10, 0, 1, 0, 1, 0, 1, gamma=5.0 / 3.0)
    o = orc.Oracle(case)
    _, p, _ = case.zeros()
    p[:] = 1e-16
    cons, prims, aux = o.prims_to_all(p)
    assert np.abs(cons).max() <= 1e-15
    assert np.all(aux[0] == 3.5)
    assert np.abs(aux[1] - 1.0).max() <= 1e-15
    assert np.all(aux[2] == 1.5)
    assert np.all(aux[3] == math.sqrt(10.0 / 21.0))
    assert np.abs(aux[4:]).max() <= 1e-15


# --------------------------------------------------------------------- dt
def test_timestep_closed_forms():
    # test_simulation.cc:34-72: 100x10x2 on [0,1]x[-0.5,0.5]x[-0.1,0.1], cfl 0.5 -> 0.002886751346
    case = orc.Case(100, 10, 2, 0.0, 1.0, -0.5, 0.5, -0.1, 0.1, cfl=0.5, Ng=4)
    o = orc.Oracle(case)
    assert abs(o.timestep(0.0, 100) - 0.002886751346) < 1e-11
    assert abs(o.timestep(0.0, 0) - 0.1 * 0.002886751346) < 1e-12
    assert abs(o.timestep(0.0, 20) - 0.25 * 0.002886751346) < 1e-12
    assert o.timestep(1.0, 100, end_time=1.0001) == pytest.approx(1e-4)
    # 2-D KH 64^2: 0.1 * 0.4 * (1/64) / sqrt(2)
    o2 = orc.Oracle(orc.Case(**KH2D))
    assert o2.timestep(0.0, 0) == 0.1 * (0.4 * (1.0 / 64) / math.sqrt(2))


# -------------------------------------------------------------------- BCs
def test_boundary_index_patterns():
    # test_boundaryConds.cc: periodic copies nx+i -> i and Ng+i -> nx+Ng+i; outflow copies the edge cell
    case = orc.Case(10, 10, 10, 0, 1, 0, 1, 0, 1, periodic=(1, 0, 1))
    o = orc.Oracle(case)
    Nx, Ny, Nz = case.shape
    a = np.arange(Nx * Ny * Nz, dtype=np.float64).reshape(1, Nx, Ny, Nz).copy()
    b = a.copy()
    o.apply_bcs(b)
    Ng, nx, ny = 4, 10, 10
    for i in range(Ng):
        assert np.array_equal(b[0, i, Ng:-Ng, Ng:-Ng], a[0, nx + i, Ng:-Ng, Ng:-Ng])
        assert np.array_equal(b[0, nx + Ng + i, Ng:-Ng, Ng:-Ng], a[0, Ng + i, Ng:-Ng, Ng:-Ng])
        assert np.array_equal(b[0, Ng:-Ng, i, Ng:-Ng], a[0, Ng:-Ng, Ng, Ng:-Ng])
        assert np.array_equal(b[0, Ng:-Ng, ny + Ng + i, Ng:-Ng], a[0, Ng:-Ng, ny + Ng - 1, Ng:-Ng])
    # corner: composition of the per-axis maps
    assert b[0, 0, 0, 0] == a[0, nx + 0, Ng, 10 + 0]


# -------------------------------------------------------------------- C2P
def test_c2p_round_trip_and_reference_vectors(golden):
    z = golden("srmhd_c2p_vectors")
    n = int(z["n"][0])
    case = orc.Case(n, 0, 0, 0.0, 1.0, 0.0, 1.0, periodic=(0, 0, 0))
    o = orc.Oracle(case)
    p, a = z["guess_prims"].copy(), z["guess_aux"].copy()
    fails, st = o.c2p(z["cons"], p, a)
    assert fails == 0 and st.failures == 0
    sl = (slice(None), slice(4, -4))
    pt, pr = z["prims_true"][sl], z["prims_ref"][sl]
    got = p[sl]
    scale = np.abs(pt).max(axis=1, keepdims=True)
    # round trip to the exact primitives
    assert (np.abs(got - pt) / np.maximum(scale, 1e-300)).max() < 1e-11
    # the reference's hybrd roots are within its own tolerance of the same root
    assert (np.abs(got - pr) / np.maximum(scale, 1e-300)).max() < 1e-7


def test_srmhd_newton_special_cases():
    L = orc.lib()
    import ctypes

    g = 4.0 / 3.0
    # rest state (v = 0): Z = rho h
    rho, p = 0.7, 0.3
    h = 1 + p / rho * g / (g - 1)
    D, tau = rho, rho * h - p - rho
    x = (ctypes.c_double * 2)(0.2, 1.0)
    its = L.oc_srmhd_newton(D, 0.0, 0.0, 0.0, tau, 0.0, 0.0, 0.0, g, x)
    assert its >= 0 and abs(x[0]) < 1e-13 and abs(x[1] - rho * h) < 1e-12
    # hydro (B = 0): closed form cons from prims, root equals rho h W^2
    v = 0.6
    W = 1 / math.sqrt(1 - v * v)
    D = rho * W
    S = rho * h * W * W * v
    tau = rho * h * W * W - p - D
    x = (ctypes.c_double * 2)(0.3, 2.0)
    its = L.oc_srmhd_newton(D, S, 0.0, 0.0, tau, 0.0, 0.0, 0.0, g, x)
    assert its >= 0 and abs(x[0] - v * v) < 1e-13 and abs(x[1] - rho * h * W * W) < 1e-12


# ------------------------------------------------- reference trajectories
def test_kh2d_fnet_matches_reference(golden):
    z = golden("kh2d_64_weno3_rk2")
    o = orc.Oracle(orc.Case(**KH2D, weno_order=3, periodic=(1, 0, 1)))
    c, p, a = o.set(z["prims0"])
    f = parity.interior(o.fnet(c, p, a), 4, 2)
    assert np.array_equal(f, z["fnet0"])


def test_kh3d_fnet_matches_reference(golden):
    z = golden("kh3d_16_weno5_rk2")
    o = orc.Oracle(orc.Case(**KH3D, weno_order=5, periodic=(1, 0, 1)))
    c, p, a = o.set(z["prims0"])
    f = parity.interior(o.fnet(c, p, a), 4, 3)
    assert np.array_equal(f, z["fnet0"])


def test_kh2d_ten_steps_match_reference(golden):
    z = golden("kh2d_64_weno3_rk2")
    o = orc.Oracle(orc.Case(**KH2D, weno_order=3, periodic=(1, 0, 1)))
    c, p, a, dts, st = o.run(z["prims0"], 10)
    assert np.array_equal(np.array(dts), z["dts"])
    assert st.failures == 0
    for got, ref, names in ((c, z["cons10"], parity.SRMHD_CONS), (p, z["prims10"], parity.SRMHD_PRIMS),
                            (a, z["aux10"], parity.SRMHD_AUX)):
        rows = parity.field_errors(parity.interior(got, 4, 2), parity.interior(ref, 4, 2), names)
        print(parity.report("oracle vs reference, KH 2-D 64^2 WENO3 RK2, 10 steps", rows))
        for n, dm, fm, rf, rg, pw in rows:
            assert rf <= 1e-9, n
            if fm == 0.0:
                assert dm == 0.0, n


def test_kh3d_weno5_steps_match_reference(golden):
    z = golden("kh3d_16_weno5_rk2")
    o = orc.Oracle(orc.Case(**KH3D, weno_order=5, periodic=(1, 0, 1)))
    c, p, a, dts, st = o.run(z["prims0"], 3)
    for got, ref, names in ((c, z["cons3"], parity.SRMHD_CONS), (p, z["prims3"], parity.SRMHD_PRIMS)):
        rows = parity.field_errors(parity.interior(got, 4, 3), ref, names)
        print(parity.report("oracle vs reference, KH 3-D 16^3 WENO5 RK2, 3 steps", rows))
        for n, dm, fm, rf, rg, pw in rows:
            assert rg <= 1e-9, n


def test_quirk_flag_changes_only_phi_z_flux():
    case = orc.Case(**KH3D, weno_order=5)
    o1 = orc.Oracle(case)
    case0 = orc.Case(**KH3D, weno_order=5, quirks=0)
    o0 = orc.Oracle(case0)
    rng = np.random.default_rng(0)
    c, p, a = case.zeros()
    p[0] = 1.0
    p[4] = 1.0
    p[5:8] = rng.uniform(-0.1, 0.1, size=p[5:8].shape)
    c1, p1, a1 = o1.prims_to_all(p)
    f1 = o1.flux_vector(c1, p1, a1, 2)
    f0 = o0.flux_vector(c1, p1, a1, 2)
    assert np.array_equal(f1[:8], f0[:8])
    assert np.array_equal(f1[8], p1[6]) and np.array_equal(f0[8], p1[7])


def test_ot2d_ssp2_steps_match_reference(golden):
    """SSP2 IMEX with the resistive model: the oracle solves each implicit stage
    on its stiff part (E and the recovery's Z jointly) instead of the
    reference's 14-unknown hybrd (SSP2.cc:85,131); the roots agree."""
    z = golden("ot2d_32_weno3_ssp2")
    o = orc.Oracle(orc.Case(32, 32, 0, 0.0, 1.0, 0.0, 1.0, weno_order=3, periodic=(1, 1, 1), model=orc.SRRMHD,
                            sigma=1e3, cp=0.1, cfl=0.5))
    c, p, a = o.set(z["prims0"])
    assert np.array_equal(parity.interior(o.fnet(c, p, a), 4, 2), z["fnet0"])
    c, p, a, dts, st = o.run(z["prims0"], 3, integrator="ssp2")
    assert np.array_equal(np.array(dts), z["dts"]) and st.failures == 0
    for got, ref, names in ((c, z["cons3"], parity.SRRMHD_CONS), (p, z["prims3"], parity.SRRMHD_PRIMS)):
        rows = parity.field_errors(parity.interior(got, 4, 2), parity.interior(ref, 4, 2), names)
        print(parity.report("oracle vs reference, OT 2-D 32^2 SRRMHD SSP2, 3 steps", rows))
        for n, dm, fm, rf, rg, pw in rows:
            assert rg <= 1e-9, n


# ------------------------------------------------------ max wavespeed bound
def _state(rng, gam=4.0 / 3.0):
    """Random SRMHD state and the aux entries the bound reads (h, c, b^2, v^2)
    from the primsToAll formulas (srmhd.cc:573-673)."""
    rho = 10.0 ** rng.uniform(-2, 1)
    p = rho * 10.0 ** rng.uniform(-2, 1)
    d = rng.normal(size=3)
    v = d / np.linalg.norm(d) * rng.uniform(0.0, 0.95)
    B = rng.normal(size=3) * math.sqrt(p) * 10.0 ** rng.uniform(-1, 1)
    W = 1.0 / math.sqrt(1.0 - v @ v)
    h = 1.0 + p / rho * gam / (gam - 1.0)
    c = math.sqrt(gam * p / (rho * h))
    b0 = W * (B @ v)
    b2 = (B @ B + b0 * b0) / (W * W)
    return rho, v, h, c, b2


def test_max_speed_special_cases():
    L = orc.lib()
    # hydro, motion along the axis: relativistic addition (v + cs) / (1 + v cs)
    for v, cs in ((0.0, 0.3), (0.6, 0.5), (-0.9, 0.2), (0.3, 0.57)):
        h = 2.0
        got = L.oc_srmhd_max_speed(1.0, v, v * v, h, cs, 0.0)
        assert abs(got - (abs(v) + cs) / (1.0 + abs(v) * cs)) < 1e-15
    # fluid at rest: the fast speed a, a^2 = cs^2 + va^2 - cs^2 va^2
    rho, h, cs, b2 = 0.7, 1.9, 0.4, 0.8
    va2 = b2 / (rho * h + b2)
    a = math.sqrt(cs * cs + va2 - cs * cs * va2)
    assert abs(L.oc_srmhd_max_speed(rho, 0.0, 0.0, h, cs, b2) - a) < 1e-15


def test_max_speed_is_the_boosted_envelope():
    """Brute force: signals of the fluid-frame fast speed a in every direction,
    carried to the lab frame by relativistic velocity addition
    u = (u'/W + v + W/(W+1) (v.u') v) / (1 + v.u'); the largest |u_d| over the
    directions is the bound (up to the sampling of directions)."""
    L = orc.lib()
    rng = np.random.default_rng(11)
    n = rng.normal(size=(400000, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    for _ in range(12):
        rho, v, h, c, b2 = _state(rng)
        va2 = b2 / (rho * h + b2)
        a = math.sqrt(c * c + va2 - c * c * va2)
        W = 1.0 / math.sqrt(1.0 - v @ v)
        up = a * n
        vu = up @ v
        u = (up / W + v[None, :] + (W / (W + 1.0)) * vu[:, None] * v[None, :]) / (1.0 + vu)[:, None]
        for d in range(3):
            brute = np.abs(u[:, d]).max()
            got = L.oc_srmhd_max_speed(rho, v[d], v @ v, h, c, b2)
            assert got < 1.0
            assert brute <= got * (1.0 + 1e-12)
            assert got - brute < 2e-4 * got, (d, got, brute)


def test_max_wavespeed_grid_reduction():
    """The grid maximum is the largest per-cell bound over interior cells and
    axes the grid has; ghost cells are not read."""
    g = orc.Case(model=orc.SRMHD, **KH2D)
    o = orc.Oracle(g)
    from method_amd import initdata as idt

    p0 = idt.kh_prims(idt.kh2d_grid(64), shear=0.7, bz=0.3)
    cons, prims, aux = o.set(p0)
    prims_g = prims.copy()
    prims_g[1, :4] = 0.99  # a ghost layer may hold anything
    got = o.max_wavespeed(prims_g, aux)
    L = orc.lib()
    I = (slice(4, -4), slice(4, -4), 0)
    want = [max(L.oc_srmhd_max_speed(r, vd, v2, h, c, b2) for r, vd, v2, h, c, b2 in
                zip(prims[0][I].ravel(), prims[1 + d][I].ravel(), aux[9][I].ravel(), aux[0][I].ravel(),
                    aux[3][I].ravel(), aux[8][I].ravel())) for d in range(2)]
    assert got[0] == want[0] and got[1] == want[1] and got[2] == 0.0


def test_rksplit_steps_match_reference(golden):
    """RKSplit (rkSplit.cc:31-53) on KH 48^2 with a divergent Bx, cp 0.1:
    8 steps against the reference build; the split phi source must act (the
    RK2 run differs in phi far beyond the tolerance)."""
    z = golden("kh2d_48_weno3_rksplit")
    grid = dict(nx=48, ny=48, nz=0, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0)
    o = orc.Oracle(orc.Case(weno_order=3, periodic=(1, 0, 1), cp=0.1, **grid))
    oc, op, oa, dts, _ = o.run(z["prims0"], 8, integrator="rksplit")
    assert np.array_equal(np.array(dts), z["dts"])
    for got, ref, names in ((oc, z["cons8"], parity.SRMHD_CONS), (op, z["prims8"], parity.SRMHD_PRIMS)):
        rows = parity.field_errors(parity.interior(got, 4, 2), parity.interior(ref, 4, 2), names)
        assert max(r[4] for r in rows) <= 1e-9
    rc, _, _, _, _ = orc.Oracle(orc.Case(weno_order=3, periodic=(1, 0, 1), cp=0.1, **grid)).run(z["prims0"], 8)
    assert np.abs(rc[8] - z["cons8"][8]).max() > 1e-5


RKPLUS_FIXTURES = [("kh2d_32_weno3_rk2b", orc.SRMHD, "rk2b"), ("kh2d_32_weno3_rk3", orc.SRMHD, "rk3"),
                   ("kh2d_32_weno3_rk4", orc.SRMHD, "rk4"), ("kh2d_32_weno3_rk4_10", orc.SRMHD, "rk4_10"),
                   ("ot2d_32_weno3_rk4", orc.SRRMHD, "rk4")]


@pytest.mark.parametrize("name, model, integrator", RKPLUS_FIXTURES)
def test_rkplus_steps_match_reference(golden, name, model, integrator):
    """RK2B, RK3, RK4 and RK4_10 (RKPlus.cc:56-726), rhs = source - F, on the
    divergent-Bx KH state and on resistive Orszag-Tang, against the reference
    build after the first and the last step."""
    z = golden(name)
    nx, ny, nz, Ng, weno, steps = (int(v) for v in z["meta"])
    b = z["bounds"]
    gam, cfl, sig, cp = (float(v) for v in z["params"])
    case = orc.Case(model=model, weno_order=weno, periodic=tuple(int(v) for v in z["periodic"]), gamma=gam,
                    cfl=cfl, sigma=sig, cp=cp, nx=nx, ny=ny, nz=nz, xmin=b[0], xmax=b[1], ymin=b[2], ymax=b[3])
    o = orc.Oracle(case)
    cons, prims, aux = o.set(z["prims0"])
    cn = parity.SRMHD_CONS if model == orc.SRMHD else parity.SRRMHD_CONS
    pn = parity.SRMHD_PRIMS if model == orc.SRMHD else parity.SRRMHD_PRIMS
    t = 0.0
    for it in range(steps):
        dt = o.timestep(t, it)
        assert dt == z["dts"][it]
        fails, _ = o.step(cons, prims, aux, dt, integrator)
        assert fails == 0
        t += dt
        if it + 1 in (1, steps):
            for got, ref, names in ((cons, z[f"cons{it + 1}"], cn), (prims, z[f"prims{it + 1}"], pn)):
                rows = parity.field_errors(parity.interior(got, 4, 2), ref, names)
                assert max(r[4] for r in rows) <= 1e-9, parity.report(f"{name} step {it + 1}", rows)


def test_rkplus_orders_differ_and_rk4_10_reset_matters(golden):
    """The four schemes give different states after one step (so no
    integrator silently falls back to another), and RK4_10's reset of the stage
    primitives at its second preparation (RKPlus.cc:612-640) changes the
    result: the oracle restated without it (fresh cons2prim of the new stage
    state) moves away from the reference."""
    z = golden("kh2d_32_weno3_rk2b")
    grid = dict(nx=32, ny=32, nz=0, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0)
    outs = {}
    for integ in ("rk2b", "rk3", "rk4", "rk4_10"):
        o = orc.Oracle(orc.Case(weno_order=3, periodic=(1, 0, 1), cp=0.1, **grid))
        c, p, a, dts, _ = o.run(z["prims0"], 1, integrator=integ)
        outs[integ] = c
    names = list(outs)
    for m in range(len(names)):
        for n in range(m + 1, len(names)):
            assert np.abs(outs[names[m]] - outs[names[n]]).max() > 1e-12, (names[m], names[n])
    z10 = golden("kh2d_32_weno3_rk4_10")
    rows = {}
    for quirks in (1, 0):
        o = orc.Oracle(orc.Case(weno_order=3, periodic=(1, 0, 1), cp=0.1, quirks=quirks, **grid))
        c, p, a, dts, _ = o.run(z10["prims0"], 1, integrator="rk4_10")
        rows[quirks] = parity.field_errors(parity.interior(c, 4, 2), z10["cons1"], parity.SRMHD_CONS)
    assert max(r[4] for r in rows[1]) <= 1e-9
    assert max(r[4] for r in rows[0]) > 1e-7


EXAMPLE_FIXTURES = [("briowu_100_rksplit", orc.SRMHD, "rksplit", 1e10), ("briowu_100_ssp2", orc.SRRMHD, "ssp2", 1e10),
                    ("ot_64_rksplit", orc.SRMHD, "rksplit", 1e10), ("khrandom_64_rksplit", orc.SRMHD, "rksplit", 1e10),
                    ("briowu_100_srrmhd_rk3", orc.SRRMHD, "rk3", 1e10)]
# the set-ups of the reference's RK2 / RKSplit serial-output tests
# (Tests/CPU/Src/test_rk2.cc, test_rkSplit.cc), end time 0.004
REFTEST_FIXTURES = [("reftest_rk2_ot40_periodic", orc.SRMHD, "rk2", 0.004),
                    ("reftest_rk2_ot40_outflow", orc.SRMHD, "rk2", 0.004),
                    ("reftest_rk2_ot40_outflow_evolve", orc.SRMHD, "rk2", 0.004),
                    ("reftest_rk2_khrandom40_flow", orc.SRMHD, "rk2", 0.004),
                    ("reftest_rk2_briowu40_periodic", orc.SRMHD, "rk2", 0.004),
                    ("reftest_rksplit_khrandom40_periodic", orc.SRMHD, "rksplit", 0.004),
                    ("reftest_rksplit_briowu40_flow", orc.SRMHD, "rksplit", 0.004),
                    ("reftest_ssp2_kh64x16_flow", orc.SRRMHD, "ssp2", 0.05)]


def example_case(z, model):
    nx, ny, nz, Ng, weno, steps = (int(v) for v in z["meta"])
    b = z["bounds"]
    gam, cfl, sig, cp = (float(v) for v in z["params"])
    grid = dict(nx=nx, ny=ny, nz=nz, xmin=b[0], xmax=b[1], ymin=b[2], ymax=b[3], zmin=b[4], zmax=b[5])
    per = tuple(int(v) for v in z["periodic"])
    return orc.Case(model=model, weno_order=weno, periodic=per, gamma=gam, cfl=cfl, sigma=sig, cp=cp, **grid), steps


@pytest.mark.parametrize("name, model, integrator, end_time", EXAMPLE_FIXTURES + REFTEST_FIXTURES)
def test_reference_examples_match_reference(golden, name, model, integrator, end_time):
    """The reference's serial examples (1-D Brio-Wu ideal and resistive,
    Orszag-Tang, random KH) at their own parameters, and the set-ups of its
    RK2 / RKSplit serial-output tests (Orszag-Tang, random KH and 2-D Brio-Wu
    under periodic, outflow and Flow boundaries), from the reference's
    post-set state, against the reference build."""
    z = golden(name)
    case, steps = example_case(z, model)
    o = orc.Oracle(case)
    oc, op, oa, dts, st = o.run(z["prims0"], steps, integrator=integrator, end_time=end_time)
    assert np.array_equal(np.array(dts), z["dts"]) and st.failures == 0
    dims = 1 if case.ny == 0 else 2
    cn = parity.SRMHD_CONS if model == orc.SRMHD else parity.SRRMHD_CONS
    pn = parity.SRMHD_PRIMS if model == orc.SRMHD else parity.SRRMHD_PRIMS
    for got, ref, names in ((oc, z[f"cons{steps}"], cn), (op, z[f"prims{steps}"], pn)):
        rows = parity.field_errors(parity.interior(got, 4, dims), parity.interior(ref, 4, dims), names)
        assert max(r[4] for r in rows) <= 1e-9


@pytest.mark.parametrize("quirks", [0, 1])
def test_c2p_retry_from_neighbour_means(quirks):
    """Cells whose Newton solve fails are solved again from the mean root of
    their converged face neighbours (srmhd.cc:465-507). Without the quirk the
    retry recovers the state that made the conserved fields; with it the retry
    solves the last interior cell's equations (the reference passes that
    cell's arguments), so a retried cell takes the last cell's root: its
    Lorentz factor equals the last cell's."""
    o = orc.Oracle(orc.Case(weno_order=3, periodic=(1, 0, 1), quirks=quirks, **parity.C2P_RETRY_GRID))
    cons, prims, aux, gp, ga = parity.c2p_retry_case(o)
    fails, st = o.c2p(cons, gp, ga)
    assert fails == len(parity.C2P_RETRY_CELLS) and st.failures == fails
    bad = tuple(np.array(parity.C2P_RETRY_CELLS).T)
    if not quirks:
        sl = (slice(None), slice(4, -4), slice(4, -4), 0)
        scale = np.abs(prims[sl]).max(axis=(1, 2), keepdims=True)
        assert (np.abs(gp[sl] - prims[sl]) / np.maximum(scale, 1e-300)).max() < 1e-12
        assert np.allclose(ga[sl], aux[sl], rtol=1e-12, atol=1e-14)
    else:
        W_last = aux[1, 4 + 15, 4 + 15, 0]
        assert np.allclose(ga[1][bad + (0,)], W_last, rtol=1e-12, atol=0)
        # away from the bad cells the recovery is untouched
        mask = np.ones((16, 16), bool)
        for i, j in parity.C2P_RETRY_CELLS:
            mask[i - 4, j - 4] = False
        got, want = gp[:, 4:-4, 4:-4, 0][:, mask], prims[:, 4:-4, 4:-4, 0][:, mask]
        assert np.allclose(got, want, rtol=1e-12, atol=1e-14)
        # and the bad cells are not recovered (the quirk is visible)
        assert np.abs(gp[1][bad + (0,)] - prims[1][bad + (0,)]).max() > 1e-3


REGIME_SRC = ["regime_src_currentsheet_128", "regime_src_briowu_128", "regime_src_kh2d_32_weno5",
              "regime_src_kh3d_12_weno3"]


def regime_case(z, regime=1):
    nx, ny, nz, Ng, weno, _ = (int(v) for v in z["meta"])
    b = z["bounds"]
    gam, cfl, sig, cp = (float(v) for v in z["params"])
    return orc.Case(nx=nx, ny=ny, nz=nz, xmin=b[0], xmax=b[1], ymin=b[2], ymax=b[3], zmin=b[4], zmax=b[5],
                    gamma=gam, cfl=cfl, sigma=sig, cp=cp, weno_order=weno,
                    periodic=tuple(int(v) for v in z["periodic"]), regime=regime)


@pytest.mark.parametrize("name", REGIME_SRC)
def test_regime_source_matches_reference(golden, name):
    """REGIME::sourceExtension (REGIME.cc:83-138) of the reference's set-up
    state, interior cells: 1-D current sheet (sigma 1e4) and Brio-Wu (sigma 50),
    2-D KH with WENO5 (whose E-flux reconstruction leaves zero faces next to
    the grid edge that reach the first interior cells) and 3-D KH."""
    z = golden(name)
    case = regime_case(z)
    o = orc.Oracle(case)
    cons, prims, aux = o.set(z["prims0"])
    src = np.zeros_like(cons)
    o.L.oc_regime_source(ctypes.byref(o.c), orc._p(cons), orc._p(prims), orc._p(aux), orc._p(src))
    dims = 1 if case.ny == 0 else (2 if case.nz == 0 else 3)
    got = parity.interior(src, 4, dims)
    ref = z["src"]
    assert np.abs(ref).max() > 0
    scale = np.abs(ref).max(axis=tuple(range(1, ref.ndim)), keepdims=True)
    assert (np.abs(got - ref) / np.maximum(scale, 1e-300)).max() <= 1e-13
    assert np.all(got[8] == 0.0)


def test_regime_source_of_uniform_state_is_zero():
    """A uniform state at rest has E = 0, q = 0 and K = 0: no resistive source."""
    grid = dict(nx=16, ny=16, nz=0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0)
    o = orc.Oracle(orc.Case(sigma=100.0, periodic=(1, 1, 1), regime=1, **grid))
    p0 = np.zeros((8, 24, 24, 1))
    p0[0], p0[4], p0[5], p0[6], p0[7] = 1.0, 0.5, 0.3, -0.2, 0.4
    cons, prims, aux = o.set(p0)
    src = np.zeros_like(cons)
    o.L.oc_regime_source(ctypes.byref(o.c), orc._p(cons), orc._p(prims), orc._p(aux), orc._p(src))
    assert np.abs(src[:, 4:-4, 4:-4]).max() == 0.0


# (fixture, integrator, tolerance). The current sheet (sigma 1e4) gets 5e-9:
# the reference's cons2prim stops at xtol 1.49e-7 (srmhd.cc:401), which leaves
# root differences that the same run without REGIME shows as 2.6e-10 in Sx,
# and the stiff REGIME source at sigma 1e4 amplifies them into Sx, a field
# 2000 times smaller than tau (its other fields agree to 2.3e-13). The
# extension's source itself equals the reference's bit for bit.
REGIME_RUNS = [("currentsheet_128_regime_rksplit", "rksplit", 5e-9), ("briowu_128_regime_rksplit", "rksplit", 1e-9),
               ("kh2d_32_weno5_regime_rk2b", "rk2b", 1e-9), ("kh3d_12_weno3_regime_rksplit", "rksplit", 1e-9),
               ("kh2d_32_weno5_regime_rk3", "rk3", 1e-9), ("kh2d_32_weno5_regime_rk4", "rk4", 1e-9),
               ("kh2d_32_weno5_regime_rk4_10", "rk4_10", 1e-9)]


@pytest.mark.parametrize("name, integrator, tol", REGIME_RUNS)
def test_regime_runs_match_reference(golden, name, integrator, tol):
    """RKSplit and the RKPlus schemes with the REGIME extension against the
    reference build: its two REGIME examples at their own parameters, 2-D
    WENO5 (RK2B, RK3, RK4, RK4_10) and 3-D."""
    z = golden(name)
    case = regime_case(z)
    steps = int(z["meta"][5])
    o = orc.Oracle(case)
    oc, op, oa, dts, st = o.run(z["prims0"], steps, integrator=integrator)
    assert np.array_equal(np.array(dts), z["dts"]) and st.failures == 0
    dims = 1 if case.ny == 0 else (2 if case.nz == 0 else 3)
    for got, ref, names in ((oc, z[f"cons{steps}"], parity.SRMHD_CONS), (op, z[f"prims{steps}"], parity.SRMHD_PRIMS)):
        got = parity.interior(got, 4, dims)
        if ref.shape != got.shape:
            ref = parity.interior(ref, 4, dims)
        rows = parity.field_errors(got, ref, names)
        assert max(r[4] for r in rows) <= tol, parity.report(name, rows)
    # the extension acts: the same run without it differs
    plain = orc.Oracle(regime_case(z, regime=0)).run(z["prims0"], steps, integrator=integrator)[0]
    assert np.abs(plain - oc).max() > 1e-8


# Known answers of the reference's REGIME tests (Tests/CPU/Src/test_REGIME.cc:36-353,
# DataAssignment1D/2D/3D): a state linear along one axis (v and B components
# 0.41 s, 0.22 s with s the coordinate, the rest constant), sigma 50, gamma 2;
# E, q = d E / d s, alpha and every dw/ds_b entry at the cell where s = 1.
# Entries: (l, m, expected, tolerance) as the reference's EXPECT_NEAR.
REGIME_KAT = {
    "1D": [
        (1, 0, 7.984099454632758e-06, 1e-11), (1, 1, 0.005702928181880542, 1e-08), (1, 2, -0.0038019521212536945, 1e-08),
        (2, 0, -0.005702928181880542, 1e-08), (2, 1, 8.37052978575575e-06, 1e-11), (2, 2, 5.79645496684485e-07, 1e-12),
        (3, 0, 0.0038019521212536945, 1e-08), (3, 1, 5.79645496684485e-07, 1e-12), (3, 2, 8.853567699659486e-06, 1e-11),
        (5, 0, 0.0, 1e-15), (5, 1, -0.008725791205832886, 1e-08), (5, 2, 0.007217193838855632, 1e-08),
        (6, 0, 0.008813616281088111, 1e-08), (6, 1, 0.0, 1e-15), (6, 2, 0.0, 1e-15),
        (7, 0, -0.007085456225972795, 1e-08), (7, 1, 0.0, 1e-15), (7, 2, 0.0, 1e-15),
        (8, 0, -0.01728160055115316, 1e-08), (8, 1, 0.0, 1e-15), (8, 2, 0.0, 1e-15),
        (9, 0, 0.0, 1e-15), (9, 1, -0.01811802983929816, 1e-08), (9, 2, -0.0012546439322175, 1e-09),
        (10, 0, 0.0, 1e-15), (10, 1, -0.0012546439322175, 1e-09), (10, 2, -0.01916356644947941, 1e-08),
        (0, 0, 0.0, 1e-15), (0, 1, 0.0, 1e-15), (0, 2, 0.0, 1e-15),
        (4, 0, 0.0, 1e-15), (4, 1, 0.0, 1e-15), (4, 2, 0.0, 1e-15),
        (11, 0, 0.0, 1e-15), (11, 1, 0.0, 1e-15), (11, 2, 0.0, 1e-15),
    ],
    "2D": [
        (1, 0, 8.853567699659485e-06, 1e-11), (1, 1, 0.0038019521212536945, 1e-08), (1, 2, 5.796454966844849e-07, 1e-12),
        (2, 0, -0.0038019521212536945, 1e-08), (2, 1, 7.984099454632758e-06, 1e-11), (2, 2, 0.005702928181880542, 1e-08),
        (3, 0, 5.796454966844848e-07, 1e-11), (3, 1, -0.005702928181880542, 1e-08), (3, 2, 8.370529785755748e-06, 1e-11),
        (5, 0, 0.0, 1e-15), (5, 1, -0.007085456225972795, 1e-08), (5, 2, 0.0, 1e-15),
        (6, 0, 0.007217193838855632, 1e-08), (6, 1, 0.0, 1e-15), (6, 2, -0.008725791205832886, 1e-08),
        (7, 0, 0.0, 1e-15), (7, 1, 0.008813616281088111, 1e-08), (7, 2, 0.0, 1e-15),
        (8, 0, -0.01916356644947941, 1e-08), (8, 1, 0.0, 1e-15), (8, 2, -0.0012546439322175, 1e-09),
        (9, 0, 0.0, 1e-15), (9, 1, -0.01728160055115316, 1e-08), (9, 2, 0.0, 1e-10),
        (10, 0, -0.0012546439322175, 1e-09), (10, 1, 0.0, 1e-10), (10, 2, -0.01811802983929816, 1e-08),
        (0, 0, 0.0, 1e-15), (0, 1, 0.0, 1e-15), (0, 2, 0.0, 1e-15),
        (4, 0, 0.0, 1e-15), (4, 1, 0.0, 1e-15), (4, 2, 0.0, 1e-15),
        (11, 0, 0.0, 1e-15), (11, 1, 0.0, 1e-15), (11, 2, 0.0, 1e-15),
    ],
    "3D": [
        (1, 0, 8.370529785755748e-06, 1e-11), (1, 1, 5.796454966844848e-07, 1e-12), (1, 2, -0.005702928181880542, 1e-08),
        (2, 0, 5.796454966844849e-07, 1e-12), (2, 1, 8.853567699659485e-06, 1e-11), (2, 2, 0.0038019521212536945, 1e-08),
        (3, 0, 0.005702928181880542, 1e-08), (3, 1, -0.0038019521212536945, 1e-08), (3, 2, 7.984099454632758e-06, 1e-11),
        (5, 0, 0.0, 1e-15), (5, 1, 0.0, 1e-15), (5, 2, 0.008813616281088111, 1e-08),
        (6, 0, 0.0, 1e-15), (6, 1, 0.0, 1e-15), (6, 2, -0.007085456225972795, 1e-08),
        (7, 0, -0.008725791205832886, 1e-08), (7, 1, 0.007217193838855632, 1e-08), (7, 2, 0.0, 1e-15),
        (8, 0, -0.01811802983929816, 1e-08), (8, 1, -0.0012546439322175, 1e-09), (8, 2, 0.0, 1e-15),
        (9, 0, -0.0012546439322175, 1e-09), (9, 1, -0.01916356644947941, 1e-08), (9, 2, 0.0, 1e-15),
        (10, 0, 0.0, 1e-10), (10, 1, 0.0, 1e-10), (10, 2, -0.01728160055115316, 1e-08),
        (0, 0, 0.0, 1e-15), (0, 1, 0.0, 1e-15), (0, 2, 0.0, 1e-15),
        (4, 0, 0.0, 1e-15), (4, 1, 0.0, 1e-15), (4, 2, 0.0, 1e-15),
        (11, 0, 0.0, 1e-15), (11, 1, 0.0, 1e-15), (11, 2, 0.0, 1e-15),
    ],
}


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_regime_fields_match_reference_known_answers(axis):
    """REGIME::set_vars / set_dwdsb on the reference tests' linear states
    (prims set on every padded cell, no boundary conditions applied)."""
    name = ("1D", "2D", "3D")[axis]
    n = (100, 10, 10)[axis]
    lo, hi = ((0.01, 2.01), (0.1, 2.1), (0.1, 2.1))[axis]
    grid = dict(nx=n if axis == 0 else 10, ny=0 if axis == 0 else 10, nz=10 if axis == 2 else 0,
                xmin=lo if axis == 0 else 0.1, xmax=hi if axis == 0 else 2.1, ymin=0.1 if axis else 0.0,
                ymax=2.1 if axis else 1.0, zmin=0.1 if axis == 2 else 0.0, zmax=2.1 if axis == 2 else 1.0)
    case = orc.Case(gamma=2.0, sigma=50.0, periodic=(1, 1, 1), **grid)
    shape = case.shape
    N = shape[axis]
    dx = (hi - lo) / n
    s = lo + (np.arange(N) + 0.5 - 4) * dx   # cell centres, simulation.cc:42-50
    mid = N // 2 - 1
    assert abs(s[mid] - 1.0) < 1e-15
    sb = s.reshape([-1 if a == axis else 1 for a in range(3)])
    p = np.zeros((8,) + shape)
    # the reference's three set-ups are one state rotated: along x (vy, By
    # linear), along y (vz, Bz linear), along z (vx, Bx linear)
    lin_v, const_v, lin_b, const_b = [(2, 3, 6, 7), (3, 1, 7, 5), (1, 2, 5, 6)][axis]
    p[0], p[4] = 0.1, 0.66
    p[lin_v] = 0.41 * sb
    p[const_v] = 0.51
    p[lin_b] = 0.22 * sb
    p[const_b] = 0.33
    ntot = int(np.prod(shape))
    E, q, alpha = np.zeros(3 * ntot), np.zeros(ntot), np.zeros(ntot)
    dw, K = np.zeros(36 * ntot), np.zeros(3 * ntot)
    o = orc.Oracle(case)
    o.L.oc_regime_fields(ctypes.byref(o.c), orc._p(p), orc._p(E), orc._p(q), orc._p(alpha), orc._p(dw), orc._p(K))
    dims = axis + 1
    cell = [mid if a < dims else 0 for a in range(3)]
    c = np.ravel_multi_index(cell, shape)
    e_exp = -(0.41 * 0.33 - 0.51 * 0.22)
    assert abs(E[axis * ntot + c] - e_exp) <= 1e-15
    assert abs(q[c] - e_exp) <= 1e-15
    assert abs(alpha[c] - 1.3825277485504367e-07) <= 1e-13
    for l, m, want, tol in REGIME_KAT[name]:
        assert abs(dw[(l * 3 + m) * ntot + c] - want) <= tol, (name, l, m)


def test_current_sheet_generator_matches_reference(golden):
    """idt.current_sheet_prims equals the reference's CurrentSheetSingleFluid
    state (initFunc.cc:191-226) on the interior, bit for bit."""
    from method_amd import initdata as idt

    z = golden("regime_src_currentsheet_128")
    g = idt.Grid(128, 0, 0, -3.0, 3.0, -1.0, 1.0, -1.5, 1.5)
    p = idt.current_sheet_prims(g, 1e4)
    assert np.array_equal(p[:, 4:-4], z["prims0"][:, 4:-4])


def _sheet_run(grid, direction, end_time=0.1):
    """The reference's REGIME symmetry set-ups (test_REGIME.cc:943-1088):
    current sheet, gamma 2, sigma 100, cfl 0.4, cp 0.1, WENO3, Outflow,
    RKSplit + REGIME, evolved to t = 0.1."""
    from method_amd import initdata as idt

    g = idt.Grid(**grid)
    o = orc.Oracle(orc.Case(gamma=2.0, sigma=100.0, cfl=0.4, cp=0.1, periodic=(0, 0, 0), regime=1, **grid))
    cons, prims, aux = o.set(idt.current_sheet_prims(g, 100.0, direction))
    t, it = 0.0, 0
    while t < end_time:
        dt = o.timestep(t, it, end_time)
        fails, _ = o.step(cons, prims, aux, dt, "rksplit")
        assert fails == 0
        t += dt
        it += 1
    return cons, prims


def test_regime_sheet_is_independent_of_y_and_z():
    """YAxisSymmetries / ZAxisSymmetries: a sheet along x stays uniform in y
    (2-D) and in y and z (3-D) to 1e-15."""
    _, p2 = _sheet_run(dict(nx=10, ny=10, nz=0, xmin=-3.0, xmax=3.0, ymin=-1.0, ymax=1.0), 0)
    inner = p2[:, 4:-4, 4:-4, 0]
    assert np.abs(np.diff(inner, axis=2)).max() <= 1e-15
    _, p3 = _sheet_run(dict(nx=10, ny=10, nz=10, xmin=-3.0, xmax=3.0, ymin=-1.0, ymax=1.0, zmin=-1.0, zmax=1.0), 0)
    inner = p3[:, 4:-4, 4:-4, 4:-4]
    assert np.abs(np.diff(inner, axis=2)).max() <= 1e-15 and np.abs(np.diff(inner, axis=3)).max() <= 1e-15
    assert np.abs(np.diff(inner[6], axis=0)).max() > 1e-3  # the sheet varies along x


def test_regime_sheet_rotations_agree():
    """RotSymmetries (test_REGIME.cc:1031-1088): the sheet along x, y and z
    on [-3, 3]^3 gives the same field, with indices and components rotated."""
    grid = dict(nx=10, ny=10, nz=10, xmin=-3.0, xmax=3.0, ymin=-3.0, ymax=3.0, zmin=-3.0, zmax=3.0)
    ca, _ = _sheet_run(grid, 0)
    cb, _ = _sheet_run(grid, 1)
    cc, _ = _sheet_run(grid, 2)
    a = ca[6, 4:-4, 4:-4, 4:-4]
    b = np.transpose(cb[7, 4:-4, 4:-4, 4:-4], (1, 2, 0))   # dB[7, k, i, j] at (i, j, k)
    c = np.transpose(cc[5, 4:-4, 4:-4, 4:-4], (2, 0, 1))   # dC[5, j, k, i] at (i, j, k)
    assert np.abs(a - b).max() <= 1e-15 and np.abs(a - c).max() <= 1e-15


# ------------------------------------------- FVS::fluxReconstruction dumps
@pytest.mark.parametrize("name, grid, order, periodic", [
    ("frecon_ot3d_weno3", dict(nx=10, ny=12, nz=8, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0, Ng=3,
                               gamma=5.0 / 3.0, cfl=0.5), 3, (0, 0, 0)),
    ("frecon_kh2d_weno5", dict(nx=16, ny=12, nz=0, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0), 5, (1, 0, 1)),
])
def test_flux_reconstruction_matches_reference(golden, name, grid, order, periodic):
    """The oracle's FVS::fluxReconstruction of the reference's own f and cons
    equals the reference's frecon bit for bit on every axis (test_fvs.cc's
    calls), including the zeros outside [shift, N - shift)."""
    z = golden(name)
    o = orc.Oracle(orc.Case(weno_order=order, periodic=periodic, **grid))
    for d in range(3):
        got = o.flux_reconstruction(z["cons0"], z[f"f{d}"], d)
        assert np.array_equal(got, z[f"frecon{d}"]), d
    # and the fixture's f is the oracle's flux vector of the post-set state
    for d in range(3 if grid["nz"] else 2):
        c, p, a = o.set(z["prims0"])
        f = o.flux_vector(c, p, a, d)
        assert np.allclose(f, z[f"f{d}"], rtol=0, atol=1e-14 * np.abs(z[f"f{d}"]).max()), d


# ----------------------------------------------- 3-D resistive trajectories
SRRMHD3D = dict(nx=12, ny=14, nz=10, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0)


def srrmhd3d_case(z, name):
    periodic = tuple(int(b) for b in z["periodic"])
    return orc.Case(model=orc.SRRMHD, weno_order=int(z["weno"]), periodic=periodic, sigma=float(z["sigma"]), cp=0.1,
                    cfl=0.5, **SRRMHD3D)


def srrmhd3d_prims0(case, z):
    """The fixture's interior primitives with ghosts holding their boundary images."""
    _, p, _ = case.zeros()
    p[:, 4:-4, 4:-4, 4:-4] = z["prims0"]
    orc.Oracle(case).apply_bcs(p)
    return p


@pytest.mark.parametrize("name", ["srrmhd3d_12x14x10_ssp2_periodic", "srrmhd3d_12x14x10_ssp2_flow",
                                  "srrmhd3d_12x14x10_rksplit_periodic", "srrmhd3d_12x14x10_rksplit_flow"])
def test_srrmhd_3d_matches_reference(golden, name):
    """3-D resistive runs of the reference build (SSP2 at sigma 50 and RKSplit
    at sigma 10, periodic and Flow, 3 steps): the oracle's z-direction flux
    quirk (srrmhd.cc:163-171) and its 3-D stencils agree with the reference to
    1e-9 of each field's group scale, dt bit for bit. The same runs without
    the quirk move away from the reference, so the quirk is exercised."""
    z = golden(name)
    case = srrmhd3d_case(z, name)
    integ = "ssp2" if "ssp2" in name else "rksplit"
    p0 = srrmhd3d_prims0(case, z)
    c, p, a, dts, st = orc.Oracle(case).run(p0, 3, integrator=integ)
    assert np.array_equal(np.array(dts), z["dts"])
    worst = 0.0
    for got, ref, names in ((c, z["cons3"], parity.SRRMHD_CONS), (p, z["prims3"], parity.SRRMHD_PRIMS)):
        rows = parity.field_errors(parity.interior(got, 4, 3), ref, names)
        print(parity.report(f"oracle vs reference, {name}", rows))
        for n, dm, fm, rf, rg, pw in rows:
            assert rg <= 1e-9, n
            worst = max(worst, rg)
    case0 = srrmhd3d_case(z, name)
    case0.quirks = 0
    c0, _, _, _, _ = orc.Oracle(case0).run(p0, 3, integrator=integ)
    rows = parity.field_errors(parity.interior(c0, 4, 3), z["cons3"], parity.SRRMHD_CONS)
    assert max(r[4] for r in rows) > 1e3 * max(worst, 1e-16)
