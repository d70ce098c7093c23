"""The reference-API operators outside the fused step, through the C ABI:
FVS::fluxReconstruction (fluxVectorSplitting.cc:4-56, flux.h:72) and the
single-cell model calls Model::getPrimitiveVarsSingleCell /
sourceTermSingleCell (model.h:55-85), against the reference build's dumps,
the reference's own stationary test and the CPU oracle. Needs an MI355X."""
import sys

import numpy as np
import pytest

from method_amd import engine as eng
from method_amd import initdata as idt
from oracle import oracle as orc
from tests import parity

pytestmark = pytest.mark.gpu

# (fixture, grid, model constants, boundary) of the reference frecon dumps
FRECON = {
    "frecon_ot3d_weno3": (dict(nx=10, ny=12, nz=8, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0,
                               Ng=3, gamma=5.0 / 3.0, cfl=0.5, weno_order=3),
                          (eng.OUTFLOW, eng.OUTFLOW, eng.OUTFLOW)),
    "frecon_kh2d_weno5": (dict(nx=16, ny=12, nz=0, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, gamma=4.0 / 3.0,
                               cfl=0.4, weno_order=5), (eng.PERIODIC, eng.OUTFLOW, eng.PERIODIC)),
}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel_by_field(got, ref):
    """max |got - ref| of each variable over that variable's max |ref| (0 when both vanish)."""
    out = []
    for v in range(ref.shape[0]):
        s = np.abs(ref[v]).max()
        d = np.abs(got[v] - ref[v]).max()
        out.append(d / s if s > 0 else (0.0 if d == 0 else np.inf))
    return np.array(out)


@pytest.mark.parametrize("name", sorted(FRECON))
def test_flux_reconstruction_of_given_arrays_matches_reference(golden, name):
    """The reference's own f and cons in, its frecon out, every axis: the GPU's
    face fluxes agree to 1e-10 of each variable's scale, and are exactly 0
    outside [shift, N - shift) along the axis. (The oracle, whose WENO is the
    reference's expression, agrees bit for bit. The engine's one-division
    first-difference form differs by rounding, which the weights amplify where
    a smoothness indicator's square is near eps = 1e-16: there the reference's
    expanded indicator itself carries a relative error near 1e-8, and the
    faces of the KH state differ by up to 2e-11.)"""
    z = golden(name)
    grid, bc = FRECON[name]
    say(f"gpu: FVS::fluxReconstruction of the reference's arrays, {name}")
    with eng.Engine(eng.Config(bc=bc, **grid)) as e:
        shift = (grid["weno_order"] + 1) // 2
        for d in range(3):
            got = e.flux_reconstruction(d, cons=z["cons0"], f=z[f"f{d}"])
            ref = z[f"frecon{d}"]
            rel = _rel_by_field(got, ref)
            say(f"  dir {d}: max rel by variable {rel.max():.3e}")
            assert rel.max() <= 1e-10, (d, rel)
            n = got.shape[1 + d]
            idx = np.arange(n)
            outside = (idx < shift) | (idx >= n - shift)
            sl = [slice(None)] * 4
            sl[1 + d] = outside
            assert np.all(got[tuple(sl)] == 0.0) and np.all(ref[tuple(sl)] == 0.0)
        # a subset of the variables leaves the others untouched (vars < Ncons)
        part = e.flux_reconstruction(0, cons=z["cons0"], f=z["f0"], vars=5)
        assert part.shape[0] == 5 and np.array_equal(part, e.flux_reconstruction(0, cons=z["cons0"], f=z["f0"])[:5])


@pytest.mark.parametrize("name", sorted(FRECON))
def test_flux_reconstruction_of_resident_state_matches_reference(golden, name):
    """Simulation::set on the reference's post-set primitives, then the
    reconstruction of the resident state's own flux vector (test_fvs.cc's
    fluxVector + fluxReconstruction) against the reference's, every axis."""
    z = golden(name)
    grid, bc = FRECON[name]
    say(f"gpu: fluxReconstruction of the resident state, {name}")
    with eng.Engine(eng.Config(bc=bc, **grid)) as e:
        e.set_prims(z["prims0"])
        for d in range(3):
            got = e.flux_reconstruction(d)
            rel = _rel_by_field(got, z[f"frecon{d}"])
            say(f"  dir {d}: max rel by variable {rel.max():.3e}")
            assert rel.max() <= 1e-10, (d, rel)


def test_flux_reconstruction_stationary():
    """FluxVectorSplittingStationary of Tests/CPU/Src/test_srmhd.cc:27-93:
    rho = 0.5 at rest, every other primitive 0, 10^3 periodic, gamma 5/3,
    cp 0.5; primsToAll, then fluxReconstruction with the (zero) f the test
    passes is 0 to 1e-15 in every interior cell, for every axis."""
    cfg = eng.Config(nx=10, ny=10, nz=10, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0, cfl=1.0,
                     Ng=4, gamma=5.0 / 3.0, sigma=1000.0, cp=0.5, weno_order=3, bc=(eng.PERIODIC,) * 3)
    with eng.Engine(cfg) as e:
        p = e.zeros(8)
        p[0] = 0.5
        e.set_prims(p)
        c, _, _ = e.download()
        f = np.zeros_like(c)
        for d in range(3):
            fr = e.flux_reconstruction(d, cons=c, f=f)
            assert np.abs(parity.interior(fr, 4, 3)).max() <= 1e-15, d
            # the resident state's own flux vector: the rest-frame flux is the pressure only
            fr2 = e.flux_reconstruction(d)
            assert np.abs(parity.interior(fr2, 4, 3)).max() <= 1e-15, d


def test_flux_reconstruction_rejects_bad_arguments():
    with eng.Engine(eng.Config(nx=8, ny=8, nz=0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0)) as e:
        c = e.zeros(9)
        with pytest.raises(eng.EngineError, match="dir"):
            e.flux_reconstruction(3, cons=c, f=c)
        rc = e.L.mgpu_flux_reconstruction(e._h, eng._dp(c), None, 0, -1, eng._dp(c))
        assert rc < 0 and "neither" in e.L.mgpu_last_error().decode()
        # dir 2 on a 2-D grid: an empty loop range, all zero as in the reference
        assert np.all(e.flux_reconstruction(2, cons=c + 1.0, f=c + 2.0) == 0.0)


def _cells(a):
    """[var][i][j][k] interior -> [cell][var]."""
    return np.ascontiguousarray(a.reshape(a.shape[0], -1).T)


def test_c2p_cells_srmhd_matches_reference_vectors(golden):
    """getPrimitiveVarsSingleCell on the 2,000 reference cons2prim vectors
    (random states over the learned-guess training ranges; srmhd.cc:297-388):
    the roots agree with the true prims to 1e-11 and with the reference's
    hybrd (xtol 1.49e-8) to 1e-7, with no failure."""
    z = golden("srmhd_c2p_vectors")
    n = int(z["n"][0])
    cons, gp, ga, pt, pr = (_cells(z[k][:, 4:-4, 0, 0]) for k in ("cons", "guess_prims", "guess_aux", "prims_true",
                                                                    "prims_ref"))
    with eng.Engine(eng.Config(nx=8, ny=8, nz=0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0)) as e:
        rc, p, a, st = e.c2p_cells(cons, gp, ga)
    assert rc == 0 and st["failures"] == 0 and st["cells"] == n
    scale = np.abs(pt).max(axis=0, keepdims=True)
    say(f"c2p cells: vs true {np.max(np.abs(p - pt) / scale):.3e}, vs hybrd {np.max(np.abs(p - pr) / scale):.3e}, "
        f"{st['iterations'] / n:.2f} Newton steps/cell")
    assert (np.abs(p - pt) / scale).max() < 1e-11
    assert (np.abs(p - pr) / scale).max() < 1e-7
    # the aux holds v^2 of the recovered velocity and W of the root, equal to the Newton tolerance
    assert np.allclose(a[:, 9], np.sum(p[:, 1:4] ** 2, axis=1), rtol=1e-14, atol=0)
    assert np.allclose(a[:, 1], 1.0 / np.sqrt(1.0 - a[:, 9]), rtol=1e-10, atol=0)


def test_c2p_cells_srmhd_equals_grid_c2p():
    """The cell list and the grid cons2prim share the Newton and recovery code:
    with the same guesses (the grid takes h from p, the cell list from aux h,
    set here from p the same way) the results agree to rounding."""
    g = idt.kh2d_grid(16)
    p0 = idt.kh_prims(g, seed=3)
    cfg = eng.Config(nx=16, ny=16, nz=0, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0)
    with eng.Engine(cfg) as e:
        e.set_prims(p0)
        c, p, a = e.download()
        # perturb the guess so the solve takes steps
        gp, ga = p.copy(), a.copy()
        gp[1:4] *= 0.97
        gp[4] *= 1.05
        gam = 4.0 / 3.0
        ga[0] = 1.0 + gp[4] / gp[0] * (gam / (gam - 1.0))
        e.upload(cons=c, prims=gp, aux=ga)
        rc, st = e.c2p()
        _, pg, ag = e.download()
        rc2, pc, ac, st2 = e.c2p_cells(_cells(parity.interior(c, 4, 2)), _cells(parity.interior(gp, 4, 2)),
                                       _cells(parity.interior(ga, 4, 2)))
    assert rc == rc2 == 0 and st["cells"] == st2["cells"] == 256 and st2["iterations"] > 0
    for got, ref in ((pc, _cells(parity.interior(pg, 4, 2))), (ac, _cells(parity.interior(ag, 4, 2)))):
        scale = np.maximum(np.abs(ref).max(axis=0), 1e-300)
        assert (np.abs(got - ref) / scale).max() <= 1e-13


def test_c2p_cells_srrmhd_matches_oracle():
    """SRRMHD getPrimitiveVarsSingleCell (srrmhd.cc:256-312): the secant from
    each cell's aux rhohWsq, on the resistive OT state with that guess moved
    by 1 %, against the oracle's full-grid cons2prim of the same arrays (which
    runs the same single-cell routine per cell)."""
    grid = dict(nx=32, ny=32, nz=0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0)
    case = orc.Case(model=orc.SRRMHD, weno_order=3, periodic=(1, 1, 1), sigma=1e3, cp=0.1, cfl=0.5, **grid)
    o = orc.Oracle(case)
    c, p, a = o.set(idt.ot_prims_resistive(idt.ot2d_grid(32)))
    a[10] *= 1.01
    op, oa = p.copy(), a.copy()
    o.c2p(c, op, oa)
    with eng.Engine(eng.Config(model=eng.SRRMHD, integrator=eng.SSP2, sigma=1e3, cp=0.1, cfl=0.5,
                               bc=(eng.PERIODIC,) * 3, **grid)) as e:
        rc, gp, ga, st = e.c2p_cells(_cells(parity.interior(c, 4, 2)), _cells(parity.interior(p, 4, 2)),
                                     _cells(parity.interior(a, 4, 2)))
    assert rc == 0 and st["failures"] == 0 and st["cells"] == 1024 and st["iterations"] > 0
    ref_p, ref_a = _cells(parity.interior(op, 4, 2)), _cells(parity.interior(oa, 4, 2))
    for got, ref in ((gp, ref_p), (ga, ref_a)):
        scale = np.maximum(np.abs(ref).max(axis=0), 1e-300)
        if ref is ref_a:
            # the current J (aux 4..6) of this ideal-field state is rounding noise of size sigma |E| eps
            scale[4:7] = 1e3 * np.abs(ref_p[:, 8:11]).max()
        assert (np.abs(got - ref) / scale).max() <= 1e-12


def test_source_cells_match_grid_source():
    """sourceTermSingleCell per cell against the full-grid Model::sourceTerm of
    the oracle on the same state, both models (srmhd.cc:212-250,
    srrmhd.cc:194-254); SRRMHD's current comes from each cell's aux."""
    # SRMHD with a non-zero phi
    grid = dict(nx=16, ny=16, nz=0, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0)
    o = orc.Oracle(orc.Case(weno_order=3, cp=0.3, **grid))
    c, p, a = o.set(idt.kh_prims(idt.kh2d_grid(16), seed=2))
    c[8] = np.linspace(-1.0, 1.0, c[8].size).reshape(c[8].shape)
    os_ = o.source(c, p, a)
    with eng.Engine(eng.Config(cp=0.3, **grid)) as e:
        got = e.source_cells(_cells(c))
    assert np.array_equal(got, _cells(os_))
    # SRRMHD
    gr = dict(nx=32, ny=32, nz=0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0)
    o = orc.Oracle(orc.Case(model=orc.SRRMHD, weno_order=3, periodic=(1, 1, 1), sigma=10.0, cp=0.2, cfl=0.5, **gr))
    c, p, a = o.set(idt.ot_prims_resistive(idt.ot2d_grid(32)))
    c[11] += 0.01
    c[12] -= 0.02
    os_ = o.source(c, p, a)
    with eng.Engine(eng.Config(model=eng.SRRMHD, integrator=eng.SSP2, sigma=10.0, cp=0.2, cfl=0.5,
                               bc=(eng.PERIODIC,) * 3, **gr)) as e:
        got = e.source_cells(_cells(c), _cells(a))
    ref = _cells(os_)
    scale = np.maximum(np.abs(ref).max(axis=0), 1e-300)
    assert (np.abs(got - ref) / scale).max() <= 1e-15


def test_max_wavespeed_refused_for_srrmhd():
    with eng.Engine(eng.Config(model=eng.SRRMHD, integrator=eng.SSP2, nx=8, ny=8, nz=0, xmin=0.0, xmax=1.0,
                               ymin=0.0, ymax=1.0, sigma=10.0, bc=(eng.PERIODIC,) * 3)) as e:
        with pytest.raises(eng.EngineError, match="light modes"):
            e.max_wavespeed()


def test_batch_rejects_an_instance_that_cannot_step():
    for end in (0.0, -1.0, float("nan"), float("inf")):
        with pytest.raises(eng.EngineError, match="end_time"):
            eng.Batch([eng.Instance(nx=16, ny=16, end_time=end)], weno_order=5)
