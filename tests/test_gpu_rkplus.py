"""The RKPlus integrators (RK2B, RK3, RK4, RK4_10; RKPlus.cc:56-726) on the
GPU through the C ABI, against the reference build's trajectories and the
oracle. Needs an MI355X."""
import sys

import numpy as np
import pytest

from method_amd import engine as eng
from oracle import oracle as orc
from tests import parity

pytestmark = pytest.mark.gpu

INTEG = {"rk2b": eng.RK2B, "rk3": eng.RK3, "rk4": eng.RK4, "rk4_10": eng.RK4_10}
FIXTURES = [("kh2d_32_weno3_rk2b", eng.SRMHD, "rk2b"), ("kh2d_32_weno3_rk3", eng.SRMHD, "rk3"),
            ("kh2d_32_weno3_rk4", eng.SRMHD, "rk4"), ("kh2d_32_weno3_rk4_10", eng.SRMHD, "rk4_10"),
            ("ot2d_32_weno3_rk4", eng.SRRMHD, "rk4"), ("briowu_100_srrmhd_rk3", eng.SRRMHD, "rk3")]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def config_of(z, model, integrator, quirks=1):
    nx, ny, nz, Ng, weno, steps = (int(v) for v in z["meta"])
    b = z["bounds"]
    gam, cfl, sig, cp = (float(v) for v in z["params"])
    bc = tuple(eng.PERIODIC if int(v) else eng.OUTFLOW for v in z["periodic"])
    return eng.Config(nx=nx, ny=ny, nz=nz, xmin=b[0], xmax=b[1], ymin=b[2], ymax=b[3], zmin=b[4], zmax=b[5],
                      gamma=gam, sigma=sig, cp=cp, cfl=cfl, model=model, weno_order=weno,
                      integrator=INTEG[integrator], bc=bc, quirks=quirks), steps


@pytest.mark.parametrize("name, model, integrator", FIXTURES)
def test_rkplus_matches_reference(golden, name, model, integrator):
    z = golden(name)
    cfg, steps = config_of(z, model, integrator)
    dims = 1 if cfg.ny == 0 else 2
    say(f"gpu: {integrator} {name} {steps} steps vs reference")
    cn = parity.SRMHD_CONS if model == eng.SRMHD else parity.SRRMHD_CONS
    pn = parity.SRMHD_PRIMS if model == eng.SRMHD else parity.SRRMHD_PRIMS
    with eng.Engine(cfg) as e:
        e.set_prims(z["prims0"])
        for s in range(steps):
            rc, st = e.update_time(stats=True)
            assert rc == 0 and st["failures"] == 0
            assert e.time()[1] == z["dts"][s]
        c, p, a = e.download()
    for got, ref, names in ((c, z[f"cons{steps}"], cn), (p, z[f"prims{steps}"], pn)):
        if ref.shape != got.shape:  # fixtures stored as interior cells
            got = parity.interior(got, 4, dims)
        else:
            got, ref = parity.interior(got, 4, dims), parity.interior(ref, 4, dims)
        rows = parity.field_errors(got, ref, names)
        say(parity.report(f"gpu vs reference, {integrator} {name}", rows))
        for n, dm, fm, rf, rg, pw in rows:
            assert rg <= 1e-9, f"{n}: {rg:.3e}"


@pytest.mark.parametrize("quirks", [1, 0])
def test_rk4_10_prims_reset_matches_oracle(golden, quirks):
    """RK4_10 with and without the reference's reset of the stage primitives
    at its second preparation (RKPlus.cc:612-640), against the oracle."""
    z = golden("kh2d_32_weno3_rk4_10")
    cfg, _ = config_of(z, eng.SRMHD, "rk4_10", quirks=quirks)
    say(f"gpu: RK4_10 quirks={quirks} vs oracle")
    o = orc.Oracle(orc.Case(weno_order=3, periodic=(1, 0, 1), cp=0.1, quirks=quirks, nx=32, ny=32, nz=0,
                            xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0))
    oc, op, oa, _, _ = o.run(z["prims0"], 2, integrator="rk4_10")
    with eng.Engine(cfg) as e:
        e.set_prims(z["prims0"])
        for _ in range(2):
            assert e.update_time() == 0
        c, p, a = e.download()
    for got, ref, names in ((c, oc, parity.SRMHD_CONS), (p, op, parity.SRMHD_PRIMS)):
        rows = parity.field_errors(parity.interior(got, 4, 2), parity.interior(ref, 4, 2), names)
        say(parity.report(f"RK4_10 quirks={quirks}", rows))
        for n, dm, fm, rf, rg, pw in rows:
            assert rg <= 1e-11, f"{n}: {rg:.3e}"
