"""The reference's serial example set-ups (Examples/BrioWuShockTube/*,
OrszagTangVortex, KelvinHelmholtz random) on the GPU engine through the C ABI:
1-D and 2-D grids, RKSplit and SSP2, from the reference's post-set state,
against the reference build. Needs an MI355X."""
import sys

import numpy as np
import pytest

from method_amd import engine as eng
from tests import parity

pytestmark = pytest.mark.gpu

CASES = [("briowu_100_rksplit", eng.SRMHD, eng.RKSPLIT, 0.4), ("briowu_100_ssp2", eng.SRRMHD, eng.SSP2, 0.4),
         ("ot_64_rksplit", eng.SRMHD, eng.RKSPLIT, 1.0), ("khrandom_64_rksplit", eng.SRMHD, eng.RKSPLIT, 3.0),
         # the set-ups of the reference's RK2 / RKSplit serial-output tests (Tests/CPU/Src/test_rk2.cc,
         # test_rkSplit.cc): 40x40, periodic / outflow / Flow, end time 0.004
         ("reftest_rk2_ot40_periodic", eng.SRMHD, eng.RK2, 0.004), ("reftest_rk2_ot40_outflow", eng.SRMHD, eng.RK2, 0.004),
         ("reftest_rk2_ot40_outflow_evolve", eng.SRMHD, eng.RK2, 0.004),
         ("reftest_rk2_khrandom40_flow", eng.SRMHD, eng.RK2, 0.004),
         ("reftest_rk2_briowu40_periodic", eng.SRMHD, eng.RK2, 0.004),
         ("reftest_rksplit_khrandom40_periodic", eng.SRMHD, eng.RKSPLIT, 0.004),
         ("reftest_rksplit_briowu40_flow", eng.SRMHD, eng.RKSPLIT, 0.004),
         # test_imex.cc SSP2FlowKHSingleFluid: resistive KH 64x16, sigma 0, Flow, evolve to 0.05 (34 steps)
         ("reftest_ssp2_kh64x16_flow", eng.SRRMHD, eng.SSP2, 0.05)]


@pytest.mark.parametrize("name, model, integrator, end_time", CASES)
def test_example_matches_reference(golden, name, model, integrator, end_time):
    z = golden(name)
    nx, ny, nz, Ng, weno, steps = (int(v) for v in z["meta"])
    b = z["bounds"]
    gam, cfl, sig, cp = (float(v) for v in z["params"])
    bc = tuple(eng.PERIODIC if p else eng.OUTFLOW for p in z["periodic"])
    cfg = eng.Config(nx=nx, ny=ny, nz=nz, xmin=b[0], xmax=b[1], ymin=b[2], ymax=b[3], zmin=b[4], zmax=b[5],
                     gamma=gam, cfl=cfl, sigma=sig, cp=cp, model=model, integrator=integrator, weno_order=weno, bc=bc)
    print(f"gpu: {name} {nx}x{ny} {steps} steps", flush=True)
    sys.stdout.flush()
    with eng.Engine(cfg) as e:
        e.set_prims(z["prims0"])
        for s in range(steps):
            rc, st = e.update_time(end_time, stats=True)
            assert rc == 0 and st["failures"] == 0
            assert e.time()[1] == z["dts"][s]
        c, p, a = e.download()
    dims = 1 if ny == 0 else 2
    cn = parity.SRMHD_CONS if model == eng.SRMHD else parity.SRRMHD_CONS
    pn = parity.SRMHD_PRIMS if model == eng.SRMHD else parity.SRRMHD_PRIMS
    for got, ref, names in ((c, z[f"cons{steps}"], cn), (p, z[f"prims{steps}"], pn)):
        rows = parity.field_errors(parity.interior(got, Ng, dims), parity.interior(ref, Ng, dims), names)
        print(parity.report(f"gpu vs reference, {name}", rows), flush=True)
        for n_, dm, fm, rf, rg, pw in rows:
            assert rg <= 1e-9, f"{n_}: {rg:.3e}"
