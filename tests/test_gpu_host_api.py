"""The C++ host API (method_amd/host, same classes and signatures as the
reference's Project/CPU/Include) driving the engine end to end, as a user's
main.cc does. Needs an MI355X."""
import os
import subprocess

import numpy as np
import pytest

from method_amd import initdata as idt
from tests import parity

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "method_amd", "lib", "kh2d_main")


def _run(tmp_path, steps, prims_in=None):
    out = tmp_path / "state.bin"
    cmd = [EXE, str(out), str(steps)] + ([str(prims_in)] if prims_in is not None else [])
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    shape = (72, 72, 1)
    n = 72 * 72
    a = np.fromfile(out)
    return a[:9 * n].reshape((9,) + shape), a[9 * n:17 * n].reshape((8,) + shape), a[17 * n:].reshape((13,) + shape)


def test_host_api_kh2d_matches_reference(golden, tmp_path):
    z = golden("kh2d_64_weno3_rk2")
    pin = tmp_path / "prims0.bin"
    np.ascontiguousarray(z["prims0"]).tofile(pin)
    c, p, a = _run(tmp_path, 10, pin)
    for got, ref, names in ((c, z["cons10"], parity.SRMHD_CONS), (p, z["prims10"], parity.SRMHD_PRIMS),
                            (a, z["aux10"], parity.SRMHD_AUX)):
        rows = parity.field_errors(parity.interior(got, 4, 2), parity.interior(ref, 4, 2), names)
        print(parity.report("host API (C++) vs reference, KH 2-D 64^2, 10 steps", rows), flush=True)
        for n_, dm, fm, rf, rg, pw in rows:
            assert rf <= 1e-9, n_
            if fm == 0.0:
                assert dm == 0.0, n_


def test_host_initial_data_matches_numpy_generator(tmp_path):
    c, p, a = _run(tmp_path, 0)
    ref = idt.kh_prims(idt.kh2d_grid(64), seed=1234)
    inner = (slice(None), slice(4, -4), slice(4, -4), slice(None))
    assert np.max(np.abs(p[inner] - ref[inner])) <= 1e-15


EXAMPLE = os.path.join(ROOT, "method_amd", "lib", "kh_example_main")


def _parse_field(text: bytes, nvar: int, nx: int, ny: int):
    lines = text.decode().split("\n")
    vals = np.array([float(v) for line in lines[1:] for v in line.split()])
    return lines[0], vals.reshape((nvar, nx, ny))


def test_reference_kh_example_output_matches_reference(golden, tmp_path):
    """The reference's KH example set-up (KHInstabilitySingleFluid, SRMHD, WENO3,
    Periodic, RKSplit, SerialSaveData) run through the host API for 5
    updateTime calls: the text files of saveAll against the reference build's.
    Headers, domain and constants (t and dt included) are identical; every
    field value is within 1e-9 of its field's scale."""
    z = golden("kh_example_32_rksplit_save")
    n, steps = (int(v) for v in z["meta"])
    print(f"run: {EXAMPLE} {n} {steps}", flush=True)
    r = subprocess.run([EXAMPLE, str(n), str(steps)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    final = tmp_path / "Data" / "Final"
    for key, rel in (("domain", "Domain/domain.dat"), ("constants", "Constants/constants.dat")):
        assert (final / rel).read_bytes() == bytes(z["text_" + key]), rel
    for key, rel, nvar in (("cons", "Conserved/cons.dat", 9), ("prims", "Primitive/prims.dat", 8),
                           ("aux", "Auxiliary/aux.dat", 13)):
        h_got, got = _parse_field((final / rel).read_bytes(), nvar, n, n)
        h_ref, ref = _parse_field(bytes(z["text_" + key]), nvar, n, n)
        assert h_got == h_ref
        for v in range(nvar):
            scale = max(np.abs(ref[v]).max(), 1e-300)
            err = np.abs(got[v] - ref[v]).max() / scale
            assert err <= 1e-9 or (scale == 1e-300 and np.abs(got[v]).max() == 0.0), f"{key}[{v}]: {err:.3e}"


def test_host_api_srrmhd_ssp2_matches_reference(golden, tmp_path):
    """SRRMHD + SSP2 through the host API (OT 32^2, sigma 1e3, 3 steps) against
    the reference build."""
    z = golden("ot2d_32_weno3_ssp2")
    exe = os.path.join(ROOT, "method_amd", "lib", "ot2d_ssp2_main")
    pin = tmp_path / "prims0.bin"
    np.ascontiguousarray(z["prims0"]).tofile(pin)
    out = tmp_path / "state.bin"
    print(f"run: {exe} {out} 3 {pin}", flush=True)
    r = subprocess.run([exe, str(out), "3", str(pin)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    shape = (40, 40, 1)
    n = 40 * 40
    a = np.fromfile(out)
    c, p = a[:14 * n].reshape((14,) + shape), a[14 * n:25 * n].reshape((11,) + shape)
    for got, ref, names in ((c, z["cons3"], parity.SRRMHD_CONS), (p, z["prims3"], parity.SRRMHD_PRIMS)):
        rows = parity.field_errors(parity.interior(got, 4, 2), parity.interior(ref, 4, 2), names)
        print(parity.report("host API SRRMHD SSP2 vs reference, OT 32^2, 3 steps", rows), flush=True)
        for n_, dm, fm, rf, rg, pw in rows:
            assert rg <= 1e-9, f"{n_}: {rg:.3e}"


RKPLUS_EXE = os.path.join(ROOT, "method_amd", "lib", "rkplus_main")


def _run_rkplus(tmp_path, integ, prims0, steps):
    pin, out = tmp_path / "prims0.bin", tmp_path / f"{integ}.bin"
    np.ascontiguousarray(prims0).tofile(pin)
    cmd = [RKPLUS_EXE, integ, str(pin), str(out), str(steps)]
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return np.fromfile(out)


@pytest.mark.parametrize("integ", ["rk2b", "rk3", "rk4", "rk4_10"])
def test_host_api_rkplus_matches_reference(golden, tmp_path, integ):
    """RK2B / RK3 / RK4 / RK4_10 built through the host API's RKPlus classes,
    4 steps of the divergent-Bx KH 32^2 state, against the reference build."""
    z = golden(f"kh2d_32_weno3_{integ}")
    a = _run_rkplus(tmp_path, integ, z["prims0"], 4)
    shape, n = (40, 40, 1), 40 * 40
    c, p = a[:9 * n].reshape((9,) + shape), a[9 * n:17 * n].reshape((8,) + shape)
    for got, ref, names in ((c, z["cons4"], parity.SRMHD_CONS), (p, z["prims4"], parity.SRMHD_PRIMS)):
        rows = parity.field_errors(parity.interior(got, 4, 2), ref, names)
        print(parity.report(f"host API {integ} vs reference, KH 32^2, 4 steps", rows), flush=True)
        for n_, dm, fm, rf, rg, pw in rows:
            assert rg <= 1e-9, n_


def test_host_api_rkplus_rhs_matches_oracle(golden, tmp_path):
    """RKPlus::rhs (RKPlus.cc:15-50) through the host API: source - F of the
    set-up state, interior cells, against the oracle's FVS::F and
    Model::sourceTerm."""
    from oracle import oracle as orc

    z = golden("kh2d_32_weno3_rk4")
    rhs = _run_rkplus(tmp_path, "rk4", z["prims0"], 0).reshape((9, 40, 40, 1))
    o = orc.Oracle(orc.Case(weno_order=3, periodic=(1, 0, 1), cp=0.1, nx=32, ny=32, nz=0, xmin=-0.5, xmax=0.5,
                            ymin=-1.0, ymax=1.0))
    cons, prims, aux = o.set(z["prims0"])
    want = o.source(cons, prims, aux) - o.fnet(cons, prims, aux)
    rows = parity.field_errors(parity.interior(rhs, 4, 2), parity.interior(want, 4, 2), parity.SRMHD_CONS)
    print(parity.report("host API RKPlus::rhs vs oracle", rows), flush=True)
    # the bar of the net-flux parity tests: the oracle forms the smoothness
    # indicators in the reference's expanded form, which where beta^2 is near
    # eps moves the WENO weights by ~1e-7 relative
    for n_, dm, fm, rf, rg, pw in rows:
        assert rg <= 1e-9, n_


REGIME_EXE = os.path.join(ROOT, "method_amd", "lib", "regime_main")
# case -> (run fixture, source fixture, integrator, padded shape, dims, steps, tolerance); the current
# sheet's bar is the reference cons2prim's tolerance amplified by the stiff source at sigma 1e4
REGIME_CASES = {
    "currentsheet": ("currentsheet_128_regime_rksplit", "regime_src_currentsheet_128", (136, 1, 1), 1, 20, 5e-9),
    "briowu": ("briowu_128_regime_rksplit", "regime_src_briowu_128", (136, 1, 1), 1, 20, 1e-9),
    "kh2d": ("kh2d_32_weno5_regime_rk2b", "regime_src_kh2d_32_weno5", (40, 40, 1), 2, 3, 1e-9),
}


def _run_regime(tmp_path, case, steps, prims0=None):
    out = tmp_path / f"{case}_{steps}.bin"
    cmd = [REGIME_EXE, case, str(out), str(steps)]
    if prims0 is not None:
        pin = tmp_path / "prims0.bin"
        np.ascontiguousarray(prims0).tofile(pin)
        cmd.append(str(pin))
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # the driver's last line gives the final time (Simulation::updateTime reports t = ... too)
    return np.fromfile(out), float(r.stdout.rsplit("t = ", 1)[1].split()[0])


@pytest.mark.parametrize("case", sorted(REGIME_CASES))
def test_host_api_regime_run_matches_reference(golden, tmp_path, case):
    """REGIME handed to RKSplit (the reference's two REGIME example set-ups,
    CurrentSheetSingleFluid and BrioWuSingleFluid built by the host API) or
    to RK2B (2-D WENO5 KH), against the reference build's steps."""
    run, _, shape, dims, steps, tol = REGIME_CASES[case]
    z = golden(run)
    a, t = _run_regime(tmp_path, case, steps, z["prims0"] if case == "kh2d" else None)
    n = int(np.prod(shape))
    c, p = a[:9 * n].reshape((9,) + shape), a[9 * n:17 * n].reshape((8,) + shape)
    assert abs(t - float(np.sum(z["dts"][:steps]))) <= 1e-14
    for got, ref, names in ((c, z[f"cons{steps}"], parity.SRMHD_CONS), (p, z[f"prims{steps}"], parity.SRMHD_PRIMS)):
        got = parity.interior(got, 4, dims)
        if ref.shape != got.shape:
            ref = parity.interior(ref, 4, dims)
        rows = parity.field_errors(got, ref, names)
        print(parity.report(f"host API REGIME {case}, {steps} steps vs reference", rows), flush=True)
        for n_, dm, fm, rf, rg, pw in rows:
            assert rg <= tol, f"{n_}: {rg:.3e}"


@pytest.mark.parametrize("case", sorted(REGIME_CASES))
def test_host_api_regime_source_and_fields(golden, tmp_path, case):
    """REGIME::sourceExtension of the set-up state against the reference
    build; E from set_vars is -v x B of the host's primitives; q, alpha, K and
    dwdsb from set_vars / set_K / set_dwdsb equal the engine's fields for the
    same state."""
    from method_amd import engine as eng

    run, srcname, shape, dims, _, _ = REGIME_CASES[case]
    zs, zr = golden(srcname), golden(run)
    a, _ = _run_regime(tmp_path, case, 0, zr["prims0"] if case == "kh2d" else None)
    n = int(np.prod(shape))
    parts, off = {}, 0
    for name, k in (("src", 9), ("E", 3), ("q", 1), ("alpha", 1), ("K", 3), ("dwdsb", 36)):
        parts[name] = a[off:off + k * n].reshape((k,) + shape)
        off += k * n
    assert off == a.size
    rows = parity.field_errors(parity.interior(parts["src"], 4, dims), zs["src"], parity.SRMHD_CONS)
    print(parity.report(f"host API REGIME::sourceExtension {case} vs reference", rows), flush=True)
    for n_, dm, fm, rf, rg, pw in rows:
        assert rf <= 1e-9, f"{n_}: {rf:.3e}"
        if fm == 0.0:
            assert dm == 0.0, f"{n_} must stay exactly zero"
    P = zs["prims0"]
    v, B = parity.interior(P[1:4], 4, dims), parity.interior(P[5:8], 4, dims)
    assert np.allclose(parity.interior(parts["E"], 4, dims), -np.cross(v, B, axis=0), rtol=0,
                       atol=1e-14 * max(1.0, np.abs(B).max()))
    nx, ny, nz = (int(x) for x in zs["meta"][:3])
    b = zs["bounds"]
    gam, cfl, sig, cp = (float(x) for x in zs["params"])
    bc = tuple(eng.PERIODIC if int(x) else eng.OUTFLOW for x in zs["periodic"])
    cfg = eng.Config(nx=nx, ny=ny, nz=nz, xmin=b[0], xmax=b[1], ymin=b[2], ymax=b[3], zmin=b[4], zmax=b[5],
                     gamma=gam, sigma=sig, cp=cp, cfl=cfl, weno_order=int(zs["meta"][4]), bc=bc)
    with eng.Engine(cfg) as e:
        e.set_prims(P)
        e.set_model_extension(eng.EXT_REGIME)
        E, q, alpha, dw, K = e.regime_fields()
    assert np.array_equal(parts["E"], E)
    assert np.array_equal(parts["q"][0], q)
    assert np.array_equal(parts["alpha"][0], alpha)
    assert np.array_equal(parts["K"], K)
    assert np.array_equal(parts["dwdsb"], dw.reshape((36,) + shape))


RHS_MAIN = r"""
// Tests/CPU/Src/test_RKPlus.cc CalculatesRHSWell{NoREGIME,REGIME}, through the host API
#include <cmath>
#include <cstdio>
#include <vector>
#include "REGIME.h"
#include "RKPlus.h"
#include "boundaryConds.h"
#include "fluxVectorSplitting.h"
#include "initFunc.h"
#include "serialEnv.h"
#include "simulation.h"
#include "srmhd.h"
#include "weno.h"
int main(int argc, char **argv) {
  const bool with_regime = argc > 1 && argv[1][0] == '1';
  SerialEnv env(&argc, &argv, 1, 1, 1, 1);
  Data d(10, 10, 10, 0, 1, 0, 1, 0, 1, 0.004, &env, 0.5, 3);
  Weno3 weno(&d);
  SRMHD model(&d);
  FVS fluxMethod(&d, &weno, &model);
  REGIME modelExtension(&d, &fluxMethod);
  Flow bcs(&d);
  Simulation sim(&d, &env);
  OTVortexSingleFluid init(&d);
  RK2B timeInt(&d, &model, &bcs, &fluxMethod, with_regime ? &modelExtension : nullptr);
  sim.set(&init, &model, &timeInt, &bcs, &fluxMethod);
  const size_t n = (size_t)d.Nx * d.Ny * d.Nz * d.Ncons;
  std::vector<double> srcR(n, 0.0), srcM(n, 0.0), flux(n, 0.0), rhs(n, 0.0);
  if (with_regime) modelExtension.sourceExtension(d.cons, d.prims, d.aux, srcR.data());
  model.sourceTerm(d.cons, d.prims, d.aux, srcM.data());
  fluxMethod.F(d.cons, d.prims, d.aux, d.f, flux.data());
  timeInt.rhs(d.cons, d.prims, d.aux, rhs.data());
  double worst = 0.0, scale = 0.0, reg = 0.0;
  for (int var = 0; var < d.Ncons; var++)
    for (int i = d.is; i < d.ie; i++)
      for (int j = d.js; j < d.je; j++)
        for (int k = d.ks; k < d.ke; k++) {
          const size_t c = ((size_t)(var * d.Nx + i) * d.Ny + j) * d.Nz + k;
          const double want = srcM[c] + srcR[c] - flux[c];
          worst = std::fmax(worst, std::fabs(rhs[c] - want));
          scale = std::fmax(scale, std::fabs(want));
          reg = std::fmax(reg, std::fabs(srcR[c]));
        }
  printf("%.3e %.3e %.3e\n", worst, scale, reg);
  return 0;
}
"""


@pytest.mark.parametrize("with_regime", [0, 1])
def test_host_api_rkplus_rhs_reference_test(tmp_path, with_regime):
    """The reference's RKPlus rhs tests (Tests/CPU/Src/test_RKPlus.cc:50-145)
    on their own set-up (3-D Orszag-Tang 10^3, Ng 3, WENO3, Flow, RK2B, with
    and without REGIME), through the host API: rhs = model source (+ REGIME
    source) - F at the tests' 1e-15."""
    import shutil

    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    lib = os.path.join(ROOT, "method_amd", "lib")
    src, exe = tmp_path / "rhs.cc", tmp_path / "rhs"
    src.write_text(RHS_MAIN)
    subprocess.run(["g++", "-std=c++14", "-O2", "-I", os.path.join(ROOT, "method_amd", "host", "include"), "-I",
                    os.path.join(ROOT, "include"), str(src), "-o", str(exe), "-L", lib, "-lmethod_host",
                    "-lmethod_gpu", f"-Wl,-rpath,{lib}"], check=True, capture_output=True, text=True)
    print("run:", exe, with_regime, flush=True)
    r = subprocess.run([str(exe), str(with_regime)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    worst, scale, reg = (float(x) for x in r.stdout.split()[-3:])
    print(f"rhs vs source + extension - F: {worst:.3e} (scale {scale:.3e}, REGIME source {reg:.3e})", flush=True)
    assert scale > 0 and worst <= 1e-15
    assert (reg > 0) == bool(with_regime)
