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
