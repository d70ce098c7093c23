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
