"""The C++ host API's set-up logic on the CPU (no GPU needed): a test
program built with g++ against the host headers and libmethod_host.so, checking
the reference's own Simulation.dataInitialisation known answers."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "method_amd", "lib")
SRC = os.path.join(ROOT, "tests", "host_cpu", "test_host_data.cc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_data_initialisation(tmp_path):
    if not os.path.exists(os.path.join(LIB, "libmethod_host.so")):
        pytest.fail("libmethod_host.so not built: run `python -m method_amd.build`")
    exe = tmp_path / "test_host_data"
    cmd = ["g++", "-std=c++14", "-O1", "-I", os.path.join(ROOT, "method_amd", "host", "include"), "-I",
           os.path.join(ROOT, "include"), SRC, "-o", str(exe), "-L", LIB, "-lmethod_host", "-lmethod_gpu",
           f"-Wl,-rpath,{LIB}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK" in r.stdout
