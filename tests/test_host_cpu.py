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


FILES = (("cons", "Conserved/cons.dat"), ("prims", "Primitive/prims.dat"), ("aux", "Auxiliary/aux.dat"),
         ("domain", "Domain/domain.dat"), ("constants", "Constants/constants.dat"))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_kh_initial_state_and_text_output_match_reference(tmp_path, golden):
    """KHInstabilitySingleFluid equals the reference's initial state bit for
    bit, and SerialSaveData::saveAll of the reference's post-set state writes
    byte-identical files to the reference's own saveAll (the reference KH
    example set-up, 32 x 32)."""
    import numpy as np

    z = golden("kh_example_32_rksplit_save")
    if not os.path.exists(os.path.join(LIB, "libmethod_host.so")):
        pytest.fail("libmethod_host.so not built: run `python -m method_amd.build`")
    src = os.path.join(ROOT, "tests", "host_cpu", "test_host_output.cc")
    exe = tmp_path / "test_host_output"
    subprocess.run(["g++", "-std=c++14", "-O1", "-I", os.path.join(ROOT, "method_amd", "host", "include"), "-I",
                    os.path.join(ROOT, "include"), src, "-o", str(exe), "-L", LIB, "-lmethod_host", "-lmethod_gpu",
                    f"-Wl,-rpath,{LIB}"], check=True, capture_output=True, text=True)
    for key in ("cons0", "prims0", "aux0"):
        np.ascontiguousarray(z[key]).tofile(tmp_path / f"{key}.bin")
    r = subprocess.run([str(exe), str(tmp_path / "kh_prims.bin"), str(tmp_path / "cons0.bin"),
                        str(tmp_path / "prims0.bin"), str(tmp_path / "aux0.bin")], capture_output=True, text=True,
                       cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    prims = np.fromfile(tmp_path / "kh_prims.bin").reshape(z["prims0"].shape)
    # the reference's dump follows Simulation::set, whose Periodic::apply has
    # replaced the ghost cells by periodic images; the formula covers the interior
    assert np.array_equal(prims[:, 4:-4, 4:-4], z["prims0"][:, 4:-4, 4:-4])
    for key, rel in FILES:
        got = (tmp_path / "Data" / "Final" / rel).read_bytes()
        assert got == bytes(z["text0_" + key]), f"{rel} differs from the reference's output"


REF_EXAMPLE = "/root/reference/Examples/KelvinHelmholtz/SingleFluidIdealSerial/main.cc"


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists(REF_EXAMPLE),
                    reason="needs g++ and the reference tree")
def test_reference_example_main_builds_against_host_api(tmp_path):
    """The reference's own serial KH example main.cc, unmodified and read in
    place, compiles and links against the host headers and libraries (it is
    not run here: that needs a GPU)."""
    exe = tmp_path / "kh_example_ref_main"
    r = subprocess.run(["g++", "-std=c++14", "-O1", "-I", os.path.join(ROOT, "method_amd", "host", "include"), "-I",
                        os.path.join(ROOT, "include"), REF_EXAMPLE, "-o", str(exe), "-L", LIB, "-lmethod_host",
                        "-lmethod_gpu", f"-Wl,-rpath,{LIB}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    assert exe.exists()


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("case, fixture", [("briowu", "briowu_100_rksplit"), ("briowu_srrmhd", "briowu_100_ssp2"),
                                           ("ot", "ot_64_rksplit"), ("khrandom", "khrandom_64_rksplit")])
def test_initial_states_match_reference(tmp_path, golden, case, fixture):
    """BrioWuSingleFluid, OTVortexSingleFluid and KHRandomInstabilitySingleFluid
    reproduce the reference's initial states bit for bit (interior cells; the
    reference's state follows Simulation::set, whose Bcs::apply rewrites the
    ghost cells)."""
    import numpy as np

    z = golden(fixture)
    exe = tmp_path / "test_host_inits"
    src = os.path.join(ROOT, "tests", "host_cpu", "test_host_inits.cc")
    subprocess.run(["g++", "-std=c++14", "-O1", "-I", os.path.join(ROOT, "method_amd", "host", "include"), "-I",
                    os.path.join(ROOT, "include"), src, "-o", str(exe), "-L", LIB, "-lmethod_host", "-lmethod_gpu",
                    f"-Wl,-rpath,{LIB}"], check=True, capture_output=True, text=True)
    out = tmp_path / "prims.bin"
    r = subprocess.run([str(exe), case, str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    ref = z["prims0"]
    got = np.fromfile(out).reshape(ref.shape)
    sl = (slice(None), slice(4, -4)) + ((slice(None),) if ref.shape[2] == 1 else (slice(4, -4),)) + (slice(None),)
    assert np.array_equal(got[sl], ref[sl])


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists("/root/reference/Examples"),
                    reason="needs g++ and the reference tree")
@pytest.mark.parametrize("example", ["KelvinHelmholtz/SingleFluidIdealRandomSerial",
                                     "OrszagTangVortex/SingleFluidIdealSerial",
                                     "BrioWuShockTube/SingleFluidIdealSerial",
                                     "BrioWuShockTube/SingleFluidResistiveSerial"])
def test_more_reference_examples_build(tmp_path, example):
    """More of the reference's serial example mains, unmodified, compile and
    link against the host API."""
    src = os.path.join("/root/reference/Examples", example, "main.cc")
    r = subprocess.run(["g++", "-std=c++14", "-O1", "-I", os.path.join(ROOT, "method_amd", "host", "include"), "-I",
                        os.path.join(ROOT, "include"), src, "-o", str(tmp_path / "m"), "-L", LIB, "-lmethod_host",
                        "-lmethod_gpu", f"-Wl,-rpath,{LIB}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
