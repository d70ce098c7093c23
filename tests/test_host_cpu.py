"""The C++ host API's set-up logic on the CPU (no GPU needed): a test
program built with g++ against the host headers and libmethod_host.so, checking
the reference's own Simulation.dataInitialisation known answers."""
import os
import shutil
import subprocess

import numpy as np
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
                                     "BrioWuShockTube/SingleFluidResistiveSerial",
                                     "REGIME/CurrentSheetSerial",
                                     "REGIME/BrioWuSerial"])
def test_more_reference_examples_build(tmp_path, example):
    """More of the reference's serial example mains, unmodified, compile and
    link against the host API."""
    src = os.path.join("/root/reference/Examples", example, "main.cc")
    r = subprocess.run(["g++", "-std=c++14", "-O1", "-I", os.path.join(ROOT, "method_amd", "host", "include"), "-I",
                        os.path.join(ROOT, "include"), src, "-o", str(tmp_path / "m"), "-L", LIB, "-lmethod_host",
                        "-lmethod_gpu", f"-Wl,-rpath,{LIB}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


INIT_MAIN = r"""
#include <cstdio>
#include "boundaryConds.h"
#include "initFunc.h"
#include "serialEnv.h"
#include "simData.h"
#include "simulation.h"
#include "srmhd.h"
int main(int argc, char **argv) {
  SerialEnv env(&argc, &argv, 1, 1, 1);
  Data a(128, 0, 0, -3.0, 3.0, -1.0, 1.0, -1.5, 1.5, 7.0, &env, 0.2, 4, 2.0, 10000);
  SRMHD ma(&a);
  Outflow oa(&a);
  Simulation sa(&a, &env);
  CurrentSheetSingleFluid ia(&a);
  Data b(128, 0, 0, 0.0, 1.0, -1.0, 1.0, -1.5, 1.5, 0.4, &env, 0.15, 4, 2.0, 50);
  SRMHD mb(&b);
  Outflow ob(&b);
  Simulation sb(&b, &env);
  BrioWuSingleFluid ib(&b);
  FILE *f = fopen(argv[1], "wb");
  fwrite(a.prims, sizeof(double), (size_t)a.Nprims * a.Nx, f);
  fwrite(b.prims, sizeof(double), (size_t)b.Nprims * b.Nx, f);
  fclose(f);
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_regime_example_initial_data_matches_reference(golden, tmp_path):
    """CurrentSheetSingleFluid and BrioWuSingleFluid of the host API, on the
    reference's two REGIME example set-ups, equal the reference build's
    initial primitives bit for bit (no GPU call is made)."""
    src, exe, out = tmp_path / "m.cc", tmp_path / "m", tmp_path / "p.bin"
    src.write_text(INIT_MAIN)
    subprocess.run(["g++", "-std=c++14", "-O2", "-I", os.path.join(ROOT, "method_amd", "host", "include"), "-I",
                    os.path.join(ROOT, "include"), str(src), "-o", str(exe), "-L", LIB, "-lmethod_host",
                    "-lmethod_gpu", f"-Wl,-rpath,{LIB}"], check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(out).reshape(2, 8, 136, 1, 1)
    assert np.array_equal(got[0], golden("currentsheet_128_regime_rksplit")["prims0"])
    assert np.array_equal(got[1], golden("briowu_128_regime_rksplit")["prims0"])


WENO_MAIN = r"""
#include <cstdio>
#include "fluxVectorSplitting.h"
#include "serialEnv.h"
#include "srmhd.h"
int main(int argc, char **argv) {
  SerialEnv env(&argc, &argv, 1, 1, 1);
  Data d(16, 0, 0, 0.0, 1.0, -1.0, 1.0, -1.5, 1.5, 0.4, &env, 0.15, 4, 2.0, 50);
  SRMHD m(&d);
  Weno3 w3(&d);
  Weno5 w5(&d);
  FVS a(&d, &w3, &m);
  printf("%d ", d.gpu.wenoOrder);
  FVS b(&d, &w5, &m);
  printf("%d\n", d.gpu.wenoOrder);
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_fvs_takes_its_own_weno_order(tmp_path):
    """With several WENO objects alive, the engine reconstructs with the order
    of the one the FVS object was given (no GPU call is made)."""
    src, exe = tmp_path / "w.cc", tmp_path / "w"
    src.write_text(WENO_MAIN)
    subprocess.run(["g++", "-std=c++14", "-O2", "-I", os.path.join(ROOT, "method_amd", "host", "include"), "-I",
                    os.path.join(ROOT, "include"), str(src), "-o", str(exe), "-L", LIB, "-lmethod_host",
                    "-lmethod_gpu", f"-Wl,-rpath,{LIB}"], check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.split() == ["3", "5"]


BIND_MAIN = r"""
#include <cstdio>
#include "boundaryConds.h"
#include "fluxVectorSplitting.h"
#include "rkSplit.h"
#include "serialEnv.h"
#include "srmhd.h"
int main(int argc, char **argv) {
  SerialEnv env(&argc, &argv, 1, 1, 1);
  Data d(16, 0, 0, 0.0, 1.0, -1.0, 1.0, -1.5, 1.5, 0.4, &env, 0.15, 4, 2.0, 50);
  SRMHD m(&d);
  Weno3 w3(&d);
  Weno5 w5(&d);
  FVS f3(&d, &w3, &m);
  Outflow out(&d);
  Periodic per(&d);
  RKSplit a(&d, &m, &out, &f3);
  Flow flow(&d);
  a.bind();
  printf("%d %d %d %d %d\n", d.gpu.integrator, d.gpu.wenoOrder, d.gpu.bc[0], d.gpu.bc[1], d.gpu.bc[2]);
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_integrator_binds_its_own_objects(tmp_path):
    """With several boundary and WENO objects alive, an integrator's bind()
    (run by every step and Simulation::set) makes the engine use the ones it
    was handed (no GPU call is made)."""
    src, exe = tmp_path / "b.cc", tmp_path / "b"
    src.write_text(BIND_MAIN)
    subprocess.run(["g++", "-std=c++14", "-O2", "-I", os.path.join(ROOT, "method_amd", "host", "include"), "-I",
                    os.path.join(ROOT, "include"), str(src), "-o", str(exe), "-L", LIB, "-lmethod_host",
                    "-lmethod_gpu", f"-Wl,-rpath,{LIB}"], check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    from method_amd import engine as eng
    want = [eng.RKSPLIT, 3] + [eng.OUTFLOW] * 3
    assert [int(x) for x in r.stdout.split()] == want


EXT_MAIN = r"""
#include <cstdio>
#include <stdexcept>
#include "REGIME.h"
#include "RKPlus.h"
#include "boundaryConds.h"
#include "fluxVectorSplitting.h"
#include "rkSplit.h"
#include "serialEnv.h"
#include "simulation.h"
#include "srmhd.h"
class OtherExtension : public ModelExtension {
 public:
  OtherExtension(Data *d) : ModelExtension(d) { sourceExists = true; }
};
int main(int argc, char **argv) {
  SerialEnv env(&argc, &argv, 1, 1, 1);
  Data d(16, 0, 0, 0.0, 1.0, -1.0, 1.0, -1.5, 1.5, 0.4, &env, 0.15, 4, 2.0, 50);
  SRMHD m(&d);
  Weno3 w(&d);
  FVS f(&d, &w, &m);
  Outflow b(&d);
  Simulation sim(&d, &env);
  OtherExtension other(&d);
  int rejected = 0;
  try { RKSplit t(&d, &m, &b, &f, &other); } catch (const std::invalid_argument &) { ++rejected; }
  try { RK3 t(&d, &m, &b, &f, &other); } catch (const std::invalid_argument &) { ++rejected; }
  REGIME r(&d, &f);
  RKSplit ok(&d, &m, &b, &f, &r);
  int logic = 0;
  try { r.set_Dx(d.cons, d.prims, d.aux); } catch (const std::logic_error &) { ++logic; }
  try { r.set_dfydw(d.cons, d.prims, d.aux); } catch (const std::logic_error &) { ++logic; }
  printf("%d %d %d %d\n", rejected, logic, (int)r.sourceExists, d.sourceExtension != nullptr);
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_model_extension_errors(tmp_path):
    """A model extension other than REGIME is refused by RKSplit and the
    RKPlus schemes; REGIME's per-axis matrix setters throw std::logic_error;
    REGIME's constructor sets sourceExists and allocates Data::sourceExtension
    as the reference's does (no GPU call is made)."""
    src, exe = tmp_path / "x.cc", tmp_path / "x"
    src.write_text(EXT_MAIN)
    subprocess.run(["g++", "-std=c++14", "-O2", "-I", os.path.join(ROOT, "method_amd", "host", "include"), "-I",
                    os.path.join(ROOT, "include"), str(src), "-o", str(exe), "-L", LIB, "-lmethod_host",
                    "-lmethod_gpu", f"-Wl,-rpath,{LIB}"], check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.split() == ["2", "2", "1", "1"]
