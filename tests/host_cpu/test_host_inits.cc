// CPU check of the host API's initial states against the reference's: builds
// the set-up of one of the reference's serial examples and writes the padded
// prims its InitialFunc produces. Usage: test_host_inits <case> <prims_out.bin>
// with case = briowu | briowu_srrmhd | ot | khrandom. Built and run by
// tests/test_host_cpu.py.
#include <cstdio>
#include <cstring>

#include "boundaryConds.h"
#include "initFunc.h"
#include "serialEnv.h"
#include "simData.h"
#include "simulation.h"
#include "srmhd.h"
#include "srrmhd.h"

static int dump(Data &d, const char *path) {
  FILE *f = std::fopen(path, "wb");
  if (!f) return 3;
  std::fwrite(d.prims, sizeof(double), (size_t)d.Nx * d.Ny * d.Nz * d.Nprims, f);
  std::fclose(f);
  return 0;
}

int main(int argc, char *argv[]) {
  if (argc < 3) return 2;
  const char *c = argv[1];
  SerialEnv env(0, nullptr, 1, 1, 1);
  if (!std::strcmp(c, "briowu") || !std::strcmp(c, "briowu_srrmhd")) {
    Data data(100, 0, 0, 0.0, 1.0, -1.0, 1.0, 0.0, 1.0, 0.4, &env, 0.4, 4, 2.0, 10);
    Model *model = nullptr;
    if (!std::strcmp(c, "briowu")) model = new SRMHD(&data);
    else model = new SRRMHD(&data);
    Outflow bcs(&data);
    Simulation sim(&data, &env);
    BrioWuSingleFluid init(&data);
    const int rc = dump(data, argv[2]);
    delete model;
    return rc;
  }
  if (!std::strcmp(c, "ot")) {
    Data data(64, 64, 0, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0, 1.0, &env, 0.6, 4, 5.0 / 3.0, 10, 1.0);
    SRMHD model(&data);
    Periodic bcs(&data);
    Simulation sim(&data, &env);
    OTVortexSingleFluid init(&data);
    return dump(data, argv[2]);
  }
  if (!std::strcmp(c, "khrandom")) {
    Data data(64, 64, 0, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0, 3.0, &env, 0.6, 4, 4.0 / 3.0, 10, 1.0);
    SRMHD model(&data);
    Periodic bcs(&data);
    Simulation sim(&data, &env);
    KHRandomInstabilitySingleFluid init(&data, 1, 10);
    return dump(data, argv[2]);
  }
  return 2;
}
