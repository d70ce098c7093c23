// CPU checks of the host API's initial data and text output, which touch no
// GPU. Usage: test_host_output <prims_out.bin> <cons0.bin> <prims0.bin> <aux0.bin>
//  1. KHInstabilitySingleFluid(mag = 1) on the reference KH example's 32 x 32
//     grid; its padded prims are written to prims_out.bin.
//  2. The given state is loaded into Data and SerialSaveData::saveAll writes
//     ./Data/Final/{Conserved,Primitive,Auxiliary,Domain,Constants}.
// Built and run by tests/test_host_cpu.py, which compares both with the
// reference build's own outputs.
#include <cstdio>
#include <cstdlib>

#include "boundaryConds.h"
#include "initFunc.h"
#include "serialEnv.h"
#include "serialSaveData.h"
#include "simData.h"
#include "simulation.h"
#include "srmhd.h"

static bool read_all(const char *path, double *dst, size_t n) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  const size_t got = std::fread(dst, sizeof(double), n, f);
  std::fclose(f);
  return got == n;
}

int main(int argc, char *argv[]) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s prims_out.bin cons0.bin prims0.bin aux0.bin\n", argv[0]);
    return 2;
  }
  SerialEnv env(0, nullptr, 1, 1, 1);
  Data data(32, 32, 0, -0.5, 0.5, -1.0, 1.0, 0.0, 1.0, 3.0, &env, 0.6, 4, 4.0 / 3.0, 10, 1.0, -100, 100, 10, 50);
  SRMHD model(&data);
  Periodic bcs(&data);
  Simulation sim(&data, &env);
  KHInstabilitySingleFluid init(&data, 1);
  const size_t n = (size_t)data.Nx * data.Ny * data.Nz;
  FILE *f = std::fopen(argv[1], "wb");
  if (!f) return 3;
  std::fwrite(data.prims, sizeof(double), n * data.Nprims, f);
  std::fclose(f);

  if (!read_all(argv[2], data.cons, n * data.Ncons) || !read_all(argv[3], data.prims, n * data.Nprims) ||
      !read_all(argv[4], data.aux, n * data.Naux))
    return 4;
  SerialSaveData save(&data, &env);
  save.saveAll();
  std::printf("OK\n");
  return 0;
}
