// Resistive Orszag-Tang driver written against the host API as the
// reference's example mains are: Data, SRRMHD, WENO3, FVS, Periodic,
// Simulation, initial data, SSP2, set, then updateTime calls.
//
//   ot2d_ssp2_main <out.bin> <steps> <prims_in.bin> [n]
//
// n x n cells (default 32) on [0,1]^2, gamma 4/3, sigma 1e3, cp 0.1, cfl 0.5.
// Initial primitives (11, reference layout, float64) come from prims_in.bin;
// cons, prims and aux are written to out.bin after the last step.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "SSP2.h"
#include "boundaryConds.h"
#include "fluxVectorSplitting.h"
#include "initFunc.h"
#include "serialEnv.h"
#include "simData.h"
#include "simulation.h"
#include "srrmhd.h"
#include "weno.h"

int main(int argc, char *argv[]) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s out.bin steps prims_in.bin [n]\n", argv[0]);
    return 2;
  }
  const int steps = std::atoi(argv[2]);
  const int n = argc > 4 ? std::atoi(argv[4]) : 32;
  SerialEnv env(&argc, &argv, 1, 1, 1);
  Data data(n, n, 0, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0, 1e10, &env, 0.5, 4, 4.0 / 3.0, 1e3, 0.1);
  SRRMHD model(&data);
  Weno3 weno(&data);
  FVS fluxMethod(&data, &weno, &model);
  Periodic bcs(&data);
  Simulation sim(&data, &env);
  const size_t cells = (size_t)data.Nx * data.Ny * data.Nz;
  std::vector<double> p0(cells * data.Nprims);
  FILE *f = std::fopen(argv[3], "rb");
  if (!f || std::fread(p0.data(), sizeof(double), p0.size(), f) != p0.size()) {
    std::fprintf(stderr, "cannot read %s\n", argv[3]);
    return 3;
  }
  std::fclose(f);
  FromArray init(&data, p0.data());
  SSP2 timeInt(&data, &model, &bcs, &fluxMethod);
  sim.set(&init, &model, &timeInt, &bcs, &fluxMethod);
  for (int s = 0; s < steps; ++s) sim.updateTime();
  f = std::fopen(argv[1], "wb");
  if (!f) return 4;
  std::fwrite(data.cons, sizeof(double), cells * data.Ncons, f);
  std::fwrite(data.prims, sizeof(double), cells * data.Nprims, f);
  std::fwrite(data.aux, sizeof(double), cells * data.Naux, f);
  std::fclose(f);
  return 0;
}
