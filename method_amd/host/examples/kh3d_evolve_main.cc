// The benchmark workload driven through the host API as a METHOD program
// drives it: ideal SRMHD 3-D Kelvin-Helmholtz (KHShearLayers, seed 1234),
// WENO5, FVS, RK2, cfl 0.4, periodic x,z / outflow y.
//
//   kh3d_evolve_main [n] [steps]
//
// n^3 cells (default 256). After Simulation::set and 5 warm-up updateTime
// calls it times `steps` (default 50) more updateTime calls with the state kept
// on the device between them, as Simulation::evolve keeps it, and prints the
// zone-update rate; the final copy of the state back to the host is timed
// apart.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "boundaryConds.h"
#include "fluxVectorSplitting.h"
#include "initFunc.h"
#include "RK2.h"
#include "serialEnv.h"
#include "simData.h"
#include "simulation.h"
#include "srmhd.h"
#include "weno.h"

int main(int argc, char *argv[]) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 256;
  const int steps = argc > 2 ? std::atoi(argv[2]) : 50;
  SerialEnv env(&argc, &argv, 1, 1, 1);
  Data data(n, n, n, -0.5, 0.5, -1.0, 1.0, 0.0, 1.0, 1e10, &env, 0.4, 4, 4.0 / 3.0, 0.0, 0.1, -1.0e4, 1.0e4, 10,
            1000000000);
  SRMHD model(&data);
  Weno5 weno(&data);
  FVS fluxMethod(&data, &weno, &model);
  PerAxisBcs bcs(&data, 1, 0, 1);
  Simulation sim(&data, &env);
  KHShearLayers init(&data, 1234);
  RK2 timeInt(&data, &model, &bcs, &fluxMethod);
  sim.set(&init, &model, &timeInt, &bcs, &fluxMethod);

  data.gpu.mirrorEachStep = false;  // as inside Simulation::evolve
  for (int s = 0; s < 5; ++s) sim.updateTime();
  mgpu_sync(data.gpu.engine);
  const auto t0 = std::chrono::steady_clock::now();
  for (int s = 0; s < steps; ++s) sim.updateTime();
  mgpu_sync(data.gpu.engine);
  const auto t1 = std::chrono::steady_clock::now();
  sim.syncHost();
  const auto t2 = std::chrono::steady_clock::now();
  const double sec = std::chrono::duration<double>(t1 - t0).count();
  const double cells = (double)n * n * n;
  printf("kh3d_evolve_main %d^3: %d steps in %.4f s, %.4f ms/step, %.4e zone-updates/s; final copy to host %.3f s; "
         "t = %.10g, iterations %d\n",
         n, steps, sec, 1e3 * sec / steps, cells * steps / sec, std::chrono::duration<double>(t2 - t1).count(),
         data.t, data.iters);
  return 0;
}
