// CPU checks of the host API's set-up logic, which touches no GPU: the
// reference's Simulation.dataInitialisation known answers
// (Tests/CPU/Src/test_simulation.cc:16-77) and the dt rule of updateTime
// (simulation.cc:113-124). Built and run by tests/test_host_cpu.py.
#include <cmath>
#include <cstdio>
#include <stdexcept>

#include "boundaryConds.h"
#include "serialEnv.h"
#include "simData.h"
#include "simulation.h"
#include "srmhd.h"

static int failures = 0;
#define CHECK(cond)                                                     \
  do {                                                                  \
    if (!(cond)) {                                                      \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);     \
      ++failures;                                                       \
    }                                                                   \
  } while (0)

template <class F>
static bool throws_runtime_error(F f) {
  try {
    f();
  } catch (const std::runtime_error &) {
    return true;
  }
  return false;
}

int main() {
  {
    SerialEnv env(0, nullptr, 1, 1, 1);
    Data d(100, 10, 2, 0, 1, -0.5, 0.5, -0.1, 0.1, 0.8, &env);
    Periodic bcs(&d);
    CHECK(throws_runtime_error([&] { Simulation sim(&d, &env); }));  // no model
  }
  {
    SerialEnv env(0, nullptr, 1, 1, 1);
    Data d(100, 10, 2, 0, 1, -0.5, 0.5, -0.1, 0.1, 0.8, &env);
    SRMHD model(&d);
    CHECK(throws_runtime_error([&] { Simulation sim(&d, &env); }));  // no Bcs
  }
  SerialEnv env(0, nullptr, 1, 1, 1);
  Data data(100, 10, 2, 0, 1, -0.5, 0.5, -0.1, 0.1, 0.8, &env);
  SRMHD model(&data);
  Periodic bcs(&data);
  Simulation sim(&data, &env);
  Data *d = sim.data;
  CHECK(d->nx == 100 && d->ny == 10 && d->nz == 2 && d->Ng == 4);
  CHECK(d->Nx == 108 && d->Ny == 18 && d->Nz == 10);
  CHECK(d->xmin == 0.0 && d->xmax == 1.0 && d->ymin == -0.5 && d->ymax == 0.5);
  CHECK(d->zmin == -0.1 && d->zmax == 0.1 && d->endTime == 0.8);
  CHECK(d->cfl == 0.5 && d->gamma == 5.0 / 3 && d->sigma == 1000.0);
  CHECK(d->memSet == 1 && d->Ncons == 9 && d->Nprims == 8 && d->Naux == 13);
  CHECK(d->alphaX == 1.0 && d->alphaY == 1.0 && d->alphaZ == 1.0);
  CHECK(d->t == 0.0 && d->iters == 0);
  CHECK(std::fabs(d->dt - 0.002886751346) < 1e-11);
  CHECK(d->dx == 0.01 && d->dy == 0.1 && d->dz == 0.1);
  CHECK(d->x[4] == 0.005 && d->x[1] == -0.025 && d->x[99 + 4] == 0.995);
  CHECK(d->y[4] == -0.45 && d->y[1] == -0.75);
  CHECK(d->id(3, 5, 6, 7) == ((3L * d->Nx + 5) * d->Ny + 6) * d->Nz + 7);

  // 2-D: Nz collapses to 1 and dz is infinite, so dt comes from x and y
  SerialEnv env2(0, nullptr, 1, 1, 1);
  Data data2(64, 64, 0, -0.5, 0.5, -1.0, 1.0, 0.0, 0.0, 1.0, &env2, 0.4, 4, 4.0 / 3.0);
  SRMHD model2(&data2);
  Periodic bcs2(&data2);
  Simulation sim2(&data2, &env2);
  CHECK(data2.Nz == 1 && data2.dims == 2);
  CHECK(std::fabs(data2.dt - 0.4 * (1.0 / 64) / std::sqrt(2.0)) < 1e-17);

  std::printf("%s (%d failures)\n", failures ? "FAIL" : "OK", failures);
  return failures ? 1 : 0;
}
