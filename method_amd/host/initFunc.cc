// Initial data of the host API (initFunc.cc:10-40 and the workload variants).
#include "initFunc.h"

#include <cmath>
#include <cstring>
#include <stdexcept>

InitialFunc::InitialFunc(Data *data) : data(data) {
  Data *d = data;
  if (d->memSet != 1)
    throw std::runtime_error("Must construct simulation class before implementing initial state. Need to allocate arrays.");
  const size_t n = (size_t)d->Nx * d->Ny * d->Nz;
  std::memset(d->cons, 0, sizeof(double) * n * d->Ncons);
  std::memset(d->f, 0, sizeof(double) * n * d->Ncons);
  std::memset(d->fnet, 0, sizeof(double) * n * d->Ncons);
  std::memset(d->source, 0, sizeof(double) * n * d->Ncons);
  std::memset(d->prims, 0, sizeof(double) * n * d->Nprims);
  std::memset(d->aux, 0, sizeof(double) * n * d->Naux);
  d->gpu.deviceCurrent = false;
}

FromArray::FromArray(Data *data, const double *prims) : InitialFunc(data) {
  std::memcpy(data->prims, prims, sizeof(double) * (size_t)data->Nx * data->Ny * data->Nz * data->Nprims);
}

// splitmix64 finaliser (Steele, Lea & Flood 2014).
static uint64_t splitmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double initUniform01(uint64_t seed, uint64_t index, int stream) {
  const uint64_t key = seed + 0x9E3779B97F4A7C15ULL * (2 * index + (uint64_t)stream + 1);
  return (double)(splitmix64(key) >> 11) * (1.0 / 9007199254740992.0);
}

KHShearLayers::KHShearLayers(Data *data, uint64_t seed, double shear, double bz, double width, double rhoIn,
                             double rhoOut, double amp, double pressure)
    : InitialFunc(data) {
  Data *d = data;
  if (d->Nprims != 8) throw std::invalid_argument("KHShearLayers is an ideal SRMHD (8 primitive) initial state");
  const double hs = 0.5 * (rhoIn + rhoOut), hd = 0.5 * (rhoIn - rhoOut);
  for (int i = 0; i < d->Nx; i++)
    for (int j = 0; j < d->Ny; j++)
      for (int k = 0; k < d->Nz; k++) {
        const double t = std::tanh((std::fabs(d->y[j]) - 0.5) / width);
        const uint64_t n = ((uint64_t)i * d->Ny + j) * d->Nz + k;
        d->prims[ID(0, i, j, k)] = hs - hd * t;
        d->prims[ID(1, i, j, k)] = shear * t;
        d->prims[ID(2, i, j, k)] = amp * (2.0 * initUniform01(seed, n, 0) - 1.0);
        d->prims[ID(3, i, j, k)] = d->dims == 3 ? amp * (2.0 * initUniform01(seed, n, 1) - 1.0) : 0.0;
        d->prims[ID(4, i, j, k)] = pressure;
        d->prims[ID(7, i, j, k)] = bz;
      }
}
