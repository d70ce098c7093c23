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

KHInstabilitySingleFluid::KHInstabilitySingleFluid(Data *data, int mag) : InitialFunc(data) {
  Data *d = data;
  if (d->Nprims > 15)
    throw std::invalid_argument(
        "Trying to implement a single fluid initial state on incorrect model.\n\tModel has wrong number of primitive "
        "variables to be single fluid model.");
  if (d->gamma != 4.0 / 3.0) throw std::invalid_argument("Expected the index gamma = 4/3\n");
  if (d->xmin != -0.5 || d->xmax != 0.5) throw std::invalid_argument("Domain has incorrect values. Expected x E [-0.5, 0.5]\n");
  if (d->ymin != -1.0 || d->ymax != 1.0) throw std::invalid_argument("Domain has incorrect values. Expected y E [-1.0, 1.0]\n");
  const double pi = 3.14159265358979323846;
  const double sig = 0.1, vShear = 0.5, A0 = 0.1, a = 0.01, rho0 = 0.55, rho1 = 0.45, B0 = 0.1;
  for (int i = 0; i < d->Nx; i++)
    for (int j = 0; j < d->Ny; j++)
      for (int k = 0; k < d->Nz; k++) {
        double *P = d->prims;
        P[ID(4, i, j, k)] = 1.0;
        if (mag) P[ID(7, i, j, k)] = B0;
        // the layer at y = +0.5 for y > 0, its mirror image below
        const double s = d->y[j] > 0 ? 1.0 : -1.0;
        const double yl = d->y[j] - s * 0.5;
        const double t = std::tanh(yl / a);
        P[ID(0, i, j, k)] = d->y[j] > 0 ? rho0 + rho1 * t : rho0 - rho1 * t;
        P[ID(1, i, j, k)] = d->y[j] > 0 ? vShear * t : -vShear * t;
        const double bump = A0 * vShear * std::sin(2 * pi * d->x[i]) * (std::exp(-std::pow(yl, 2) / (sig * sig)));
        P[ID(2, i, j, k)] = d->y[j] > 0 ? bump : -bump;
        if (d->Ncons > 9) {
          P[ID(8, i, j, k)] = -(P[ID(2, i, j, k)] * P[ID(7, i, j, k)] - P[ID(3, i, j, k)] * P[ID(6, i, j, k)]);
          P[ID(9, i, j, k)] = -(P[ID(3, i, j, k)] * P[ID(5, i, j, k)] - P[ID(1, i, j, k)] * P[ID(7, i, j, k)]);
          P[ID(10, i, j, k)] = -(P[ID(1, i, j, k)] * P[ID(6, i, j, k)] - P[ID(2, i, j, k)] * P[ID(5, i, j, k)]);
        }
      }
}
