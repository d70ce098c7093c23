// Boundary conditions of the host API (boundaryConds.cc:5-730).
#include "boundaryConds.h"

#include "gpu_link.h"

void Bcs::setAxes(int px, int py, int pz) {
  axisRule[0] = px ? MGPU_BC_PERIODIC : MGPU_BC_OUTFLOW;
  axisRule[1] = py ? MGPU_BC_PERIODIC : MGPU_BC_OUTFLOW;
  axisRule[2] = pz ? MGPU_BC_PERIODIC : MGPU_BC_OUTFLOW;
  bind();
}

void Bcs::bind() {
  for (int a = 0; a < 3; ++a) data->gpu.bc[a] = axisRule[a];
}

// Ghost cells of the given arrays take their boundary images: the arrays go
// up to the engine and come back with every ghost cell filled.
void Bcs::apply(double *cons, double *prims, double *aux) {
  bind();
  mgpu_engine *e = method_host::engine(data);
  method_host::check(mgpu_upload_state(e, cons, prims, aux), "Bcs::apply");
  method_host::check(mgpu_download_state(e, cons, prims, aux), "Bcs::apply");
  data->gpu.deviceCurrent = (cons == data->cons && prims == data->prims && aux == data->aux);
}
