// Ideal SRMHD model of the host API (srmhd.cc:25-673): the full-grid methods
// upload the given host arrays, run the engine's kernels and copy back.
#include "srmhd.h"

#include <stdexcept>

#include "gpu_link.h"

using method_host::check;
using method_host::engine;

SRMHD::SRMHD() : Model(), smartGuesses(0) {
  Ncons = 9;
  Nprims = 8;
  Naux = 13;
}

SRMHD::SRMHD(Data *data) : Model(data), smartGuesses(0) {
  Ncons = data->Ncons = 9;
  Nprims = data->Nprims = 8;
  Naux = data->Naux = 13;
  data->gpu.model = MGPU_SRMHD;
  const char *c[] = {"D", "Sx", "Sy", "Sz", "tau", "Bx", "By", "Bz", "phi"};
  const char *p[] = {"rho", "vx", "vy", "vz", "p", "Bx", "By", "Bz"};
  const char *a[] = {"h", "W", "e", "c", "b0", "bx", "by", "bz", "bsq", "vsq", "BS", "Bsq", "Ssq"};
  for (auto s : c) data->consLabels.push_back(s);
  for (auto s : p) data->primsLabels.push_back(s);
  for (auto s : a) data->auxLabels.push_back(s);
}

SRMHD::~SRMHD() {}

void SRMHD::fluxVector(double *cons, double *prims, double *aux, double *f, const int dir) {
  (void)aux;
  mgpu_engine *e = engine(data);
  check(mgpu_upload_state(e, cons, prims, nullptr), "SRMHD::fluxVector");
  check(mgpu_flux_vector(e, dir, f), "SRMHD::fluxVector");
  data->gpu.deviceCurrent = false;
}

// One cell through the engine's cell-list kernels (srmhd.cc:212-223).
void SRMHD::sourceTermSingleCell(double *cons, double *prims, double *aux, double *source, int, int, int) {
  check(mgpu_source_cells(engine(data), 1, cons, prims, aux, source), "SRMHD::sourceTermSingleCell");
}

void SRMHD::sourceTerm(double *cons, double *prims, double *aux, double *source) {
  (void)prims;
  (void)aux;
  mgpu_engine *e = engine(data);
  check(mgpu_upload_state(e, cons, nullptr, nullptr), "SRMHD::sourceTerm");
  check(mgpu_source(e, source), "SRMHD::sourceTerm");
  data->gpu.deviceCurrent = false;
}

// One cell through the engine's cell-list kernels (srmhd.cc:297-388): the
// start comes from the cell's prims and aux h, and a cell that does not
// converge throws as the reference's does.
void SRMHD::getPrimitiveVarsSingleCell(double *cons, double *prims, double *aux, int, int, int) {
  if (check(mgpu_c2p_cells(engine(data), 1, cons, prims, aux, nullptr), "SRMHD::getPrimitiveVarsSingleCell") > 0)
    throw std::runtime_error("C2P could not converge.\n");
}

// Failures are reported on stdout as the reference's retry path does
// (srmhd.cc:495-498); the state is still written.
void SRMHD::getPrimitiveVars(double *cons, double *prims, double *aux) {
  mgpu_engine *e = engine(data);
  check(mgpu_upload_state(e, cons, prims, aux), "SRMHD::getPrimitiveVars");
  mgpu_stats st;
  int rc = check(mgpu_c2p(e, &st), "SRMHD::getPrimitiveVars");
  if (rc > 0) printf("Smart guessing did not work for %d cells\n", rc);
  check(mgpu_download_state(e, nullptr, prims, aux), "SRMHD::getPrimitiveVars");
  data->gpu.deviceCurrent = false;
}

void SRMHD::primsToAll(double *cons, double *prims, double *aux) {
  mgpu_engine *e = engine(data);
  check(mgpu_upload_state(e, nullptr, prims, nullptr), "SRMHD::primsToAll");
  check(mgpu_prims_to_all(e), "SRMHD::primsToAll");
  check(mgpu_download_state(e, cons, nullptr, aux), "SRMHD::primsToAll");
  data->gpu.deviceCurrent = false;
}
