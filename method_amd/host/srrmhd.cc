// Resistive SRMHD model of the host API (srrmhd.cc:24-581): the full-grid
// methods upload the given host arrays, run the engine's kernels and copy back.
#include "srrmhd.h"

#include <cstdio>
#include <stdexcept>

#include "gpu_link.h"

using method_host::check;
using method_host::engine;

SRRMHD::SRRMHD() : Model() {
  Ncons = 14;
  Nprims = 11;
  Naux = 17;
}

SRRMHD::SRRMHD(Data *data) : Model(data) {
  Ncons = data->Ncons = 14;
  Nprims = data->Nprims = 11;
  Naux = data->Naux = 17;
  data->gpu.model = MGPU_SRRMHD;
  const char *c[] = {"D", "Sx", "Sy", "Sz", "tau", "Bx", "By", "Bz", "Ex", "Ey", "Ez", "psi", "phi", "qch"};
  const char *p[] = {"rho", "vx", "vy", "vz", "p", "Bx", "By", "Bz", "Ex", "Ey", "Ez"};
  const char *a[] = {"h",   "W",       "e",  "c",     "Jx",    "Jy",    "Jz",     "Bsq",   "Esq",
                     "vsq", "rhohWsq", "vE", "Sbarx", "Sbary", "Sbarz", "Sbarsq", "tauBar"};
  for (auto s : c) data->consLabels.push_back(s);
  for (auto s : p) data->primsLabels.push_back(s);
  for (auto s : a) data->auxLabels.push_back(s);
}

void SRRMHD::fluxVector(double *cons, double *prims, double *aux, double *f, const int dir) {
  (void)aux;
  mgpu_engine *e = engine(data);
  check(mgpu_upload_state(e, cons, prims, nullptr), "SRRMHD::fluxVector");
  check(mgpu_flux_vector(e, dir, f), "SRRMHD::fluxVector");
  data->gpu.deviceCurrent = false;
}

// One cell through the engine's cell-list kernels (srrmhd.cc:194-207): the
// current comes from the cell's aux.
void SRRMHD::sourceTermSingleCell(double *cons, double *prims, double *aux, double *source, int, int, int) {
  check(mgpu_source_cells(engine(data), 1, cons, prims, aux, source), "SRRMHD::sourceTermSingleCell");
}

// The stiff source needs the current J, formed from cons, prims and sigma.
void SRRMHD::sourceTerm(double *cons, double *prims, double *aux, double *source) {
  (void)aux;
  mgpu_engine *e = engine(data);
  check(mgpu_upload_state(e, cons, prims, nullptr), "SRRMHD::sourceTerm");
  check(mgpu_source(e, source), "SRRMHD::sourceTerm");
  data->gpu.deviceCurrent = false;
}

// One cell through the engine's cell-list kernels (srrmhd.cc:256-312): the
// secant starts from the cell's aux rhohWsq; a cell that does not converge
// is reported and throws as the reference's secant does (srrmhd.cc:574-579).
void SRRMHD::getPrimitiveVarsSingleCell(double *cons, double *prims, double *aux, int i, int j, int k) {
  if (check(mgpu_c2p_cells(engine(data), 1, cons, prims, aux, nullptr), "SRRMHD::getPrimitiveVarsSingleCell") > 0) {
    printf("Original C2P could not converge in cell (%d, %d, %d)\n", i, j, k);
    throw std::runtime_error("C2P could not converge.\n");
  }
}

void SRRMHD::getPrimitiveVars(double *cons, double *prims, double *aux) {
  mgpu_engine *e = engine(data);
  check(mgpu_upload_state(e, cons, prims, aux), "SRRMHD::getPrimitiveVars");
  mgpu_stats st;
  const int rc = check(mgpu_c2p(e, &st), "SRRMHD::getPrimitiveVars");
  if (rc > 0) printf("C2P did not converge for %d cells\n", rc);
  check(mgpu_download_state(e, nullptr, prims, aux), "SRRMHD::getPrimitiveVars");
  data->gpu.deviceCurrent = false;
}

void SRRMHD::primsToAll(double *cons, double *prims, double *aux) {
  mgpu_engine *e = engine(data);
  check(mgpu_upload_state(e, nullptr, prims, nullptr), "SRRMHD::primsToAll");
  check(mgpu_prims_to_all(e), "SRRMHD::primsToAll");
  check(mgpu_download_state(e, cons, nullptr, aux), "SRRMHD::primsToAll");
  data->gpu.deviceCurrent = false;
}
