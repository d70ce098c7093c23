// REGIME, the resistive extension of ideal SRMHD (Project/CPU/Include/REGIME.h,
// Src/REGIME.cc; Wright & Hawke 2019). Handing it to RKSplit or an RKPlus
// integrator switches the engine's extension source on (MGPU_EXT_REGIME), so
// every step adds it on the GPU. sourceExtension() evaluates it for given
// host arrays; set_vars / set_dwdsb / set_K fill the public per-cell fields
// (E, q, alpha, dwdsb, K over the padded grid, the reference's layouts) from
// the engine. The per-axis matrices (df/dw, M, D) are formed inside the
// engine's kernels: their setters throw std::logic_error.
#ifndef REGIME_H
#define REGIME_H

#include "flux.h"
#include "modelExtension.h"

class REGIME : public ModelExtension {
 public:
  double *dfxdw, *dfydw, *dfzdw;  // [8][12][cell], not filled (formed in the kernels)
  double *dwdsb;                  // [12][3][cell]
  double *E;                      // [3][cell]
  double *q;                      // [cell]
  double *K;                      // [3][cell]
  double *Mx, *My, *Mz;           // [8][3][cell], not filled
  double *fbx, *fby, *fbz;        // [3][cell], not filled
  double *diffuX, *diffuY, *diffuZ;  // [8][cell], not filled
  double *alpha;                  // [cell]
  FluxMethod *fluxMethod = nullptr;

  REGIME();
  REGIME(Data *data, FluxMethod *fluxMethod);
  virtual ~REGIME();
  // REGIME::sourceExtension (REGIME.cc:83-138) of the given state: variables
  // 0..7 over the interior cells into source (Ncons x padded cells).
  void sourceExtension(double *cons, double *prims, double *aux, double *source);
  void set_vars(double *cons, double *prims, double *aux);
  void set_Dx(double *cons, double *prims, double *aux);
  void set_Dy(double *cons, double *prims, double *aux);
  void set_Dz(double *cons, double *prims, double *aux);
  void set_K(double *cons, double *prims, double *aux);
  void set_dfxdw(double *cons, double *prims, double *aux);
  void set_dfydw(double *cons, double *prims, double *aux);
  void set_dfzdw(double *cons, double *prims, double *aux);
  void set_dwdsb(double *cons, double *prims, double *aux);

 private:
  void fields(double *prims, double *E, double *q, double *alpha, double *dwdsb, double *K);
};

#endif
