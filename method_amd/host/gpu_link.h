// Glue between the host API objects and the engine's C ABI.
#ifndef METHOD_HOST_GPU_LINK_H
#define METHOD_HOST_GPU_LINK_H

#include "method_gpu.h"
#include "simData.h"

namespace method_host {

// Creates (or re-creates after a configuration change) the engine for `d`.
mgpu_engine *engine(Data *d);
// Gives the engine the model extension d's integrator was handed, once.
mgpu_engine *with_extension(Data *d, mgpu_engine *e);
// Throws std::runtime_error with the engine's message when rc < 0.
int check(int rc, const char *what);
// Destroys the engine owned by `d`.
void release(Data *d);

}  // namespace method_host

#endif
