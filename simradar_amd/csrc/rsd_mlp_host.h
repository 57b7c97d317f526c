// rsd_mlp_host.h -- internal link between the surrogate network (rsd_mlp.hip)
// and the engine (rsd_engine.hip): the device view of a network's parameters.
#ifndef RSD_MLP_HOST_H
#define RSD_MLP_HOST_H

#include "rsd.h"
#include "rsd_mlp.h"

extern "C" int rsd_mlp_device_view(const rsd_mlp *m, MlpDev *out, int *device);

#endif
