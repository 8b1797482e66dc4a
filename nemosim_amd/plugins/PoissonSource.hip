/* HIP neuron-model plugin libPoissonSource.so (ABI: include/nemo_hip_plugin.h). */
#include "../csrc/device/models.h"
#define NEMO_PLUGIN_MODEL nemo_dev::MODEL_POISSON_SOURCE
#include "plugin_kernel.h"
