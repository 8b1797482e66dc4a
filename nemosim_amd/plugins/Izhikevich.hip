/* HIP neuron-model plugin libIzhikevich.so (ABI: include/nemo_hip_plugin.h). */
#include "../csrc/device/models.h"
#define NEMO_PLUGIN_MODEL nemo_dev::MODEL_IZHIKEVICH
#include "plugin_kernel.h"
