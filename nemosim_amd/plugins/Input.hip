/* HIP neuron-model plugin libInput.so (ABI: include/nemo_hip_plugin.h). */
#include "../csrc/device/models.h"
#define NEMO_PLUGIN_MODEL nemo_dev::MODEL_INPUT
#include "plugin_kernel.h"
