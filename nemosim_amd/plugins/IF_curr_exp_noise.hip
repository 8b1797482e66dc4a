/* HIP neuron-model plugin libIF_curr_exp_noise.so (ABI: include/nemo_hip_plugin.h). */
#include "../csrc/device/models.h"
#define NEMO_PLUGIN_MODEL nemo_dev::MODEL_IF_CURR_EXP_NOISE
#include "plugin_kernel.h"
