#ifndef NEMO_HIP_PLUGIN_H
#define NEMO_HIP_PLUGIN_H

/*
 * nemo_hip_plugin.h -- neuron-model plugin ABI of the HIP backend.
 *
 * The counterpart of NeMo's CUDA plugin ABI (cuda_update_neurons_t /
 * cuda_init_neurons_t, src/nemo/cuda/plugins/neuron_model.h:26-62) and CPU
 * plugin ABI (src/nemo/cpu/plugins/neuron_model.h:25-52). A model is a shared
 * library lib<Model>.so in <plugindir>/hip/, described by <plugindir>/<Model>.ini
 * (the same descriptor format as src/nemo/plugins/<Model>.ini: parameters=,
 * state-variables=, membrane-potential=, [rng] normal=). The backend loads it
 * with dlopen and resolves the two symbols below.
 *
 * The plugin updates the neurons [start, end) of one type, given as
 * simulation-local indices; all arrays are device pointers and all work is
 * queued on `stream` (a hipStream_t). Contract, as in NeMo's plugins:
 *   - currents: currentE/currentI hold this cycle's synaptic input (already
 *     saturated and converted from Q11.20), currentExt the external input;
 *   - the plugin writes fired bits (one bit per simulation-local neuron,
 *     32 per word) for its range and shifts recentFiring (bit 0 = this cycle);
 *   - forced firing (fstim bits) fires a neuron regardless of its state.
 * Parameters and state are structure-of-arrays: parameter k of type-local
 * neuron i is param[k*paramStride + i]; state likewise; the RNG is 4 planes
 * of rngStride words (type-local index).
 */

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces cuda_update_neurons_t (src/nemo/cuda/plugins/neuron_model.h:26-44). */
typedef int hip_update_neurons_t(
		void* stream,
		unsigned cycle,
		unsigned start, unsigned end,
		float* param, size_t paramStride,
		float* state, size_t stateStride,
		uint32_t* rng, size_t rngStride,
		const uint32_t* fstim,
		const float* currentE,
		const float* currentI,
		const float* currentExt,
		uint64_t* recentFiring,
		uint32_t* fired);

/* Replaces cuda_init_neurons_t (src/nemo/cuda/plugins/neuron_model.h:52-62). */
typedef int hip_init_neurons_t(
		void* stream,
		unsigned start, unsigned end,
		float* param, size_t paramStride,
		float* state, size_t stateStride,
		uint32_t* rng, size_t rngStride);

/* Symbols a plugin exports */
int hip_update_neurons(void* stream, unsigned cycle, unsigned start, unsigned end,
		float* param, size_t paramStride, float* state, size_t stateStride,
		uint32_t* rng, size_t rngStride, const uint32_t* fstim,
		const float* currentE, const float* currentI, const float* currentExt,
		uint64_t* recentFiring, uint32_t* fired);

int hip_init_neurons(void* stream, unsigned start, unsigned end,
		float* param, size_t paramStride, float* state, size_t stateStride,
		uint32_t* rng, size_t rngStride);

#ifdef __cplusplus
}
#endif

#endif
