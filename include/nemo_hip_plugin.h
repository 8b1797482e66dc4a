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
 * state-variables=, membrane-potential=, history=, [rcm] ..., [rng] normal=).
 * The backend loads it with dlopen and resolves the two symbols below.
 *
 * The plugin updates the neurons [start, end) of one type, given as
 * simulation-local indices; all arrays are device pointers and all work is
 * queued on `stream` (a hipStream_t). Contract, as in NeMo's plugins:
 *   - currents: currentE/currentI hold this cycle's synaptic input (already
 *     saturated and converted from Q11.20), currentExt the external input;
 *   - the plugin writes fired bits (one bit per simulation-local neuron,
 *     32 per word) for its range and shifts recentFiring (bit 0 = this cycle);
 *     a model that never fires (Kuramoto) leaves both alone;
 *   - forced firing (fstim bits) fires a neuron regardless of its state.
 * Parameters and state are structure-of-arrays: parameter k of type-local
 * neuron i is param[k*paramStride + i]. State variable k of cycle c is
 * state[(c mod history) * stateHistoryStride + k*stateVarStride + i], where
 * history is the descriptor's history= (1 unless given): the plugin reads
 * the plane of `cycle` and writes the plane of cycle+1, as NeMo's CPU
 * plugins do (src/nemo/cpu/Neurons.cpp:121-139); `cycle` is the
 * simulation's elapsed cycle count (nemo_elapsed_simulation), which
 * nemo_reset_timer sets back to 0. The RNG is 4 planes of rngStride words
 * (type-local index).
 */

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reverse connectivity of the simulation: for each simulation-local target
 * t, its incoming synapses are entries start[t] .. start[t+1]-1, in the
 * order of the network's synapse iterator (sources ascending by neuron
 * index, each source's synapses in the order they were added), like the
 * reverse matrix of src/nemo/construction/RCM.cpp:150-171. Only the fields
 * a model's descriptor asks for ([rcm] sources=, delays=, forward=,
 * weights=) are present; the others are NULL. All pointers are device
 * pointers; the struct itself is host memory, valid for the duration of the
 * call. NULL when no neuron type of the network asks for the matrix.
 * Replaces the rcm_dt* argument of the CUDA plugins
 * (src/nemo/cuda/plugins/neuron_model.h:44) and the void* rcm_ptr of the
 * CPU plugins (src/nemo/cpu/plugins/neuron_model.h:38). */
typedef struct nemo_hip_rcm {
	uint32_t targets;          /* simulation-local neurons (start has targets+1 entries) */
	uint32_t pad;
	size_t synapses;
	const uint32_t* start;
	const uint32_t* source;    /* simulation-local source neuron */
	const uint32_t* delay;     /* delay in cycles (1..) */
	const float* weight;       /* the weight as given to addSynapse */
	const uint32_t* forward;   /* index of the synapse's record in the forward matrix */
} nemo_hip_rcm;

/* Replaces cuda_update_neurons_t (src/nemo/cuda/plugins/neuron_model.h:26-44)
 * and cpu_update_neurons_t (src/nemo/cpu/plugins/neuron_model.h:25-38). */
typedef int hip_update_neurons_t(
		void* stream,
		unsigned cycle,
		unsigned start, unsigned end,
		float* param, size_t paramStride,
		float* state, size_t stateHistoryStride, size_t stateVarStride,
		uint32_t* rng, size_t rngStride,
		const uint32_t* fstim,
		const float* currentE,
		const float* currentI,
		const float* currentExt,
		uint64_t* recentFiring,
		uint32_t* fired,
		const nemo_hip_rcm* rcm);

/* Replaces cuda_init_neurons_t (src/nemo/cuda/plugins/neuron_model.h:52-62)
 * and cpu_init_neurons_t (src/nemo/cpu/plugins/neuron_model.h:47-51). Called
 * once at construction, after the RNG is seeded, with the initial state in
 * plane 0; a model with history fills the planes of the cycles before 0. */
typedef int hip_init_neurons_t(
		void* stream,
		unsigned start, unsigned end,
		float* param, size_t paramStride,
		float* state, size_t stateHistoryStride, size_t stateVarStride,
		uint32_t* rng, size_t rngStride);

/* Symbols a plugin exports */
int hip_update_neurons(void* stream, unsigned cycle, unsigned start, unsigned end,
		float* param, size_t paramStride, float* state, size_t stateHistoryStride, size_t stateVarStride,
		uint32_t* rng, size_t rngStride, const uint32_t* fstim,
		const float* currentE, const float* currentI, const float* currentExt,
		uint64_t* recentFiring, uint32_t* fired, const nemo_hip_rcm* rcm);

int hip_init_neurons(void* stream, unsigned start, unsigned end,
		float* param, size_t paramStride, float* state, size_t stateHistoryStride, size_t stateVarStride,
		uint32_t* rng, size_t rngStride);

#ifdef __cplusplus
}
#endif

#endif
