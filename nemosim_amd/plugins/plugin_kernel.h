/*
 * plugin_kernel.h -- shared body of the HIP neuron-model plugins
 * (lib<Model>.so, ABI in include/nemo_hip_plugin.h). Each plugin source
 * defines NEMO_PLUGIN_MODEL to one of the nemo_dev model ids and includes
 * this file; the per-neuron arithmetic is the same device code the fused
 * step kernel runs (nemosim_amd/csrc/device/models.h).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../csrc/device/models.h"
#include "nemo_hip_plugin.h"

using namespace nemo_dev;

/* One thread per neuron of the type range: src/nemo/cuda/plugins/Izhikevich.cu:162-234
 * and src/nemo/cuda/plugins/IF_curr_exp.cu:57-273 are the CUDA counterparts;
 * the arithmetic follows the CPU plugins (see models.h). */
__global__ void plugin_update_kernel(unsigned start, unsigned end,
		float* param, size_t pstride, float* state, size_t sstride,
		uint32_t* rngp, size_t rstride, const uint32_t* fstim,
		const float* cE, const float* cI, const float* cX,
		uint64_t* recent, uint32_t* fired)
{
	const unsigned nl = blockIdx.x * blockDim.x + threadIdx.x;
	if(nl >= end - start) {
		return;
	}
	const unsigned ng = start + nl;
	float p[MAX_PARAM];
	float s[MAX_STATE];
	for(int k = 0; k < MAX_PARAM; ++k) p[k] = 0.0f;
	for(int k = 0; k < MAX_STATE; ++k) s[k] = 0.0f;
	const ModelInfoDev mi = model_info_dev(NEMO_PLUGIN_MODEL);
	for(int k = 0; k < mi.nparam; ++k) p[k] = param[size_t(k) * pstride + nl];
	for(int k = 0; k < mi.nstate; ++k) s[k] = state[size_t(k) * sstride + nl];
	Rng r;
	for(int k = 0; k < 4; ++k) r.s[k] = rngp[size_t(k) * rstride + nl];
	const bool fs = fstim ? (((fstim[ng >> 5] >> (ng & 31u)) & 1u) != 0u) : false;
	const float x = cX ? cX[ng] : 0.0f;
	const bool f = update_neuron(NEMO_PLUGIN_MODEL, p, s, r, cE[ng], cI[ng], x, fs);
	for(int k = 0; k < mi.nstate; ++k) state[size_t(k) * sstride + nl] = s[k];
	if(mi.normalRng) {
		for(int k = 0; k < 4; ++k) rngp[size_t(k) * rstride + nl] = r.s[k];
	}
	recent[ng] = (recent[ng] << 1) | (f ? 1ull : 0ull);
	if(f) {
		atomicOr(&fired[ng >> 5], 1u << (ng & 31u));
	}
}

extern "C" __attribute__((visibility("default")))
int hip_update_neurons(void* stream, unsigned cycle, unsigned start, unsigned end,
		float* param, size_t paramStride, float* state, size_t stateHistoryStride, size_t stateStride,
		uint32_t* rng, size_t rngStride, const uint32_t* fstim,
		const float* currentE, const float* currentI, const float* currentExt,
		uint64_t* recentFiring, uint32_t* fired, const nemo_hip_rcm* rcm)
{
	/* history length 1: the cycle selects no state plane; no reverse matrix */
	(void) cycle; (void) stateHistoryStride; (void) rcm;
	if(end <= start) return 0;
	const unsigned n = end - start;
	hipLaunchKernelGGL(plugin_update_kernel, dim3((n + 255) / 256), dim3(256), 0,
			static_cast<hipStream_t>(stream), start, end, param, paramStride, state, stateStride,
			rng, rngStride, fstim, currentE, currentI, currentExt, recentFiring, fired);
	return int(hipGetLastError());
}

/* src/nemo/cpu/plugins/default_init.c: nothing to initialise for models
 * with a history of one cycle */
extern "C" __attribute__((visibility("default")))
int hip_init_neurons(void* stream, unsigned start, unsigned end,
		float* param, size_t paramStride, float* state, size_t stateHistoryStride, size_t stateStride,
		uint32_t* rng, size_t rngStride)
{
	(void) stream; (void) start; (void) end; (void) param; (void) paramStride;
	(void) state; (void) stateHistoryStride; (void) stateStride; (void) rng; (void) rngStride;
	return 0;
}

static hip_update_neurons_t* const check_update = &hip_update_neurons;
static hip_init_neurons_t* const check_init = &hip_init_neurons;
