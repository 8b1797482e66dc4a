/*
 * plugin_abi.h -- POD descriptor of one network for the modular (plugin)
 * step kernels in plugin_step.hip, shared by host C++ and HIP.
 */
#ifndef NEMOSIM_PLUGIN_ABI_H
#define NEMOSIM_PLUGIN_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct PluginStepDev {
	uint32_t n;            /* neurons */
	uint32_t D;            /* ring depth = maximum delay */
	uint32_t words32;      /* 32-bit raster words per cycle (even: whole 64-bit words) */
	uint32_t pad;
	uint64_t cycle0;       /* cycle at the start of the run */

	long long* acc;        /* [2][D][n] pending E (plane 0) and I (plane 1) current, Q11.20 */
	float* curE;           /* [n] this cycle's input, converted */
	float* curI;
	float* curX;           /* [n] external current */
	uint32_t* fstim;       /* [words32] forced-firing bits */
	uint32_t* raster;      /* [cycles][words32] fired bits; the plugin's `fired` row */
	uint32_t* list;        /* [n] fired neurons of the cycle */
	uint32_t* nfired;      /* their count */
	const uint32_t* off;   /* [n+1] synapses of source s: syn[off[s] .. off[s+1]) */
	const uint32_t* syn;   /* records {target | (delay-1) << 26, Q11.20 weight} */
	uint64_t* recent;      /* [n] firing history */
	unsigned long long* events;

	const uint32_t* rcm_start;
	const uint32_t* rcm_sd;
	int32_t* rcm_acc;
	const void* stdp;

	const uint32_t* fs_ptr;
	const uint32_t* fs_idx;
	const uint32_t* is_ptr;
	const uint32_t* is_idx;
	const float* is_val;
} PluginStepDev;

int nemo_dev_pm_begin(const PluginStepDev* s, uint32_t k, void* stream);
int nemo_dev_pm_stimulus(const PluginStepDev* s, uint32_t k, uint32_t entries, void* stream);
int nemo_dev_pm_compact(const PluginStepDev* s, uint32_t k, void* stream);
int nemo_dev_pm_scatter(const PluginStepDev* s, uint32_t k, void* stream);
int nemo_dev_pm_stdp(const PluginStepDev* s, void* stream);
int nemo_dev_pm_replay(const PluginStepDev* s, void* stream);

#ifdef __cplusplus
}
#endif

#endif
