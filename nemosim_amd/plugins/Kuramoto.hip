/*
 * Kuramoto.hip -- HIP plugin for Kuramoto phase oscillators
 * (lib<Kuramoto>.so, ABI in include/nemo_hip_plugin.h).
 *
 * The model of src/nemo/cpu/plugins/Kuramoto.cpp:85-123: each cycle, the
 * phase theta of an oscillator with natural frequency f moves by one
 * fourth-order Runge-Kutta step of
 *     d theta / dt = f + sum_j w_j sin(theta_j - theta),
 * where theta_j is the phase of source j at cycle t - (delay_j - 1), held
 * fixed over the four stages, and the result is wrapped into [0, 2*pi).
 * The arithmetic is the reference CPU plugin's, operation for operation, in
 * IEEE single precision without contraction (sums in reverse-matrix order,
 * the host C library's sinf restated in csrc/device/libm_f32.h, fmodf exact),
 * so the phases agree bit for bit with the CPU backend.
 *
 * One lane per oscillator: the four stages of a target depend on each
 * other and each is an ordered sum over its incoming couplings, so the
 * parallelism is across oscillators. Phase history: plane c mod 32 of the
 * state holds the phase of cycle c (the descriptor's history=32).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../csrc/device/libm_f32.h"
#include "nemo_hip_plugin.h"

namespace {

constexpr unsigned HISTORY = 32u;   /* history= of Kuramoto.ini */

/* sum of w * sin(theta_source - x) over the couplings [r0, r1) of one
 * target, in order (src/nemo/cpu/plugins/Kuramoto.cpp:18-29) */
__device__ inline float coupling(const uint32_t* src, const uint32_t* dly, const float* w, uint32_t r0, uint32_t r1,
		const float* state, size_t hstride, unsigned cycle, unsigned start, float x)
{
	float sum = 0.0f;
	for(uint32_t r = r0; r < r1; ++r) {
		const unsigned plane = (cycle - (dly[r] - 1u)) % HISTORY;
		const float sp = state[size_t(plane) * hstride + (src[r] - start)];
		sum += w[r] * nemo_libm::sinf_glibc(sp - x);
	}
	return sum;
}

__global__ void kuramoto_update_kernel(unsigned start, unsigned end, unsigned cycle,
		const float* freq, float* state, size_t hstride, const uint32_t* rstart,
		const uint32_t* src, const uint32_t* dly, const float* w)
{
	const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= end - start) {
		return;
	}
	const float* phase0 = state + size_t(cycle % HISTORY) * hstride;
	float* phase1 = state + size_t((cycle + 1u) % HISTORY) * hstride;
	const float f = freq[i];
	const float theta = phase0[i];
	const uint32_t r0 = rstart[start + i], r1 = rstart[start + i + 1u];
	const float k0 = f + coupling(src, dly, w, r0, r1, state, hstride, cycle, start, theta);
	const float k1 = f + coupling(src, dly, w, r0, r1, state, hstride, cycle, start, theta + k0 * 0.5f);
	const float k2 = f + coupling(src, dly, w, r0, r1, state, hstride, cycle, start, theta + k1 * 0.5f);
	const float k3 = f + coupling(src, dly, w, r0, r1, state, hstride, cycle, start, theta + k2);
	const float next = theta + (k0 + 2.0f * k1 + 2.0f * k2 + k3) / 6.0f;
	phase1[i] = nemo_libm::kuramoto_wrap(next);
}

/* src/nemo/cpu/plugins/Kuramoto.cpp:133-152: the planes of cycles -1 .. -31
 * from the initial phase in plane 0, running the uncoupled model backwards */
__global__ void kuramoto_init_kernel(unsigned n, const float* freq, float* state, size_t hstride)
{
	const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n) {
		return;
	}
	const float f = freq[i];
	for(unsigned t = 0; t + 1u < HISTORY; ++t) {
		const float p0 = state[size_t((0u - t) % HISTORY) * hstride + i];
		state[size_t((0u - (t + 1u)) % HISTORY) * hstride + i] = nemo_libm::kuramoto_wrap(p0 - f);
	}
}

} // namespace

extern "C" __attribute__((visibility("default")))
int hip_update_neurons(void* stream, unsigned cycle, unsigned start, unsigned end,
		float* param, size_t paramStride, float* state, size_t stateHistoryStride, size_t stateVarStride,
		uint32_t* rng, size_t rngStride, const uint32_t* fstim,
		const float* currentE, const float* currentI, const float* currentExt,
		uint64_t* recentFiring, uint32_t* fired, const nemo_hip_rcm* rcm)
{
	/* no synaptic current, no firing, no RNG (src/nemo/cpu/plugins/Kuramoto.cpp:85-99) */
	(void) paramStride; (void) stateVarStride; (void) rng; (void) rngStride; (void) fstim;
	(void) currentE; (void) currentI; (void) currentExt; (void) recentFiring; (void) fired;
	if(end <= start) return 0;
	if(!rcm || !rcm->start || (rcm->synapses && (!rcm->source || !rcm->delay || !rcm->weight))) {
		return int(hipErrorInvalidValue);   /* the descriptor asks for sources, delays and weights */
	}
	const unsigned n = end - start;
	hipLaunchKernelGGL(kuramoto_update_kernel, dim3((n + 255u) / 256u), dim3(256), 0,
			static_cast<hipStream_t>(stream), start, end, cycle, param, state, stateHistoryStride,
			rcm->start, rcm->source, rcm->delay, rcm->weight);
	return int(hipGetLastError());
}

extern "C" __attribute__((visibility("default")))
int hip_init_neurons(void* stream, unsigned start, unsigned end,
		float* param, size_t paramStride, float* state, size_t stateHistoryStride, size_t stateVarStride,
		uint32_t* rng, size_t rngStride)
{
	(void) paramStride; (void) stateVarStride; (void) rng; (void) rngStride;
	if(end <= start) return 0;
	const unsigned n = end - start;
	hipLaunchKernelGGL(kuramoto_init_kernel, dim3((n + 255u) / 256u), dim3(256), 0,
			static_cast<hipStream_t>(stream), n, param, state, stateHistoryStride);
	return int(hipGetLastError());
}

[[maybe_unused]] static hip_update_neurons_t* const check_update = &hip_update_neurons;
[[maybe_unused]] static hip_init_neurons_t* const check_init = &hip_init_neurons;
