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
 * One wave per oscillator: the four stages of a target depend on each other
 * and each is an ordered float sum over its incoming couplings, so the
 * lanes share the loads and the sines of a stage and the sum runs through
 * them in order (through LDS, as 64 dependent register adds). Phase history: plane c mod 32 of the state holds the phase
 * of cycle c (the descriptor's history=32).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../csrc/device/libm_f32.h"
#include "nemo_hip_plugin.h"

namespace {

constexpr unsigned HISTORY = 32u;   /* history= of Kuramoto.ini */

constexpr unsigned CHUNKS = 4u;     /* chunks of 64 couplings a wave holds in registers */

/* s + p(lane 0) + p(lane 1) + ... + p(lane cnt-1), each addition rounded to
 * float, in lane order: the reference's sequential sum over one chunk of 64
 * couplings. The products go through the wave's 64 floats of LDS (`buf`)
 * and every lane reads them all back (broadcast reads, four at a time), so
 * the chain is 64 dependent adds on registers. cnt (1..64) is wave-uniform;
 * every lane ends with the same s. A wave's LDS operations complete in
 * order, so no barrier is needed between the store and the loads. */
__device__ inline float ordered_sum(float s, float p, unsigned cnt, float* buf, unsigned lane)
{
	buf[lane] = p;
	__builtin_amdgcn_wave_barrier();
	const float4* b4 = reinterpret_cast<const float4*>(buf);
	if(cnt == 64u) {
#pragma unroll
		for(int q = 0; q < 16; ++q) {
			const float4 v = b4[q];
			s = s + v.x;
			s = s + v.y;
			s = s + v.z;
			s = s + v.w;
		}
	} else {
#pragma unroll
		for(int q = 0; q < 16; ++q) {
			if(unsigned(4 * q) < cnt) {
				const float4 v = b4[q];
				s = s + v.x;
				if(unsigned(4 * q + 1) < cnt) s = s + v.y;
				if(unsigned(4 * q + 2) < cnt) s = s + v.z;
				if(unsigned(4 * q + 3) < cnt) s = s + v.w;
			}
		}
	}
	__builtin_amdgcn_wave_barrier();   /* the reads are done before the next chunk's store */
	return s;
}

/* weight and source phase of the couplings rb + c*64 + lane, c < CHUNKS
 * (zero past rend); the phase of the source at cycle t - (delay - 1) */
__device__ inline void load_chunks(float (&sp)[CHUNKS], float (&wt)[CHUNKS], uint32_t rb, uint32_t rend,
		unsigned lane, const uint32_t* src, const uint32_t* dly, const float* w, const float* state,
		size_t hstride, unsigned cycle, unsigned start)
{
#pragma unroll
	for(unsigned c = 0; c < CHUNKS; ++c) {
		const uint32_t r = rb + c * 64u + lane;
		sp[c] = 0.0f;
		wt[c] = 0.0f;
		if(r < rend) {
			const unsigned plane = (cycle - (dly[r] - 1u)) % HISTORY;
			sp[c] = state[size_t(plane) * hstride + (src[r] - start)];
			wt[c] = w[r];
		}
	}
}

/* s plus the couplings held in sp/wt (cnt of them from the first chunk on,
 * at most CHUNKS*64), each w * sin(theta_source - x) rounded to float
 * (src/nemo/cpu/plugins/Kuramoto.cpp:18-29), summed in order */
__device__ inline float sum_chunks(float s, const float (&sp)[CHUNKS], const float (&wt)[CHUNKS], float x, uint32_t cnt,
		float* buf, unsigned lane)
{
#pragma unroll
	for(unsigned c = 0; c < CHUNKS; ++c) {
		if(c * 64u >= cnt) {
			break;
		}
		const float p = wt[c] * nemo_libm::sinf_glibc(sp[c] - x);
		s = ordered_sum(s, p, min(64u, cnt - c * 64u), buf, lane);
	}
	return s;
}

/* One wave per oscillator: its lanes load 64 couplings at a time (the row of
 * the reverse matrix is contiguous, so the loads are coalesced) and evaluate
 * their sines together; the sum then runs through the lanes in order. Rows
 * of up to CHUNKS*64 couplings stay in registers over the four stages;
 * longer rows are read again for each stage. */
__global__ void __launch_bounds__(256) kuramoto_update_kernel(unsigned start, unsigned end, unsigned cycle,
		const float* freq, float* state, size_t hstride, const uint32_t* rstart,
		const uint32_t* src, const uint32_t* dly, const float* w)
{
	__shared__ __align__(16) float sumBuf[4][64];   /* one row per wave of the 256-thread block */
	const unsigned lane = threadIdx.x & 63u;
	float* buf = sumBuf[threadIdx.x >> 6];
	const unsigned o = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);   /* wave-uniform */
	if(o >= end - start) {
		return;
	}
	const float* phase0 = state + size_t(cycle % HISTORY) * hstride;
	float* phase1 = state + size_t((cycle + 1u) % HISTORY) * hstride;
	const float f = freq[o];
	const float theta = phase0[o];
	const uint32_t r0 = rstart[start + o], r1 = rstart[start + o + 1u];
	float k0, k1, k2, k3;
	if(r1 - r0 <= CHUNKS * 64u) {
		float sp[CHUNKS], wt[CHUNKS];
		load_chunks(sp, wt, r0, r1, lane, src, dly, w, state, hstride, cycle, start);
		const uint32_t cnt = r1 - r0;
		k0 = f + sum_chunks(0.0f, sp, wt, theta, cnt, buf, lane);
		k1 = f + sum_chunks(0.0f, sp, wt, theta + k0 * 0.5f, cnt, buf, lane);
		k2 = f + sum_chunks(0.0f, sp, wt, theta + k1 * 0.5f, cnt, buf, lane);
		k3 = f + sum_chunks(0.0f, sp, wt, theta + k2, cnt, buf, lane);
	} else {
		auto stage = [&](float x) {
			float s = 0.0f;
			for(uint32_t rb = r0; rb < r1; rb += CHUNKS * 64u) {
				float sp[CHUNKS], wt[CHUNKS];
				load_chunks(sp, wt, rb, r1, lane, src, dly, w, state, hstride, cycle, start);
				s = sum_chunks(s, sp, wt, x, r1 - rb, buf, lane);
			}
			return s;
		};
		k0 = f + stage(theta);
		k1 = f + stage(theta + k0 * 0.5f);
		k2 = f + stage(theta + k1 * 0.5f);
		k3 = f + stage(theta + k2);
	}
	if(lane == 0) {
		const float next = theta + (k0 + 2.0f * k1 + 2.0f * k2 + k3) / 6.0f;
		phase1[o] = nemo_libm::kuramoto_wrap(next);
	}
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
	hipLaunchKernelGGL(kuramoto_update_kernel, dim3((n + 3u) / 4u), dim3(256), 0,
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
