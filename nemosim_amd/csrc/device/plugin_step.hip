/*
 * plugin_step.hip -- the modular step: one cycle as a sequence of kernels
 * around the neuron-model plugin's hip_update_neurons (the structure of
 * NeMo's CUDA step, src/nemo/cuda/Simulation.cpp:336-430: gather, plugin
 * update, scatter, STDP). It runs any model a plugin provides and any network
 * size that fits in device memory; the fused persistent kernel
 * (step_kernel.hip) is the fast path for the built-in models.
 *
 * Pending current lives in a global ring of int64 accumulators
 * acc[2][D][n] (E, I); spikes are added with 64-bit global atomics at fire
 * time into slot (t + delay) mod D, so sums equal NeMo's int64 sums exactly
 * (src/nemo/cpu/Simulation.cpp:292-307) and are clamped at conversion
 * (src/nemo/fixedpoint.cpp:41-49).
 *
 * Forward records here carry the absolute target: {target | (delay-1) << 26,
 * Q11.20 weight}; a source's synapses are syn[off[s] .. off[s+1]).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "models.h"
#include "plugin_abi.h"

using namespace nemo_dev;

namespace {

__device__ inline float sat_to_float(long long v)
{
	if(v > 2147483647ll) v = 2147483647ll;
	if(v < -2147483648ll) v = -2147483648ll;
	return float(int32_t(v)) / 1048576.0f;
}

/* Start of cycle: convert and clear this cycle's ring slot, clear the
 * external current, forced-firing bits and the cycle's fired bits. */
__global__ void pm_begin_kernel(PluginStepDev s, uint32_t k)
{
	const uint64_t t = s.cycle0 + k;
	const uint32_t slot = uint32_t(t % s.D);
	long long* accE = s.acc + size_t(slot) * s.n;
	long long* accI = s.acc + size_t(s.D + slot) * s.n;
	uint32_t* fired = s.raster + size_t(k) * s.words32;
	const size_t stride = size_t(gridDim.x) * blockDim.x;
	for(size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < s.n; i += stride) {
		s.curE[i] = sat_to_float(accE[i]);
		s.curI[i] = sat_to_float(accI[i]);
		accE[i] = 0;
		accI[i] = 0;
		s.curX[i] = 0.0f;
	}
	for(size_t w = size_t(blockIdx.x) * blockDim.x + threadIdx.x; w < s.words32; w += stride) {
		s.fstim[w] = 0u;
		fired[w] = 0u;
	}
	if(blockIdx.x == 0 && threadIdx.x == 0) {
		*s.nfired = 0u;
	}
}

/* Stimulus of cycle k (deduplicated by the host) */
__global__ void pm_stimulus_kernel(PluginStepDev s, uint32_t k)
{
	const size_t stride = size_t(gridDim.x) * blockDim.x;
	const size_t tid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if(s.fs_ptr) {
		for(size_t e = s.fs_ptr[k] + tid; e < s.fs_ptr[k + 1]; e += stride) {
			uint32_t i = s.fs_idx[e];
			atomicOr(&s.fstim[i >> 5], 1u << (i & 31u));
		}
	}
	if(s.is_ptr) {
		for(size_t e = s.is_ptr[k] + tid; e < s.is_ptr[k + 1]; e += stride) {
			s.curX[s.is_idx[e]] = s.is_val[e];
		}
	}
}

/* Fired bits of cycle k -> list of fired neurons (any order: delivery is
 * order-independent; the ordered output is the bit raster). */
__global__ void pm_compact_kernel(PluginStepDev s, uint32_t k)
{
	const uint32_t* fired = s.raster + size_t(k) * s.words32;
	const size_t stride = size_t(gridDim.x) * blockDim.x;
	for(size_t w = size_t(blockIdx.x) * blockDim.x + threadIdx.x; w < s.words32; w += stride) {
		uint32_t b = fired[w];
		if(b) {
			uint32_t pos = atomicAdd(s.nfired, uint32_t(__popc(b)));
			while(b) {
				uint32_t bit = uint32_t(__ffs(b) - 1);
				b &= b - 1;
				s.list[pos++] = uint32_t(w * 32u + bit);
			}
		}
	}
}

/* One wave per fired source: its synapses into ring slots (t + d) mod D. */
__global__ void pm_scatter_kernel(PluginStepDev s, uint32_t k)
{
	const uint64_t t = s.cycle0 + k;
	const uint32_t base = uint32_t(t % s.D);
	const uint32_t lane = threadIdx.x & 63u;
	const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
	const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
	const uint32_t nf = *s.nfired;
	unsigned long long ev = 0;
	for(uint32_t j = wave; j < nf; j += nwaves) {
		const uint32_t src = s.list[j];
		const uint32_t a = s.off[src], b = s.off[src + 1];
		ev += b - a;
		for(uint32_t q = a + lane; q < b; q += 64u) {
			const uint32_t x = s.syn[2 * size_t(q)];
			const int32_t w = int32_t(s.syn[2 * size_t(q) + 1]);
			const uint32_t tgt = x & 0x3ffffffu;
			uint32_t sl = base + (x >> 26) + 1u;
			sl = (sl >= s.D) ? sl - s.D : sl;
			const uint32_t plane = (w >= 0) ? 0u : s.D;
			atomicAdd(reinterpret_cast<unsigned long long*>(&s.acc[size_t(plane + sl) * s.n + tgt]),
					(unsigned long long)(long long) w);
		}
	}
	if(lane == 0 && ev) {
		atomicAdd(s.events, ev);
	}
}

/* STDP: targets whose history has bit postLen set walk their plastic
 * in-synapses (src/nemo/ConnectivityMatrix.cpp:193-227). One wave per
 * qualifying target. */
__global__ void pm_stdp_kernel(PluginStepDev s)
{
	const StdpParams& sp = *reinterpret_cast<const StdpParams*>(s.stdp);
	const uint32_t lane = threadIdx.x & 63u;
	const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
	const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
	for(uint32_t tgt = wave; tgt < s.n; tgt += nwaves) {
		if(!((s.recent[tgt] >> sp.postLen) & 1ull)) continue;
		const uint32_t r0 = s.rcm_start[tgt], r1 = s.rcm_start[tgt + 1];
		for(uint32_t r = r0 + lane; r < r1; r += 64u) {
			const uint32_t sd = s.rcm_sd[r];
			const uint64_t pre = s.recent[sd & 0x3ffffffu] >> ((sd >> 26) + 1u);
			const int32_t wd = stdp_weight_change(sp, pre);
			if(wd != 0) {
				s.rcm_acc[r] = int32_t(uint32_t(s.rcm_acc[r]) + uint32_t(wd));
			}
		}
	}
}

/* After weights change (applyStdp): rebuild the ring from the firing
 * history, since NeMo reads a weight when the spike arrives
 * (src/nemo/cpu/Simulation.cpp:258-307). The ring has been zeroed; a spike
 * fired at cycle0-1-k with delay d > k is still pending and lands in slot
 * (cycle0-1-k+d) mod D. One wave per source with pending spikes. */
__global__ void pm_replay_kernel(PluginStepDev s)
{
	const uint32_t lane = threadIdx.x & 63u;
	const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
	const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
	const uint64_t pend = (s.D >= 64u) ? ~0ull : ((1ull << s.D) - 1ull);
	const uint32_t last = uint32_t((s.cycle0 + s.D - 1u) % s.D);
	for(uint32_t src = wave; src < s.n; src += nwaves) {
		const uint64_t h = s.recent[src] & pend;
		if(!h) continue;
		const uint32_t a = s.off[src], b = s.off[src + 1];
		for(uint32_t q = a + lane; q < b; q += 64u) {
			const uint32_t x = s.syn[2 * size_t(q)];
			const int32_t w = int32_t(s.syn[2 * size_t(q) + 1]);
			const uint32_t tgt = x & 0x3ffffffu;
			const uint32_t d = (x >> 26) + 1u;
			const uint32_t plane = (w >= 0) ? 0u : s.D;
			uint64_t hb = h & ((d >= 64u) ? ~0ull : ((1ull << d) - 1ull));
			while(hb) {
				const uint32_t kk = uint32_t(__builtin_ctzll(hb));
				hb &= hb - 1ull;
				const uint32_t sl = (last + d - kk) % s.D;
				atomicAdd(reinterpret_cast<unsigned long long*>(&s.acc[size_t(plane + sl) * s.n + tgt]),
						(unsigned long long)(long long) w);
			}
		}
	}
}

unsigned grid(size_t n, unsigned block, unsigned cap)
{
	size_t g = (n + block - 1) / block;
	if(g < 1) g = 1;
	if(g > cap) g = cap;
	return unsigned(g);
}

} // namespace

extern "C" {

int nemo_dev_pm_begin(const PluginStepDev* s, uint32_t k, void* stream)
{
	hipLaunchKernelGGL(pm_begin_kernel, dim3(grid(s->n, 256, 2048)), dim3(256), 0,
			static_cast<hipStream_t>(stream), *s, k);
	return int(hipGetLastError());
}

int nemo_dev_pm_stimulus(const PluginStepDev* s, uint32_t k, uint32_t entries, void* stream)
{
	if(entries == 0) return 0;
	hipLaunchKernelGGL(pm_stimulus_kernel, dim3(grid(entries, 256, 256)), dim3(256), 0,
			static_cast<hipStream_t>(stream), *s, k);
	return int(hipGetLastError());
}

int nemo_dev_pm_compact(const PluginStepDev* s, uint32_t k, void* stream)
{
	hipLaunchKernelGGL(pm_compact_kernel, dim3(grid(s->words32, 256, 1024)), dim3(256), 0,
			static_cast<hipStream_t>(stream), *s, k);
	return int(hipGetLastError());
}

int nemo_dev_pm_scatter(const PluginStepDev* s, uint32_t k, void* stream)
{
	hipLaunchKernelGGL(pm_scatter_kernel, dim3(2048), dim3(256), 0,
			static_cast<hipStream_t>(stream), *s, k);
	return int(hipGetLastError());
}

int nemo_dev_pm_replay(const PluginStepDev* s, void* stream)
{
	hipLaunchKernelGGL(pm_replay_kernel, dim3(2048), dim3(256), 0,
			static_cast<hipStream_t>(stream), *s);
	return int(hipGetLastError());
}

int nemo_dev_pm_stdp(const PluginStepDev* s, void* stream)
{
	hipLaunchKernelGGL(pm_stdp_kernel, dim3(grid(size_t(s->n) * 64, 256, 4096)), dim3(256), 0,
			static_cast<hipStream_t>(stream), *s);
	return int(hipGetLastError());
}

} // extern "C"
