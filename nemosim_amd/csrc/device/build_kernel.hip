/*
 * Network construction on the device: the forward matrix bucketed by
 * (source, target partition), with each bucket's delay-1 records first, the
 * reverse matrix of plastic synapses and the record index of every synapse.
 * The host restatement of the same steps is Simulation::Simulation
 * (csrc/host/simulation.cpp); NeMo builds these on the host, synapse by
 * synapse (src/nemo/ConnectivityMatrix.cpp:98-161, src/nemo/construction/
 * RCM.cpp:111-170).
 *
 * Records land in their bucket through an atomic cursor, so their order
 * inside a bucket is not fixed. Nothing observable depends on it: currents
 * and STDP accumulators are integer sums, and every synapse's record index
 * is kept for the queries.
 */
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "build_abi.h"

namespace {

constexpr int BLOCK = 256;

inline unsigned grid_for(size_t n)
{
	size_t g = (n + BLOCK - 1) / BLOCK;
	return unsigned(g < 65536 ? (g ? g : 1) : 65536);
}

/* global -> local ids, Q11.20 weights (src/nemo/fixedpoint.cpp:18-29,
 * truncation), per-target bounds of the E and I sums (with a plastic
 * synapse's bound at its STDP limit), and flags: bit 0 = a synapse the host
 * must look at (invalid id, weight outside Q11.20 or not finite: the host's
 * own loop then raises exactly its error), bit 1 = local sources not in
 * ascending order */
__global__ void build_map_kernel(const uint32_t* S, const uint32_t* Tg, const float* W, const uint8_t* Pl,
		size_t nsyn, const uint32_t* dense, uint32_t minG, size_t ndense, int stdp, int32_t maxE, int32_t maxI,
		uint32_t* lsrc, uint32_t* ltgt, int32_t* wfx, unsigned long long* boundE, unsigned long long* boundI,
		uint32_t* flags)
{
	for(size_t i = size_t(blockIdx.x) * BLOCK + threadIdx.x; i < nsyn; i += size_t(gridDim.x) * BLOCK) {
		const uint32_t s = S[i], t = Tg[i];
		const uint32_t ls = (s >= minG && size_t(s - minG) < ndense) ? dense[s - minG] : 0u;
		const uint32_t lt = (t >= minG && size_t(t - minG) < ndense) ? dense[t - minG] : 0u;
		const float w = W[i];
		if(!ls || !lt || !isfinite(w) || fabsf(truncf(w)) >= 2048.0f) {
			atomicOr(flags, 1u);
			continue;
		}
		lsrc[i] = ls - 1u;
		ltgt[i] = lt - 1u;
		const int32_t x = int32_t(w * 1048576.0f);
		wfx[i] = x;
		int64_t wb = x;
		if(stdp && Pl[i]) {
			if(x > 0) wb = x > maxE ? x : maxE;
			else if(x < 0) wb = x < maxI ? x : maxI;
		}
		if(x >= 0) atomicAdd(boundE + (lt - 1u), (unsigned long long) wb);
		else atomicAdd(boundI + (lt - 1u), (unsigned long long) wb);
		if(i > 0) {
			const uint32_t p = S[i - 1];
			const uint32_t lp = (p >= minG && size_t(p - minG) < ndense) ? dense[p - minG] : 0u;
			if(lp > ls) atomicOr(flags, 2u);
		}
	}
}

/* bit 2 of flags: some target's E or I sum can leave the int32 range */
__global__ void build_saturation_kernel(const unsigned long long* boundE, const unsigned long long* boundI,
		uint32_t n, uint32_t* flags)
{
	for(uint32_t l = blockIdx.x * BLOCK + threadIdx.x; l < n; l += gridDim.x * BLOCK) {
		const int64_t e = int64_t(boundE[l]), i = int64_t(boundI[l]);
		if(e > int64_t(INT32_MAX) || i < int64_t(INT32_MIN)) atomicOr(flags, 4u);
	}
}

__global__ void build_count_kernel(const uint32_t* lsrc, const uint32_t* ltgt, const uint8_t* Dl, size_t nsyn,
		uint32_t T, uint32_t P1, uint32_t* count, uint32_t* nearCount)
{
	for(size_t i = size_t(blockIdx.x) * BLOCK + threadIdx.x; i < nsyn; i += size_t(gridDim.x) * BLOCK) {
		const size_t b = size_t(lsrc[i]) * P1 + ltgt[i] / T;
		atomicAdd(count + b, 1u);
		if(Dl[i] == 1) atomicAdd(nearCount + b, 1u);
	}
}

/* pass 0: the delay-1 records, pass 1: the others; record {target within
 * the partition | delay << 16, weight} */
__global__ void build_scatter_kernel(const uint32_t* lsrc, const uint32_t* ltgt, const uint8_t* Dl,
		const int32_t* wfx, size_t nsyn, uint32_t T, uint32_t P1, int pass, uint32_t* cursor, uint32_t* syn,
		uint32_t* fwd)
{
	for(size_t i = size_t(blockIdx.x) * BLOCK + threadIdx.x; i < nsyn; i += size_t(gridDim.x) * BLOCK) {
		const uint32_t d = Dl[i];
		if((d == 1u) != (pass == 0)) continue;
		const uint32_t lt = ltgt[i];
		const uint32_t p = lt / T;
		const uint32_t pos = atomicAdd(cursor + size_t(lsrc[i]) * P1 + p, 1u);
		syn[2 * size_t(pos)] = (lt - p * T) | (d << 16);
		syn[2 * size_t(pos) + 1] = uint32_t(wfx[i]);
		fwd[i] = pos;
	}
}

/* one atomic per block (a grid of at most 1024 blocks) */
__global__ void build_max_kernel(const uint32_t* v, size_t n, uint32_t* out)
{
	__shared__ uint32_t wmax[BLOCK / 64];
	uint32_t m = 0;
	for(size_t i = size_t(blockIdx.x) * BLOCK + threadIdx.x; i < n; i += size_t(gridDim.x) * BLOCK) {
		m = v[i] > m ? v[i] : m;
	}
	for(int d = 32; d >= 1; d >>= 1) {
		const uint32_t y = __shfl_xor(m, d, 64);
		m = y > m ? y : m;
	}
	if((threadIdx.x & 63u) == 0u) wmax[threadIdx.x >> 6] = m;
	__syncthreads();
	if(threadIdx.x == 0) {
		for(int w = 1; w < BLOCK / 64; ++w) m = wmax[w] > m ? wmax[w] : m;
		atomicMax(out, m);
	}
}

__global__ void build_near_pack_kernel(uint32_t* off, const uint32_t* nearCount, size_t nb, uint32_t shift)
{
	for(size_t k = size_t(blockIdx.x) * BLOCK + threadIdx.x; k < nb; k += size_t(gridDim.x) * BLOCK) {
		off[k] |= nearCount[k] << shift;
	}
}

/* out[s] = in[s * stride] for s <= n */
__global__ void build_gather_kernel(const uint32_t* in, uint32_t stride, uint32_t n, uint32_t* out)
{
	for(uint32_t s = blockIdx.x * BLOCK + threadIdx.x; s <= n; s += gridDim.x * BLOCK) {
		out[s] = in[size_t(s) * stride];
	}
}

__global__ void build_rcm_count_kernel(const uint32_t* ltgt, const uint8_t* Pl, size_t nsyn, uint32_t* count)
{
	for(size_t i = size_t(blockIdx.x) * BLOCK + threadIdx.x; i < nsyn; i += size_t(gridDim.x) * BLOCK) {
		if(Pl[i]) atomicAdd(count + ltgt[i], 1u);
	}
}

/* reverse record: source | (delay - 1) << 26, and the forward record index */
__global__ void build_rcm_scatter_kernel(const uint32_t* lsrc, const uint32_t* ltgt, const uint8_t* Dl,
		const uint8_t* Pl, const uint32_t* fwd, size_t nsyn, uint32_t* cursor, uint32_t* rcmSd, uint32_t* rcmFwd)
{
	for(size_t i = size_t(blockIdx.x) * BLOCK + threadIdx.x; i < nsyn; i += size_t(gridDim.x) * BLOCK) {
		if(!Pl[i]) continue;
		const uint32_t r = atomicAdd(cursor + ltgt[i], 1u);
		rcmSd[r] = lsrc[i] | (uint32_t(Dl[i] - 1u) << 26);
		rcmFwd[r] = fwd[i];
	}
}

} // namespace

extern "C" {

int nemo_dev_build_map(const uint32_t* S, const uint32_t* Tg, const float* W, const uint8_t* Pl, size_t nsyn,
		const uint32_t* dense, uint32_t minG, size_t ndense, int stdp, int32_t maxE, int32_t maxI,
		uint32_t* lsrc, uint32_t* ltgt, int32_t* wfx, int64_t* boundE, int64_t* boundI, uint32_t n,
		uint32_t* flags, void* stream)
{
	hipStream_t st = static_cast<hipStream_t>(stream);
	hipLaunchKernelGGL(build_map_kernel, dim3(grid_for(nsyn)), dim3(BLOCK), 0, st, S, Tg, W, Pl, nsyn, dense, minG,
			ndense, stdp, maxE, maxI, lsrc, ltgt, wfx, reinterpret_cast<unsigned long long*>(boundE),
			reinterpret_cast<unsigned long long*>(boundI), flags);
	hipLaunchKernelGGL(build_saturation_kernel, dim3(grid_for(n)), dim3(BLOCK), 0, st,
			reinterpret_cast<const unsigned long long*>(boundE), reinterpret_cast<const unsigned long long*>(boundI),
			n, flags);
	return int(hipGetLastError());
}

int nemo_dev_build_count(const uint32_t* lsrc, const uint32_t* ltgt, const uint8_t* Dl, size_t nsyn, uint32_t T,
		uint32_t P1, uint32_t* count, uint32_t* nearCount, void* stream)
{
	hipLaunchKernelGGL(build_count_kernel, dim3(grid_for(nsyn)), dim3(BLOCK), 0, static_cast<hipStream_t>(stream),
			lsrc, ltgt, Dl, nsyn, T, P1, count, nearCount);
	return int(hipGetLastError());
}

int nemo_dev_exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, void* temp, size_t* tempBytes,
		void* stream)
{
	return int(hipcub::DeviceScan::ExclusiveSum(temp, *tempBytes, in, out, n, static_cast<hipStream_t>(stream)));
}

int nemo_dev_build_scatter(const uint32_t* lsrc, const uint32_t* ltgt, const uint8_t* Dl, const int32_t* wfx,
		size_t nsyn, uint32_t T, uint32_t P1, int pass, uint32_t* cursor, uint32_t* syn, uint32_t* fwd, void* stream)
{
	hipLaunchKernelGGL(build_scatter_kernel, dim3(grid_for(nsyn)), dim3(BLOCK), 0, static_cast<hipStream_t>(stream),
			lsrc, ltgt, Dl, wfx, nsyn, T, P1, pass, cursor, syn, fwd);
	return int(hipGetLastError());
}

int nemo_dev_build_max(const uint32_t* v, size_t n, uint32_t* out, void* stream)
{
	const unsigned g = grid_for(n);
	hipLaunchKernelGGL(build_max_kernel, dim3(g < 1024u ? g : 1024u), dim3(BLOCK), 0, static_cast<hipStream_t>(stream),
			v, n, out);
	return int(hipGetLastError());
}

int nemo_dev_build_near_pack(uint32_t* off, const uint32_t* nearCount, size_t nb, uint32_t shift, void* stream)
{
	hipLaunchKernelGGL(build_near_pack_kernel, dim3(grid_for(nb)), dim3(BLOCK), 0, static_cast<hipStream_t>(stream),
			off, nearCount, nb, shift);
	return int(hipGetLastError());
}

int nemo_dev_build_gather(const uint32_t* in, uint32_t stride, uint32_t n, uint32_t* out, void* stream)
{
	hipLaunchKernelGGL(build_gather_kernel, dim3(grid_for(size_t(n) + 1)), dim3(BLOCK), 0,
			static_cast<hipStream_t>(stream), in, stride, n, out);
	return int(hipGetLastError());
}

int nemo_dev_build_rcm_count(const uint32_t* ltgt, const uint8_t* Pl, size_t nsyn, uint32_t* count, void* stream)
{
	hipLaunchKernelGGL(build_rcm_count_kernel, dim3(grid_for(nsyn)), dim3(BLOCK), 0, static_cast<hipStream_t>(stream),
			ltgt, Pl, nsyn, count);
	return int(hipGetLastError());
}

int nemo_dev_build_rcm_scatter(const uint32_t* lsrc, const uint32_t* ltgt, const uint8_t* Dl, const uint8_t* Pl,
		const uint32_t* fwd, size_t nsyn, uint32_t* cursor, uint32_t* rcmSd, uint32_t* rcmFwd, void* stream)
{
	hipLaunchKernelGGL(build_rcm_scatter_kernel, dim3(grid_for(nsyn)), dim3(BLOCK), 0,
			static_cast<hipStream_t>(stream), lsrc, ltgt, Dl, Pl, fwd, nsyn, cursor, rcmSd, rcmFwd);
	return int(hipGetLastError());
}

} // extern "C"
