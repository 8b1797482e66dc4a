/* Launchers of the construction kernels (build_kernel.hip): device pointers
 * and sizes, the stream as void*, hipError_t returned as int. Internal to
 * the library. */
#ifndef NEMOSIM_DEVICE_BUILD_ABI_H
#define NEMOSIM_DEVICE_BUILD_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Bits of the flags word of nemo_dev_build_map */
#define NEMO_BUILD_HOST_CHECK 1u   /* a synapse needs the host's validation (invalid id or weight) */
#define NEMO_BUILD_UNORDERED  2u   /* local sources are not in ascending order */
#define NEMO_BUILD_SATURATE   4u   /* some target's E or I sum can leave the int32 range */

/* Synapse i -> local source and target, Q11.20 weight; per-target bounds
 * (n targets, zeroed by the caller) and the flags above (zeroed by the
 * caller). Plastic synapses count at their STDP limit (maxE / maxI) when
 * stdp is set. */
int nemo_dev_build_map(const uint32_t* S, const uint32_t* Tg, const float* W, const uint8_t* Pl, size_t nsyn,
		const uint32_t* dense, uint32_t minG, size_t ndense, int stdp, int32_t maxE, int32_t maxI,
		uint32_t* lsrc, uint32_t* ltgt, int32_t* wfx, int64_t* boundE, int64_t* boundI, uint32_t n,
		uint32_t* flags, void* stream);

/* Records per (source, target partition) bucket, count[s * P1 + p], and of
 * them the delay-1 records, nearCount[s * P1 + p] (both zeroed by the
 * caller) */
int nemo_dev_build_count(const uint32_t* lsrc, const uint32_t* ltgt, const uint8_t* Dl, size_t nsyn, uint32_t T,
		uint32_t P1, uint32_t* count, uint32_t* nearCount, void* stream);

/* out = exclusive prefix sum of in[0..n); with temp == NULL only sets
 * *tempBytes */
int nemo_dev_exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, void* temp, size_t* tempBytes,
		void* stream);

/* Records of one pass (0: delay 1, 1: the rest) into their buckets at
 * cursor[s * P1 + p]++; fwd[i] = record index of synapse i */
int nemo_dev_build_scatter(const uint32_t* lsrc, const uint32_t* ltgt, const uint8_t* Dl, const int32_t* wfx,
		size_t nsyn, uint32_t T, uint32_t P1, int pass, uint32_t* cursor, uint32_t* syn, uint32_t* fwd, void* stream);

/* *out = max(*out, max of v[0..n)) */
int nemo_dev_build_max(const uint32_t* v, size_t n, uint32_t* out, void* stream);

/* off[k] |= nearCount[k] << shift for k < nb */
int nemo_dev_build_near_pack(uint32_t* off, const uint32_t* nearCount, size_t nb, uint32_t shift, void* stream);

/* out[s] = in[s * stride] for s = 0..n */
int nemo_dev_build_gather(const uint32_t* in, uint32_t stride, uint32_t n, uint32_t* out, void* stream);

/* Plastic in-synapses per target, count[t] (zeroed by the caller) */
int nemo_dev_build_rcm_count(const uint32_t* ltgt, const uint8_t* Pl, size_t nsyn, uint32_t* count, void* stream);

/* Reverse records of plastic synapses at cursor[target]++ */
int nemo_dev_build_rcm_scatter(const uint32_t* lsrc, const uint32_t* ltgt, const uint8_t* Dl, const uint8_t* Pl,
		const uint32_t* fwd, size_t nsyn, uint32_t* cursor, uint32_t* rcmSd, uint32_t* rcmFwd, void* stream);

#ifdef __cplusplus
}
#endif

#endif
