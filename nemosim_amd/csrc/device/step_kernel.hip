/*
 * step_kernel.hip -- the simulation step on gfx950.
 *
 * One persistent kernel runs many 1 ms cycles of one or more independent
 * networks. Each workgroup owns one partition of T target neurons of one
 * network, and every workgroup of a launch is resident at once (the host
 * sizes launches by the device's measured co-residency; the workgroups of
 * a network exchange fired lists every cycle):
 *
 *   - the partition's neuron parameters, state, RNG and 64-cycle firing
 *     history stay in registers for the whole launch (one neuron per thread);
 *   - incoming synaptic current for the next D cycles is accumulated in an
 *     LDS ring [D][T] of Q11.20 int32 (separate excitatory and inhibitory
 *     planes, as NeMo's CPU backend sums them, src/nemo/cpu/Simulation.cpp:
 *     292-307); LDS integer atomics make the sum order-independent, so the
 *     result equals the CPU's int64 sum whenever no partial sum leaves the
 *     int32 range -- networks where one might are run with returning atomics
 *     and sticky overflow bits (src/nemo/cuda/fixedpoint.cu:69-81 does the
 *     same), which saturates exactly like wfx_toFloat.
 *
 * Cycle t of partition p (the CPU's fire(), src/nemo/cpu/Simulation.cpp:94-111):
 *   1. current_t = ring[t mod D]; neuron update (Izhikevich / IF_curr_exp,
 *      per-neuron xorshift + Box-Muller); firing history shift;
 *   2. fired-list compaction: wavefront ballot + prefix over waves, giving
 *      the partition's fired neurons in ascending order, and the 64-bit
 *      ballot words as the step's firing raster;
 *   3. publish the partition's fired count tagged with the cycle, then wait
 *      for all partitions' counts of the cycle (the only grid-level sync;
 *      no counter atomics, no L2 write-back or invalidate);
 *   4. delivery: every workgroup walks the fired lists of all partitions and,
 *      for each fired source, reads only its own bucket of the source's
 *      synapses (all delays at once) and adds each weight into ring slot
 *      (t + delay) mod D -- the delay-bucketed spike queue lives in LDS;
 *   5. STDP: targets that fired postLen cycles ago walk their plastic
 *      in-synapses and accumulate weight changes
 *      (src/nemo/ConnectivityMatrix.cpp:193-227).
 *
 * Blocks are assigned to partitions by the host so that each XCD (blocks
 * b, b+8, ...) gets a contiguous range of partitions: the synapse buckets of
 * one source are then contiguous per XCD and each XCD's L2 fetches only its
 * slice of every row.
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <stdint.h>

#include "models.h"
#include "step_abi.h"

using namespace nemo_dev;

static_assert(sizeof(StdpParams) <= NEMO_STDP_LDS_BYTES, "STDP tables must fit their LDS area");

namespace {

__device__ inline uint32_t lds_prefix_waves(uint32_t* wsum, uint32_t wave, uint32_t nwaves, uint32_t* total)
{
	uint32_t before = 0, all = 0;
	for(uint32_t w = 0; w < nwaves; ++w) {
		uint32_t c = wsum[w];
		before += (w < wave) ? c : 0u;
		all += c;
	}
	*total = all;
	return before;
}

/* Cross-workgroup data (fired lists, their tagged counts, the STDP firing
 * history) is exchanged with agent-scope relaxed atomic loads and stores.
 * On gfx950 these are sc1 accesses, coherent across the XCDs' L2s, so no L2
 * write-back or invalidate is needed. A writer drains its stores
 * (s_waitcnt; __syncthreads alone waits only for LDS traffic) before
 * publishing the count that names them. */
#if defined(__HIP_DEVICE_COMPILE__)
#define NEMO_GLOBAL __attribute__((address_space(1)))
#define NEMO_LDS __attribute__((address_space(3)))
#else
#define NEMO_GLOBAL   /* host pass: the kernels are only parsed there */
#define NEMO_LDS
#endif

/* The network's arrays as global-address-space pointers. Pointers read from
 * the network descriptor are generic, and generic (flat) accesses also count
 * against the LDS wait counter, so every LDS wait would also wait for all
 * outstanding memory loads; global accesses keep the two independent. */
template<typename V>
__device__ inline NEMO_GLOBAL V* glb(V* p)
{
	return (NEMO_GLOBAL V*) p;
}

template<typename V>
__device__ inline void st_agent(NEMO_GLOBAL V* p, V v)
{
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template<typename V>
__device__ inline V ld_agent(const NEMO_GLOBAL V* p)
{
	return __hip_atomic_load(const_cast<NEMO_GLOBAL V*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* device-coherent add (two's-complement wrap, as NeMo's int32 accumulator) */
__device__ inline void add_agent(NEMO_GLOBAL int32_t* p, int32_t v)
{
	(void) __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* Wave 0 waits until every partition of the network has published its
 * fired count for cycle t (count tagged with the low 32 bits of t; the
 * buffer of the same parity last held cycle t-2), then writes the exclusive
 * prefix of the counts to pref[0..P]. This is the network's only grid-level
 * synchronisation per cycle. Each lane polls up to GATHER_CH partitions
 * (lane, lane+64, ...) with all their loads in flight together, and polls
 * again only those not yet published. The wait is bounded: after spinLimit
 * polls the status word is flagged, and a flagged status makes every later
 * wait of the launch fall through, so all waves run to completion. */
constexpr uint32_t GATHER_CH = 4u;

/* Flags a timed-out wait in the status word; the first one also records
 * where it happened (control block words 4..7, see step_abi.h). */
__device__ inline void report_timeout(uint32_t* status, uint32_t kind, uint32_t cycle, uint32_t part,
		uint32_t awaited)
{
	const uint32_t old = atomicOr(status, kind);
	if(old == 0u) {
		status[3] = cycle;
		status[4] = part;
		status[5] = awaited;
		status[6] = ld_agent(glb(status - 1));
	}
}

__device__ inline void gather_counts(const NEMO_GLOBAL uint64_t* fc, uint32_t P, uint32_t tag, uint32_t* pref,
		uint32_t lane, uint32_t spinLimit, uint32_t* status, uint32_t part)
{
	uint32_t carry = 0;
	for(uint32_t c0 = 0; c0 < P; c0 += 64u * GATHER_CH) {
		uint32_t cnt[GATHER_CH];
		uint32_t pending = 0;
#pragma unroll
		for(uint32_t i = 0; i < GATHER_CH; ++i) {
			cnt[i] = 0u;
			if(c0 + 64u * i + lane < P) pending |= 1u << i;
		}
		uint32_t spins = 0;
		while(pending) {
			uint64_t v[GATHER_CH];
#pragma unroll
			for(uint32_t i = 0; i < GATHER_CH; ++i) {
				v[i] = ((pending >> i) & 1u) ? ld_agent(fc + c0 + 64u * i + lane) : 0ull;
			}
#pragma unroll
			for(uint32_t i = 0; i < GATHER_CH; ++i) {
				if(((pending >> i) & 1u) && uint32_t(v[i] >> 32) == tag) {
					cnt[i] = uint32_t(v[i]);
					pending &= ~(1u << i);
				}
			}
			if(!pending) break;
			/* the status word is read every 16th poll only: each read is
			 * another round trip in the wait */
			if((++spins & 15u) == 0u) {
				if(ld_agent(glb(status)) != 0u) break;
				if(spins > spinLimit) {
					report_timeout(status, NEMO_STATUS_BARRIER_TIMEOUT, tag, part,
							c0 + 64u * uint32_t(__builtin_ctz(pending)) + lane);
					break;
				}
			}
			__builtin_amdgcn_s_sleep(1);
		}
#pragma unroll
		for(uint32_t i = 0; i < GATHER_CH; ++i) {
			uint32_t x = cnt[i];
#pragma unroll
			for(uint32_t d = 1; d < 64u; d <<= 1) {
				uint32_t y = __shfl_up(x, d, 64);
				if(lane >= d) x += y;
			}
			const uint32_t q = c0 + 64u * i + lane;
			if(q < P) pref[q + 1] = carry + x;
			carry += __shfl(x, 63, 64);
		}
	}
	if(lane == 0) pref[0] = 0u;
}

/* A fired-list entry of cycle tag: (tag << 32) | neuron, given its first
 * read v. The count that names it may become visible first, so a stale entry
 * (of an earlier cycle) is read again, boundedly, as gather_counts waits. */
__device__ inline uint32_t fired_entry(const NEMO_GLOBAL uint64_t* e, uint64_t v, uint32_t tag,
		uint32_t spinLimit, uint32_t* status, uint32_t part, uint32_t src_part)
{
	uint32_t spins = 0;
	while(uint32_t(v >> 32) != tag) {
		if((++spins & 15u) == 0u) {
			if(ld_agent(glb(status)) != 0u) {
				return 0u;
			}
			if(spins > spinLimit) {
				report_timeout(status, NEMO_STATUS_ENTRY_TIMEOUT, tag, part, src_part);
				return 0u;
			}
		}
		__builtin_amdgcn_s_sleep(1);
		v = ld_agent(e);
	}
	return uint32_t(v);
}

/* Loads of a synapse record and of a bucket bound in the delivery loops */
__device__ inline uint2 ld_rec(const NEMO_GLOBAL uint2* syn, uint32_t q)
{
	return syn[q];
}

__device__ inline uint32_t ld_off(const NEMO_GLOBAL uint32_t* o)
{
	return *o;
}

/* With deferred delivery the bucket start off[s][p] carries, in its top
 * 4 bits, the number of delay-1 records at the head of the bucket. */
constexpr uint32_t OFF_NEAR_SHIFT = 28u;

/* Delivery rounds whose record loads are issued together (fewer with the
 * saturating accumulation, whose returning atomics need more registers),
 * and records per lane per round (lane i of a group reads records i and i+G) */
/* rounds of record loads in flight per batch: fewer where the variant has
 * fewer registers left (saturating accumulation; any model, 10 parameters
 * and 4 state variables per neuron) */
template<bool SAT, bool IZO> struct DeliverRounds { static constexpr uint32_t value = SAT ? 4u : (IZO ? 6u : 5u); };
constexpr uint32_t DELIVER_KPL = 2u;

/* Partition whose fired list holds flat index j: the q with
 * pref[q] <= j < pref[q+1] (pref is the exclusive prefix of counts) */
__device__ inline uint32_t partition_of(const uint32_t* pref, uint32_t P, uint32_t j)
{
	uint32_t lo = 0, hi = P;
	while(hi - lo > 1u) {
		const uint32_t mid = (lo + hi) >> 1;
		if(pref[mid] <= j) lo = mid; else hi = mid;
	}
	return lo;
}

/* Add a Q11.20 weight into ring slot sl of local target tl: the excitatory
 * plane for w >= 0, the inhibitory plane otherwise (as the CPU's split,
 * src/nemo/cpu/Simulation.cpp:299). In saturating mode the returning atomic
 * detects the first partial sum to leave int32 and sets a sticky bit. */
template<bool SAT>
__device__ inline void ring_add(int32_t* ringE, int32_t* ringI, uint32_t* ovf, uint32_t D, uint32_t T,
		uint32_t fwords, uint32_t sl, uint32_t tl, int32_t w)
{
	int32_t* acc = (w >= 0) ? ringE : ringI;
	if(SAT) {
		int32_t old = __hip_atomic_fetch_add(&acc[sl * T + tl], w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
		int64_t sum = int64_t(old) + int64_t(w);
		if(sum > int64_t(INT32_MAX) || sum < int64_t(INT32_MIN)) {
			uint32_t plane = (w >= 0) ? 0u : D;
			atomicOr(&ovf[(plane + sl) * fwords + (tl >> 5)], 1u << (tl & 31u));
		}
	} else {
		__hip_atomic_fetch_add(&acc[sl * T + tl], w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
	}
}

/* One pass of record delivery for up to 128 sources held one per lane
 * (a0/b0 for sources 0..63, a1/b1 for 64..127 of the batch): groups of
 * G = 2^glog lanes share one source, lane i of a group reads records i and
 * i+G of its bucket, so a load instruction touches a few cache lines
 * instead of 64; DELIVER_RU rounds of record loads are issued before their
 * LDS atomics. Rounds [rbeg, rend) of the pass_rounds(nsrc) are done here,
 * so a pass can be split; records from 2G on are left to the caller. */
__device__ inline uint32_t pass_rounds(uint32_t nsrc, uint32_t glog)
{
	return (nsrc + (64u >> glog) - 1u) >> (6u - glog);
}

template<bool SAT, bool IZO>
__device__ inline void deliver_pass(uint32_t a0, uint32_t b0, uint32_t a1, uint32_t b1, uint32_t nsrc,
		uint32_t glog, uint32_t lane, uint32_t base, const NEMO_GLOBAL uint2* syn,
		int32_t* ringE, int32_t* ringI, uint32_t* ovf, uint32_t D, uint32_t T, uint32_t fwords,
		uint32_t rbeg, uint32_t rend)
{
	const uint32_t G = 1u << glog;
	const uint32_t NG = 64u >> glog;                 /* sources per round */
	const uint32_t grp = lane >> glog;
	const uint32_t gl = lane & (G - 1u);
	const uint32_t rounds = rend;
	constexpr uint32_t DELIVER_RU = DeliverRounds<SAT, IZO>::value;
	for(uint32_t r0 = rbeg; r0 < rounds; r0 += DELIVER_RU) {    /* wave-uniform */
		uint2 rec[DELIVER_RU][DELIVER_KPL];
		uint32_t ok = 0;
#pragma unroll
		for(uint32_t u = 0; u < DELIVER_RU; ++u) {
			const uint32_t k = (r0 + u) * NG + grp;      /* source within the batch */
			const bool hi = ((r0 + u) * NG) >= 64u;     /* wave-uniform */
			const uint32_t ak = __shfl(hi ? a1 : a0, int(k & 63u), 64);
			const uint32_t bk = __shfl(hi ? b1 : b0, int(k & 63u), 64);
			const bool live = (r0 + u < rounds) && (k < nsrc);
#pragma unroll
			for(uint32_t i = 0; i < DELIVER_KPL; ++i) {
				const uint32_t q = ak + gl + i * G;
				const bool valid = live && (q < bk);
				/* the first load of a round is issued by every lane (record 0
				 * stands in for absent ones); the later ones only when some
				 * lane of the wave has a record there: buckets rarely reach
				 * past G records, and the vector memory unit's time goes by
				 * load instructions, not by the records they carry */
				if(i == 0 || __any(valid)) {   /* wave-uniform */
					rec[u][i] = ld_rec(syn, valid ? q : 0u);
				} else {
					rec[u][i] = make_uint2(0u, 0u);
				}
				ok |= uint32_t(valid) << (u * DELIVER_KPL + i);
			}
		}
#pragma unroll
		for(uint32_t u = 0; u < DELIVER_RU; ++u) {
#pragma unroll
			for(uint32_t i = 0; i < DELIVER_KPL; ++i) {
				if((ok >> (u * DELIVER_KPL + i)) & 1u) {
					const uint2 r = rec[u][i];
					uint32_t sl = base + (r.x >> 16);
					sl = (sl >= D) ? sl - D : sl;
					ring_add<SAT>(ringE, ringI, ovf, D, T, fwords, sl, r.x & 0xffffu, int32_t(r.y));
				}
			}
		}
	}
}

/* Delivers the records [a, b) of up to 128 sources held one per lane into
 * the ring, relative to base slot `base`. The first pass takes up to 2G
 * records of every bucket. The remainders of longer buckets are then packed
 * into the first lanes (each lane pushes its range to the lane of its rank
 * among the long ones) and delivered by further passes, so a long bucket
 * costs one more round of loads for the whole wave, not one per bucket.
 * All lanes of the wave must call it together (nsrc is wave-uniform). */
template<bool SAT, bool IZO>
__device__ inline void deliver_ranges(uint32_t a0, uint32_t b0, uint32_t a1, uint32_t b1, uint32_t nsrc,
		uint32_t glog, uint32_t lane, uint32_t base, const NEMO_GLOBAL uint2* syn,
		int32_t* ringE, int32_t* ringI, uint32_t* ovf, uint32_t D, uint32_t T, uint32_t fwords)
{
	const uint32_t covered = DELIVER_KPL << glog;   /* records per bucket taken by a pass */
	if(lane >= nsrc) { a0 = b0 = 0u; }
	if(64u + lane >= nsrc) { a1 = b1 = 0u; }
	for(;;) {
		deliver_pass<SAT, IZO>(a0, b0, a1, b1, nsrc, glog, lane, base, syn, ringE, ringI, ovf, D, T, fwords,
				0u, pass_rounds(nsrc, glog));
		/* what the pass left */
		a0 = (b0 - a0 > covered) ? a0 + covered : b0;
		a1 = (b1 - a1 > covered) ? a1 + covered : b1;
		const uint64_t m0 = __ballot(a0 < b0);
		const uint64_t m1 = __ballot(a1 < b1);
		const uint32_t n0 = uint32_t(__popcll(m0));
		const uint32_t nl = n0 + uint32_t(__popcll(m1));
		if(nl == 0u) {
			return;
		}
		if(nl > 63u) {
			/* too many to pack into one register: another pass as they are */
			nsrc = 128u;
			continue;
		}
		/* pack: a long source goes to the lane of its rank; the others to
		 * lane 63, which no rank reaches */
		const uint32_t below = uint32_t(__popcll(m0 & ((1ull << lane) - 1ull)));
		const uint32_t below1 = uint32_t(__popcll(m1 & ((1ull << lane) - 1ull)));
		const int dst0 = int(((a0 < b0) ? below : 63u) * 4u);
		const int dst1 = int(((a1 < b1) ? n0 + below1 : 63u) * 4u);
		const uint32_t pa0 = uint32_t(__builtin_amdgcn_ds_permute(dst0, int(a0)));
		const uint32_t pb0 = uint32_t(__builtin_amdgcn_ds_permute(dst0, int(b0)));
		const uint32_t pa1 = uint32_t(__builtin_amdgcn_ds_permute(dst1, int(a1)));
		const uint32_t pb1 = uint32_t(__builtin_amdgcn_ds_permute(dst1, int(b1)));
		a0 = (lane < n0) ? pa0 : pa1;
		b0 = (lane < n0) ? pb0 : pb1;
		if(lane >= nl) { a0 = b0 = 0u; }
		a1 = b1 = 0u;
		nsrc = nl;
	}
}

constexpr uint32_t STDP_CHUNK = 64u;   /* synapses per STDP work item (one per lane of a wave) */

/* Work list of one cycle, built by one wave for the targets j0 + lane
 * (< Fc) of the list described by pref: the chunks of STDP_CHUNK synapses
 * of each target's reverse row [r0, r0+len) that belong to partition p of
 * P (chunk c of the j-th target goes to partition (j + c) mod P) are
 * appended to `work`, and the reverse records of the first `cache` chunks
 * of the list are loaded straight into LDS (sdc, 64 words per chunk: one
 * wave instruction each, with every lane active). Called by whole waves. */
__device__ inline void stdp_list_build(uint32_t j0, uint32_t Fc, const uint32_t* pref, uint32_t P, uint32_t p,
		uint32_t T, const NEMO_GLOBAL uint64_t* srw, const NEMO_GLOBAL uint32_t* rsd, uint32_t lane,
		uint32_t cache, uint32_t* cnt, uint2* work, uint32_t* sdc)
{
	const uint32_t j = j0 + lane;
	uint32_t r0 = 0, len = 0, ch = 0, nch = 0;
	if(j < Fc) {
		const uint32_t q = partition_of(pref, P, j);
		const uint64_t rl = ld_agent(srw + size_t(q) * T + (j - pref[q]));
		r0 = uint32_t(rl);
		len = uint32_t(rl >> 32);
		nch = (len + STDP_CHUNK - 1u) / STDP_CHUNK;
		ch = (p + P - j % P) % P;
	}
	while(__any(ch < nch)) {   /* wave-uniform */
		const bool has = ch < nch;
		const uint32_t r = r0 + ch * STDP_CHUNK;
		const uint32_t l = has ? min(STDP_CHUNK, len - ch * STDP_CHUNK) : 0u;
		uint32_t slot = ~0u;
		if(has) {
			slot = atomicAdd(cnt, 1u);
			if(slot < NEMO_STDP_WORK_CAP) {
				work[slot] = make_uint2(r, l);
			}
		}
		uint64_t m = __ballot(has && slot < cache);
		while(m) {
			const uint32_t src = uint32_t(__ffsll((unsigned long long)m) - 1);
			const uint32_t sl = __builtin_amdgcn_readlane(slot, src);
			const uint32_t rs = __builtin_amdgcn_readlane(r, src);
			const uint32_t ls = __builtin_amdgcn_readlane(l, src);   /* >= 1 */
			__builtin_amdgcn_global_load_lds((NEMO_GLOBAL void*)(rsd + rs + min(lane, ls - 1u)),
					(NEMO_LDS void*)(sdc + sl * STDP_CHUNK), 4, 0, 0);
			m &= m - 1u;
		}
		ch += P;
	}
}

/* Accumulation over a work list of chunks: a wave takes the next STDP_K
 * chunks from the shared counter `next`, one synapse per lane of each; the reverse records of all of them
 * are loaded together (from LDS for the first `ncached` chunks, whose
 * records were loaded there during delivery), then the sources' histories. The same synapse is
 * accumulated by another partition (on another XCD) when its target appears
 * at another position of a later cycle's list, so the accumulator is
 * updated with a device-scope atomic add, not a cached read and write. */
template<bool IZO>
__device__ inline void stdp_accumulate(const uint2* work, uint32_t nw, uint32_t* next,
		uint32_t lane, const NEMO_GLOBAL uint64_t* rf, const NEMO_GLOBAL uint32_t* rsd,
		const uint32_t* sdc, uint32_t ncached, NEMO_GLOBAL int32_t* racc, const StdpParams& sp)
{
	constexpr uint32_t STDP_K = IZO ? 8u : 4u;   /* fewer registers left in the generic variant */
	for(;;) {
		uint32_t w0 = 0;
		if(lane == 0) w0 = atomicAdd(next, STDP_K);
		w0 = __builtin_amdgcn_readfirstlane(w0);   /* wave-uniform */
		if(w0 >= nw) break;
		uint32_t r[STDP_K], sd[STDP_K];
		uint64_t pre[STDP_K];
		bool ok[STDP_K];
#pragma unroll
		for(uint32_t k = 0; k < STDP_K; ++k) {
			const uint2 wk = (w0 + k < nw) ? work[w0 + k] : make_uint2(0u, 0u);
			ok[k] = lane < wk.y;
			r[k] = wk.x + lane;
		}
#pragma unroll
		for(uint32_t k = 0; k < STDP_K; ++k) {
			/* unconditional loads (entry 0 stands in for absent ones) */
			sd[k] = (w0 + k < ncached) ? sdc[(w0 + k) * 64u + lane] : rsd[ok[k] ? r[k] : 0u];
		}
#pragma unroll
		for(uint32_t k = 0; k < STDP_K; ++k) {
			pre[k] = ld_agent(rf + (sd[k] & 0x3ffffffu)) >> ((sd[k] >> 26) + 1u);
		}
#pragma unroll
		for(uint32_t k = 0; k < STDP_K; ++k) {
			const int32_t wd = ok[k] ? stdp_weight_change(sp, pre[k]) : 0;
			if(wd != 0) {
				add_agent(racc + r[k], wd);
			}
		}
	}
}

/* ---- exchange waves and far waves (networks with deferred delivery).
 *
 * After publishing cycle t, the first exWaves waves of a workgroup run the
 * latency-critical chain of the cycle: wait for every partition's count,
 * read the fired-list entries, the bucket bounds of this partition (loaded
 * straight into LDS) and the delay-1 records, whose weights go to slot t+1
 * before the next update. The rest of every bucket (delays >= 2) is left as
 * a range in the LDS far list. Meanwhile the other waves deliver the far
 * list of cycle t-1 into slots t+1 .. t-1+D, which the update of t+1 does
 * not read beyond slot t+1, and the delay-2 records of cycle t-1, bound for
 * slot t+1, are done before the cycle's closing barrier. */
/* fired sources per exchange lane per pass (fewer where the variant has
 * fewer registers left) */
template<bool SAT, bool IZO> struct ExSources { static constexpr uint32_t value = (SAT && !IZO) ? 4u : 6u; };
constexpr uint32_t EX_NEAR = 4u;   /* delay-1 records per lane loaded together */

template<bool SAT, bool IZO>
__device__ inline unsigned long long exchange_phase(const NetDev& nd, uint32_t p, uint64_t t, uint32_t k,
		uint32_t slot, uint32_t flb, uint32_t wave, uint32_t lane, uint32_t* pref, uint32_t* exFlag,
		uint32_t* FA, uint32_t* FB, uint32_t* fcnt, uint32_t* pringDst, const NEMO_GLOBAL uint2* syn,
		const NEMO_GLOBAL uint32_t* off, uint32_t offMask, int32_t* ringE, int32_t* ringI, uint32_t* ovf,
		uint32_t fwords, const uint32_t* listDone, uint32_t listTarget)
{
	constexpr uint32_t EX_K = ExSources<SAT, IZO>::value;
	const uint32_t P = nd.P, T = nd.T, D = nd.D, exw = nd.exWaves, cap = nd.farCap;
	const uint32_t tag = uint32_t(t);
	/* wave 0 waits for the counts; the other exchange waves for wave 0 */
	if(wave == 0) {
		gather_counts(glb(nd.fc) + flb * P, P, tag, pref, lane, nd.spinLimit, nd.status, p);
		if(lane == 0) {
			__hip_atomic_store(exFlag, k + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
		}
	}
	while(__hip_atomic_load(exFlag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != k + 1u) {
		__builtin_amdgcn_s_sleep(1);
	}
	const uint32_t F = pref[P];
	const uint32_t lim = min(F, cap);
	const uint32_t nEl = exw * 64u;
	const uint32_t el = wave * 64u + lane;
	if(el == 0) *fcnt = lim;
	if(pringDst) {
		/* the slot still holds the prefix the far waves build this cycle's
		 * STDP list from: wait until they have */
		while(__hip_atomic_load(listDone, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < listTarget) {
			__builtin_amdgcn_s_sleep(1);
		}
		for(uint32_t i = el; i <= P; i += nEl) pringDst[i] = pref[i];
	}
	const NEMO_GLOBAL uint64_t* fl = glb(nd.fl) + size_t(flb) * P * T;
	const uint32_t s1 = (slot + 1u >= D) ? slot + 1u - D : slot + 1u;   /* slot of cycle t+1 */
	unsigned long long events = 0;
	for(uint32_t base = 0; base < lim; base += nEl * EX_K) {   /* wave-uniform */
		/* the fired-list entries: EX_K sources per lane, loads in flight together */
		uint64_t ev[EX_K];
		uint32_t eo[EX_K];
#pragma unroll
		for(uint32_t i = 0; i < EX_K; ++i) {
			const uint32_t j = base + i * nEl + el;
			eo[i] = 0u;
			ev[i] = 0ull;
			if(j < lim) {
				const uint32_t q = partition_of(pref, P, j);
				eo[i] = q * T + (j - pref[q]);
				ev[i] = ld_agent(fl + eo[i]);
			}
		}
		uint32_t src[EX_K];
#pragma unroll
		for(uint32_t i = 0; i < EX_K; ++i) {
			const uint32_t j = base + i * nEl + el;
			src[i] = (j < lim) ? fired_entry(fl + eo[i], ev[i], tag, nd.spinLimit, nd.status, p, eo[i] / T) : 0u;
		}
		/* the bucket bounds off[s][p], off[s][p+1], straight into the far
		 * list slots of the sources (one wave instruction per 64 sources;
		 * lanes past the end load the row of neuron 0 into slots that are
		 * not read) */
#pragma unroll
		for(uint32_t i = 0; i < EX_K; ++i) {
			const uint32_t jw = base + i * nEl + wave * 64u;   /* wave-uniform */
			if(jw < lim) {
				const NEMO_GLOBAL uint32_t* row = off + size_t(src[i]) * (P + 1u) + p;
				__builtin_amdgcn_global_load_lds((NEMO_GLOBAL void*) row, (NEMO_LDS void*)(FA + jw), 4, 0, 0);
				__builtin_amdgcn_global_load_lds((NEMO_GLOBAL void*)(row + 1), (NEMO_LDS void*)(FB + jw), 4, 0, 0);
			}
		}
		__builtin_amdgcn_s_waitcnt(0);
		/* delay-1 records (first in the bucket, their count in the top bits
		 * of the start); the rest of the bucket stays in the far list */
		uint32_t na[EX_K], nn[EX_K];
		uint32_t R = 0;
#pragma unroll
		for(uint32_t i = 0; i < EX_K; ++i) {
			const uint32_t j = base + i * nEl + el;
			na[i] = 0u;
			nn[i] = 0u;
			if(j < lim) {
				const uint32_t ar = FA[j];
				const uint32_t a = ar & offMask;
				const uint32_t b = FB[j] & offMask;
				na[i] = a;
				nn[i] = ar >> OFF_NEAR_SHIFT;
				events += b - a;
				FA[j] = a + nn[i];
				FB[j] = b;
			}
			R += nn[i];
		}
		for(uint32_t r0 = 0; __any(r0 < R); r0 += EX_NEAR) {   /* wave-uniform */
			uint2 rec[EX_NEAR];
#pragma unroll
			for(uint32_t u = 0; u < EX_NEAR; ++u) {
				uint32_t r = r0 + u;
				uint32_t q = 0u;
				bool found = false;
#pragma unroll
				for(uint32_t i = 0; i < EX_K; ++i) {
					if(!found) {
						if(r < nn[i]) {
							q = na[i] + r;
							found = true;
						} else {
							r -= nn[i];
						}
					}
				}
				rec[u] = ld_rec(syn, found ? q : 0u);
			}
#pragma unroll
			for(uint32_t u = 0; u < EX_NEAR; ++u) {
				if(r0 + u < R) {
					ring_add<SAT>(ringE, ringI, ovf, D, T, fwords, s1, rec[u].x & 0xffffu, int32_t(rec[u].y));
				}
			}
		}
	}
	/* sources beyond the far list's capacity: every record now */
	for(uint32_t jb = cap + wave * 128u; jb < F; jb += exw * 128u) {   /* wave-uniform */
		const uint32_t nsrc = min(128u, F - jb);
		uint32_t a0 = 0, b0 = 0, a1 = 0, b1 = 0;
		const uint32_t j0 = jb + lane;
		const uint32_t j1 = jb + 64u + lane;
		const bool v0 = lane < nsrc;
		const bool v1 = 64u + lane < nsrc;
		const uint32_t q0 = v0 ? partition_of(pref, P, j0) : 0u;
		const uint32_t q1 = v1 ? partition_of(pref, P, j1) : 0u;
		const uint32_t o0 = q0 * T + (j0 - pref[q0]);
		const uint32_t o1 = q1 * T + (j1 - pref[q1]);
		uint64_t e0 = 0, e1 = 0;
		if(v0) e0 = ld_agent(fl + o0);
		if(v1) e1 = ld_agent(fl + o1);
		uint32_t x0 = 0, x1 = 0;
		if(v0) x0 = fired_entry(fl + o0, e0, tag, nd.spinLimit, nd.status, p, q0);
		if(v1) x1 = fired_entry(fl + o1, e1, tag, nd.spinLimit, nd.status, p, q1);
		if(v0) {
			const NEMO_GLOBAL uint32_t* o = off + size_t(x0) * (P + 1u) + p;
			a0 = ld_off(o) & offMask;
			b0 = ld_off(o + 1) & offMask;
		}
		if(v1) {
			const NEMO_GLOBAL uint32_t* o = off + size_t(x1) * (P + 1u) + p;
			a1 = ld_off(o) & offMask;
			b1 = ld_off(o + 1) & offMask;
		}
		events += (b0 - a0) + (b1 - a1);
		deliver_ranges<SAT, IZO>(a0, b0, a1, b1, nsrc, nd.glog2, lane, slot, syn, ringE, ringI, ovf, D, T, fwords);
	}
	return events;
}

/* The far list of an earlier cycle (ranges FA[j]..FB[j] of its Fp fired
 * sources, relative to that cycle's slot `base`), delivered by nw waves,
 * wave w taking a contiguous share. Called by whole waves. */
template<bool SAT, bool IZO>
__device__ inline void far_phase(const NetDev& nd, uint32_t Fp, const uint32_t* FA, const uint32_t* FB,
		uint32_t w, uint32_t nw, uint32_t lane, uint32_t base, const NEMO_GLOBAL uint2* syn,
		int32_t* ringE, int32_t* ringI, uint32_t* ovf, uint32_t fwords)
{
	const uint32_t share = (Fp + nw - 1u) / nw;
	const uint32_t j0 = min(Fp, w * share);
	const uint32_t j1 = min(Fp, j0 + share);
	for(uint32_t jb = j0; jb < j1; jb += 128u) {   /* wave-uniform */
		const uint32_t nsrc = min(128u, j1 - jb);
		uint32_t a0 = 0, b0 = 0, a1 = 0, b1 = 0;
		if(lane < nsrc) {
			a0 = FA[jb + lane];
			b0 = FB[jb + lane];
		}
		if(64u + lane < nsrc) {
			a1 = FA[jb + 64u + lane];
			b1 = FB[jb + 64u + lane];
		}
		deliver_ranges<SAT, IZO>(a0, b0, a1, b1, nsrc, nd.glog2, lane, base, syn, ringE, ringI, ovf,
				nd.D, nd.T, fwords);
	}
}

/* IZO: every neuron of the network is an Izhikevich neuron, so only its 5
 * parameters and 2 state variables are held in registers. Always inlined
 * into the kernels: as a call, the network descriptor is copied to scratch
 * (about 500 B per lane), which the compiler chooses by itself once the
 * loop grows past its inlining threshold. */
template<bool SAT, bool IZO>
__device__ __forceinline__ void run_partition(const NetDev& nd, uint32_t p)
{
	constexpr int NP = IZO ? 5 : MAX_PARAM;
	constexpr int NS = IZO ? 2 : MAX_STATE;
	extern __shared__ __align__(16) uint8_t lds[];

	const uint32_t T = nd.T;
	const uint32_t D = nd.D;
	const uint32_t P = nd.P;
	const uint32_t n_total = nd.n;
	const uint32_t tid = threadIdx.x;
	const uint32_t nthreads = blockDim.x;
	const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);   /* wave-uniform, in a scalar register */
	const uint32_t lane = tid & 63u;
	const uint32_t nwaves = nthreads >> 6;
	const uint32_t n = p * T + tid;
	const bool active = (tid < T) && (n < n_total);

	int32_t* ringE = reinterpret_cast<int32_t*>(lds);
	int32_t* ringI = ringE + D * T;
	float* ext = reinterpret_cast<float*>(ringI + D * T);
	uint32_t* fst = reinterpret_cast<uint32_t*>(ext + T);
	uint32_t* pref = fst + (T >> 5);
	uint32_t* wsum = pref + ((P + 1u + 63u) & ~63u);
	uint32_t* ovf = wsum + 64;                                 /* [2][D][T/32] when SAT */
	const uint32_t fwords = T >> 5;
	/* with STDP: the STDP tables, then the row start and the prefix of the
	 * row lengths of each target processed in the STDP phase */
	uint8_t* stdpArea = reinterpret_cast<uint8_t*>(ovf + (SAT ? 2u * D * fwords : 0u));
	StdpParams* spl = reinterpret_cast<StdpParams*>(stdpArea);
	uint2* work = reinterpret_cast<uint2*>(stdpArea + NEMO_STDP_LDS_BYTES);   /* [cap] {row index, length} */
	uint32_t* nwork = reinterpret_cast<uint32_t*>(work + NEMO_STDP_WORK_CAP);  /* [2], by cycle parity */
	uint32_t* spref = nwork + 16;   /* prefix of the counts of cycle t-postLen, [P+1] */
	uint32_t* sdcache = spref + ((P + 1u + 63u) & ~63u);   /* [stdpCache][64] reverse records */
	/* with NEMO_NET_PREF_RING: the prefixes of the counts of the last postLen
	 * cycles, [postLen][(P+1) rounded up to 64], slot t mod postLen */
	uint32_t* pringArea = sdcache + nd.stdpCache * 64u;
	const uint32_t PW = (P + 1u + 63u) & ~63u;
	const bool defer = (nd.flags & NEMO_NET_DEFER) != 0u;
	const uint32_t offMask = defer ? ((1u << OFF_NEAR_SHIFT) - 1u) : ~0u;
	const NEMO_GLOBAL uint2* syn = glb(reinterpret_cast<const uint2*>(nd.syn));
	const NEMO_GLOBAL uint32_t* off = glb(nd.off);

	/* count the workgroups of the network that have started (diagnostics),
	 * and keep the first start on the device wall clock (control block
	 * words 8..9; with the last end in 10..11, the network's time in a
	 * batched launch, whose graph initialises both) */
	if(tid == 0) {
		atomicAdd(nd.status - 1, 1u);
		atomicMin(reinterpret_cast<unsigned long long*>(nd.status + 7), (unsigned long long) wall_clock64());
	}

	/* restore the pending-current ring: 1 KB per wave instruction straight
	 * into LDS (16 bytes per lane), every load in flight at once; the ring
	 * is 8 D T bytes, a multiple of 512 since T is a multiple of 64, so at
	 * most half a KB is left for ordinary loads */
	const int32_t* gring = nd.ring + size_t(p) * 2u * D * T;
	const bool replay = (nd.flags & NEMO_NET_REPLAY) != 0u;
	const uint32_t ringWords = 2u * D * T;
	if(replay) {
		for(uint32_t i = tid; i < ringWords; i += nthreads) {
			ringE[i] = 0;
		}
	} else {
		const uint32_t chunks = ringWords / 256u;
		for(uint32_t c = wave; c < chunks; c += nwaves) {   /* wave-uniform */
			__builtin_amdgcn_global_load_lds((NEMO_GLOBAL void*)(gring + c * 256u + lane * 4u),
					(NEMO_LDS void*)(ringE + c * 256u), 16, 0, 0);
		}
		for(uint32_t i = chunks * 256u + tid; i < ringWords; i += nthreads) {
			ringE[i] = gring[i];
		}
		__builtin_amdgcn_s_waitcnt(0);
	}
	if(SAT) {
		for(uint32_t i = tid; i < 2u * D * fwords; i += nthreads) {
			ovf[i] = 0u;
		}
	}
	if(replay) {
		/* Weights changed (applyStdp) since the ring was filled. NeMo reads a
		 * weight when the spike arrives (src/nemo/cpu/Simulation.cpp:258-307),
		 * so rebuild the ring from the firing history: every spike fired at
		 * cycle0-1-k whose synapse delay d > k has not arrived yet and lands
		 * in slot (cycle0-1-k+d) mod D, with the current weight. */
		__syncthreads();
		const uint64_t pend = (D >= 64u) ? ~0ull : ((1ull << D) - 1ull);
		const uint32_t last = uint32_t((nd.cycle0 + D - 1u) % D);    /* slot of cycle0-1 */
		for(uint32_t s = tid; s < n_total; s += nthreads) {
			const uint64_t h = glb(nd.recent)[s] & pend;
			if(!h) continue;
			const NEMO_GLOBAL uint32_t* o = off + size_t(s) * (P + 1u) + p;
			const uint32_t a = o[0] & offMask, b = o[1] & offMask;
			for(uint32_t q = a; q < b; ++q) {
				const uint2 rec = syn[q];
				const uint32_t d = rec.x >> 16;
				uint64_t hb = h & ((d >= 64u) ? ~0ull : ((1ull << d) - 1ull));
				while(hb) {
					const uint32_t kk = uint32_t(__builtin_ctzll(hb));
					hb &= hb - 1ull;
					ring_add<SAT>(ringE, ringI, ovf, D, T, fwords, (last + d - kk) % D, rec.x & 0xffffu, int32_t(rec.y));
				}
			}
		}
	}

	/* neuron registers */
	int model = 0;
	float prm[NP];
	float st[NS];
	Rng rng;
	uint64_t recent = 0;
#pragma unroll
	for(int k = 0; k < NP; ++k) prm[k] = 0.0f;
#pragma unroll
	for(int k = 0; k < NS; ++k) st[k] = 0.0f;
	rng.s[0] = rng.s[1] = rng.s[2] = rng.s[3] = 0u;
	if(active) {
		model = nd.model[n];
#pragma unroll
		for(int k = 0; k < NP; ++k) prm[k] = nd.param[size_t(k) * n_total + n];
#pragma unroll
		for(int k = 0; k < NS; ++k) st[k] = nd.state[size_t(k) * n_total + n];
#pragma unroll
		for(int k = 0; k < 4; ++k) rng.s[k] = nd.rng[size_t(k) * n_total + n];
		recent = nd.recent[n];
	}

	const bool stdp = (nd.flags & NEMO_NET_STDP) != 0u;
	const StdpParams* sp = reinterpret_cast<const StdpParams*>(nd.stdp);
	const uint32_t postLen = stdp ? uint32_t(sp->postLen) : 0u;
	if(stdp) {
		/* the tables to LDS: they are read for every plastic synapse */
		const uint32_t* src = reinterpret_cast<const uint32_t*>(sp);
		uint32_t* dst = reinterpret_cast<uint32_t*>(spl);
		for(uint32_t i = tid; i < uint32_t(sizeof(StdpParams) / 4u); i += nthreads) dst[i] = src[i];
	}
	/* Early STDP work lists: with postLen >= 2 the work of cycle t is listed
	 * right after this partition publishes cycle t, from the fired lists and
	 * reverse-row table of cycle t-postLen, which this partition has seen
	 * published (its gather of cycle t-1) -- so only the accumulation itself
	 * waits for cycle t's exchange. */
	const bool early = stdp && postLen >= 2u;
	const bool pring = early && (nd.flags & NEMO_NET_PREF_RING);
	/* exchange waves: the far list after the optional STDP areas */
	const uint32_t exw = (defer && nd.exWaves < nwaves) ? nd.exWaves : 0u;
	const uint32_t fcap = nd.farCap;
	uint32_t* farA = stdp ? pringArea + (pring ? postLen * PW : 0u) : reinterpret_cast<uint32_t*>(stdpArea);
	uint32_t* farB = farA + 2u * fcap;
	uint32_t* fcnt = farB + 2u * fcap;      /* [2], by cycle parity */
	uint32_t* exFlag = fcnt + 2;
	/* with exchange waves and the early STDP list: far waves that have built
	 * their share of the cycle's list (records staged), counted over the
	 * launch (the word after the flag) */
	uint32_t* listDone = exFlag + 1;
	const uint32_t nfar = nwaves - exw;
	if(exw && tid == 0) {
		fcnt[0] = fcnt[1] = 0u;
		*exFlag = 0u;
		*listDone = 0u;
	}
	uint32_t rslot = pring ? uint32_t(nd.cycle0 % postLen) : 0u;   /* t mod postLen */
	if(stdp && tid == 0) {
		nwork[0] = nwork[1] = 0u;   /* list lengths, by cycle parity */
		nwork[2] = nwork[3] = 0u;   /* chunks taken from them */
	}
	if(pring) {
		/* the prefixes of the launch's first postLen cycles' lists, one
		 * earlier cycle per wave at a time */
		for(uint32_t i = wave; i < postLen; i += nwaves) {
			if(nd.cycle0 >= postLen - i) {
				const uint64_t tc = nd.cycle0 - postLen + i;
				gather_counts(glb(nd.fc) + uint32_t(tc % nd.flRing) * P, P, uint32_t(tc),
						pringArea + uint32_t(tc % postLen) * PW, lane, nd.spinLimit, nd.status, p);
			}
		}
	} else if(early && wave == 0 && nd.cycle0 >= postLen) {
		const uint64_t tc = nd.cycle0 - postLen;
		gather_counts(glb(nd.fc) + uint32_t(tc % nd.flRing) * P, P, uint32_t(tc), spref, lane,
				nd.spinLimit, nd.status, p);
	}
	unsigned long long events = 0;
	/* deferred delivery: this wave's far ranges [a, b) (delays >= 2) of the
	 * previous cycle's spikes, one source per lane, kn sources (wave-uniform) */
	uint32_t ka0 = 0, kb0 = 0, ka1 = 0, kb1 = 0, kn = 0;
	/* Izhikevich-only networks draw each neuron's Gaussian input of the next
	 * cycle ahead, while the wave waits for memory (the first cycle's here):
	 * sigma * nrand() does not depend on the cycle's input, and every cycle
	 * takes exactly one draw, so the stream is the same. The last cycle of a
	 * launch draws nothing ahead, so the saved RNG state is exact. */
	float noise = 0.0f;
	if(IZO && active && nd.nsteps > 0u && prm[IZ_SIGMA] != 0.0f) {
		noise = prm[IZ_SIGMA] * nrand(rng);
	}
	/* ring slot t mod D and fired-list buffer t mod flRing, kept as counters */
	uint32_t slotK = uint32_t(nd.cycle0 % D);
	uint32_t flbK = uint32_t(nd.cycle0 % nd.flRing);
	__syncthreads();

	for(uint32_t k = 0; k < nd.nsteps; ++k) {
		const uint64_t t = nd.cycle0 + k;
		const uint32_t slot = slotK;
		const uint32_t par = uint32_t(t & 1u);
		/* fired-list and count buffer of this cycle: a ring of flRing
		 * cycles (at least 2; 3 when deferred delivery reads the previous
		 * cycle's lists while the next are written; postLen+2 with STDP,
		 * which reads the lists of cycle t-postLen) */
		const uint32_t flb = flbK;
		slotK = (slotK + 1u == D) ? 0u : slotK + 1u;
		flbK = (flbK + 1u == nd.flRing) ? 0u : flbK + 1u;
		const bool drawn = !(IZO && active && k + 1u < nd.nsteps && prm[IZ_SIGMA] != 0.0f);

		/* ---- stimulus of this step (sim-local indices, deduplicated by the host) */
		bool has_stim = false;
		uint32_t fs0 = 0, fs1 = 0, is0 = 0, is1 = 0;
		if(nd.fs_ptr) { fs0 = glb(nd.fs_ptr)[k]; fs1 = glb(nd.fs_ptr)[k + 1]; }
		if(nd.is_ptr) { is0 = glb(nd.is_ptr)[k]; is1 = glb(nd.is_ptr)[k + 1]; }
		has_stim = (fs1 > fs0) || (is1 > is0);
		if(has_stim) {
			if(tid < T) ext[tid] = 0.0f;
			if(tid < fwords) fst[tid] = 0u;
			__syncthreads();
			const uint32_t lo = p * T, hi = p * T + T;
			for(uint32_t e = fs0 + tid; e < fs1; e += nthreads) {
				uint32_t idx = glb(nd.fs_idx)[e];
				if(idx >= lo && idx < hi) {
					uint32_t l = idx - lo;
					atomicOr(&fst[l >> 5], 1u << (l & 31u));
				}
			}
			for(uint32_t e = is0 + tid; e < is1; e += nthreads) {
				uint32_t idx = glb(nd.is_idx)[e];
				if(idx >= lo && idx < hi) {
					ext[idx - lo] = glb(nd.is_val)[e];
				}
			}
			__syncthreads();
		}

		/* ---- 1. neuron update */
		bool fired = false;
		if(active) {
			const uint32_t r = slot * T + tid;
			int32_t e = ringE[r];
			int32_t i = ringI[r];
			if(SAT) {
				const uint32_t bit = 1u << (tid & 31u);
				if(ovf[slot * fwords + (tid >> 5)] & bit) e = INT32_MAX;
				if(ovf[(D + slot) * fwords + (tid >> 5)] & bit) i = INT32_MIN;
			}
			ringE[r] = 0;
			ringI[r] = 0;
			float cx = 0.0f;
			bool fs = false;
			if(has_stim) {
				cx = ext[tid];
				fs = ((fst[tid >> 5] >> (tid & 31u)) & 1u) != 0u;
			}
			if(IZO) {
				fired = izhikevich_drawn(prm, st, noise, fx20_to_float(e), fx20_to_float(i), cx, fs);
			} else {
				fired = update_neuron(model, prm, st, rng, fx20_to_float(e), fx20_to_float(i), cx, fs);
			}
			recent = (recent << 1) | (fired ? 1ull : 0ull);
			if(stdp) {
				st_agent(glb(nd.rfg) + size_t(par) * n_total + n, recent);
			}
		}

		/* ---- 2. compaction: ballot + prefix over waves */
		const uint64_t mask = __ballot(fired);
		if(lane == 0) {
			wsum[wave] = uint32_t(__popcll(mask));
			if(wave < (T >> 6)) {
				glb(nd.raster)[size_t(k) * nd.words + p * (T >> 6) + wave] = mask;
			}
		}
		__syncthreads();
		if(SAT) {
			/* the slot just consumed starts empty for cycle t + D */
			for(uint32_t w = tid; w < fwords; w += nthreads) {
				ovf[slot * fwords + w] = 0u;
				ovf[(D + slot) * fwords + w] = 0u;
			}
		}
		{
			uint32_t total;
			uint32_t base = lds_prefix_waves(wsum, wave, nwaves, &total);
			const uint64_t below = lane ? (mask & (~0ull >> (64u - lane))) : 0ull;
			const uint32_t pos = base + uint32_t(__popcll(below));   /* in the partition's fired list */
			if(fired) {
				st_agent(glb(nd.fl) + (size_t(flb) * P + p) * T + pos, (uint64_t(uint32_t(t)) << 32) | uint64_t(n));
				/* Touch every 128-byte line of this neuron's bucket-offset row
				 * (loads straight into an LDS sink, no registers held), so
				 * the other partitions, reading one entry each right after
				 * the exchange, find the row in the memory-side cache. */
				const NEMO_GLOBAL uint32_t* row = off + size_t(n) * (P + 1u);
				for(uint32_t li = 0; li < P + 1u; li += 32u) {
					__builtin_amdgcn_global_load_lds((NEMO_GLOBAL void*)(row + li), (NEMO_LDS void*)ext, 4, 0, 0);
				}
			}
			/* publish the count. Fired-list entries carry the cycle tag, so
			 * readers validate each entry and need no ordering against the
			 * count; the STDP firing history does not, so with STDP every
			 * wave's stores are drained first (s_waitcnt; __syncthreads alone
			 * waits only for LDS traffic). The reverse row of each fired
			 * neuron (the STDP lists of cycle t+postLen are built from it) is
			 * loaded meanwhile and stored after the drain: the publish of
			 * cycle t+1 drains it. */
			if(stdp) {
				uint32_t my0 = 0, my1 = 0;
				if(early && fired) {
					my0 = glb(nd.rcm_start)[n];
					my1 = glb(nd.rcm_start)[n + 1];
				}
				__builtin_amdgcn_s_waitcnt(0);
				__syncthreads();
				if(early && fired) {
					st_agent(glb(nd.srow) + (size_t(flb) * P + p) * T + pos, (uint64_t(my1 - my0) << 32) | my0);
				}
			}
			if(tid == 0) {
				st_agent(glb(nd.fc) + flb * P + p, (uint64_t(uint32_t(t)) << 32) | uint64_t(total));
			}
		}

		/* the next cycle's Gaussian input, drawn while the memory-bound
		 * phases below keep the SIMDs' other waves busy */
		if(!drawn) {
			noise = prm[IZ_SIGMA] * nrand(rng);
		}

		/* ---- deferred records (delays >= 2) of the previous cycle's spikes,
		 * delivered while the other partitions publish this cycle */
		if(!exw && defer && kn) {
			deliver_ranges<SAT, IZO>(ka0, kb0, ka1, kb1, kn, nd.glog2, lane, (slot + D - 1u) % D, syn,
					ringE, ringI, ovf, D, T, fwords);
			kn = 0;
		}

		/* ---- STDP work list of this cycle, early (see the prologue): the
		 * targets are the fired list of cycle t-postLen; chunk c of the j-th
		 * target goes to partition (j + c) mod P */
		if(early && t >= postLen) {
			const uint32_t cb = uint32_t((t - postLen) % nd.flRing);
			const uint32_t* sp = pring ? pringArea + rslot * PW : spref;
			const uint32_t Fc = sp[P];
			const NEMO_GLOBAL uint64_t* srw = glb(nd.srow) + size_t(cb) * P * T;
			/* with exchange waves the far waves build it, so that the
			 * exchange starts at once */
			const uint32_t lw0 = exw ? wave - exw : wave, lnw = exw ? nfar : nwaves;
			if(!exw || wave >= exw) {
				for(uint32_t j0 = lw0 * 64u; j0 < Fc; j0 += lnw * 64u) {   /* wave-uniform */
					stdp_list_build(j0, Fc, sp, P, p, T, srw, glb(nd.rcm_sd), lane, nd.stdpCache,
							nwork + (t & 1u), work, sdcache);
				}
			}
			/* the staged records land before the exchange barrier, or,
			 * with exchange waves, before the wave is counted: after that,
			 * any wave may read them (wave 0's gather waits for them anyway;
			 * exchange waves stage none) */
			if(exw ? wave >= exw : wave != 0) __builtin_amdgcn_s_waitcnt(0);
		}
		if(early && exw && wave >= exw && lane == 0) {
			__hip_atomic_fetch_add(listDone, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
		}

		if(exw) {
			/* ---- 3+4 with exchange waves: the exchange and this cycle's
			 * delay-1 records on waves [0, exw), the far list of cycle t-1
			 * on the others (see exchange_phase) */
			const uint32_t cb = uint32_t(t & 1u);
			if(wave < exw) {
				/* the prefix is kept for the list of cycle t+postLen (slot t
				 * mod postLen, whose cycle t-postLen list was built above) */
				events += exchange_phase<SAT, IZO>(nd, p, t, k, slot, flb, wave, lane, pref, exFlag,
						farA + cb * fcap, farB + cb * fcap, fcnt + cb, pring ? pringArea + rslot * PW : nullptr,
						syn, off, offMask, ringE, ringI, ovf, fwords, listDone, (k + 1u) * nfar);
			} else {
				far_phase<SAT, IZO>(nd, fcnt[cb ^ 1u], farA + (cb ^ 1u) * fcap, farB + (cb ^ 1u) * fcap,
						wave - exw, nwaves - exw, lane, (slot + D - 1u) % D, syn, ringE, ringI, ovf, fwords);
			}
			if(pring) {
				rslot = rslot + 1u == postLen ? 0u : rslot + 1u;
			}
		} else {
		/* ---- 3. wait for the network's other partitions */
		if(wave == 0) {
			gather_counts(glb(nd.fc) + flb * P, P, uint32_t(t), pref, lane, nd.spinLimit, nd.status, p);
		}
		__syncthreads();

		if(pring) {
			/* kept for the list of cycle t+postLen (slot t mod postLen, whose
			 * cycle t-postLen list was built above) */
			for(uint32_t i = tid; i <= P; i += nthreads) pringArea[rslot * PW + i] = pref[i];
			rslot = rslot + 1u == postLen ? 0u : rslot + 1u;
		}

		/* ---- 4. delivery of this cycle's spikes into the ring */
		{
			const uint32_t F = pref[P];
			const NEMO_GLOBAL uint64_t* fl = glb(nd.fl) + size_t(flb) * P * T;
			/* Each wave takes a contiguous share of the F fired sources and
			 * loads, for up to 128 of them at once, the fired-list entry and
			 * the bucket bounds off[s][p..p+1] (one lane per source, all loads
			 * in flight together); deliver_ranges then reads the records. */
			const uint32_t share = (F + nwaves - 1u) / nwaves;
			const uint32_t jw0 = min(F, wave * share);
			const uint32_t jw1 = min(F, jw0 + share);
			for(uint32_t jb = jw0; jb < jw1; jb += 128u) {   /* wave-uniform */
				const uint32_t nsrc = min(128u, jw1 - jb);
				uint32_t a0 = 0, b0 = 0, a1 = 0, b1 = 0;
				{
					const uint32_t j0 = jb + lane;
					const uint32_t j1 = jb + 64u + lane;
					const bool v0 = lane < nsrc;
					const bool v1 = 64u + lane < nsrc;
					const uint32_t q0 = v0 ? partition_of(pref, P, j0) : 0u;
					const uint32_t q1 = v1 ? partition_of(pref, P, j1) : 0u;
					/* entry offsets within the cycle's fired-list buffer */
					const uint32_t o0 = q0 * T + (j0 - pref[q0]);
					const uint32_t o1 = q1 * T + (j1 - pref[q1]);
					uint64_t ev0 = 0, ev1 = 0;
					if(v0) ev0 = ld_agent(fl + o0);
					if(v1) ev1 = ld_agent(fl + o1);
					uint32_t s0 = 0, s1 = 0;
					if(v0) s0 = fired_entry(fl + o0, ev0, uint32_t(t), nd.spinLimit, nd.status, p, q0);
					if(v1) s1 = fired_entry(fl + o1, ev1, uint32_t(t), nd.spinLimit, nd.status, p, q1);
					if(v0) {
						const NEMO_GLOBAL uint32_t* o = off + size_t(s0) * (P + 1u) + p;
						a0 = ld_off(o);
						b0 = ld_off(o + 1) & offMask;
					}
					if(v1) {
						const NEMO_GLOBAL uint32_t* o = off + size_t(s1) * (P + 1u) + p;
						a1 = ld_off(o);
						b1 = ld_off(o + 1) & offMask;
					}
				}
				if(defer && jb == jw0) {
					/* the records of delay 1 (first in each bucket, their
					 * count in the top bits of the start) now; the rest in
					 * the next cycle, while the exchange is waited for */
					const uint32_t n0 = a0 >> OFF_NEAR_SHIFT, n1 = a1 >> OFF_NEAR_SHIFT;
					a0 &= offMask;
					a1 &= offMask;
					events += (b0 - a0) + (b1 - a1);
					/* delay-1 records: slot t+1; two per source loaded together */
					const uint32_t m = max(n0, n1);
					for(uint32_t i = 0; i < m; i += 2u) {
						const bool ok00 = i < n0, ok01 = i + 1u < n0, ok10 = i < n1, ok11 = i + 1u < n1;
						const uint2 r00 = ld_rec(syn, ok00 ? a0 + i : 0u);
						const uint2 r01 = ld_rec(syn, ok01 ? a0 + i + 1u : 0u);
						const uint2 r10 = ld_rec(syn, ok10 ? a1 + i : 0u);
						const uint2 r11 = ld_rec(syn, ok11 ? a1 + i + 1u : 0u);
						const uint32_t sl = (slot + 1u >= D) ? slot + 1u - D : slot + 1u;
						if(ok00) ring_add<SAT>(ringE, ringI, ovf, D, T, fwords, sl, r00.x & 0xffffu, int32_t(r00.y));
						if(ok01) ring_add<SAT>(ringE, ringI, ovf, D, T, fwords, sl, r01.x & 0xffffu, int32_t(r01.y));
						if(ok10) ring_add<SAT>(ringE, ringI, ovf, D, T, fwords, sl, r10.x & 0xffffu, int32_t(r10.y));
						if(ok11) ring_add<SAT>(ringE, ringI, ovf, D, T, fwords, sl, r11.x & 0xffffu, int32_t(r11.y));
					}
					ka0 = a0 + n0;
					kb0 = b0;
					ka1 = a1 + n1;
					kb1 = b1;
					kn = nsrc;
				} else {
					a0 &= offMask;
					a1 &= offMask;
					events += (b0 - a0) + (b1 - a1);
					deliver_ranges<SAT, IZO>(a0, b0, a1, b1, nsrc, nd.glog2, lane, slot, syn,
							ringE, ringI, ovf, D, T, fwords);
				}
			}
		}
		}

		/* the row touches above write into `ext`, which the STDP phase and the
		 * next cycle's stimulus reuse: let them land first */
		__builtin_amdgcn_s_waitcnt(0);

		/* ---- 5. STDP accumulation for the targets that fired postLen cycles
		 * ago (src/nemo/ConnectivityMatrix.cpp:193-227). They are the fired
		 * list of cycle t-postLen, which every partition still holds (ring
		 * of postLen+2 cycles); their plastic in-synapse rows are cut into
		 * chunks of STDP_CHUNK, and chunk c of the j-th target of that list
		 * is processed by partition (j + c) mod P, so every partition takes
		 * an equal share whichever partition the targets belong to. With the
		 * list built early, each wave starts on it as soon as its own
		 * delivery is done, taking STDP_K chunks at a time from a shared
		 * counter (the waves that finish delivery first take more). */
		if(stdp) {
			if(early && exw) {
				/* the list is complete and its staged records have landed
				 * once every far wave has counted itself */
				while(__hip_atomic_load(listDone, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < (k + 1u) * nfar) {
					__builtin_amdgcn_s_sleep(1);
				}
			}
			/* final since the exchange barrier (or the count above) */
			const uint32_t listed = early ? nwork[t & 1u] : 0u;
			const bool late = t >= postLen && (!early || listed > NEMO_STDP_WORK_CAP);
			if(late) {
				/* the list is built here, with the targets' entries and rows
				 * read now; chunks beyond the list's capacity are done by the
				 * thread that finds them */
				const uint64_t tc = t - postLen;
				const uint32_t cb = uint32_t(tc % nd.flRing);
				__syncthreads();   /* delivery (which reads pref) done; every wave has read `listed` */
				if(wave == 0) {
					gather_counts(glb(nd.fc) + cb * P, P, uint32_t(tc), pref, lane, nd.spinLimit, nd.status, p);
				}
				if(tid == 0) nwork[t & 1u] = 0u;
				__syncthreads();
				const uint32_t Fc = pref[P];
				const NEMO_GLOBAL uint64_t* flc = glb(nd.fl) + size_t(cb) * P * T;
				const NEMO_GLOBAL uint32_t* rstart = glb(nd.rcm_start);
				const NEMO_GLOBAL uint64_t* rf = glb(nd.rfg) + size_t(par) * n_total;
				for(uint32_t j = tid; j < Fc; j += nthreads) {
					const uint32_t q = partition_of(pref, P, j);
					const NEMO_GLOBAL uint64_t* e = flc + size_t(q) * T + (j - pref[q]);
					const uint32_t tgt = fired_entry(e, ld_agent(e), uint32_t(tc), nd.spinLimit, nd.status, p, q);
					const uint32_t r0 = rstart[tgt];
					const uint32_t len = rstart[tgt + 1] - r0;
					const uint32_t nch = (len + STDP_CHUNK - 1u) / STDP_CHUNK;
					for(uint32_t ch = (p + P - j % P) % P; ch < nch; ch += P) {
						const uint32_t slot = atomicAdd(nwork + (t & 1u), 1u);
						const uint32_t r = r0 + ch * STDP_CHUNK;
						const uint32_t l = min(STDP_CHUNK, len - ch * STDP_CHUNK);
						if(slot < NEMO_STDP_WORK_CAP) {
							work[slot] = make_uint2(r, l);
						} else {
							for(uint32_t i = 0; i < l; ++i) {
								const uint32_t sd = glb(nd.rcm_sd)[r + i];
								const uint64_t pre = ld_agent(rf + (sd & 0x3ffffffu)) >> ((sd >> 26) + 1u);
								const int32_t wd = stdp_weight_change(*spl, pre);
								if(wd != 0) {
									add_agent(glb(nd.rcm_acc) + r + i, wd);
								}
							}
						}
					}
				}
				__syncthreads();
			}
			/* without the prefix ring: the counts of cycle t+1-postLen, for
			 * the next cycle's list, by the last wave (spref is no longer
			 * read in this cycle) */
			if(early && !pring && wave == nwaves - 1u && t + 1u >= postLen) {
				const uint64_t tn = t + 1u - postLen;
				gather_counts(glb(nd.fc) + uint32_t(tn % nd.flRing) * P, P, uint32_t(tn), spref, lane,
						nd.spinLimit, nd.status, p);
			}
			if(t >= postLen) {
				if(exw && wave >= exw) {
					/* the sources' histories of cycle t are complete only once
					 * every partition has published the cycle: far waves wait
					 * for the exchange waves' gather */
					while(__hip_atomic_load(exFlag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != k + 1u) {
						__builtin_amdgcn_s_sleep(1);
					}
				}
				const uint32_t nw = min(nwork[t & 1u], NEMO_STDP_WORK_CAP);
				const uint32_t ncached = late ? 0u : min(listed, nd.stdpCache);
				stdp_accumulate<IZO>(work, nw, nwork + 2u + (t & 1u), lane, glb(nd.rfg) + size_t(par) * n_total,
						glb(nd.rcm_sd), sdcache, ncached, glb(nd.rcm_acc), *spl);
			}
			if(tid == 0) {
				/* the next cycle's list and chunk counter start empty */
				nwork[(t + 1u) & 1u] = 0u;
				nwork[2u + ((t + 1u) & 1u)] = 0u;
			}
		}
		__syncthreads();
	}

	/* ---- the last cycle's deferred records */
	if(exw && nd.nsteps > 0u) {
		const uint32_t cb = uint32_t((nd.cycle0 + nd.nsteps - 1u) & 1u);
		far_phase<SAT, IZO>(nd, fcnt[cb], farA + cb * fcap, farB + cb * fcap, wave, nwaves, lane,
				uint32_t((nd.cycle0 + nd.nsteps - 1u) % D), syn, ringE, ringI, ovf, fwords);
	}
	if(defer) {
		if(kn) {
			deliver_ranges<SAT, IZO>(ka0, kb0, ka1, kb1, kn, nd.glog2, lane,
					uint32_t((nd.cycle0 + nd.nsteps - 1u) % D), syn, ringE, ringI, ovf, D, T, fwords);
		}
		__syncthreads();
	}

	/* ---- save state and the ring for the next launch */
	if(active) {
#pragma unroll
		for(int k = 0; k < NS; ++k) nd.state[size_t(k) * n_total + n] = st[k];
#pragma unroll
		for(int k = 0; k < 4; ++k) nd.rng[size_t(k) * n_total + n] = rng.s[k];
		nd.recent[n] = recent;
	}
	int32_t* gout = nd.ring + size_t(p) * 2u * D * T;
	if(SAT) {
		/* fold sticky overflow into the saved values */
		for(uint32_t i = tid; i < D * T; i += nthreads) {
			uint32_t sl = i / T, tl = i % T;
			uint32_t bit = 1u << (tl & 31u);
			int32_t e = ringE[i], ii = ringI[i];
			if(ovf[sl * fwords + (tl >> 5)] & bit) e = INT32_MAX;
			if(ovf[(D + sl) * fwords + (tl >> 5)] & bit) ii = INT32_MIN;
			gout[i] = e;
			gout[D * T + i] = ii;
		}
	} else {
		/* 16 bytes per lane and instruction (both sides 16-byte aligned) */
		const uint4* src4 = reinterpret_cast<const uint4*>(ringE);
		NEMO_GLOBAL uint4* dst4 = glb(reinterpret_cast<uint4*>(gout));
		for(uint32_t i = tid; i < ringWords / 4u; i += nthreads) {
			dst4[i] = src4[i];
		}
	}
	/* delivered-synapse count */
	for(uint32_t d = 32; d >= 1; d >>= 1) {
		events += __shfl_down(events, d, 64);
	}
	if(lane == 0 && events) {
		atomicAdd(nd.events, events);
	}
	__syncthreads();
	if(tid == 0) {
		atomicMax(reinterpret_cast<unsigned long long*>(nd.status + 9), (unsigned long long) wall_clock64());
	}
}

} // namespace

/* The step kernel in four variants: saturating accumulation or not, and
 * Izhikevich-only networks or any built-in model. Each variant is its own
 * kernel so that each gets its own register allocation; the host launches
 * the variant of the network (a batched launch packs networks of one
 * variant). */
template<bool SAT, bool IZO>
__global__ void __launch_bounds__(1024)
nemo_step_kernel(const NetDev* __restrict__ nets, const WorkItem* __restrict__ work)
{
	const WorkItem wi = work[blockIdx.x];
	const NetDev nd = nets[wi.net];
	run_partition<SAT, IZO>(nd, wi.part);
}

/* One network per launch: the descriptor travels as a kernel argument, so a
 * launch needs no copy of it to device memory first. */
template<bool SAT, bool IZO>
__global__ void __launch_bounds__(1024)
nemo_step_kernel_one(const NetDev nd, const WorkItem* __restrict__ work)
{
	run_partition<SAT, IZO>(nd, work[blockIdx.x].part);
}

static const void* step_kernel_variant(uint32_t variant)
{
	switch(variant & 3u) {
	case 0: return reinterpret_cast<const void*>(&nemo_step_kernel<false, false>);
	case 1: return reinterpret_cast<const void*>(&nemo_step_kernel<true, false>);
	case 2: return reinterpret_cast<const void*>(&nemo_step_kernel<false, true>);
	default: return reinterpret_cast<const void*>(&nemo_step_kernel<true, true>);
	}
}

static const void* step_kernel_one_variant(uint32_t variant)
{
	switch(variant & 3u) {
	case 0: return reinterpret_cast<const void*>(&nemo_step_kernel_one<false, false>);
	case 1: return reinterpret_cast<const void*>(&nemo_step_kernel_one<true, false>);
	case 2: return reinterpret_cast<const void*>(&nemo_step_kernel_one<false, true>);
	default: return reinterpret_cast<const void*>(&nemo_step_kernel_one<true, true>);
	}
}

/* The dynamic LDS limit of a kernel is raised once to the largest size
 * asked for (the attribute call is a host-side round trip per launch
 * otherwise). Launches are serialised by the host library's device lock. */
static hipError_t ensure_lds_limit(const void* kern, uint32_t ldsBytes)
{
	static const void* kerns[16] = {};
	static uint32_t limits[16] = {};
	for(int i = 0; i < 16; ++i) {
		if(kerns[i] == kern) {
			if(limits[i] >= ldsBytes) return hipSuccess;
			const hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(ldsBytes));
			if(e == hipSuccess) limits[i] = ldsBytes;
			return e;
		}
		if(kerns[i] == nullptr) {
			const hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(ldsBytes));
			if(e == hipSuccess) {
				kerns[i] = kern;
				limits[i] = ldsBytes;
			}
			return e;
		}
	}
	return hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(ldsBytes));
}

/* Batched launch, first node: clear the first 8 words of each network's
 * control block (the step kernel only ORs and adds into them). */
__global__ void nemo_batch_clear_kernel(const NetDev* __restrict__ nets, uint32_t count)
{
	for(uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
		uint32_t* ctl = nets[i].status - 1;
		for(uint32_t w = 0; w < 8u; ++w) ctl[w] = 0u;
		ctl[8] = ctl[9] = ~0u;   /* first start: none yet */
		ctl[10] = ctl[11] = 0u;  /* last end */
	}
}

/* Batched launch, last node: gather the first 12 words of each network's
 * control block (status, events, diagnostics, first start and last end)
 * into out[16*i .. 16*i+11], so the host reads all of them with one copy. */
__global__ void nemo_batch_collect_kernel(const NetDev* __restrict__ nets, uint32_t count, uint32_t* __restrict__ out)
{
	for(uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
		const uint32_t* ctl = nets[i].status - 1;
		for(uint32_t w = 0; w < 12u; ++w) out[16 * i + w] = ctl[w];
	}
}

/* applyStdp: src/nemo/ConnectivityMatrix.cpp:241-294. With reward 0 every
 * accumulator contributes fx_mul(0, acc) = 0, the same as the reference's
 * clear-then-apply. */
__global__ void nemo_apply_stdp_kernel(uint32_t* syn, const uint32_t* __restrict__ rcm_fwd,
		int32_t* __restrict__ rcm_acc, size_t nr, const StdpParams* __restrict__ sp, int32_t fx_reward)
{
	for(size_t r = size_t(blockIdx.x) * blockDim.x + threadIdx.x; r < nr; r += size_t(gridDim.x) * blockDim.x) {
		const size_t f = rcm_fwd[r];
		const int32_t w_old = int32_t(syn[2 * f + 1]);
		const int32_t w_new = stdp_updated_weight(*sp, w_old, fx_mul20(fx_reward, rcm_acc[r]));
		if(w_new != w_old) {
			syn[2 * f + 1] = uint32_t(w_new);
		}
		rcm_acc[r] = 0;
	}
}

__global__ void nemo_fill_u32_kernel(uint32_t* p, uint32_t v, size_t n)
{
	for(size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
		p[i] = v;
	}
}

__global__ void nemo_fill_u64_kernel(uint64_t* p, uint64_t v, size_t n)
{
	for(size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
		p[i] = v;
	}
}

__global__ void nemo_eval_libm_kernel(int kind, const float* in, float* out, size_t n)
{
	for(size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
		out[i] = (kind == 0) ? nemo_libm::logf_glibc(in[i])
			: (kind == 1) ? nemo_libm::sinf_rng(in[i])
			: (kind == 2) ? nemo_libm::kuramoto_wrap(in[i]) : nemo_libm::sinf_glibc(in[i]);
	}
}

__global__ void nemo_nrand_seq_kernel(const uint32_t* state4, float* out, size_t n)
{
	if(blockIdx.x == 0 && threadIdx.x == 0) {
		Rng r;
		for(int k = 0; k < 4; ++k) r.s[k] = state4[k];
		for(size_t i = 0; i < n; ++i) out[i] = nrand(r);
	}
}

/* Scratch memory of the allocation-free run (nemo_hip_run_static): the
 * stimulus schedule of a run is staged here instead of in buffers the
 * library allocates. Each device holds its own copy of the array for the
 * life of the process; host code reaches it only through the address
 * hipGetSymbolAddress returns on that device (the array's name is not a
 * device address on the host). Calls on one device share it. */
static __device__ __align__(256) uint32_t scratch[NEMO_STATIC_SCRATCH_WORDS];

static unsigned grid_for(size_t n, unsigned block)
{
	size_t g = (n + block - 1) / block;
	if(g < 1) g = 1;
	if(g > 4096) g = 4096;
	return unsigned(g);
}

extern "C" {

int nemo_dev_step_launch(const NetDev* nets_dev, const WorkItem* work_dev, uint32_t nblocks,
		uint32_t blockDim, uint32_t ldsBytes, uint32_t variant, void* stream)
{
	const void* kern = step_kernel_variant(variant);
	hipError_t e = ensure_lds_limit(kern, ldsBytes);
	if(e != hipSuccess) return int(e);
	void* args[] = { const_cast<NetDev**>(&nets_dev), const_cast<WorkItem**>(&work_dev) };
	e = hipLaunchKernel(kern, dim3(nblocks), dim3(blockDim), args, ldsBytes, static_cast<hipStream_t>(stream));
	return int(e);
}

int nemo_dev_step_launch_one(const NetDev* nd_host, const WorkItem* work_dev, uint32_t nblocks,
		uint32_t blockDim, uint32_t ldsBytes, uint32_t variant, void* stream)
{
	const void* kern = step_kernel_one_variant(variant);
	hipError_t e = ensure_lds_limit(kern, ldsBytes);
	if(e != hipSuccess) return int(e);
	void* args[] = { const_cast<NetDev*>(nd_host), const_cast<WorkItem**>(&work_dev) };
	e = hipLaunchKernel(kern, dim3(nblocks), dim3(blockDim), args, ldsBytes, static_cast<hipStream_t>(stream));
	return int(e);
}

/* The batched step as a HIP graph: clear -> step -> collect.
 * Everything a replay needs is read from the device tables nets_dev and
 * work_dev, so one instantiated graph serves every launch of the same
 * work table; the host rewrites nets_dev before each replay. */
int nemo_dev_batch_graph_create(const NetDev* nets_dev, uint32_t nnets, const WorkItem* work_dev,
		uint32_t nblocks, uint32_t blockDim, uint32_t ldsBytes, uint32_t variant, uint32_t* collect_dev,
		void** exec_out)
{
	*exec_out = nullptr;
	const void* kern = step_kernel_variant(variant);
	hipError_t e = ensure_lds_limit(kern, ldsBytes);
	if(e != hipSuccess) return int(e);
	hipGraph_t g = nullptr;
	e = hipGraphCreate(&g, 0);
	if(e != hipSuccess) return int(e);
	const uint32_t small = (nnets + 255u) / 256u;

	void* clearArgs[] = { const_cast<NetDev**>(&nets_dev), &nnets };
	hipKernelNodeParams kp = {};
	kp.func = reinterpret_cast<void*>(&nemo_batch_clear_kernel);
	kp.gridDim = dim3(small);
	kp.blockDim = dim3(256);
	kp.sharedMemBytes = 0;
	kp.kernelParams = clearArgs;
	kp.extra = nullptr;
	hipGraphNode_t nClear = nullptr, nStep = nullptr, nCollect = nullptr;
	e = hipGraphAddKernelNode(&nClear, g, nullptr, 0, &kp);

	void* stepArgs[] = { const_cast<NetDev**>(&nets_dev), const_cast<WorkItem**>(&work_dev) };
	if(e == hipSuccess) {
		kp.func = const_cast<void*>(kern);
		kp.gridDim = dim3(nblocks);
		kp.blockDim = dim3(blockDim);
		kp.sharedMemBytes = ldsBytes;
		kp.kernelParams = stepArgs;
		e = hipGraphAddKernelNode(&nStep, g, &nClear, 1, &kp);
	}
	void* collectArgs[] = { const_cast<NetDev**>(&nets_dev), &nnets, &collect_dev };
	if(e == hipSuccess) {
		kp.func = reinterpret_cast<void*>(&nemo_batch_collect_kernel);
		kp.gridDim = dim3(small);
		kp.blockDim = dim3(256);
		kp.sharedMemBytes = 0;
		kp.kernelParams = collectArgs;
		e = hipGraphAddKernelNode(&nCollect, g, &nStep, 1, &kp);
	}
	hipGraphExec_t ex = nullptr;
	if(e == hipSuccess) {
		e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
	}
	(void) hipGraphDestroy(g);
	if(e == hipSuccess) *exec_out = ex;
	return int(e);
}

int nemo_dev_graph_launch(void* exec, void* stream)
{
	return int(hipGraphLaunch(static_cast<hipGraphExec_t>(exec), static_cast<hipStream_t>(stream)));
}

int nemo_dev_graph_destroy(void* exec)
{
	return exec ? int(hipGraphExecDestroy(static_cast<hipGraphExec_t>(exec))) : 0;
}

int nemo_dev_step_max_blocks(uint32_t blockDim, uint32_t ldsBytes, uint32_t variant, uint32_t* maxBlocks)
{
	const void* kern = step_kernel_variant(variant);
	hipError_t e = ensure_lds_limit(kern, ldsBytes);
	if(e != hipSuccess) return int(e);
	int perCU = 0;
	e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kern, int(blockDim), size_t(ldsBytes));
	if(e != hipSuccess) return int(e);
	int dev = 0;
	e = hipGetDevice(&dev);
	if(e != hipSuccess) return int(e);
	int cus = 0;
	e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
	if(e != hipSuccess) return int(e);
	/* Co-residency measured on gfx950 (workgroups that count themselves in
	 * and wait, with a time bound, for the rest) falls short of the occupancy
	 * query when LDS is the limit: each workgroup takes its LDS rounded up to
	 * 512 bytes plus 512 more (32536 B: 4 per CU, not 5; 16000 B: 9, not 10;
	 * 8000 B: 18, not 20). A persistent launch needs every workgroup
	 * resident, so the count follows the measurement. This bound is what
	 * keeps launches co-resident: they are ordinary launches (a cooperative
	 * launch checks against the occupancy query, which over-reports here,
	 * and costs about 15 us more per launch), and every wait for another
	 * workgroup is bounded, so a launch that did not fit would end with an
	 * error, not hang. */
	int ldsPerCU = 0;
	e = hipDeviceGetAttribute(&ldsPerCU, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev);
	if(e != hipSuccess) return int(e);
	if(ldsBytes > 0u && ldsPerCU > 0) {
		const uint32_t alloc = ((ldsBytes + 511u) & ~511u) + 512u;
		perCU = std::min(perCU, int(uint32_t(ldsPerCU) / alloc));
	}
	*maxBlocks = uint32_t(perCU) * uint32_t(cus);
	return 0;
}

int nemo_dev_static_scratch(void** ptr)
{
	static void* addr[64] = {};
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if(e != hipSuccess) return int(e);
	if(dev < 0 || dev >= 64) return int(hipErrorInvalidDevice);
	if(!addr[dev]) {
		void* p = nullptr;
		e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if(e != hipSuccess) return int(e);
		addr[dev] = p;
	}
	*ptr = addr[dev];
	return 0;
}

int nemo_dev_apply_stdp(uint32_t* syn, const uint32_t* rcm_fwd, int32_t* rcm_acc, size_t nr,
		const void* stdp_dev, int32_t fx_reward, void* stream)
{
	if(nr == 0) return 0;
	hipLaunchKernelGGL(nemo_apply_stdp_kernel, dim3(grid_for(nr, 256)), dim3(256), 0,
			static_cast<hipStream_t>(stream), syn, rcm_fwd, rcm_acc, nr,
			static_cast<const StdpParams*>(stdp_dev), fx_reward);
	return int(hipGetLastError());
}

int nemo_dev_fill_u32(uint32_t* p, uint32_t value, size_t n, void* stream)
{
	if(n == 0) return 0;
	hipLaunchKernelGGL(nemo_fill_u32_kernel, dim3(grid_for(n, 256)), dim3(256), 0,
			static_cast<hipStream_t>(stream), p, value, n);
	return int(hipGetLastError());
}

int nemo_dev_fill_u64(uint64_t* p, uint64_t value, size_t n, void* stream)
{
	if(n == 0) return 0;
	hipLaunchKernelGGL(nemo_fill_u64_kernel, dim3(grid_for(n, 256)), dim3(256), 0,
			static_cast<hipStream_t>(stream), p, value, n);
	return int(hipGetLastError());
}

int nemo_dev_eval_libm(int kind, const float* in_dev, float* out_dev, size_t n, void* stream)
{
	if(n == 0) return 0;
	hipLaunchKernelGGL(nemo_eval_libm_kernel, dim3(grid_for(n, 256)), dim3(256), 0,
			static_cast<hipStream_t>(stream), kind, in_dev, out_dev, n);
	return int(hipGetLastError());
}

int nemo_dev_nrand_seq(const uint32_t* state4_dev, float* out_dev, size_t n, void* stream)
{
	hipLaunchKernelGGL(nemo_nrand_seq_kernel, dim3(1), dim3(64), 0,
			static_cast<hipStream_t>(stream), state4_dev, out_dev, n);
	return int(hipGetLastError());
}

} // extern "C"
