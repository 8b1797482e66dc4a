/*
 * step_abi.h -- plain-old-data structures shared by the C++ host code and
 * the HIP kernels (the internal C-ABI between them). No HIP types appear
 * here, so the host side compiles against it with any C++ compiler.
 */
#ifndef NEMOSIM_STEP_ABI_H
#define NEMOSIM_STEP_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Bits of NetDev.flags */
#define NEMO_NET_STDP      1u   /* accumulate STDP this launch */
#define NEMO_NET_SATURATE  2u   /* accumulators may leave the int32 range */
#define NEMO_NET_REPLAY    4u   /* weights changed since the ring was filled:
                                 * rebuild it from the firing history */
#define NEMO_NET_IZHIKEVICH_ONLY 16u  /* every neuron is an Izhikevich neuron */
#define NEMO_NET_DEFER     8u   /* delay-1 records first in each bucket, their
                                 * count in bits 28..31 of off[]; records of
                                 * delay >= 2 are delivered one cycle later */
#define NEMO_NET_PREF_RING 32u  /* STDP: prefixes of the fired counts of the last
                                 * postLen cycles kept in LDS */

/* Bits of the per-network status word */
#define NEMO_STATUS_BARRIER_TIMEOUT 1u   /* a fired count never arrived */
#define NEMO_STATUS_ENTRY_TIMEOUT   2u   /* a fired-list entry never arrived */

/* The per-network control block (16 words): [0] workgroups started, [1]
 * status, [2..3] delivered-synapse count, and, for the first timeout,
 * [4] cycle, [5] waiting partition, [6] awaited partition, [7] workgroups
 * started by then; [8..9] the earliest start and [10..11] the latest end
 * of its workgroups on the device wall clock (read after batched
 * launches, whose graph initialises them). */
#define NEMO_CTL_WORDS 16u

/*
 * One simulated network as seen by the step kernel. Neurons use
 * simulation-local indices 0..n-1; partition p owns neurons [p*T, p*T+T).
 * All arrays are device pointers.
 */
typedef struct NetDev {
	uint32_t n;          /* neurons */
	uint32_t P;          /* partitions (workgroups) */
	uint32_t T;          /* neurons per partition, multiple of 64, <= 1024 */
	uint32_t D;          /* delay-ring depth = maximum delay (1..64) */
	uint32_t words;      /* 64-bit raster words per step = P*T/64 */
	uint32_t flags;      /* NEMO_NET_* */
	uint32_t nsteps;     /* steps this network runs in the current launch */
	uint32_t spinLimit;  /* barrier spin bound (iterations) */
	uint32_t glog2;      /* log2 of the lanes that share one fired source in delivery (0..6) */
	uint32_t flRing;     /* fired lists and counts kept for this many cycles (buffer t mod flRing) */
	uint32_t stdpCache;  /* STDP chunks whose reverse records are staged in LDS (64 words each) */
	uint32_t exWaves;    /* with NEMO_NET_DEFER: waves of a workgroup that run the exchange and the
	                      * delay-1 delivery while the others deliver the previous cycle's far
	                      * records (0: every wave does both, one after the other) */
	uint32_t farCap;     /* fired sources per cycle whose far ranges fit the LDS far list (multiple
	                      * of 64; with exWaves > 0) */
	uint64_t cycle0;     /* cycle count at the start of the launch */

	const uint8_t* model;     /* [n] model id */
	float* param;             /* [MAX_PARAM][n] */
	float* state;             /* [MAX_STATE][n] */
	uint32_t* rng;            /* [4][n] */
	uint64_t* recent;         /* [n] firing history, bit 0 = last cycle */
	int32_t* ring;            /* [P][2][D][T] pending current (saved between launches) */

	/* forward matrix: synapses of source s bucketed by target partition;
	 * partition p's bucket is syn[off[s*(P+1)+p] .. off[s*(P+1)+p+1]).
	 * Record = {target - p*T | delay << 16, Q11.20 weight}. */
	const uint32_t* off;
	uint32_t* syn;            /* [2 * synapses] */

	/* per-step exchange between the network's workgroups, double-buffered
	 * by cycle parity */
	uint64_t* fl;             /* [flRing][P][T] fired neurons of each partition, (cycle << 32) | neuron */
	uint64_t* fc;             /* [flRing][P] (cycle << 32) | fired count; starts all-ones */
	uint64_t* rfg;            /* [2][n] firing history (STDP only) */
	uint32_t* status;         /* NEMO_STATUS_* */
	unsigned long long* events; /* delivered-synapse counter */

	/* reverse matrix of plastic synapses (STDP) */
	const uint32_t* rcm_start;   /* [n+1] */
	const uint32_t* rcm_sd;      /* source | (delay-1) << 26 */
	const uint32_t* rcm_fwd;     /* index of the synapse record */
	int32_t* rcm_acc;            /* accumulated weight change */
	const void* stdp;            /* StdpParams */
	uint64_t* srow;              /* [flRing][P][T] reverse row {start, length} of each fired
	                              * neuron, at its fired-list position (STDP only) */

	/* stimulus of this launch, CSR over steps (sim-local indices); NULL if none */
	const uint32_t* fs_ptr;
	const uint32_t* fs_idx;
	const uint32_t* is_ptr;
	const uint32_t* is_idx;
	const float* is_val;

	uint64_t* raster;         /* [nsteps][words] fired bits */
} NetDev;

/* blockIdx -> (network, partition) */
typedef struct WorkItem {
	uint32_t net;
	uint32_t part;
} WorkItem;

/* Kernel launchers (nemosim_amd/csrc/device/step_kernel.hip). They return a
 * hipError_t value as int and take a hipStream_t as void*. */
/* Step-kernel variants: bit 0 = saturating accumulation (NEMO_NET_SATURATE),
 * bit 1 = Izhikevich-only network (NEMO_NET_IZHIKEVICH_ONLY). All networks
 * of one launch must be of the launch's variant. */
static inline uint32_t nemo_step_variant(uint32_t flags)
{
	return ((flags & NEMO_NET_SATURATE) ? 1u : 0u) | ((flags & NEMO_NET_IZHIKEVICH_ONLY) ? 2u : 0u);
}
int nemo_dev_step_launch(const NetDev* nets_dev, const WorkItem* work_dev, uint32_t nblocks,
		uint32_t blockDim, uint32_t ldsBytes, uint32_t variant, void* stream);
/* One network: the descriptor is passed by value (nd_host is read at the
 * call), the work table maps blocks to its partitions. */
int nemo_dev_step_launch_one(const NetDev* nd_host, const WorkItem* work_dev, uint32_t nblocks,
		uint32_t blockDim, uint32_t ldsBytes, uint32_t variant, void* stream);
int nemo_dev_step_max_blocks(uint32_t blockDim, uint32_t ldsBytes, uint32_t variant, uint32_t* maxBlocks);
/* Batched step as an instantiated HIP graph (clear -> persistent step ->
 * collect); collect_dev receives [nnets][16] words: the first 12 words of
 * each network's control block, the last four of them the first start and
 * the last end of its workgroups on the device wall clock (64-bit each; the
 * graph initialises them before the step kernel). The graph reads
 * nets_dev/work_dev at replay time. */
int nemo_dev_batch_graph_create(const NetDev* nets_dev, uint32_t nnets, const WorkItem* work_dev,
		uint32_t nblocks, uint32_t blockDim, uint32_t ldsBytes, uint32_t variant, uint32_t* collect_dev,
		void** exec_out);
int nemo_dev_graph_launch(void* exec, void* stream);
int nemo_dev_graph_destroy(void* exec);
/* The allocation-free run's scratch array: 4-byte words, one array per
 * device. nemo_dev_static_scratch gives its device address on the current
 * device (looked up once per device with hipGetSymbolAddress). The caller
 * holds the library's device lock. */
#define NEMO_STATIC_SCRATCH_WORDS (1u << 20)
int nemo_dev_static_scratch(void** ptr);
int nemo_dev_apply_stdp(uint32_t* syn, const uint32_t* rcm_fwd, int32_t* rcm_acc, size_t nr,
		const void* stdp_dev, int32_t fx_reward, void* stream);
int nemo_dev_fill_u32(uint32_t* p, uint32_t value, size_t n, void* stream);
int nemo_dev_fill_u64(uint64_t* p, uint64_t value, size_t n, void* stream);
/* Evaluates the device math pieces on inputs in[] (for tests):
 * kind 0 = logf_glibc(in), 1 = sinf_rng(in), 2 = kuramoto_wrap(in), 3 = sinf_glibc(in) */
int nemo_dev_eval_libm(int kind, const float* in_dev, float* out_dev, size_t n, void* stream);
int nemo_dev_nrand_seq(const uint32_t* state4_dev, float* out_dev, size_t n, void* stream);

/* Bytes of LDS the STDP tables take (the StdpParams of models.h, rounded) */
#define NEMO_STDP_LDS_BYTES 576u
/* Entries of a workgroup's per-cycle STDP work list (64-synapse chunks) */
#define NEMO_STDP_WORK_CAP 1024u

/* LDS bytes the step kernel needs for partitions of T neurons, ring depth D,
 * P partitions. Optional STDP areas follow, in this order, when the host
 * finds room for them: stdpCache x 256 bytes of staged reverse records, then
 * with NEMO_NET_PREF_RING postLen x ((P+1) rounded up to 64) x 4 bytes.
 * With exchange waves the far list follows: NEMO_FAR_LDS_BYTES(farCap). */
/* far list: [2][farCap] range starts, [2][farCap] range ends, two counts
 * and the exchange waves' flag */
#define NEMO_FAR_LDS_BYTES(cap) (16u * (cap) + 16u)

static inline uint32_t nemo_step_lds_bytes(uint32_t T, uint32_t D, uint32_t P, int saturate, int stdp)
{
	uint32_t b = 0;
	b += 2u * D * T * 4u;          /* ringE, ringI */
	b += T * 4u;                   /* external current / STDP target list */
	b += (T / 32u) * 4u;           /* forced-firing bits */
	b += ((P + 1u + 63u) & ~63u) * 4u; /* partition prefix */
	b += 64u * 4u;                 /* wave sums / scratch */
	if(saturate) {
		b += 2u * D * (T / 32u) * 4u;  /* overflow bits */
	}
	if(stdp) {
		b += NEMO_STDP_LDS_BYTES;      /* STDP tables */
		b += NEMO_STDP_WORK_CAP * 8u + 64u;  /* STDP work list and its counters */
		b += ((P + 1u + 63u) & ~63u) * 4u;   /* prefix of the counts of the targets' cycle */
	}
	return b;
}

#ifdef __cplusplus
}
#endif

#endif
