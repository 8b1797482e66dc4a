#ifndef NEMO_HIP_H
#define NEMO_HIP_H

/*
 * nemo_hip.h -- extensions of the NeMo C API specific to the HIP backend.
 *
 * NeMo's API constructs networks one synapse per call and steps one cycle
 * per call with a host round trip. The functions here add the forms that a
 * GPU needs to be fed at full rate: bulk construction from arrays, runs of
 * many cycles that stay on the device, and batched runs of many independent
 * networks in one persistent launch. None of them changes the model
 * arithmetic: a run of K cycles gives the same rasters, state and weights as
 * K calls of nemo_step().
 */

#include <stddef.h>
#include <stdint.h>
#include "nemo.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ construction */

/* Bulk form of nemo_add_neuron (src/nemo.h:265-268): neuron i has index
 * idx[i] and arguments args[i*nargs .. i*nargs+nargs). */
nemo_status_t nemo_add_neurons(nemo_network_t net, unsigned type, size_t n,
		const unsigned idx[], unsigned nargs, const float args[]);

/* Bulk form of nemo_add_synapse (src/nemo.h:278-286) for whole axons:
 * source k (sources[k], or k when sources is NULL) gets synapses
 * rowptr[k] .. rowptr[k+1]-1 of the arrays, in that order, so its synapse
 * ids are (source << 32) | (previous count + position). plastic may be NULL
 * (all static). */
nemo_status_t nemo_add_synapses_csr(nemo_network_t net, size_t nsources,
		const unsigned sources[], const uint64_t rowptr[],
		const unsigned targets[], const unsigned char delays[],
		const float weights[], const unsigned char plastic[]);

/* ----------------------------------------------------------- configuration */

/* Same as nemo_set_cuda_backend: run on HIP device dev (dev < 0: device 0). */
nemo_status_t nemo_set_hip_backend(nemo_configuration_t conf, int dev);

/* Neurons per partition (workgroup) of the step kernel, rounded up to a
 * multiple of 64; 0 = automatic. Counterpart of NeMo's
 * Configuration::setCudaPartitionSize (src/nemo/Configuration.hpp:61). */
nemo_status_t nemo_set_partition_size(nemo_configuration_t conf, unsigned neurons);

/* Lanes of a wavefront that share one fired source when its synapses are
 * delivered (the delivery queue's bin: 1, 2, 4, ..., 64); 0 = automatic
 * (about 1.5x the mean synapses per source and partition). No counterpart
 * in NeMo; a tuning knob of the fused step kernel for the autotuner. */
nemo_status_t nemo_set_delivery_lanes(nemo_configuration_t conf, unsigned lanes);

/* STDP work lists in LDS (default on): with it the fused kernel keeps the
 * fired-count prefixes of the last postLen cycles and the reverse records of
 * each cycle's first STDP chunks in LDS, as far as LDS allows; off, it reads
 * both from device memory every cycle. Same results either way. No
 * counterpart in NeMo; a tuning knob of the fused step kernel. */
nemo_status_t nemo_set_stdp_lds_staging(nemo_configuration_t conf, int enable);

/* Synapse structures (forward matrix, reverse matrix, query tables) built
 * on the device (default on) or on the host. Same results either way; a
 * network with an invalid synapse is always checked on the host, which
 * raises NeMo's error. No counterpart in NeMo, whose construction is on the
 * host (src/nemo/ConnectivityMatrix.cpp:98-161). */
nemo_status_t nemo_set_device_construction(nemo_configuration_t conf, int enable);

/* Waves of each step-kernel workgroup that run the exchange between
 * partitions and the delivery of delay-1 records, while the others deliver
 * the previous cycle's records of longer delays (and, with STDP, build the
 * work lists and accumulate): n > 0 = n waves (at least one other must
 * remain; needs deferred delivery and LDS room for the far list), < 0 =
 * none (every wave does both, one after the other), 0 (default) =
 * automatic: 3 of every 8 waves for networks with STDP and at least 8
 * waves (512 neurons) per workgroup, none otherwise. Same results either
 * way. No counterpart in NeMo; a tuning knob of the fused step kernel. */
nemo_status_t nemo_set_exchange_waves(nemo_configuration_t conf, int waves);

/* Which step implementation runs a simulation:
 *   NEMO_STEP_AUTO    the fused persistent kernel when every neuron type is a
 *                     built-in model and the network fits it; else plugins
 *   NEMO_STEP_FUSED   the fused kernel, or fail at nemo_new_simulation
 *   NEMO_STEP_PLUGINS one kernel sequence per cycle around each type's
 *                     plugin (lib<Model>.so, include/nemo_hip_plugin.h)
 * Both give identical results. */
enum { NEMO_STEP_AUTO = 0, NEMO_STEP_FUSED = 1, NEMO_STEP_PLUGINS = 2 };
nemo_status_t nemo_set_step_mode(nemo_configuration_t conf, int mode);

/* ------------------------------------------------ accessors for nemo.hpp */

/* Accessors behind the C++ surface (include/nemo.hpp) that the C API of
 * src/nemo.h lacks. */
/* A copy of conf (nemo::Configuration's copy constructor,
 * src/nemo/Configuration.hpp:49); NULL on failure. */
nemo_configuration_t nemo_copy_configuration(nemo_configuration_t conf);
/* Configuration::disableLogging / loggingEnabled (src/nemo/Configuration.hpp:58-59) */
nemo_status_t nemo_disable_logging(nemo_configuration_t conf);
nemo_status_t nemo_logging_enabled(nemo_configuration_t conf, int* enabled);
/* Configuration::cudaPartitionSize (src/nemo/Configuration.hpp:62) */
nemo_status_t nemo_partition_size(nemo_configuration_t conf, unsigned* neurons);
/* Configuration::writeOnlySynapses (src/nemo/Configuration.hpp:106) */
nemo_status_t nemo_write_only_synapses(nemo_configuration_t conf, int* enabled);
/* Network::maxDelay / minWeight / maxWeight (src/nemo/Network.hpp:202-205);
 * the weights' range starts at 0 (src/nemo/NetworkImpl.cpp:28-30). */
nemo_status_t nemo_max_delay(nemo_network_t net, unsigned* delay);
nemo_status_t nemo_min_weight(nemo_network_t net, float* weight);
nemo_status_t nemo_max_weight(nemo_network_t net, float* weight);

/* ------------------------------------------------------- multi-cycle runs */

/*
 * Stimulus of a run, as CSR over the run's cycles: cycle k forces the
 * neurons fstim_idx[fstim_ptr[k] .. fstim_ptr[k+1]) to fire and assigns
 * istim_val[j] to the external current of neuron istim_idx[j] for
 * j in istim_ptr[k] .. istim_ptr[k+1] (the last assignment of a neuron in a
 * cycle wins, as nemo_step's addCurrentStimulus). Indices are user neuron
 * indices. Any of the pointer pairs may be NULL.
 */
typedef struct nemo_schedule {
	const unsigned* fstim_ptr;
	const unsigned* fstim_idx;
	const unsigned* istim_ptr;
	const unsigned* istim_idx;
	const float* istim_val;
} nemo_schedule_t;

/* Run nsteps cycles on the device. The firing raster of the run stays on
 * the device; read it with nemo_hip_read_fired / nemo_hip_read_raster.
 * events (may be NULL) receives the number of synapses delivered. */
nemo_status_t nemo_hip_run(nemo_simulation_t sim, unsigned nsteps,
		const nemo_schedule_t* schedule, unsigned long long* events);

/* No counterpart in NeMo: nemo_hip_run for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). It allocates nothing: the raster of the
 * run goes to the caller's device buffer raster_dev (nsteps x words_per_step
 * 64-bit words, words_per_step from nemo_hip_read_raster; layout as
 * there), and the stimulus schedule is staged in a fixed per-device scratch
 * area of NEMO_HIP_STATIC_SCRATCH_BYTES: a schedule that needs more 4-byte
 * words than fit there -- 2 (nsteps + 1) + forced firings + 2 x current
 * assignments -- is refused with NEMO_INVALID_INPUT before anything runs. Every
 * call on a device shares that device's scratch area, so calls that overlap
 * on one device race. Fused step kernel only (NEMO_API_UNSUPPORTED
 * otherwise). nemo_hip_read_fired / nemo_hip_read_raster refuse to read
 * such a run; events (may be NULL) as nemo_hip_run. */
#define NEMO_HIP_STATIC_SCRATCH_BYTES (4u << 20)
nemo_status_t nemo_hip_run_static(nemo_simulation_t sim, unsigned nsteps,
		const nemo_schedule_t* schedule, uint64_t* raster_dev, unsigned long long* events);

/* Fired neurons of the last run: counts[k] for each of its cycles, and the
 * user indices of all fired neurons back to back (cycle order, within a
 * cycle in the order nemo_step reports them). ids may be NULL; cap is its
 * capacity; *total receives the number of ids. */
nemo_status_t nemo_hip_read_fired(nemo_simulation_t sim, unsigned counts[],
		unsigned ids[], size_t cap, size_t* total);

/* Raw raster of the last run: words_per_step 64-bit words per cycle; bit b
 * of word w is simulation-local neuron 64*w+b. */
nemo_status_t nemo_hip_read_raster(nemo_simulation_t sim, uint64_t words[], size_t cap,
		size_t* words_per_step);

/* ------------------------------------------------------------------ batches */

/*
 * Run nsteps[i] cycles of each of n independent simulations in persistent
 * launches that hold as many of them on the device at once as fit (one
 * launch if all fit). Each simulation keeps its own state, RNG streams,
 * stimulus-free dynamics and raster, identical to running it alone;
 * events[i] (may be NULL) receives each one's delivered synapses.
 */
nemo_status_t nemo_hip_run_batch(nemo_simulation_t sims[], size_t n,
		const unsigned nsteps[], unsigned long long events[]);

/* ------------------------------------------------------------------- info */

typedef struct nemo_hip_info {
	unsigned neurons;
	unsigned partitions;         /* workgroups of the step kernel */
	unsigned partition_size;     /* neurons per partition */
	unsigned max_delay;
	unsigned saturating;         /* 1 if accumulators may leave int32 */
	unsigned stdp;
	unsigned long long synapses;
	unsigned long long plastic_synapses;
	unsigned long long device_bytes;
	unsigned lds_bytes;
	unsigned block_threads;
	float last_kernel_ms;        /* step-kernel time of the last run: HIP events for its own launches;
	                              * in a batched run, its workgroups' first start to last end */
	unsigned fused;              /* 1: fused persistent step kernel; 0: plugin step */
	unsigned deferred;           /* 1: records of delay >= 2 delivered one cycle late */
	unsigned lanes_per_source;   /* lanes sharing one fired source in delivery */
	unsigned stdp_staged_chunks; /* STDP chunks whose reverse records are staged in LDS per cycle */
	unsigned stdp_prefix_ring;   /* 1: fired-count prefixes of the last postLen cycles kept in LDS */
	unsigned device_built;       /* 1: synapse structures were built on the device */
	unsigned exchange_waves;     /* waves per workgroup running the exchange and delay-1 delivery */
	unsigned far_list_sources;   /* fired sources per cycle whose far records wait in the LDS far list */
} nemo_hip_info_t;

nemo_status_t nemo_hip_get_info(nemo_simulation_t sim, nemo_hip_info_t* info);

/* Status of the most recent API call (nemo_new_simulation returns NULL on
 * failure; this gives its status code). */
nemo_status_t nemo_hip_last_status(void);

/* nemo_apply_stdp(sims[i], rewards[i]) for i < n, under one hold of the
 * device lock, the kernels of all simulations queued before any is waited
 * for (a serving loop applies STDP to many simulations between chunks).
 * Each simulation's result equals its own nemo_apply_stdp. On an error
 * (a simulation without STDP, a reward that rounds to zero) the
 * simulations before the failing one have been updated, the others not.
 * No counterpart in NeMo (the batched form of nemo_apply_stdp,
 * src/nemo.h:133). */
nemo_status_t nemo_hip_apply_stdp_batch(nemo_simulation_t sims[], size_t n, const float rewards[]);

/* Evaluate the device's math on n host floats, for checking it against the
 * host C library: kind 0 = logf and 1 = sinf as the RNG uses them (for the
 * RNG's arguments), 2 = the Kuramoto phase wrap (fmodf by 2*pi, plus 2*pi in
 * double for a negative phase), 3 = sinf for any argument (the Kuramoto
 * model's). With on_device = 0 the same code runs on the host CPU (no GPU
 * needed). */
nemo_status_t nemo_hip_eval_libm(int kind, const float in[], float out[], size_t n, int on_device);

/* The Gaussian sequence nrand() yields from RNG state state4, computed on
 * the device. */
nemo_status_t nemo_hip_nrand_seq(const uint32_t state4[4], float out[], size_t n);

#ifdef __cplusplus
}
#endif

#endif
