#ifndef NEMOSIM_HOST_SIMULATION_HPP
#define NEMOSIM_HOST_SIMULATION_HPP

#include <chrono>
#include <string>
#include <utility>
#include <vector>

#include "common.hpp"
#include "network.hpp"
#include "../device/models.h"
#include "../device/step_abi.h"
#include "../device/plugin_abi.h"
#include "../device/build_abi.h"
#include "nemo_hip.h"
#include "nemo_hip_plugin.h"

namespace nemo {
namespace hip {

/* Configuration (src/nemo/ConfigurationImpl.hpp) */
struct ConfigurationImpl {
	bool stdp = false;
	std::vector<float> prefire, postfire;
	float minE = 0, maxE = 0, minI = 0, maxI = 0;
	int backend = NEMO_BACKEND_CUDA;
	int device = 0;
	bool writeOnlySynapses = false;
	bool logging = false;
	unsigned partitionSize = 0;   /* 0 = automatic */
	unsigned deliveryLanes = 0;   /* lanes per fired source in delivery; 0 = automatic */
	bool stdpLdsStaging = true;   /* STDP count prefixes and reverse records staged in LDS */
	bool deviceConstruction = true;   /* synapse structures built on the device */
	int stepMode = NEMO_STEP_AUTO;
	int exchangeWaves = 0;        /* 0 = automatic, < 0 = none */
	std::string description;

	/* src/nemo/StdpFunction.cpp:12-50 */
	void setStdpFunction(const std::vector<float>& pre, const std::vector<float>& post,
			float minE, float maxE, float minI, float maxI);
};

/* A device allocation that frees itself */
struct DevBuf {
	void* ptr = nullptr;
	size_t bytes = 0;
	DevBuf() = default;
	DevBuf(const DevBuf&) = delete;
	DevBuf& operator=(const DevBuf&) = delete;
	~DevBuf();
	void alloc(size_t b);
	void ensure(size_t b);   /* grow (contents not kept) */
	void upload(const void* src, size_t b);
	template<typename T> T* as() const { return static_cast<T*>(ptr); }
};

/*
 * One network on the GPU (the HIP counterpart of NeMo's
 * cpu::Simulation / cuda::Simulation behind SimulationBackend,
 * src/nemo/SimulationBackend.hpp:40-142).
 */
class Simulation
{
	public:
		Simulation(const NetworkImpl& net, const ConfigurationImpl& conf);
		~Simulation();

		/* one cycle with stimulus (src/nemo/nemo_c.cpp:440-461); the result is
		 * valid until the next step */
		const std::vector<unsigned>& step(const std::vector<unsigned>& fstim,
				const unsigned* istimIdx, const float* istimVal, size_t nistim);

		void applyStdp(float reward);
		/* applyStdp in two halves, for applying to several simulations
		 * under one hold of the device lock (the caller holds it): the
		 * checks and the kernel launch, then the wait for it */
		void applyStdpLaunch(float reward);
		void applyStdpFinish();

		/* multi-cycle runs */
		void run(unsigned nsteps, const nemo_schedule_t* schedule, unsigned long long* events);
		/* run() with no device allocation (nemo_hip_run_static) */
		void runStatic(unsigned nsteps, const nemo_schedule_t* schedule, uint64_t* rasterDev,
				unsigned long long* events);
		void readFired(unsigned counts[], unsigned ids[], size_t cap, size_t* total);
		void readRaster(uint64_t words[], size_t cap, size_t* wordsPerStep);
		void info(nemo_hip_info_t* out) const;

		/* queries (src/nemo/Simulation.hpp:130-200) */
		float getMembranePotential(unsigned g) const;
		float getNeuronState(unsigned g, unsigned var) const;
		float getNeuronParameter(unsigned g, unsigned param) const;
		void setNeuron(unsigned g, unsigned nargs, const float args[]);
		void setNeuronState(unsigned g, unsigned var, float val);
		void setNeuronParameter(unsigned g, unsigned param, float val);
		const std::vector<synapse_id>& getSynapsesFrom(unsigned source);
		unsigned getSynapseSource(synapse_id id) const;
		unsigned getSynapseTarget(synapse_id id) const;
		unsigned getSynapseDelay(synapse_id id) const;
		float getSynapseWeight(synapse_id id) const;
		unsigned char getSynapsePlastic(synapse_id id) const;

		unsigned long elapsedWallclock() const;
		unsigned long elapsedSimulation() const { return (unsigned long) m_timerCycles; }
		void resetTimer();

		/* batched launches (nemo_hip_run_batch) */
		uint32_t partitions() const { return m_P; }
		uint32_t blockThreads() const { return m_T; }
		uint32_t ldsBytes() const { return m_lds; }
		/* LDS of a batched launch: without the far list of exchange waves
		 * the library chose by itself (a batched launch runs every
		 * workgroup at its largest member's LDS, so one network's far list
		 * would take co-residency from all the others) */
		uint32_t batchLdsBytes() const { return m_exAuto ? m_ldsNoFar : m_lds; }
		/* the step-kernel time of the last run, for a batched one: from the
		 * first start to the last end of this network's workgroups */
		void setKernelMs(float ms) { m_kernelMs = ms; }
		/* prepare a launch of nsteps cycles: fills and returns the device
		 * descriptor (host copy) with per-launch fields set; batched: as
		 * one member of a batched launch (see batchLdsBytes) */
		NetDev prepareLaunch(unsigned nsteps, bool batched = false);
		/* advance the cycle count by a finished launch; throws on a status
		 * error. ctl holds the first 8 words of the control block as the
		 * launch left them, counted from zero at the launch's start (a
		 * batched launch clears them first). */
		void completeLaunch(unsigned nsteps, const uint32_t ctl[8]);
		/* the control block's words as a batched launch left them on the
		 * device: the base from which the next single launch counts */
		void setControlBase(const uint32_t ctl[8]);
		unsigned neuronCount() const { return m_n; }
		int device() const { return m_device; }
		/* step-kernel variant (nemo_step_variant): saturating, Izhikevich-only */
		uint32_t variant() const { return (m_saturate ? 1u : 0u) | (m_izOnly ? 2u : 0u); }
		/* true: the fused persistent kernel; false: plugin kernels per cycle */
		bool fused() const { return m_fused; }

		static std::vector<WorkItem> xcdWorkTable(const std::vector<std::pair<uint32_t, uint32_t> >& items);

	private:
		bool m_fused = true;
		uint32_t m_n = 0, m_P = 0, m_T = 64, m_D = 1, m_words = 0, m_lds = 0;
		uint32_t m_np = nemo_dev::MAX_PARAM, m_ns = nemo_dev::MAX_STATE;

		/* plugin path */
		struct PluginGroup {
			unsigned start, end;
			unsigned model, history;
			void* handle;
			hip_update_neurons_t* update;
		};
		std::vector<PluginGroup> m_plugins;
		/* state planes: [m_hist][m_ns][m_n]; the current plane of each model
		 * with history > 1 (by model id; src/nemo/cpu/Neurons.cpp:123) */
		uint32_t m_hist = 1;
		std::vector<uint32_t> m_plane;
		size_t stateOffset(unsigned l, unsigned var) const;
		/* reverse matrix of every synapse for models that read it */
		DevBuf d_kStart, d_kSrc, d_kDelay, d_kW, d_kFwd;
		nemo_hip_rcm m_rcmDesc = {};
		bool m_hasRcm = false;
		uint32_t m_words32 = 0;
		DevBuf d_acc, d_curE, d_curI, d_curX, d_fstim, d_list;
		std::vector<uint32_t> m_fsPtrHost, m_isPtrHost;
		void loadPlugins(const NetworkImpl& net);
		void launchPlugins(unsigned nsteps, unsigned long long* events);
		bool m_saturate = false;
		bool m_stdp = false;
		bool m_writeOnly = false;
		size_t m_nsyn = 0, m_nplastic = 0;
		uint64_t m_cycle = 0;          /* absolute cycle (ring slot, buffer parity) */
		uint64_t m_timerCycles = 0;    /* cycles since the last timer reset */
		std::chrono::steady_clock::time_point m_timerStart;
		int m_device = 0;
		void* m_stream = nullptr;
		/* polls before a wait for another workgroup is declared lost: about
		 * 0.1-0.2 us each, so seconds -- a safety net against workgroups
		 * that never run, not a bound any legitimate wait approaches */
		uint32_t m_spinLimit = 1u << 26;
		uint32_t m_glog2 = 0;
		/* fused step delivers delay >= 2 records one cycle late (NEMO_NET_DEFER) */
		bool m_defer = false;
		/* cycles of fired lists and counts kept on the device */
		uint32_t m_flRing = 2;
		uint32_t m_stdpCache = 0;    /* STDP chunks staged in LDS per cycle */
		bool m_prefRing = false;     /* STDP list prefixes of postLen cycles kept in LDS */
		uint32_t m_exWaves = 0;      /* exchange waves (NetDev.exWaves) */
		bool m_exAuto = false;       /* exchange waves chosen by the library, not the configuration */
		uint32_t m_ldsNoFar = 0;     /* m_lds without the far list */
		uint32_t m_farCap = 0;       /* far-list capacity in sources (NetDev.farCap) */
		void planExchange(const ConfigurationImpl& conf, uint32_t ldsLimit);
		/* every neuron is an Izhikevich neuron (NEMO_NET_IZHIKEVICH_ONLY) */
		bool m_izOnly = false;

		/* index maps */
		std::vector<unsigned> m_localToGlobal;
		std::vector<uint32_t> m_dense;         /* gidx - m_minG -> local + 1 (0: none) */
		unsigned m_minG = 0;
		std::vector<uint8_t> m_model;
		std::vector<unsigned> m_typeOfLocal;   /* model id per local neuron */

		/* synapse queries: per local source, flat synapse records in ordinal order */
		std::vector<uint64_t> m_qStart;
		std::vector<uint32_t> m_qTarget;       /* global target */
		std::vector<uint8_t> m_qDelay, m_qPlastic;
		std::vector<uint32_t> m_qFwd;          /* record index in the forward matrix */
		std::vector<synapse_id> m_queried;

		nemo_dev::StdpParams m_sp;

		/* device data */
		DevBuf d_model, d_param, d_state, d_rng, d_recent, d_ring;
		DevBuf d_off, d_syn;
		DevBuf d_fl, d_fc, d_rfg, d_srow, d_ctl;       /* ctl: bar, status, events */
		DevBuf d_rcmStart, d_rcmSd, d_rcmFwd, d_rcmAcc, d_stdp;
		/* the stimulus schedule of the next launch: one device block, staged
		 * through pinned host memory, and its parts */
		DevBuf d_sched;
		uint32_t* h_sched = nullptr;
		size_t m_schedHostBytes = 0;
		const uint32_t* m_sFsPtr = nullptr;
		const uint32_t* m_sFsIdx = nullptr;
		const uint32_t* m_sIsPtr = nullptr;
		const uint32_t* m_sIsIdx = nullptr;
		const float* m_sIsVal = nullptr;
		DevBuf d_raster;
		DevBuf d_work;
		size_t m_deviceBytes = 0;
		float m_kernelMs = 0.0f;
		/* HIP events around each step-kernel launch (created once) */
		void* m_ev0 = nullptr;
		void* m_ev1 = nullptr;
		/* pinned host staging: [0..15] the control block read back after a
		 * launch, then one cycle's raster words (nemo_step) */
		uint32_t* h_ctl = nullptr;
		uint64_t* h_step = nullptr;
		/* control-block words 0 (workgroups started) and 2..3 (delivered
		 * synapses) as the last launch left them: the device block is not
		 * cleared between single launches, each launch's counts are the
		 * difference */
		uint32_t m_ctlStarted = 0;
		unsigned long long m_ctlEvents = 0;
		/* the step kernel's launch: descriptor by value, no copies before it */
		void launchFused(unsigned nsteps, bool stepOut);
		void launchDescriptor(const NetDev& nd, bool stepOut);
		void ensureRaster(size_t bytes);
		NetDev descriptor(unsigned nsteps) const;
		bool translateSchedule(unsigned nsteps, const nemo_schedule_t* s, std::vector<uint32_t>& fsPtr,
				std::vector<uint32_t>& fsIdx, std::vector<uint32_t>& isPtr, std::vector<uint32_t>& isIdx,
				std::vector<float>& isVal) const;
		/* the allocation-free run: schedule staging and the raster's place */
		std::vector<uint32_t> m_stFsPtr, m_stFsIdx, m_stIsPtr, m_stIsIdx, m_stage;
		std::vector<float> m_stIsVal;
		bool m_rasterExternal = false;   /* the last run's raster is in the caller's buffer */
		void requireOwnRaster() const;
		void readControl(unsigned nsteps, unsigned long long* events);

		/* schedule of the current launch (host side, sim-local indices) */
		bool m_hasSchedule = false;
		/* weights changed since the pending-current ring was filled */
		bool m_replay = false;
		unsigned m_lastRunSteps = 0;

		std::vector<unsigned> m_fired;
		std::vector<uint64_t> m_rasterHost;

		unsigned localIdx(unsigned g) const;
		void uploadSchedule(unsigned nsteps, const nemo_schedule_t* s);
		void launch(unsigned nsteps, unsigned long long* events);
		void readScalar(const DevBuf& b, size_t offset, void* out, size_t bytes) const;
		void writeScalar(DevBuf& b, size_t offset, const void* in, size_t bytes);
		void decodeStep(const uint64_t* words, std::vector<unsigned>& out) const;
		size_t synapseRecord(synapse_id id) const;
		/* the forward matrix (d_off, d_syn, m_defer), the synapse query
		 * tables and, with STDP, the reverse matrix, built on the device from
		 * the mapped synapses (device arrays of m_nsyn entries) */
		void buildForwardOnDevice(const std::vector<unsigned>& Tg, const std::vector<unsigned char>& Dl,
				const std::vector<unsigned char>& Pl, uint32_t bflags, const uint32_t* lsrc, const uint32_t* ltgt,
				const int32_t* wfx, const uint8_t* dDl, const uint8_t* dPl);
		bool m_deviceBuilt = false;   /* synapse structures built on the device */
};

} // namespace hip
} // namespace nemo

#endif
