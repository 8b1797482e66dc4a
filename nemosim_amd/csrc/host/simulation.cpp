#include "simulation.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <random>
#include <dlfcn.h>
#include <mutex>
#include <unordered_map>

#include <hip/hip_runtime.h>

namespace nemo {
namespace hip {

/* LDS available to one workgroup of the step kernel on gfx950 */
static const uint32_t LDS_LIMIT = 160u * 1024u;

/* ------------------------------------------------------------ configuration */

void
ConfigurationImpl::setStdpFunction(const std::vector<float>& pre, const std::vector<float>& post,
		float fminE, float fmaxE, float fminI, float fmaxI)
{
	/* src/nemo/StdpFunction.cpp:12-50 */
	if(fmaxI > 0.0f || fminI > 0.0f) {
		throw exception(NEMO_INVALID_INPUT, "STDP function should not have positive limits for inhibitory synapses");
	}
	if(fmaxE < 0.0f || fminE < 0.0f) {
		throw exception(NEMO_INVALID_INPUT, "STDP function should not have negative limits for inhibitory synapses");
	}
	if(fmaxE < fminE) {
		throw exception(NEMO_INVALID_INPUT, "STDP should have maximum excitatory weight >= minimum excitatory weight");
	}
	if(fmaxI > fminI) {
		throw exception(NEMO_INVALID_INPUT, "STDP should have abs(maximum inhibitory weight) >= abs(minimum inhibitory weight)");
	}
	if(pre.size() + post.size() > 64) {
		throw exception(NEMO_INVALID_INPUT, "size of STDP window too large");
	}
	stdp = true;
	prefire = pre;
	postfire = post;
	minE = fminE; maxE = fmaxE; minI = fminI; maxI = fmaxI;
}

/* ------------------------------------------------------------------ DevBuf */

DevBuf::~DevBuf()
{
	if(ptr) {
		(void) hipFree(ptr);
	}
}

void
DevBuf::alloc(size_t b)
{
	if(ptr) {
		(void) hipFree(ptr);
		ptr = nullptr;
		bytes = 0;
	}
	if(b == 0) return;
	hipCheck(hipMalloc(&ptr, b), "hipMalloc");
	bytes = b;
}

void
DevBuf::ensure(size_t b)
{
	if(b > bytes) {
		/* geometric growth: buffers that follow a run's size are reallocated
		 * O(log) times, not once per longer run */
		alloc(std::max(b, bytes * 2));
	}
}

void
DevBuf::upload(const void* src, size_t b)
{
	alloc(b);
	if(b) {
		hipCheck(hipMemcpy(ptr, src, b, hipMemcpyHostToDevice), "hipMemcpy H2D");
	}
}

/* -------------------------------------------------------------- simulation */

/* Streams and timing events of released simulations, kept per device for
 * the next ones: creating a stream took 3.9 ms and destroying one 0.8 ms
 * in a serving run, on the threads that build and release requests. At
 * most 64 per device are kept; the pool is never torn down, so no HIP call
 * runs at process exit. */
namespace {
struct StreamSet {
	hipStream_t stream;
	hipEvent_t ev0, ev1;
};
std::mutex g_streamPoolMutex;
std::vector<StreamSet>* g_streamPool = new std::vector<StreamSet>[64];
const size_t STREAM_POOL_MAX = 64;

bool
takeStreamSet(int dev, StreamSet* out)
{
	if(dev < 0 || dev >= 64) return false;
	std::lock_guard<std::mutex> lock(g_streamPoolMutex);
	if(g_streamPool[dev].empty()) return false;
	*out = g_streamPool[dev].back();
	g_streamPool[dev].pop_back();
	return true;
}

bool
giveStreamSet(int dev, const StreamSet& set)
{
	if(dev < 0 || dev >= 64) return false;
	std::lock_guard<std::mutex> lock(g_streamPoolMutex);
	if(g_streamPool[dev].size() >= STREAM_POOL_MAX) return false;
	g_streamPool[dev].push_back(set);
	return true;
}
} // namespace

/* blockIdx b runs on XCD b mod 8 (round-robin dispatch); give each XCD a
 * contiguous run of the (network, partition) list. */
std::vector<WorkItem>
Simulation::xcdWorkTable(const std::vector<std::pair<uint32_t, uint32_t> >& items)
{
	const size_t G = items.size();
	const size_t X = 8;
	std::vector<WorkItem> work(G);
	std::vector<size_t> start(X + 1, 0);
	for(size_t x = 0; x < X; ++x) {
		size_t cnt = G / X + ((x < G % X) ? 1 : 0);
		start[x + 1] = start[x] + cnt;
	}
	for(size_t b = 0; b < G; ++b) {
		size_t x = b % X, j = b / X;
		const auto& it = items[start[x] + j];
		work[b].net = it.first;
		work[b].part = it.second;
	}
	return work;
}

static uint32_t
roundUp64(uint32_t v)
{
	return (v + 63u) & ~63u;
}

Simulation::Simulation(const NetworkImpl& net, const ConfigurationImpl& conf) :
	m_stdp(conf.stdp),
	m_writeOnly(conf.writeOnlySynapses),
	m_device(conf.device < 0 ? 0 : conf.device)
{
	std::memset(&m_sp, 0, sizeof(m_sp));
	hipCheck(hipSetDevice(m_device), "hipSetDevice");
	StreamSet set;
	if(takeStreamSet(m_device, &set)) {
		m_stream = set.stream;
		m_ev0 = set.ev0;
		m_ev1 = set.ev1;
	} else {
		hipStream_t st;
		hipCheck(hipStreamCreate(&st), "hipStreamCreate");
		m_stream = st;
		hipEvent_t ev0, ev1;
		hipCheck(hipEventCreate(&ev0), "hipEventCreate");
		m_ev0 = ev0;
		hipCheck(hipEventCreate(&ev1), "hipEventCreate");
		m_ev1 = ev1;
	}
	hipCheck(hipHostMalloc(reinterpret_cast<void**>(&h_ctl), NEMO_CTL_WORDS * 4), "hipHostMalloc");
	std::memset(h_ctl, 0, NEMO_CTL_WORDS * 4);

	/* ---- simulation-local indices: types in order, insertion order within a
	 * type (src/nemo/cpu/Simulation.cpp:58-73, src/nemo/cpu/Neurons.cpp:26-39) */
	unsigned maxG = 0;
	m_minG = ~0u;
	for(const NeuronGroup& g : net.groups) {
		for(unsigned gi : g.gidx) {
			m_localToGlobal.push_back(gi);
			m_model.push_back(uint8_t(g.model->id));
			maxG = std::max(maxG, gi);
			m_minG = std::min(m_minG, gi);
		}
	}
	m_n = uint32_t(m_localToGlobal.size());
	if(m_n >= (1u << 26)) {
		throw exception(NEMO_API_UNSUPPORTED, "the HIP backend supports up to 2^26 neurons per simulation");
	}
	if(m_n > 0) {
		m_dense.assign(size_t(maxG - m_minG) + 1, 0u);
		for(uint32_t l = 0; l < m_n; ++l) {
			m_dense[m_localToGlobal[l] - m_minG] = l + 1;
		}
	} else {
		m_minG = 0;
	}

	/* ---- delays and the STDP window (src/nemo/StdpFunction.cpp:52-60) */
	unsigned maxDelay = net.maxDelay();
	if(maxDelay > 64) {
		throw exception(NEMO_INVALID_INPUT, fmt("The network has synapses with delay ", maxDelay,
					"ms. The HIP backend supports a maximum of 64 ms"));
	}
	if(m_stdp && maxDelay + conf.postfire.size() > 64) {
		throw exception(NEMO_INVALID_INPUT, "Max delay + postfire part of STDP function must be <= 64");
	}
	m_D = std::max(1u, maxDelay);

	/* ---- STDP tables (src/nemo/StdpProcess.cpp:20-46, src/nemo/StdpFunction.cpp:73-108) */
	if(m_stdp) {
		m_sp.minE = fx_toFix(conf.minE);
		m_sp.maxE = fx_toFix(conf.maxE);
		m_sp.minI = fx_toFix(conf.minI);
		m_sp.maxI = fx_toFix(conf.maxI);
		for(size_t i = 0; i < conf.prefire.size(); ++i) m_sp.fnPre[i] = fx_toFix(conf.prefire[i]);
		for(size_t i = 0; i < conf.postfire.size(); ++i) m_sp.fnPost[i] = fx_toFix(conf.postfire[i]);
		uint64_t pot = 0, dep = 0;
		int bit = 0;
		for(auto f = conf.postfire.rbegin(); f != conf.postfire.rend(); ++f, ++bit) {
			if(*f > 0.0f) pot |= uint64_t(1) << bit;
			if(*f < 0.0f) dep |= uint64_t(1) << bit;
		}
		for(auto f = conf.prefire.begin(); f != conf.prefire.end(); ++f, ++bit) {
			if(*f > 0.0f) pot |= uint64_t(1) << bit;
			if(*f < 0.0f) dep |= uint64_t(1) << bit;
		}
		m_sp.potentiation = pot;
		m_sp.depression = dep;
		m_sp.preLen = int32_t(conf.prefire.size());
		m_sp.postLen = int32_t(conf.postfire.size());
		uint64_t preMask = m_sp.preLen >= 64 ? ~uint64_t(0) : ~(~uint64_t(0) << m_sp.preLen);
		m_sp.prePost = preMask << m_sp.postLen;
		m_sp.postPre = m_sp.postLen >= 64 ? ~uint64_t(0) : ~(~uint64_t(0) << m_sp.postLen);
	}

	/* ---- partition geometry */
	int dev = m_device;
	int ldsMax = 0;
	hipCheck(hipDeviceGetAttribute(&ldsMax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "hipDeviceGetAttribute");
	const uint32_t ldsLimit = std::min<uint32_t>(LDS_LIMIT, uint32_t(ldsMax));

	/* ---- synapse validation and saturation bound. On the device when the
	 * configuration allows (each synapse mapped to local ids, its weight
	 * converted, the bounds summed per target); the host loop does the same
	 * and raises the errors, so it runs whenever the device finds a synapse
	 * it must look at. */
	const std::vector<unsigned>& S = net.src();
	const std::vector<unsigned>& Tg = net.tgt();
	const std::vector<unsigned char>& Dl = net.delay();
	const std::vector<float>& W = net.weight();
	const std::vector<unsigned char>& Pl = net.plastic();
	m_nsyn = S.size();
	hipStream_t bs = static_cast<hipStream_t>(m_stream);
	std::unique_lock<DeviceMutex> buildLock(deviceMutex(), std::defer_lock);
	DevBuf dS, dTg, dW, dPl, dDl, dLsrc, dLtgt, dWfx;
	bool onDevice = conf.deviceConstruction && m_nsyn > 0 && m_nsyn < (size_t(1) << 32) && m_n > 0;
	uint32_t bflags = 0;
	if(onDevice) {
		buildLock.lock();
		DevBuf dDense, dBound, dFlags;
		dS.upload(S.data(), m_nsyn * 4);
		dTg.upload(Tg.data(), m_nsyn * 4);
		dW.upload(W.data(), m_nsyn * 4);
		dPl.upload(Pl.data(), m_nsyn);
		dDl.upload(Dl.data(), m_nsyn);
		dDense.upload(m_dense.data(), m_dense.size() * 4);
		dLsrc.alloc(m_nsyn * 4);
		dLtgt.alloc(m_nsyn * 4);
		dWfx.alloc(m_nsyn * 4);
		dBound.alloc(size_t(2) * m_n * 8);
		dFlags.alloc(4);
		hipCheck(nemo_dev_fill_u64(dBound.as<uint64_t>(), 0, size_t(2) * m_n, bs), "fill");
		hipCheck(nemo_dev_fill_u32(dFlags.as<uint32_t>(), 0, 1, bs), "fill");
		hipCheck(nemo_dev_build_map(dS.as<uint32_t>(), dTg.as<uint32_t>(), dW.as<float>(), dPl.as<uint8_t>(),
					m_nsyn, dDense.as<uint32_t>(), m_minG, m_dense.size(), m_stdp ? 1 : 0, m_sp.maxE, m_sp.maxI,
					dLsrc.as<uint32_t>(), dLtgt.as<uint32_t>(), dWfx.as<int32_t>(), dBound.as<int64_t>(),
					dBound.as<int64_t>() + m_n, m_n, dFlags.as<uint32_t>(), bs), "construction kernel");
		hipCheck(hipMemcpyAsync(&bflags, dFlags.ptr, 4, hipMemcpyDeviceToHost, bs), "hipMemcpyAsync D2H");
		hipCheck(hipStreamSynchronize(bs), "hipStreamSynchronize");
		buildLock.unlock();
		if(bflags & NEMO_BUILD_HOST_CHECK) {
			onDevice = false;
		}
	}
	std::vector<uint32_t> lsrc, ltgt;
	std::vector<int32_t> wfx;
	if(onDevice) {
		m_saturate = (bflags & NEMO_BUILD_SATURATE) != 0;
		if(m_stdp) {
			m_nplastic = size_t(std::count_if(Pl.begin(), Pl.end(), [](unsigned char c) { return c != 0; }));
		}
	} else {
		lsrc.resize(m_nsyn);
		ltgt.resize(m_nsyn);
		wfx.resize(m_nsyn);
		std::vector<int64_t> boundE(m_n, 0), boundI(m_n, 0);
		for(size_t i = 0; i < m_nsyn; ++i) {
			unsigned s = S[i], t = Tg[i];
			uint32_t ls = (s >= m_minG && size_t(s - m_minG) < m_dense.size()) ? m_dense[s - m_minG] : 0u;
			if(!ls) {
				throw exception(NEMO_INVALID_INPUT, fmt("Invalid synapse source neuron ", s));
			}
			uint32_t lt = (t >= m_minG && size_t(t - m_minG) < m_dense.size()) ? m_dense[t - m_minG] : 0u;
			if(!lt) {
				throw exception(NEMO_INVALID_INPUT, fmt("Invalid synapse target neuron ", t, " (source: ", s, ")"));
			}
			lsrc[i] = ls - 1;
			ltgt[i] = lt - 1;
			int32_t w = fx_toFix(W[i]);   /* src/nemo/ConnectivityMatrix.cpp:101 */
			wfx[i] = w;
			int64_t wb = w;
			if(m_stdp && Pl[i]) {
				if(w > 0) wb = std::max<int64_t>(w, m_sp.maxE);
				else if(w < 0) wb = std::min<int64_t>(w, m_sp.maxI);
			}
			if(w >= 0) boundE[lt - 1] += wb; else boundI[lt - 1] += wb;
			if(m_stdp && Pl[i]) m_nplastic += 1;
		}
		for(uint32_t l = 0; l < m_n; ++l) {
			if(boundE[l] > int64_t(INT32_MAX) || boundI[l] < int64_t(INT32_MIN)) {
				m_saturate = true;
				break;
			}
		}
	}

	/* ---- which step runs this network: the fused kernel needs built-in
	 * models, an LDS ring that fits and all partitions co-resident */
	bool allBuiltin = true;
	for(const NeuronGroup& g : net.groups) {
		if(!g.gidx.empty() && !g.model->builtin) allBuiltin = false;
		m_np = std::max(m_np, g.model->nparam);
		m_ns = std::max(m_ns, g.model->nstate);
	}
	/* models that keep state history or read the reverse matrix (Kuramoto).
	 * src/nemo/construction/RCM.cpp:90-97 refuses STDP with such a type.
	 * The reference's CPU Kuramoto plugin indexes its arrays with
	 * simulation-wide indices (src/nemo/cpu/plugins/Kuramoto.cpp:105-110) and
	 * its CUDA plugin asserts that the type fills the whole network
	 * (src/nemo/cuda/plugins/Kuramoto.cu:270): a type reading the reverse
	 * matrix must be the network's only type here. A delay d reads the
	 * plane of cycle t-(d-1), which a history of H cycles holds while
	 * d <= H-1 (src/nemo/cuda/plugins/Kuramoto.cu:116). */
	{
		bool rcmType = false, otherType = false;
		unsigned rcmHistory = 0;
		for(const NeuronGroup& g : net.groups) {
			if(g.gidx.empty()) continue;
			m_hist = std::max(m_hist, g.model->history);
			if(g.model->usesRcm()) {
				rcmType = true;
				if(g.model->rcmDelays) rcmHistory = std::max(rcmHistory, g.model->history);
			} else {
				otherType = true;
			}
		}
		if(m_plane.size() <= 255) m_plane.assign(256, 0u);
		if(rcmType && m_stdp) {
			throw exception(NEMO_API_UNSUPPORTED,
					"The current version does not support a mix of STDP with neuron types which require reverse connectivity for other purposes");
		}
		if(rcmType && otherType) {
			throw exception(NEMO_API_UNSUPPORTED,
					"a neuron type that reads the reverse connectivity (e.g. Kuramoto) must be the only type in the network");
		}
		if(rcmHistory > 1 && net.maxDelay() >= rcmHistory) {
			throw exception(NEMO_INVALID_INPUT, fmt("delays of up to ", rcmHistory - 1,
						" cycles are supported for a model with a state history of ", rcmHistory,
						" cycles; the network has a delay of ", net.maxDelay()));
		}
		m_hasRcm = rcmType;
	}
	std::string whyNotFused;
	if(conf.stepMode == NEMO_STEP_PLUGINS) {
		whyNotFused = "plugin step requested";
	} else if(!allBuiltin) {
		whyNotFused = "a neuron type has no fused implementation";
	}

	/* T: neurons per partition. The automatic choice takes at most one
	 * partition per compute unit (256), and for mid-size networks fewer,
	 * wider partitions: their cycle is bound by the exchange between
	 * partitions. Measured per cycle: about 40 partitions of up to 512
	 * neurons (10k x 100: 7.2 -> 6.3 us; 50k x 1000: 10.4 -> 9.7 us), and
	 * with rows of 500 synapses or more, from 1.5k to 12k neurons, about 13
	 * partitions (2k x 1000: 8.0 -> 7.2 us; 5k x 1000: 9.6 -> 7.1 us; 10k x
	 * 1000: 10.2 -> 8.6 us). The LDS ring [D][T] bounds it from above. */
	uint32_t T;
	if(conf.partitionSize) {
		T = std::min(1024u, roundUp64(conf.partitionSize));
	} else {
		const uint32_t perCU = roundUp64(std::max<uint32_t>(1u, (m_n + 255u) / 256u));
		const bool longRows = m_n >= 1500u && m_n <= 12000u && m_nsyn / m_n >= 500u;
		const uint32_t parts = longRows ? 13u : 40u;
		const uint32_t wide = std::min(longRows ? 1024u : 512u,
				roundUp64(std::max<uint32_t>(1u, (m_n + parts - 1u) / parts)));
		T = std::max(64u, std::min(1024u, std::max(perCU, wide)));
	}
	auto ldsFor = [&](uint32_t t) {
		uint32_t p = (m_n + t - 1) / t;
		return nemo_step_lds_bytes(t, m_D, std::max(1u, p), m_saturate ? 1 : 0, m_stdp ? 1 : 0);
	};
	while(T > 64 && ldsFor(T) > ldsLimit) {
		T -= 64;
	}
	if(whyNotFused.empty() && ldsFor(T) > ldsLimit) {
		whyNotFused = "maximum delay too large for the LDS delay ring";
	}
	m_T = T;
	m_P = (m_n + T - 1) / T;
	m_lds = ldsFor(T);
	m_izOnly = !m_model.empty() && std::all_of(m_model.begin(), m_model.end(),
			[](uint8_t m) { return m == uint8_t(nemo_dev::MODEL_IZHIKEVICH); });
	if(whyNotFused.empty() && m_P > 0) {
		uint32_t maxBlocks = 0;
		hipCheck(nemo_dev_step_max_blocks(m_T, m_lds, variant(), &maxBlocks), "occupancy query");
		if(m_P > maxBlocks) {
			whyNotFused = fmt("network of ", m_n, " neurons needs ", m_P,
					" co-resident workgroups of ", m_T, " neurons; the device holds ", maxBlocks);
		}
		/* STDP, with LDS left over while the workgroups stay co-resident:
		 * first the prefixes of the fired counts of the last postLen cycles
		 * (the work lists' targets; otherwise gathered again each cycle),
		 * then the reverse records of up to 64 chunks (256 bytes each) of
		 * every cycle's work list */
		auto fits = [&](uint32_t lds) {
			if(lds > ldsLimit) return false;
			hipCheck(nemo_dev_step_max_blocks(m_T, lds, variant(), &maxBlocks), "occupancy query");
			return m_P <= maxBlocks;
		};
		const uint32_t postLen = uint32_t(m_sp.postLen);
		if(m_stdp && conf.stdpLdsStaging && whyNotFused.empty() && postLen >= 2) {
			const uint32_t ring = postLen * ((m_P + 1u + 63u) & ~63u) * 4u;
			if(fits(m_lds + ring)) {
				m_prefRing = true;
				m_lds += ring;
			}
		}
		for(uint32_t c = 64; m_stdp && conf.stdpLdsStaging && whyNotFused.empty() && c >= 8; c /= 2) {
			if(fits(m_lds + c * 256u)) {
				m_stdpCache = c;
				m_lds += c * 256u;
				break;
			}
		}
	}
	if(!whyNotFused.empty()) {
		if(conf.stepMode == NEMO_STEP_FUSED) {
			throw exception(NEMO_API_UNSUPPORTED, "fused step kernel unavailable: " + whyNotFused);
		}
		if(conf.stepMode == NEMO_STEP_AUTO) {
			/* a different kernel sequence with its own speed: say why */
			fprintf(stderr, "nemosim HIP: %u neurons run on the plugin step (%s)\n", m_n, whyNotFused.c_str());
		}
		/* plugin path: one bucket per source holding all its synapses, with
		 * absolute targets */
		m_fused = false;
		m_P = 1;
		m_T = 0;
		m_lds = 0;
		m_words32 = 2 * ((m_n + 63) / 64);
		m_words = m_words32 / 2;
	} else {
		m_words = m_P * (T / 64u);
		/* lanes per fired source in delivery, from the mean bucket length
		 * (synapses of one source into one partition): a power of two about
		 * 1.5x of it for short buckets, so that one pass (two records per
		 * lane) covers almost every bucket (200k x 1000, buckets of 4: 8
		 * lanes beat 4 and 2); for long buckets no more than half of it,
		 * and at least 8, since every round serves only 64/lanes sources
		 * and a wave with many fired sources runs as many rounds. Measured
		 * per cycle, the capped count against 1.5x: 20k x 1000 (buckets of
		 * 25) 8.6 against 10.9 us; IF_curr_exp_noise at 333 Hz, 3k x 300
		 * (12.5) 28.1 against 48.8, 12k x 550 (42) 57.8 against 74.1,
		 * 50k x 1000 (10) 145.5 against 165.8; 3k x 300 at 6 Hz 6.7
		 * against 6.1. */
		const double lavg = m_n ? double(m_nsyn) / (double(m_n) * double(m_P)) : 0.0;
		m_glog2 = 0;
		if(conf.deliveryLanes) {
			while((1u << m_glog2) < conf.deliveryLanes) ++m_glog2;
		} else {
			while(m_glog2 < 6 && double(1u << m_glog2) < 1.5 * lavg) {
				++m_glog2;
			}
			uint32_t half = 0;
			while(half < 6 && double(1u << half) < 0.5 * lavg) {
				++half;
			}
			m_glog2 = std::min(m_glog2, std::max(3u, half));
		}
	}

	if(onDevice && !m_fused) {
		/* the plugin step's layout is built on the host, from the mapped synapses */
		lsrc.resize(m_nsyn);
		ltgt.resize(m_nsyn);
		wfx.resize(m_nsyn);
		buildLock.lock();
		hipCheck(hipMemcpy(lsrc.data(), dLsrc.ptr, m_nsyn * 4, hipMemcpyDeviceToHost), "hipMemcpy D2H");
		hipCheck(hipMemcpy(ltgt.data(), dLtgt.ptr, m_nsyn * 4, hipMemcpyDeviceToHost), "hipMemcpy D2H");
		hipCheck(hipMemcpy(wfx.data(), dWfx.ptr, m_nsyn * 4, hipMemcpyDeviceToHost), "hipMemcpy D2H");
		buildLock.unlock();
		onDevice = false;
	}
	m_deviceBuilt = onDevice;
	const uint32_t P1 = m_P + 1;
	const uint32_t NEAR_SHIFT = 28, NEAR_MAX = 15;
	std::vector<uint32_t> off, syn;
	std::vector<uint32_t> rcmStart, rcmSd, rcmFwd;
	if(onDevice) {
		buildForwardOnDevice(Tg, Dl, Pl, bflags, dLsrc.as<uint32_t>(), dLtgt.as<uint32_t>(), dWfx.as<int32_t>(),
				dDl.as<uint8_t>(), dPl.as<uint8_t>());
	} else {
		/* ---- forward matrix, bucketed by (source, target partition) */
		std::vector<uint32_t> nearCount;
		off.assign(size_t(m_n) * P1 + 1, 0u);
		for(size_t i = 0; i < m_nsyn; ++i) {
			off[size_t(lsrc[i]) * P1 + (m_fused ? ltgt[i] / T : 0u)] += 1;
		}
		uint64_t run = 0;
		for(size_t k = 0; k < off.size(); ++k) {
			uint64_t c = off[k];
			off[k] = uint32_t(run);
			run += c;
		}
		if(run >= (uint64_t(1) << 32)) {
			throw exception(NEMO_API_UNSUPPORTED, "more than 2^32 synapses in one simulation");
		}
		std::vector<uint32_t> cursor(off.begin(), off.end());
		syn.assign(2 * m_nsyn, 0u);
		std::vector<uint32_t> fwd(m_nsyn);
		/* fused step: the delay-1 records of each bucket come first (pass 0),
		 * so that the kernel can deliver them in the cycle and the rest one
		 * cycle later */
		for(int pass = 0; pass < (m_fused ? 2 : 1); ++pass) {
			for(size_t i = 0; i < m_nsyn; ++i) {
				if(m_fused && ((Dl[i] == 1) != (pass == 0))) continue;
				uint32_t lt = ltgt[i];
				uint32_t p = m_fused ? lt / T : 0u;
				uint32_t pos = cursor[size_t(lsrc[i]) * P1 + p]++;
				if(m_fused) {
					syn[2 * size_t(pos)] = (lt - p * T) | (uint32_t(Dl[i]) << 16);
				} else {
					syn[2 * size_t(pos)] = lt | (uint32_t(Dl[i] - 1) << 26);
				}
				syn[2 * size_t(pos) + 1] = uint32_t(wfx[i]);
				fwd[i] = pos;
			}
			if(m_fused && pass == 0) {
				/* cursor - off = delay-1 records per bucket; the deferred form
				 * keeps that count in the top bits of the bucket start */
				m_defer = run < (uint64_t(1) << NEAR_SHIFT);
				for(size_t k = 0; k + 1 < off.size() && m_defer; ++k) {
					if(cursor[k] - off[k] > NEAR_MAX) m_defer = false;
				}
				if(m_defer) {
					for(size_t k = 0; k + 1 < off.size(); ++k) {
						nearCount.push_back(cursor[k] - off[k]);
					}
				}
			}
		}
		if(m_defer) {
			for(size_t k = 0; k < nearCount.size(); ++k) {
				off[k] |= nearCount[k] << NEAR_SHIFT;
			}
			std::vector<uint32_t>().swap(nearCount);
		}
		if(!m_fused) {
			/* rows per source: off[s] .. off[s+1] */
			std::vector<uint32_t> rows(size_t(m_n) + 1, 0u);
			for(uint32_t s = 0; s <= m_n; ++s) rows[s] = off[size_t(s) * 2];
			off.swap(rows);
		}
		std::vector<uint32_t>().swap(cursor);
		std::vector<int32_t>().swap(wfx);

		/* ---- synapse query tables, ordinal order per source */
		if(!m_writeOnly) {
			m_qStart.assign(size_t(m_n) + 1, 0);
			for(size_t i = 0; i < m_nsyn; ++i) m_qStart[lsrc[i] + 1] += 1;
			for(uint32_t l = 0; l < m_n; ++l) m_qStart[l + 1] += m_qStart[l];
			std::vector<uint64_t> qc(m_qStart.begin(), m_qStart.end() - 1);
			m_qTarget.resize(m_nsyn);
			m_qDelay.resize(m_nsyn);
			m_qPlastic.resize(m_nsyn);
			m_qFwd.resize(m_nsyn);
			for(size_t i = 0; i < m_nsyn; ++i) {
				uint64_t q = qc[lsrc[i]]++;
				m_qTarget[q] = Tg[i];
				m_qDelay[q] = Dl[i];
				m_qPlastic[q] = Pl[i];
				m_qFwd[q] = fwd[i];
			}
		}

		/* ---- reverse matrix of every synapse for a model that reads it
		 * (src/nemo/construction/RCM.cpp:150-171 with STDP off): per target,
		 * in the order of the network's synapse iterator -- sources ascending
		 * by neuron index, each source's synapses in insertion order
		 * (src/nemo/NetworkImpl.cpp:311-325, src/nemo/ConnectivityMatrix.cpp:79-87) */
		if(m_hasRcm) {
			std::vector<size_t> order(m_nsyn);
			for(size_t i = 0; i < m_nsyn; ++i) order[i] = i;
			std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return S[a] < S[b]; });
			std::vector<uint32_t> kStart(size_t(m_n) + 1, 0u), kSrc(m_nsyn), kDelay(m_nsyn), kFwd(m_nsyn);
			std::vector<float> kW(m_nsyn);
			for(size_t i = 0; i < m_nsyn; ++i) kStart[ltgt[i] + 1] += 1;
			for(uint32_t l = 0; l < m_n; ++l) kStart[l + 1] += kStart[l];
			std::vector<uint32_t> kc(kStart.begin(), kStart.end() - 1);
			for(size_t k = 0; k < m_nsyn; ++k) {
				const size_t i = order[k];
				const uint32_t r = kc[ltgt[i]]++;
				kSrc[r] = lsrc[i];
				kDelay[r] = Dl[i];
				kW[r] = W[i];
				kFwd[r] = fwd[i];
			}
			bool wantSrc = false, wantDelay = false, wantW = false, wantFwd = false;
			for(const NeuronGroup& g : net.groups) {
				if(g.gidx.empty()) continue;
				wantSrc = wantSrc || g.model->rcmSources;
				wantDelay = wantDelay || g.model->rcmDelays;
				wantW = wantW || g.model->rcmWeights;
				wantFwd = wantFwd || g.model->rcmForward;
			}
			std::lock_guard<DeviceMutex> uploadLock(deviceMutex());
			std::memset(&m_rcmDesc, 0, sizeof(m_rcmDesc));
			m_rcmDesc.targets = m_n;
			m_rcmDesc.synapses = m_nsyn;
			d_kStart.upload(kStart.data(), kStart.size() * 4);
			m_rcmDesc.start = d_kStart.as<uint32_t>();
			if(wantSrc && m_nsyn) { d_kSrc.upload(kSrc.data(), m_nsyn * 4); m_rcmDesc.source = d_kSrc.as<uint32_t>(); }
			if(wantDelay && m_nsyn) { d_kDelay.upload(kDelay.data(), m_nsyn * 4); m_rcmDesc.delay = d_kDelay.as<uint32_t>(); }
			if(wantW && m_nsyn) { d_kW.upload(kW.data(), m_nsyn * 4); m_rcmDesc.weight = d_kW.as<float>(); }
			if(wantFwd && m_nsyn) { d_kFwd.upload(kFwd.data(), m_nsyn * 4); m_rcmDesc.forward = d_kFwd.as<uint32_t>(); }
		}

		/* ---- reverse matrix of plastic synapses (src/nemo/construction/RCM.cpp:150-171) */
		if(m_stdp) {
			rcmStart.assign(size_t(m_n) + 1, 0);
			for(size_t i = 0; i < m_nsyn; ++i) if(Pl[i]) rcmStart[ltgt[i] + 1] += 1;
			for(uint32_t l = 0; l < m_n; ++l) rcmStart[l + 1] += rcmStart[l];
			std::vector<uint32_t> rc(rcmStart.begin(), rcmStart.end() - 1);
			rcmSd.resize(m_nplastic);
			rcmFwd.resize(m_nplastic);
			for(size_t i = 0; i < m_nsyn; ++i) {
				if(!Pl[i]) continue;
				uint32_t r = rc[ltgt[i]]++;
				rcmSd[r] = lsrc[i] | (uint32_t(Dl[i] - 1) << 26);
				rcmFwd[r] = fwd[i];
			}
		}
		std::vector<uint32_t>().swap(lsrc);
		std::vector<uint32_t>().swap(ltgt);
		std::vector<uint32_t>().swap(fwd);
	}
	dS.alloc(0); dTg.alloc(0); dW.alloc(0); dPl.alloc(0); dDl.alloc(0); dLsrc.alloc(0); dLtgt.alloc(0); dWfx.alloc(0);

	/* ---- neuron arrays */
	std::vector<float> param(size_t(m_np) * m_n, 0.0f);
	/* plane 0 holds the initial state; the others, for models with history,
	 * are filled by the plugin's init */
	std::vector<float> state(size_t(m_hist) * m_ns * m_n, 0.0f);
	{
		uint32_t l = 0;
		for(const NeuronGroup& g : net.groups) {
			for(size_t i = 0; i < g.gidx.size(); ++i, ++l) {
				for(unsigned k = 0; k < g.model->nparam; ++k) param[size_t(k) * m_n + l] = g.param[k][i];
				for(unsigned k = 0; k < g.model->nstate; ++k) state[size_t(k) * m_n + l] = g.state[k][i];
			}
		}
	}
	/* per-neuron RNG seeds, src/nemo/RNG.cpp:38-63: neuron k takes draws
	 * 4k .. 4k+3 of the default-seeded MT19937, each mapped to draw >> 1 */
	std::vector<uint32_t> rng(size_t(4) * m_n);
	{
		std::mt19937 mt;
		for(uint32_t l = 0; l < m_n; ++l) {
			for(unsigned k = 0; k < 4; ++k) rng[size_t(k) * m_n + l] = uint32_t(mt()) >> 1;
		}
	}

	/* ---- upload */
	std::unique_lock<DeviceMutex> devLock(deviceMutex());
	d_model.upload(m_model.data(), m_model.size());
	d_param.upload(param.data(), param.size() * sizeof(float));
	d_state.upload(state.data(), state.size() * sizeof(float));
	d_rng.upload(rng.data(), rng.size() * sizeof(uint32_t));
	d_recent.alloc(size_t(m_n) * 8);
	d_ring.alloc(size_t(m_P) * 2 * m_D * m_T * 4);
	if(!m_deviceBuilt) {
		d_off.upload(off.data(), off.size() * sizeof(uint32_t));
		if(syn.size() < 2) {
			/* the delivery loop reads record 0 in place of an absent record */
			syn.resize(2, 0u);
		}
		d_syn.upload(syn.data(), syn.size() * sizeof(uint32_t));
	}
	planExchange(conf, ldsLimit);

	/* fired lists and counts are kept for flRing cycles: 2 (parity), 3 when
	 * deferred delivery reads the previous cycle's, postLen + 2 with STDP,
	 * whose accumulation reads those of cycle t - postLen */
	m_flRing = m_stdp ? std::max(3u, uint32_t(m_sp.postLen) + 2u) : (m_defer ? 3u : 2u);
	d_fl.alloc(size_t(m_flRing) * m_P * m_T * 8);
	d_fc.alloc(size_t(m_flRing) * m_P * 8 + 8);
	d_ctl.alloc(64);
	if(m_stdp) {
		d_rfg.alloc(size_t(2) * m_n * 8);
		if(!m_deviceBuilt) {
			d_rcmStart.upload(rcmStart.data(), rcmStart.size() * 4);
			d_rcmSd.upload(rcmSd.data(), rcmSd.size() * 4);
			d_rcmFwd.upload(rcmFwd.data(), rcmFwd.size() * 4);
		}
		d_rcmAcc.alloc(std::max<size_t>(m_nplastic, 1) * 4);
		d_stdp.upload(&m_sp, sizeof(m_sp));
		if(m_fused) {
			/* reverse row {start, length} of each fired neuron, beside its
			 * fired-list entry (the next STDP work lists are built from it) */
			d_srow.alloc(size_t(m_flRing) * m_P * m_T * 8);
		}
	}
	hipStream_t s = static_cast<hipStream_t>(m_stream);
	if(d_recent.bytes) hipCheck(nemo_dev_fill_u64(d_recent.as<uint64_t>(), 0, m_n, s), "fill");
	if(d_ring.bytes) hipCheck(nemo_dev_fill_u32(d_ring.as<uint32_t>(), 0, d_ring.bytes / 4, s), "fill");
	if(d_rfg.bytes) hipCheck(nemo_dev_fill_u64(d_rfg.as<uint64_t>(), 0, size_t(2) * m_n, s), "fill");
	if(d_rcmAcc.bytes) hipCheck(nemo_dev_fill_u32(d_rcmAcc.as<uint32_t>(), 0, d_rcmAcc.bytes / 4, s), "fill");
	/* tagged fired counts: all-ones never matches a cycle's tag */
	if(d_fc.bytes) hipCheck(nemo_dev_fill_u64(d_fc.as<uint64_t>(), ~0ull, d_fc.bytes / 8, s), "fill");
	/* fired-list entries are tagged with their cycle: no entry is valid yet */
	if(d_fl.bytes) hipCheck(nemo_dev_fill_u64(d_fl.as<uint64_t>(), ~0ull, d_fl.bytes / 8, s), "fill");
	hipCheck(nemo_dev_fill_u32(d_ctl.as<uint32_t>(), 0, 16, s), "fill");

	if(m_fused) {
		std::vector<std::pair<uint32_t, uint32_t> > items;
		for(uint32_t p = 0; p < m_P; ++p) items.push_back(std::make_pair(0u, p));
		std::vector<WorkItem> work = xcdWorkTable(items);
		d_work.upload(work.data(), work.size() * sizeof(WorkItem));
	} else {
		/* pending current as int64 ring [2][D][n], converted inputs, bits */
		d_acc.alloc(std::max<size_t>(size_t(2) * m_D * m_n, 1) * 8);
		d_curE.alloc(std::max<size_t>(m_n, 1) * 4);
		d_curI.alloc(std::max<size_t>(m_n, 1) * 4);
		d_curX.alloc(std::max<size_t>(m_n, 1) * 4);
		d_fstim.alloc(std::max<size_t>(m_words32, 2) * 4);
		d_list.alloc(std::max<size_t>(m_n, 1) * 4 + 4);
		hipCheck(nemo_dev_fill_u64(d_acc.as<uint64_t>(), 0, d_acc.bytes / 8, s), "fill");
		loadPlugins(net);
	}
	hipCheck(hipHostMalloc(reinterpret_cast<void**>(&h_step), std::max<size_t>(m_words, 1) * 8), "hipHostMalloc");
	ensureRaster(8);
	hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize");
	devLock.unlock();

	m_deviceBytes = d_model.bytes + d_param.bytes + d_state.bytes + d_rng.bytes + d_recent.bytes
		+ d_ring.bytes + d_off.bytes + d_syn.bytes + d_fl.bytes + d_fc.bytes + d_rfg.bytes + d_srow.bytes
		+ d_rcmStart.bytes + d_rcmSd.bytes + d_rcmFwd.bytes + d_rcmAcc.bytes + d_acc.bytes;

	if(conf.logging) {
		fprintf(stdout, "nemosim HIP: %u neurons, %zu synapses, %s step, %u partitions x %u, max delay %u, "
				"%s accumulation, %.1f MB device\n", m_n, m_nsyn, m_fused ? "fused" : "plugin",
				m_P, m_T, m_D, m_saturate ? "saturating" : "int32", double(m_deviceBytes) / 1e6);
	}
	resetTimer();
}

void
Simulation::buildForwardOnDevice(const std::vector<unsigned>& Tg, const std::vector<unsigned char>& Dl,
		const std::vector<unsigned char>& Pl, uint32_t bflags, const uint32_t* lsrc, const uint32_t* ltgt,
		const int32_t* wfx, const uint8_t* dDl, const uint8_t* dPl)
{
	std::lock_guard<DeviceMutex> lock(deviceMutex());
	hipStream_t st = static_cast<hipStream_t>(m_stream);
	const uint32_t P1 = m_P + 1;
	const size_t nb = size_t(m_n) * P1;
	const uint32_t NEAR_SHIFT = 28, NEAR_MAX = 15;
	auto scan = [&](const uint32_t* in, uint32_t* out, size_t n) {
		size_t tb = 0;
		hipCheck(nemo_dev_exclusive_scan_u32(in, out, n, nullptr, &tb, st), "scan (size)");
		DevBuf temp;
		temp.alloc(std::max<size_t>(tb, 16));
		hipCheck(nemo_dev_exclusive_scan_u32(in, out, n, temp.ptr, &tb, st), "scan");
		hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
	};

	/* bucket sizes and their delay-1 part, then bucket starts */
	DevBuf dCount, dNear, dCursor, dFwd;
	dCount.alloc((nb + 1) * 4);
	dNear.alloc(nb * 4);
	hipCheck(nemo_dev_fill_u32(dCount.as<uint32_t>(), 0, nb + 1, st), "fill");
	hipCheck(nemo_dev_fill_u32(dNear.as<uint32_t>(), 0, nb, st), "fill");
	hipCheck(nemo_dev_build_count(lsrc, ltgt, dDl, m_nsyn, m_T, P1, dCount.as<uint32_t>(), dNear.as<uint32_t>(), st),
			"construction kernel");
	d_off.alloc((nb + 1) * 4);
	scan(dCount.as<uint32_t>(), d_off.as<uint32_t>(), nb + 1);
	dCount.alloc(0);

	/* records: the delay-1 records of every bucket first */
	dCursor.alloc(nb * 4);
	hipCheck(hipMemcpyAsync(dCursor.ptr, d_off.ptr, nb * 4, hipMemcpyDeviceToDevice, st), "hipMemcpyAsync D2D");
	d_syn.alloc(std::max<size_t>(m_nsyn, 1) * 8);
	dFwd.alloc(m_nsyn * 4);
	for(int pass = 0; pass < 2; ++pass) {
		hipCheck(nemo_dev_build_scatter(lsrc, ltgt, dDl, wfx, m_nsyn, m_T, P1, pass, dCursor.as<uint32_t>(),
					d_syn.as<uint32_t>(), dFwd.as<uint32_t>(), st), "construction kernel");
	}
	dCursor.alloc(0);

	/* per-source query starts (before the delay-1 counts go into off) */
	if(!m_writeOnly && !(bflags & NEMO_BUILD_UNORDERED)) {
		DevBuf dq;
		dq.alloc((size_t(m_n) + 1) * 4);
		hipCheck(nemo_dev_build_gather(d_off.as<uint32_t>(), P1, m_n, dq.as<uint32_t>(), st), "construction kernel");
		std::vector<uint32_t> q(size_t(m_n) + 1);
		hipCheck(hipMemcpyAsync(q.data(), dq.ptr, q.size() * 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync D2H");
		hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
		m_qStart.assign(q.begin(), q.end());
	}

	/* deferred delivery: every bucket holds at most NEAR_MAX delay-1 records
	 * and the record indices leave the top bits free */
	m_defer = uint64_t(m_nsyn) < (uint64_t(1) << NEAR_SHIFT);
	if(m_defer) {
		DevBuf dMax;
		dMax.alloc(4);
		uint32_t mx = 0;
		hipCheck(nemo_dev_fill_u32(dMax.as<uint32_t>(), 0, 1, st), "fill");
		hipCheck(nemo_dev_build_max(dNear.as<uint32_t>(), nb, dMax.as<uint32_t>(), st), "construction kernel");
		hipCheck(hipMemcpyAsync(&mx, dMax.ptr, 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync D2H");
		hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
		m_defer = mx <= NEAR_MAX;
	}
	if(m_defer) {
		hipCheck(nemo_dev_build_near_pack(d_off.as<uint32_t>(), dNear.as<uint32_t>(), nb, NEAR_SHIFT, st),
				"construction kernel");
	}
	dNear.alloc(0);

	/* synapse query tables, ordinal order per source: with the sources in
	 * ascending local order that is the network's own order */
	if(!m_writeOnly) {
		std::vector<uint32_t> fwd(m_nsyn);
		hipCheck(hipMemcpyAsync(fwd.data(), dFwd.ptr, m_nsyn * 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync D2H");
		std::vector<uint32_t> ls;
		if(bflags & NEMO_BUILD_UNORDERED) {
			ls.resize(m_nsyn);
			hipCheck(hipMemcpyAsync(ls.data(), lsrc, m_nsyn * 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync D2H");
		}
		hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
		if(bflags & NEMO_BUILD_UNORDERED) {
			m_qStart.assign(size_t(m_n) + 1, 0);
			for(size_t i = 0; i < m_nsyn; ++i) m_qStart[ls[i] + 1] += 1;
			for(uint32_t l = 0; l < m_n; ++l) m_qStart[l + 1] += m_qStart[l];
			std::vector<uint64_t> qc(m_qStart.begin(), m_qStart.end() - 1);
			m_qTarget.resize(m_nsyn);
			m_qDelay.resize(m_nsyn);
			m_qPlastic.resize(m_nsyn);
			m_qFwd.resize(m_nsyn);
			for(size_t i = 0; i < m_nsyn; ++i) {
				uint64_t q = qc[ls[i]]++;
				m_qTarget[q] = Tg[i];
				m_qDelay[q] = Dl[i];
				m_qPlastic[q] = Pl[i];
				m_qFwd[q] = fwd[i];
			}
		} else {
			m_qTarget.assign(Tg.begin(), Tg.end());
			m_qDelay.assign(Dl.begin(), Dl.end());
			m_qPlastic.assign(Pl.begin(), Pl.end());
			m_qFwd.swap(fwd);
		}
	}

	/* reverse matrix of plastic synapses (src/nemo/construction/RCM.cpp:150-171) */
	if(m_stdp) {
		DevBuf dRc;
		dRc.alloc((size_t(m_n) + 1) * 4);
		hipCheck(nemo_dev_fill_u32(dRc.as<uint32_t>(), 0, size_t(m_n) + 1, st), "fill");
		hipCheck(nemo_dev_build_rcm_count(ltgt, dPl, m_nsyn, dRc.as<uint32_t>(), st), "construction kernel");
		d_rcmStart.alloc((size_t(m_n) + 1) * 4);
		scan(dRc.as<uint32_t>(), d_rcmStart.as<uint32_t>(), size_t(m_n) + 1);
		hipCheck(hipMemcpyAsync(dRc.ptr, d_rcmStart.ptr, size_t(m_n) * 4, hipMemcpyDeviceToDevice, st),
				"hipMemcpyAsync D2D");
		d_rcmSd.alloc(std::max<size_t>(m_nplastic, 1) * 4);
		d_rcmFwd.alloc(std::max<size_t>(m_nplastic, 1) * 4);
		hipCheck(nemo_dev_build_rcm_scatter(lsrc, ltgt, dDl, dPl, dFwd.as<uint32_t>(), m_nsyn, dRc.as<uint32_t>(),
					d_rcmSd.as<uint32_t>(), d_rcmFwd.as<uint32_t>(), st), "construction kernel");
	}
	hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
}

/* Load lib<Model>.so for every neuron type (src/nemo/cuda/Neurons.cpp:43-44,
 * src/nemo/Plugin.cpp:84-121): the descriptor's own directory first, then
 * the plugin path; each directory's plugins live in its hip/ subdirectory. */
void
Simulation::loadPlugins(const NetworkImpl& net)
{
	unsigned base = 0;
	for(const NeuronGroup& g : net.groups) {
		unsigned cnt = unsigned(g.gidx.size());
		if(cnt == 0) continue;
		std::vector<std::string> dirs;
		if(!g.model->dir.empty()) dirs.push_back(g.model->dir);
		for(const std::string& d : pluginPaths()) dirs.push_back(d);
		void* h = nullptr;
		std::string tried;
		for(const std::string& d : dirs) {
			std::string path = d + "/hip/lib" + g.model->name + ".so";
			h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
			if(h) break;
			tried += " " + path;
		}
		if(!h) {
			throw exception(NEMO_DL_ERROR, fmt("No HIP plugin for neuron model ", g.model->name, " (tried", tried, ")"));
		}
		PluginGroup pg;
		pg.start = base;
		pg.end = base + cnt;
		pg.model = g.model->id;
		pg.history = g.model->history;
		pg.handle = h;
		pg.update = reinterpret_cast<hip_update_neurons_t*>(dlsym(h, "hip_update_neurons"));
		hip_init_neurons_t* init = reinterpret_cast<hip_init_neurons_t*>(dlsym(h, "hip_init_neurons"));
		if(!pg.update || !init) {
			dlclose(h);
			throw exception(NEMO_DL_ERROR, fmt("Plugin for ", g.model->name, " lacks hip_update_neurons / hip_init_neurons"));
		}
		m_plugins.push_back(pg);
		hipCheck(init(m_stream, pg.start, pg.end, d_param.as<float>() + pg.start, m_n,
					d_state.as<float>() + pg.start, size_t(m_ns) * m_n, m_n, d_rng.as<uint32_t>() + pg.start, m_n),
				"hip_init_neurons");
		base += cnt;
	}
}

Simulation::~Simulation()
{
	/* Without the device lock: nothing here dispatches a kernel (the lock
	 * keeps kernels away from the start of a persistent launch), and a
	 * release must not queue behind the uploads of simulations being built.
	 * The buffers are freed by their destructors after this body. */
	/* the stream and events go back to the pool once the stream is idle */
	bool pooled = false;
	if(m_stream) {
		(void) hipStreamSynchronize(static_cast<hipStream_t>(m_stream));
		if(m_ev0 && m_ev1) {
			pooled = giveStreamSet(m_device, StreamSet{static_cast<hipStream_t>(m_stream),
					static_cast<hipEvent_t>(m_ev0), static_cast<hipEvent_t>(m_ev1)});
		}
		if(!pooled) (void) hipStreamDestroy(static_cast<hipStream_t>(m_stream));
	}
	if(!pooled) {
		if(m_ev0) (void) hipEventDestroy(static_cast<hipEvent_t>(m_ev0));
		if(m_ev1) (void) hipEventDestroy(static_cast<hipEvent_t>(m_ev1));
	}
	if(h_ctl) (void) hipHostFree(h_ctl);
	if(h_step) (void) hipHostFree(h_step);
	if(h_sched) (void) hipHostFree(h_sched);
	for(PluginGroup& pg : m_plugins) {
		if(pg.handle) dlclose(pg.handle);
	}
}

unsigned
Simulation::localIdx(unsigned g) const
{
	if(g >= m_minG && size_t(g - m_minG) < m_dense.size()) {
		uint32_t l = m_dense[g - m_minG];
		if(l) return l - 1;
	}
	throw exception(NEMO_INVALID_INPUT, fmt("Non-existing neuron index ", g));
}

/* Stimulus schedule to device CSR in simulation-local indices; forced
 * firing is idempotent and a current assignment keeps the last value per
 * neuron and cycle (src/nemo/cpu/Simulation.cpp:140-163). */
bool
Simulation::translateSchedule(unsigned nsteps, const nemo_schedule_t* s, std::vector<uint32_t>& fsPtr,
		std::vector<uint32_t>& fsIdx, std::vector<uint32_t>& isPtr, std::vector<uint32_t>& isIdx,
		std::vector<float>& isVal) const
{
	fsPtr.clear();
	fsIdx.clear();
	isPtr.clear();
	isIdx.clear();
	isVal.clear();
	if(!s) return false;
	bool anyF = s->fstim_ptr && s->fstim_idx;
	bool anyI = s->istim_ptr && s->istim_idx && s->istim_val;
	if(!anyF && !anyI) return false;
	fsPtr.assign(size_t(nsteps) + 1, 0);
	isPtr.assign(size_t(nsteps) + 1, 0);
	std::unordered_map<uint32_t, size_t> last;
	for(unsigned k = 0; k < nsteps; ++k) {
		if(anyF) {
			for(unsigned j = s->fstim_ptr[k]; j < s->fstim_ptr[k + 1]; ++j) {
				fsIdx.push_back(localIdx(s->fstim_idx[j]));
			}
		}
		fsPtr[k + 1] = uint32_t(fsIdx.size());
		if(anyI) {
			last.clear();
			for(unsigned j = s->istim_ptr[k]; j < s->istim_ptr[k + 1]; ++j) {
				uint32_t l = localIdx(s->istim_idx[j]);
				auto it = last.find(l);
				if(it == last.end()) {
					last[l] = isIdx.size();
					isIdx.push_back(l);
					isVal.push_back(s->istim_val[j]);
				} else {
					isVal[it->second] = s->istim_val[j];
				}
			}
		}
		isPtr[k + 1] = uint32_t(isIdx.size());
	}
	return true;
}

void
Simulation::uploadSchedule(unsigned nsteps, const nemo_schedule_t* s)
{
	std::lock_guard<DeviceMutex> devLock(deviceMutex());
	m_hasSchedule = false;
	std::vector<uint32_t> fsPtr, fsIdx, isPtr, isIdx;
	std::vector<float> isVal;
	if(!translateSchedule(nsteps, s, fsPtr, fsIdx, isPtr, isIdx, isVal)) return;
	/* one pinned staging block, one copy, no wait: the launch that reads the
	 * schedule follows on the same stream and ends with a synchronisation,
	 * before this staging block is written again */
	const size_t words = fsPtr.size() + isPtr.size() + fsIdx.size() + 2 * isIdx.size();
	if(words * 4 > m_schedHostBytes) {
		const size_t bytes = std::max(words * 4, size_t(2) * m_schedHostBytes + 4096);
		if(h_sched) (void) hipHostFree(h_sched);
		h_sched = nullptr;
		m_schedHostBytes = 0;
		hipCheck(hipHostMalloc(reinterpret_cast<void**>(&h_sched), bytes), "hipHostMalloc");
		m_schedHostBytes = bytes;
	}
	uint32_t* w = h_sched;
	w = std::copy(fsPtr.begin(), fsPtr.end(), w);
	w = std::copy(isPtr.begin(), isPtr.end(), w);
	w = std::copy(fsIdx.begin(), fsIdx.end(), w);
	w = std::copy(isIdx.begin(), isIdx.end(), w);
	if(!isVal.empty()) std::memcpy(w, isVal.data(), isVal.size() * 4);
	d_sched.ensure(words * 4);
	hipCheck(hipMemcpyAsync(d_sched.ptr, h_sched, words * 4, hipMemcpyHostToDevice,
				static_cast<hipStream_t>(m_stream)), "H2D");
	const uint32_t* base = d_sched.as<uint32_t>();
	m_sFsPtr = base;
	m_sIsPtr = m_sFsPtr + fsPtr.size();
	m_sFsIdx = m_sIsPtr + isPtr.size();
	m_sIsIdx = m_sFsIdx + fsIdx.size();
	m_sIsVal = reinterpret_cast<const float*>(m_sIsIdx + isIdx.size());
	m_fsPtrHost.swap(fsPtr);
	m_isPtrHost.swap(isPtr);
	m_hasSchedule = true;
}

/* Exchange waves (NetDev.exWaves): with deferred delivery and LDS left for
 * the far list while the workgroups stay co-resident, that many waves of
 * each workgroup (at least one other must remain) run the exchange and the
 * delay-1 records, the rest the far records of the previous cycle. The far
 * list takes up to 1536 sources per cycle, fewer as LDS allows, and at
 * least 256; the sources of a cycle beyond it are delivered whole by the
 * exchange waves. The configuration gives the count, or none (< 0), or
 * leaves it to this rule (0): with STDP and at least 8 waves per
 * workgroup, 3 of every 8 waves. STDP adds the work-list build and the
 * accumulation to the waves that do not wait on the exchange (50k x 1000:
 * 21.6 -> 19.6 ms per simulated second at 3 of 8, 20.4 at 1, 20.0 at 2 and
 * 4); without STDP the far records alone gain less than the exchange loses
 * queued behind them. Batched launches leave out exchange waves chosen by
 * this rule (prepareLaunch). */
void
Simulation::planExchange(const ConfigurationImpl& conf, uint32_t ldsLimit)
{
	m_exWaves = 0;
	m_farCap = 0;
	m_exAuto = false;
	m_ldsNoFar = m_lds;
	const uint32_t nw = m_T / 64u;
	if(!m_fused || !m_defer || m_P == 0 || conf.exchangeWaves < 0) return;
	uint32_t ex = uint32_t(conf.exchangeWaves);
	if(conf.exchangeWaves == 0) {
		if(!m_stdp || nw < 8u) return;
		ex = (3u * nw + 4u) / 8u;
		m_exAuto = true;
	}
	if(ex + 1u > nw) return;
	for(uint32_t cap = 1536; cap >= 256; cap -= 64) {
		const uint32_t lds = m_lds + NEMO_FAR_LDS_BYTES(cap);
		if(lds > ldsLimit) continue;
		uint32_t maxBlocks = 0;
		hipCheck(nemo_dev_step_max_blocks(m_T, lds, variant(), &maxBlocks), "occupancy query");
		if(m_P <= maxBlocks) {
			m_exWaves = ex;
			m_farCap = cap;
			m_lds = lds;
			return;
		}
	}
}

NetDev
Simulation::prepareLaunch(unsigned nsteps, bool batched)
{
	hipCheck(hipSetDevice(m_device), "hipSetDevice");
	ensureRaster(size_t(nsteps) * m_words * 8);
	m_rasterExternal = false;
	NetDev nd = descriptor(nsteps);
	if(batched && m_exAuto) {
		/* every launch delivers its last cycle's far records before it
		 * ends, so launches with and without exchange waves can follow
		 * each other */
		nd.exWaves = 0;
		nd.farCap = 0;
	}
	return nd;
}

/* The raster buffer holds at least `bytes`. It starts at 1024 cycles (at
 * most 32 MB) and doubles when a run needs more, so runs of varying length
 * do not reallocate it (hipMalloc and hipFree synchronise with the device
 * and cost tens of microseconds). */
void
Simulation::ensureRaster(size_t bytes)
{
	bytes = std::max<size_t>(bytes, 8);
	if(bytes <= d_raster.bytes) return;
	const size_t perCycle = std::max<size_t>(size_t(m_words) * 8, 8);
	size_t cap = std::max<size_t>(d_raster.bytes * 2, perCycle * std::max<size_t>(1, std::min<size_t>(1024,
					(size_t(32) << 20) / perCycle)));
	d_raster.alloc(std::max(cap, bytes));
}

/* The launch descriptor: network arrays, per-launch fields, the schedule
 * uploaded for this launch (if any) and the raster buffer d_raster. */
NetDev
Simulation::descriptor(unsigned nsteps) const
{
	NetDev nd;
	std::memset(&nd, 0, sizeof(nd));
	nd.n = m_n;
	nd.P = m_P;
	nd.T = m_T;
	nd.D = m_D;
	nd.words = m_words;
	nd.flags = (m_stdp ? NEMO_NET_STDP : 0u) | (m_saturate ? NEMO_NET_SATURATE : 0u)
		| (m_replay ? NEMO_NET_REPLAY : 0u) | (m_defer ? NEMO_NET_DEFER : 0u)
		| (m_izOnly ? NEMO_NET_IZHIKEVICH_ONLY : 0u) | (m_prefRing ? NEMO_NET_PREF_RING : 0u);
	nd.nsteps = nsteps;
	nd.spinLimit = m_spinLimit;
	nd.glog2 = m_glog2;
	nd.flRing = m_flRing;
	nd.stdpCache = m_stdpCache;
	nd.exWaves = m_exWaves;
	nd.farCap = m_farCap;
	nd.cycle0 = m_cycle;
	nd.model = d_model.as<uint8_t>();
	nd.param = d_param.as<float>();
	nd.state = d_state.as<float>();
	nd.rng = d_rng.as<uint32_t>();
	nd.recent = d_recent.as<uint64_t>();
	nd.ring = d_ring.as<int32_t>();
	nd.off = d_off.as<uint32_t>();
	nd.syn = d_syn.as<uint32_t>();
	nd.fl = d_fl.as<uint64_t>();
	nd.fc = d_fc.as<uint64_t>();
	nd.rfg = d_rfg.as<uint64_t>();
	nd.status = d_ctl.as<uint32_t>() + 1;
	nd.events = reinterpret_cast<unsigned long long*>(d_ctl.as<uint32_t>() + 2);
	nd.rcm_start = d_rcmStart.as<uint32_t>();
	nd.rcm_sd = d_rcmSd.as<uint32_t>();
	nd.rcm_fwd = d_rcmFwd.as<uint32_t>();
	nd.rcm_acc = d_rcmAcc.as<int32_t>();
	nd.stdp = d_stdp.ptr;
	nd.srow = d_srow.as<uint64_t>();
	if(m_hasSchedule) {
		nd.fs_ptr = m_sFsPtr;
		nd.fs_idx = m_sFsIdx;
		nd.is_ptr = m_sIsPtr;
		nd.is_idx = m_sIsIdx;
		nd.is_val = m_sIsVal;
	}
	nd.raster = d_raster.as<uint64_t>();
	return nd;
}

/* The control block as a single launch left it (h_ctl, already copied
 * back): its counts relative to the previous launch's, then the launch's
 * completion. After an error the block is cleared on the device. */
void
Simulation::readControl(unsigned nsteps, unsigned long long* events)
{
	uint32_t raw[8], rel[8];
	std::memcpy(raw, h_ctl, sizeof(raw));
	std::memcpy(rel, raw, sizeof(rel));
	unsigned long long e;
	std::memcpy(&e, &raw[2], sizeof(e));
	e -= m_ctlEvents;
	std::memcpy(&rel[2], &e, sizeof(e));
	rel[0] = raw[0] - m_ctlStarted;
	rel[7] = raw[7] - m_ctlStarted;
	if(events) *events = e;
	setControlBase(raw);
	completeLaunch(nsteps, rel);
}

void
Simulation::setControlBase(const uint32_t ctl[8])
{
	if(ctl[1] != 0u) {
		/* a failed launch: start the next one from a clear block */
		hipStream_t st = static_cast<hipStream_t>(m_stream);
		hipCheck(nemo_dev_fill_u32(d_ctl.as<uint32_t>(), 0, NEMO_CTL_WORDS, st), "fill");
		hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
		m_ctlStarted = 0;
		m_ctlEvents = 0;
		return;
	}
	m_ctlStarted = ctl[0];
	std::memcpy(&m_ctlEvents, &ctl[2], sizeof(m_ctlEvents));
}

void
Simulation::completeLaunch(unsigned nsteps, const uint32_t ctl[8])
{
	m_cycle += nsteps;
	m_timerCycles += nsteps;
	m_lastRunSteps = nsteps;
	m_hasSchedule = false;
	m_replay = false;
	if(ctl[1] & (NEMO_STATUS_BARRIER_TIMEOUT | NEMO_STATUS_ENTRY_TIMEOUT)) {
		throw exception(NEMO_CUDA_ERROR, fmt("step kernel: a wait for another workgroup timed out (",
					(ctl[1] & NEMO_STATUS_BARRIER_TIMEOUT) ? "fired count" : "fired-list entry",
					"; cycle ", ctl[4], ", partition ", ctl[5], " waiting for partition ", ctl[6],
					"; ", ctl[7], " of ", m_P, " workgroups had started, ", ctl[0], " by the end)"));
	}
}

/* The modular step: per cycle, convert the due ring slot, apply stimulus,
 * run each type's plugin, compact the fired bits, scatter the fired
 * sources' synapses into the int64 ring, accumulate STDP
 * (src/nemo/cuda/Simulation.cpp:336-430 has the same phase structure). */
void
Simulation::launchPlugins(unsigned nsteps, unsigned long long* events)
{
	std::lock_guard<DeviceMutex> devLock(deviceMutex());
	hipStream_t st = static_cast<hipStream_t>(m_stream);
	hipCheck(hipSetDevice(m_device), "hipSetDevice");
	ensureRaster(size_t(nsteps) * m_words32 * 4);
	hipCheck(nemo_dev_fill_u32(d_ctl.as<uint32_t>(), 0, 16, st), "fill");
	m_rasterExternal = false;
	PluginStepDev s;
	std::memset(&s, 0, sizeof(s));
	s.n = m_n;
	s.D = m_D;
	s.words32 = m_words32;
	s.cycle0 = m_cycle;
	s.acc = d_acc.as<long long>();
	s.curE = d_curE.as<float>();
	s.curI = d_curI.as<float>();
	s.curX = d_curX.as<float>();
	s.fstim = d_fstim.as<uint32_t>();
	s.raster = d_raster.as<uint32_t>();
	s.list = d_list.as<uint32_t>();
	s.nfired = d_list.as<uint32_t>() + m_n;
	s.off = d_off.as<uint32_t>();
	s.syn = d_syn.as<uint32_t>();
	s.recent = d_recent.as<uint64_t>();
	s.events = reinterpret_cast<unsigned long long*>(d_ctl.as<uint32_t>() + 2);
	s.rcm_start = d_rcmStart.as<uint32_t>();
	s.rcm_sd = d_rcmSd.as<uint32_t>();
	s.rcm_acc = d_rcmAcc.as<int32_t>();
	s.stdp = d_stdp.ptr;
	if(m_hasSchedule) {
		s.fs_ptr = m_sFsPtr;
		s.fs_idx = m_sFsIdx;
		s.is_ptr = m_sIsPtr;
		s.is_idx = m_sIsIdx;
		s.is_val = m_sIsVal;
	}
	hipEvent_t e0 = static_cast<hipEvent_t>(m_ev0), e1 = static_cast<hipEvent_t>(m_ev1);
	hipCheck(hipEventRecord(e0, st), "hipEventRecord");
	if(m_replay) {
		/* weights changed: rebuild the pending current from the history */
		hipCheck(nemo_dev_fill_u64(d_acc.as<uint64_t>(), 0, d_acc.bytes / 8, st), "fill");
		hipCheck(nemo_dev_pm_replay(&s, st), "plugin step: replay");
		m_replay = false;
	}
	for(unsigned k = 0; k < nsteps; ++k) {
		hipCheck(nemo_dev_pm_begin(&s, k, st), "plugin step: begin");
		if(m_hasSchedule) {
			uint32_t entries = (m_fsPtrHost[k + 1] - m_fsPtrHost[k]) + (m_isPtrHost[k + 1] - m_isPtrHost[k]);
			hipCheck(nemo_dev_pm_stimulus(&s, k, entries, st), "plugin step: stimulus");
		}
		uint32_t* firedRow = s.raster + size_t(k) * m_words32;
		/* the cycle a plugin sees is the timer's (src/nemo/cpu/Simulation.cpp:98) */
		const unsigned tcycle = unsigned(m_timerCycles + k);
		for(const PluginGroup& pg : m_plugins) {
			hipCheck(pg.update(m_stream, tcycle, pg.start, pg.end,
						d_param.as<float>() + pg.start, m_n, d_state.as<float>() + pg.start, size_t(m_ns) * m_n, m_n,
						d_rng.as<uint32_t>() + pg.start, m_n, s.fstim, s.curE, s.curI, s.curX,
						s.recent, firedRow, m_hasRcm ? &m_rcmDesc : nullptr), "hip_update_neurons");
		}
		hipCheck(nemo_dev_pm_compact(&s, k, st), "plugin step: compact");
		hipCheck(nemo_dev_pm_scatter(&s, k, st), "plugin step: scatter");
		if(m_stdp) {
			hipCheck(nemo_dev_pm_stdp(&s, st), "plugin step: stdp");
		}
	}
	hipCheck(hipEventRecord(e1, st), "hipEventRecord");
	uint32_t ctl[8];
	hipCheck(hipMemcpyAsync(h_ctl, d_ctl.ptr, sizeof(ctl), hipMemcpyDeviceToHost, st), "D2H");
	hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
	std::memcpy(ctl, h_ctl, sizeof(ctl));
	setControlBase(ctl);
	float ms = 0.0f;
	hipCheck(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
	m_kernelMs = ms;
	/* the last update of the run, of timer cycle c, made plane (c+1) mod H
	 * current (src/nemo/cpu/Neurons.cpp:123) */
	for(const PluginGroup& pg : m_plugins) {
		if(pg.history > 1) {
			m_plane.at(pg.model) = (unsigned(m_timerCycles + nsteps - 1) + 1u) % pg.history;
		}
	}
	m_cycle += nsteps;
	m_timerCycles += nsteps;
	m_lastRunSteps = nsteps;
	m_hasSchedule = false;
	if(events) {
		unsigned long long e;
		std::memcpy(&e, &ctl[2], sizeof(e));
		*events = e;
	}
}

void
Simulation::launch(unsigned nsteps, unsigned long long* events)
{
	std::lock_guard<DeviceMutex> devLock(deviceMutex());
	if(events) *events = 0;
	if(!m_fused && m_n > 0 && nsteps > 0) {
		launchPlugins(nsteps, events);
		return;
	}
	if(m_P == 0 || m_n == 0 || nsteps == 0) {
		m_cycle += nsteps;
		m_timerCycles += nsteps;
		m_lastRunSteps = nsteps;
		m_hasSchedule = false;
		return;
	}
	launchFused(nsteps, false);
	readControl(nsteps, events);
}

/* One launch of the fused kernel: the descriptor goes by value, HIP events
 * on the kernel's own stream time the kernel alone, and one copy brings back
 * the control block (and, for a single step, its raster words) into pinned
 * memory, followed by the launch's only synchronisation. */
void
Simulation::launchFused(unsigned nsteps, bool stepOut)
{
	launchDescriptor(prepareLaunch(nsteps), stepOut);
}

void
Simulation::launchDescriptor(const NetDev& nd, bool stepOut)
{
	hipStream_t st = static_cast<hipStream_t>(m_stream);
	hipEvent_t e0 = static_cast<hipEvent_t>(m_ev0), e1 = static_cast<hipEvent_t>(m_ev1);
	hipCheck(hipEventRecord(e0, st), "hipEventRecord");
	hipCheck(nemo_dev_step_launch_one(&nd, d_work.as<WorkItem>(), m_P, m_T, m_lds, variant(), st),
			"step kernel launch");
	hipCheck(hipEventRecord(e1, st), "hipEventRecord");
	hipCheck(hipMemcpyAsync(h_ctl, d_ctl.ptr, 32, hipMemcpyDeviceToHost, st), "D2H");
	if(stepOut && m_words) {
		hipCheck(hipMemcpyAsync(h_step, d_raster.ptr, size_t(m_words) * 8, hipMemcpyDeviceToHost, st), "D2H");
	}
	hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
	float ms = 0.0f;
	hipCheck(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
	m_kernelMs = ms;
}

static_assert(size_t(NEMO_STATIC_SCRATCH_WORDS) * 4u == size_t(NEMO_HIP_STATIC_SCRATCH_BYTES),
		"the scratch array and its published size agree");

/* nemo_hip_run without any device allocation: the stimulus schedule is
 * staged in the per-device scratch array, the raster goes to the caller's
 * device buffer, and everything else the kernel touches was allocated when
 * the simulation was built. */
void
Simulation::runStatic(unsigned nsteps, const nemo_schedule_t* schedule, uint64_t* rasterDev,
		unsigned long long* events)
{
	std::lock_guard<DeviceMutex> devLock(deviceMutex());
	if(events) *events = 0;
	if(!m_fused) {
		throw exception(NEMO_API_UNSUPPORTED,
				"the allocation-free run needs the fused step kernel; this simulation runs on the plugin step");
	}
	const bool has = translateSchedule(nsteps, schedule, m_stFsPtr, m_stFsIdx, m_stIsPtr, m_stIsIdx, m_stIsVal);
	const size_t words = has ? m_stFsPtr.size() + m_stIsPtr.size() + m_stFsIdx.size() + 2 * m_stIsIdx.size() : 0;
	if(words > NEMO_STATIC_SCRATCH_WORDS) {
		throw exception(NEMO_INVALID_INPUT, fmt("stimulus schedule of ", words,
					" words is larger than the allocation-free run's scratch of ", NEMO_STATIC_SCRATCH_WORDS,
					" 4-byte words; use nemo_hip_run"));
	}
	if(nsteps == 0 || m_P == 0 || m_n == 0) {
		m_cycle += nsteps;
		m_timerCycles += nsteps;
		m_lastRunSteps = nsteps;
		m_rasterExternal = true;
		return;
	}
	if(!rasterDev) throw exception(NEMO_INVALID_INPUT, "null raster buffer");
	hipCheck(hipSetDevice(m_device), "hipSetDevice");
	void* scr = nullptr;
	hipCheck(nemo_dev_static_scratch(&scr), "allocation-free run: scratch address");
	m_hasSchedule = false;
	NetDev nd = descriptor(nsteps);
	nd.raster = rasterDev;
	nd.fs_ptr = nullptr;
	nd.fs_idx = nullptr;
	nd.is_ptr = nullptr;
	nd.is_idx = nullptr;
	nd.is_val = nullptr;
	if(has) {
		/* one staging array, one copy: [fired ptr][current ptr][fired idx]
		 * [current idx][current value bits] */
		m_stage.resize(words);
		uint32_t* w = m_stage.data();
		w = std::copy(m_stFsPtr.begin(), m_stFsPtr.end(), w);
		w = std::copy(m_stIsPtr.begin(), m_stIsPtr.end(), w);
		w = std::copy(m_stFsIdx.begin(), m_stFsIdx.end(), w);
		w = std::copy(m_stIsIdx.begin(), m_stIsIdx.end(), w);
		if(!m_stIsVal.empty()) std::memcpy(w, m_stIsVal.data(), m_stIsVal.size() * 4);
		hipCheck(hipMemcpyAsync(scr, m_stage.data(), words * 4, hipMemcpyHostToDevice,
					static_cast<hipStream_t>(m_stream)), "H2D");
		const uint32_t* base = static_cast<const uint32_t*>(scr);
		nd.fs_ptr = base;
		nd.is_ptr = base + m_stFsPtr.size();
		nd.fs_idx = nd.is_ptr + m_stIsPtr.size();
		nd.is_idx = nd.fs_idx + m_stFsIdx.size();
		nd.is_val = reinterpret_cast<const float*>(nd.is_idx + m_stIsIdx.size());
	}
	launchDescriptor(nd, false);
	readControl(nsteps, events);
	m_rasterExternal = true;
}

void
Simulation::decodeStep(const uint64_t* words, std::vector<unsigned>& out) const
{
	out.clear();
	for(uint32_t w = 0; w < m_words; ++w) {
		uint64_t b = words[w];
		while(b) {
			uint32_t l = w * 64u + uint32_t(__builtin_ctzll(b));
			b &= b - 1;
			if(l < m_n) out.push_back(m_localToGlobal[l]);
		}
	}
}

const std::vector<unsigned>&
Simulation::step(const std::vector<unsigned>& fstim, const unsigned* istimIdx, const float* istimVal, size_t nistim)
{
	std::lock_guard<DeviceMutex> devLock(deviceMutex());
	unsigned fp[2] = { 0, unsigned(fstim.size()) };
	unsigned ip[2] = { 0, unsigned(nistim) };
	nemo_schedule_t s;
	s.fstim_ptr = fp;
	s.fstim_idx = fstim.empty() ? nullptr : fstim.data();
	s.istim_ptr = ip;
	s.istim_idx = istimIdx;
	s.istim_val = istimVal;
	if(fstim.empty()) { s.fstim_ptr = nullptr; }
	if(nistim == 0) { s.istim_ptr = nullptr; s.istim_idx = nullptr; s.istim_val = nullptr; }
	uploadSchedule(1, &s);
	if(m_fused && m_P > 0 && m_n > 0) {
		/* the raster words come back with the control block */
		launchFused(1, true);
		readControl(1, nullptr);
		decodeStep(h_step, m_fired);
		return m_fired;
	}
	launch(1, nullptr);
	m_rasterHost.resize(m_words);
	if(m_words) {
		hipCheck(hipMemcpy(m_rasterHost.data(), d_raster.ptr, size_t(m_words) * 8, hipMemcpyDeviceToHost), "D2H");
	}
	decodeStep(m_rasterHost.data(), m_fired);
	return m_fired;
}

void
Simulation::run(unsigned nsteps, const nemo_schedule_t* schedule, unsigned long long* events)
{
	uploadSchedule(nsteps, schedule);
	launch(nsteps, events);
}

void
Simulation::readFired(unsigned counts[], unsigned ids[], size_t cap, size_t* total)
{
	std::lock_guard<DeviceMutex> devLock(deviceMutex());
	requireOwnRaster();
	size_t nw = size_t(m_lastRunSteps) * m_words;
	m_rasterHost.resize(nw);
	if(nw) {
		hipCheck(hipMemcpy(m_rasterHost.data(), d_raster.ptr, nw * 8, hipMemcpyDeviceToHost), "D2H");
	}
	size_t k = 0;
	std::vector<unsigned> one;
	for(unsigned t = 0; t < m_lastRunSteps; ++t) {
		decodeStep(m_rasterHost.data() + size_t(t) * m_words, one);
		if(counts) counts[t] = unsigned(one.size());
		if(ids) {
			if(k + one.size() > cap) {
				throw exception(NEMO_BUFFER_OVERFLOW, "fired-id buffer too small");
			}
			std::copy(one.begin(), one.end(), ids + k);
		}
		k += one.size();
	}
	if(total) *total = k;
}

void
Simulation::readRaster(uint64_t words[], size_t cap, size_t* wordsPerStep)
{
	std::lock_guard<DeviceMutex> devLock(deviceMutex());
	size_t nw = size_t(m_lastRunSteps) * m_words;
	if(wordsPerStep) *wordsPerStep = m_words;
	if(words) {
		requireOwnRaster();
		if(cap < nw) throw exception(NEMO_BUFFER_OVERFLOW, "raster buffer too small");
		if(nw) hipCheck(hipMemcpy(words, d_raster.ptr, nw * 8, hipMemcpyDeviceToHost), "D2H");
	}
}

void
Simulation::requireOwnRaster() const
{
	if(m_rasterExternal) {
		throw exception(NEMO_INVALID_INPUT,
				"the last run wrote its raster to the caller's buffer (nemo_hip_run_static)");
	}
}

void
Simulation::info(nemo_hip_info_t* out) const
{
	out->neurons = m_n;
	out->partitions = m_P;
	out->partition_size = m_T;
	out->max_delay = m_D;
	out->saturating = m_saturate ? 1 : 0;
	out->stdp = m_stdp ? 1 : 0;
	out->synapses = m_nsyn;
	out->plastic_synapses = m_nplastic;
	out->device_bytes = m_deviceBytes;
	out->lds_bytes = m_lds;
	out->block_threads = m_T;
	out->last_kernel_ms = m_kernelMs;
	out->fused = m_fused ? 1u : 0u;
	out->deferred = m_defer ? 1u : 0u;
	out->lanes_per_source = m_fused ? (1u << m_glog2) : 0u;
	out->stdp_staged_chunks = m_stdpCache;
	out->stdp_prefix_ring = m_prefRing ? 1u : 0u;
	out->exchange_waves = m_exWaves;
	out->far_list_sources = m_farCap;
	out->device_built = m_deviceBuilt ? 1u : 0u;
}

/* src/nemo/ConnectivityMatrix.cpp:241-294 */
void
Simulation::applyStdp(float reward)
{
	std::lock_guard<DeviceMutex> devLock(deviceMutex());
	applyStdpLaunch(reward);
	applyStdpFinish();
}

void
Simulation::applyStdpLaunch(float reward)
{
	if(!m_stdp) {
		throw exception(NEMO_LOGIC_ERROR, "applyStdp called, but no STDP model specified");
	}
	int32_t fxr = fx_toFix(reward);
	if(fxr == 0 && reward != 0) {
		throw exception(NEMO_INVALID_INPUT, fmt("STDP reward rounded down to zero. The smallest valid reward is ",
					fx_toFloat(1)));
	}
	hipCheck(hipSetDevice(m_device), "hipSetDevice");
	hipStream_t st = static_cast<hipStream_t>(m_stream);
	hipCheck(nemo_dev_apply_stdp(d_syn.as<uint32_t>(), d_rcmFwd.as<uint32_t>(), d_rcmAcc.as<int32_t>(),
				m_nplastic, d_stdp.ptr, fxr, st), "applyStdp kernel");
}

void
Simulation::applyStdpFinish()
{
	hipCheck(hipStreamSynchronize(static_cast<hipStream_t>(m_stream)), "hipStreamSynchronize");
	/* spikes in flight now arrive with the new weights */
	m_replay = true;
}

void
Simulation::readScalar(const DevBuf& b, size_t offset, void* out, size_t bytes) const
{
	std::lock_guard<DeviceMutex> devLock(deviceMutex());
	hipCheck(hipSetDevice(m_device), "hipSetDevice");
	hipCheck(hipMemcpy(out, static_cast<const char*>(b.ptr) + offset, bytes, hipMemcpyDeviceToHost), "D2H");
}

void
Simulation::writeScalar(DevBuf& b, size_t offset, const void* in, size_t bytes)
{
	std::lock_guard<DeviceMutex> devLock(deviceMutex());
	hipCheck(hipSetDevice(m_device), "hipSetDevice");
	hipCheck(hipMemcpy(static_cast<char*>(b.ptr) + offset, in, bytes, hipMemcpyHostToDevice), "H2D");
}

float
Simulation::getNeuronState(unsigned g, unsigned var) const
{
	unsigned l = localIdx(g);
	if(var >= modelInfo(m_model[l]).nstate) {
		throw exception(NEMO_INVALID_INPUT, fmt("Invalid state variable index ", var));
	}
	float v;
	readScalar(d_state, stateOffset(l, var) * 4, &v, 4);
	return v;
}

/* element of state variable var of neuron l in its type's current plane */
size_t
Simulation::stateOffset(unsigned l, unsigned var) const
{
	const ModelInfo& m = modelInfo(m_model[l]);
	const size_t plane = m.history > 1 ? m_plane.at(m.id) : 0u;
	return (plane * m_ns + var) * m_n + l;
}

float
Simulation::getNeuronParameter(unsigned g, unsigned param) const
{
	unsigned l = localIdx(g);
	if(param >= modelInfo(m_model[l]).nparam) {
		throw exception(NEMO_INVALID_INPUT, fmt("Invalid parameter index ", param));
	}
	float v;
	readScalar(d_param, (size_t(param) * m_n + l) * 4, &v, 4);
	return v;
}

float
Simulation::getMembranePotential(unsigned g) const
{
	unsigned l = localIdx(g);
	return getNeuronState(g, modelInfo(m_model[l]).membranePotential);
}

void
Simulation::setNeuron(unsigned g, unsigned nargs, const float args[])
{
	unsigned l = localIdx(g);
	const ModelInfo& m = modelInfo(m_model[l]);
	if(nargs != m.nparam + m.nstate) {
		throw exception(NEMO_INVALID_INPUT, fmt("Unexpected number of parameters/state variables when modifying neuron. Expected ",
					m.nparam + m.nstate, ", found ", nargs));
	}
	for(unsigned k = 0; k < m.nparam; ++k) writeScalar(d_param, (size_t(k) * m_n + l) * 4, &args[k], 4);
	/* src/nemo/cpu/Neurons.cpp:66-81: the state goes to plane 0, whichever
	 * plane is current */
	for(unsigned k = 0; k < m.nstate; ++k) writeScalar(d_state, (size_t(k) * m_n + l) * 4, &args[m.nparam + k], 4);
}

void
Simulation::setNeuronState(unsigned g, unsigned var, float val)
{
	unsigned l = localIdx(g);
	if(var >= modelInfo(m_model[l]).nstate) {
		throw exception(NEMO_INVALID_INPUT, fmt("Invalid state variable index ", var));
	}
	writeScalar(d_state, stateOffset(l, var) * 4, &val, 4);
}

void
Simulation::setNeuronParameter(unsigned g, unsigned param, float val)
{
	unsigned l = localIdx(g);
	if(param >= modelInfo(m_model[l]).nparam) {
		throw exception(NEMO_INVALID_INPUT, fmt("Invalid parameter index ", param));
	}
	writeScalar(d_param, (size_t(param) * m_n + l) * 4, &val, 4);
}

size_t
Simulation::synapseRecord(synapse_id id) const
{
	if(m_writeOnly) {
		throw exception(NEMO_INVALID_INPUT, "Cannot read synapse state if simulation configured with write-only synapses");
	}
	unsigned source = unsigned(id >> 32);
	size_t k = size_t(id & 0xffffffffu);
	unsigned l;
	try {
		l = localIdx(source);
	} catch(exception&) {
		throw exception(NEMO_INVALID_INPUT, fmt("Non-existing neuron id (", source, ") in synapse query"));
	}
	if(k >= m_qStart[l + 1] - m_qStart[l]) {
		throw exception(NEMO_INVALID_INPUT, "synapse id out or range");
	}
	return size_t(m_qStart[l] + k);
}

const std::vector<synapse_id>&
Simulation::getSynapsesFrom(unsigned source)
{
	if(m_writeOnly) {
		throw exception(NEMO_INVALID_INPUT, "Cannot read synapse state if simulation configured with write-only synapses");
	}
	unsigned l;
	try {
		l = localIdx(source);
	} catch(exception&) {
		throw exception(NEMO_INVALID_INPUT, fmt("Non-existing source neuron id (", source, ") in synapse id query"));
	}
	size_t cnt = size_t(m_qStart[l + 1] - m_qStart[l]);
	m_queried.resize(cnt);
	for(size_t k = 0; k < cnt; ++k) m_queried[k] = (synapse_id(source) << 32) | synapse_id(k);
	return m_queried;
}

unsigned Simulation::getSynapseSource(synapse_id id) const { synapseRecord(id); return unsigned(id >> 32); }
unsigned Simulation::getSynapseTarget(synapse_id id) const { return m_qTarget[synapseRecord(id)]; }
unsigned Simulation::getSynapseDelay(synapse_id id) const { return m_qDelay[synapseRecord(id)]; }
unsigned char Simulation::getSynapsePlastic(synapse_id id) const { return m_qPlastic[synapseRecord(id)]; }

float
Simulation::getSynapseWeight(synapse_id id) const
{
	size_t q = synapseRecord(id);
	int32_t w;
	readScalar(d_syn, (size_t(m_qFwd[q]) * 2 + 1) * 4, &w, 4);
	return fx_toFloat(w);
}

unsigned long
Simulation::elapsedWallclock() const
{
	auto now = std::chrono::steady_clock::now();
	return (unsigned long) std::chrono::duration_cast<std::chrono::milliseconds>(now - m_timerStart).count();
}

void
Simulation::resetTimer()
{
	m_timerStart = std::chrono::steady_clock::now();
	m_timerCycles = 0;
}

} // namespace hip
} // namespace nemo
