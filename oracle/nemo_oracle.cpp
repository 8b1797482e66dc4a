/*
 * nemo_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of NeMo 0.7.2's CPU simulation backend (the parity oracle
 * for the HIP backend in nemosim_amd/). Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load this library, and only as the checker
 * or the timed CPU baseline; the product never links or calls it.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the NeMo source tree). The op order of every floating-point expression is
 * kept, and the file is compiled with -ffp-contract=off and without
 * -ffast-math, so that the arithmetic is the reference's IEEE arithmetic.
 *
 * Parity pins (what checks this restatement; see tests/):
 *   - fixed point: closed forms + edge values             (test_oracle_fixedpoint)
 *   - urand/initialiseRng: xorshift128 closed form, mt19937 known answer
 *   - nrand: host glibc logf/sinf/sqrtf are called directly (the reference
 *     calls the same functions, src/nemo/RNG.cpp:26-33)
 *   - step semantics: the reference's own closed-form tests (ring, double
 *     ring, fstim, istim, saturation, STDP closed form) and brute-force
 *     current sums on small random networks.
 *   - neuron updates (update_izhikevich, update_if_curr_exp with and without
 *     noise): replayed independently in float64/float32 from the
 *     reference's plugin sources, cycle by cycle over thousands of neurons
 *     with synaptic, external and Gaussian input (test_oracle_model_pins);
 *     each plausible mistake (float32 right-hand side, no early exit, fused
 *     multiply-add, noise order, threshold and refractory comparisons) is
 *     shown to change the results on the same inputs.
 *   - IF_curr_exp_noise is a build-defined model (no reference counterpart):
 *     IF_curr_exp plus sigma * nrand added once the total current is
 *     formed, pinned against that definition as above.
 *
 * One deliberate deviation: a spike at delay 64 (bit 63 of the firing
 * history) would make the reference's delivery loop shift a 64-bit word by
 * 64 (src/nemo/cpu/Simulation.cpp:270-275), which is undefined behaviour;
 * this restatement delivers it as delay 64.
 */

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <unordered_map>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>
#include <chrono>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

/* Status codes: src/nemo/types.h:31-61 */
enum {
	NEMO_OK = 0,
	NEMO_API_UNSUPPORTED = 5,
	NEMO_INVALID_INPUT = 6,
	NEMO_BUFFER_UNDERFLOW = 8,
	NEMO_LOGIC_ERROR = 9,
	NEMO_UNKNOWN_ERROR = 14
};

struct exception : public std::runtime_error {
	int code;
	exception(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

typedef int32_t fix_t;
typedef int64_t wfix_t;

/* src/nemo/ConfigurationImpl.cpp:22 -- fractional bits are fixed at 20 */
static const unsigned FBITS = 20;

/* src/nemo/fixedpoint.hpp:49-50 */
static const fix_t fx_min = fix_t(uint32_t(1) << 31);
static const fix_t fx_max = ~fix_t(uint32_t(1) << 31);

/* src/nemo/fixedpoint.cpp:18-29: truncating conversion, throws on overflow */
static fix_t
fx_toFix(float f, unsigned fb)
{
	if(std::abs(int(f)) >= 1 << (31 - fb)) {
		std::ostringstream msg;
		msg << "Fixed-point overflow. Value " << f
			<< " does not fit into fixed-point format Q" << 31 - fb << "." << fb;
		throw exception(NEMO_INVALID_INPUT, msg.str());
	}
	return static_cast<fix_t>(f * (1 << fb));
}

/* src/nemo/fixedpoint.cpp:33-37 */
static float
fx_toFloat(fix_t v, unsigned fb)
{
	return float(v) / float(1 << fb);
}

/* src/nemo/fixedpoint.cpp:41-49 (NEMO_WEIGHT_FIXED_POINT_SATURATION on,
 * src/nemo/CMakeLists.txt:1) */
static float
wfx_toFloat(wfix_t v, unsigned fb)
{
	v = std::min(v, wfix_t(fx_max));
	v = std::max(v, wfix_t(fx_min));
	return float(v) / float(1 << fb);
}

/* src/nemo/fixedpoint.hpp:42-46: arithmetic shift of the int64 product */
static fix_t
fx_mul(fix_t a, fix_t b, unsigned fb)
{
	int64_t r = int64_t(a) * int64_t(b);
	return fix_t(r >> fb);
}

/* ------------------------------------------------------------------ RNG */

/* src/nemo/RNG.hpp:8-10 */
struct RNG { uint32_t state[4]; };

/* src/nemo/RNG.cpp:9-18: Marsaglia xorshift128 */
static inline uint32_t
urand(RNG* rng)
{
	uint32_t t = (rng->state[0] ^ (rng->state[0] << 11));
	rng->state[0] = rng->state[1];
	rng->state[1] = rng->state[2];
	rng->state[2] = rng->state[3];
	rng->state[3] = (rng->state[3] ^ (rng->state[3] >> 19)) ^ (t ^ (t >> 8));
	return rng->state[3];
}

/* src/nemo/RNG.cpp:25-33: Box-Muller, sine branch only, host libm */
static inline float
nrand(RNG* rng)
{
	float a = urand(rng) * 1.4629180792671596810513378043098e-9f;
	float b = urand(rng) * 0.00000000023283064365386962890625f;
	float r = sqrtf(-2 * logf(1 - b));
	return sinf(a) * r;
}

/* src/nemo/RNG.cpp:38-63. boost::mt19937 with its default seed is the same
 * engine as std::mt19937 (seed 5489); boost::uniform_int<unsigned long>(0,
 * 0x7fffffff) over a 32-bit engine maps each draw to draw >> 1 (Boost >= 1.47:
 * bucket size 2, no rejection). */
static void
initialiseRng(unsigned minIdx, unsigned maxIdx, std::vector<RNG>& rngs)
{
	if(minIdx > maxIdx) {
		throw exception(NEMO_LOGIC_ERROR, "Invalid neuron range when initialising RNG");
	}
	std::mt19937 mt;
	for(uint64_t i = 0; i < 4ull * minIdx; ++i) {
		mt();
	}
	for(unsigned g = minIdx; g <= maxIdx; ++g) {
		RNG& r = rngs.at(g - minIdx);
		for(unsigned plane = 0; plane < 4; ++plane) {
			r.state[plane] = uint32_t(mt()) >> 1;
		}
	}
}

/* ---------------------------------------------------------- neuron types */

enum ModelId { MODEL_IZHIKEVICH = 0, MODEL_IF_CURR_EXP = 1, MODEL_IF_CURR_EXP_NOISE = 2,
	MODEL_POISSON_SOURCE = 3, MODEL_INPUT = 4, MODEL_KURAMOTO = 5 };
static const unsigned NMODELS = 6;

struct TypeInfo {
	const char* name;
	unsigned nparam;
	unsigned nstate;
	unsigned membranePotential;
	unsigned history;      /* cycles of state kept (planes of the state array) */
	bool rcm;              /* reads the reverse connectivity (sources, delays, weights) */
};

/* src/nemo/plugins/Izhikevich.ini:1-5, src/nemo/plugins/IF_curr_exp.ini:14-26,
 * src/nemo/plugins/PoissonSource.ini:1-4, src/nemo/plugins/Input.ini:1-4,
 * src/nemo/plugins/Kuramoto.ini:2-12; IF_curr_exp_noise is IF_curr_exp plus
 * a 10th parameter sigma. */
static const TypeInfo TYPES[NMODELS] = {
	{ "Izhikevich", 5, 2, 1, 1, false },
	{ "IF_curr_exp", 9, 4, 0, 1, false },
	{ "IF_curr_exp_noise", 10, 4, 0, 1, false },
	{ "PoissonSource", 1, 0, 0, 1, false },
	{ "Input", 0, 0, 0, 1, false },
	{ "Kuramoto", 1, 1, 0, 32, true },
};

/* Parameter / state indices: src/nemo/plugins/Izhikevich.h:6-13 */
enum { IZ_A = 0, IZ_B, IZ_C, IZ_D, IZ_SIGMA };
enum { IZ_U = 0, IZ_V };

/* src/nemo/plugins/IF_curr_exp.h:4-16 (header order, not the .ini comment) */
enum { IF_V_REST = 0, IF_V_RESET, IF_C_M, IF_TAU_M, IF_TAU_REFRAC,
	IF_TAU_SYN_E, IF_TAU_SYN_I, IF_V_THRESH, IF_I_OFFSET, IF_SIGMA };
enum { IF_V = 0, IF_IE, IF_II, IF_LASTFIRED };

/* src/nemo/cpu/plugins/Izhikevich.cpp:14-96. SoA params [nparam][n], state
 * [nstate][n] (history length 1, so current and next state coincide). */
static void
update_izhikevich(unsigned start, unsigned end,
		const float* param, size_t pstride, float* state, size_t sstride,
		unsigned fstim[], RNG rng[],
		const float currentEPSP[], const float currentIPSP[], float currentExternal[],
		uint64_t recentFiring[], unsigned fired[])
{
	const unsigned SUBSTEPS = 4;
	const float SUBSTEP_MULT = 0.25f;
	const float* a = param + IZ_A * pstride;
	const float* b = param + IZ_B * pstride;
	const float* c = param + IZ_C * pstride;
	const float* d = param + IZ_D * pstride;
	const float* sigma = param + IZ_SIGMA * pstride;
	float* uu = state + IZ_U * sstride;
	float* vv = state + IZ_V * sstride;

	int nn = int(end - start);
#pragma omp parallel for default(shared)
	for(int nl = 0; nl < nn; nl++) {
		unsigned ng = start + nl;
		float I = currentEPSP[ng] + currentIPSP[ng] + currentExternal[ng];
		currentExternal[ng] = 0.0f;
		if(sigma[nl] != 0.0f) {
			I += sigma[nl] * nrand(&rng[nl]);
		}
		fired[ng] = 0;
		float u = uu[nl];
		float v = vv[nl];
		for(unsigned t = 0; t < SUBSTEPS; ++t) {
			if(!fired[ng]) {
				/* double literals: the right-hand side and the add to v are
				 * evaluated in double and rounded to float once */
				v += SUBSTEP_MULT * ((0.04 * v + 5.0) * v + 140.0 - u + I);
				u += SUBSTEP_MULT * (a[nl] * (b[nl] * v - u));
				fired[ng] = v >= 30.0;
			}
		}
		fired[ng] |= fstim[ng];
		fstim[ng] = 0;
		recentFiring[ng] = (recentFiring[ng] << 1) | (uint64_t) fired[ng];
		if(fired[ng]) {
			v = c[nl];
			u += d[nl];
		}
		uu[nl] = u;
		vv[nl] = v;
	}
}

/* src/nemo/cpu/plugins/IF_curr_exp.cpp:27-123; with_noise adds the
 * build-defined Gaussian input (sigma = parameter 9), drawn once per cycle
 * after the total current is formed. */
/* PoissonSource, src/nemo/cpu/plugins/PoissonSource.cpp:22-29: parameter 0
 * is the firing probability per cycle; the neuron fires when the low 16 bits
 * of its next urand are below rate * 2^16, or when forced. Synaptic and
 * external input are ignored. */
static void
update_poisson_source(unsigned start, unsigned end, const float* rate, unsigned fstim[], RNG rng[],
		uint64_t recentFiring[], unsigned fired[])
{
	for(unsigned g = start; g < end; ++g) {
		const unsigned l = g - start;
		const unsigned threshold = unsigned(rate[l] * float(1 << 16));
		const unsigned draw = urand(&rng[l]) & 0xffffu;
		fired[g] = (draw < threshold || fstim[g]) ? 1u : 0u;
		fstim[g] = 0;
		recentFiring[g] = (recentFiring[g] << 1) | uint64_t(fired[g]);
	}
}

/* Input, src/nemo/cpu/plugins/Input.cpp:22-27: fires exactly when forced */
static void
update_input(unsigned start, unsigned end, unsigned fstim[], uint64_t recentFiring[], unsigned fired[])
{
	for(unsigned g = start; g < end; ++g) {
		fired[g] = fstim[g] ? 1u : 0u;
		fstim[g] = 0;
		recentFiring[g] = (recentFiring[g] << 1) | uint64_t(fired[g]);
	}
}

static void
update_if_curr_exp(unsigned start, unsigned end, bool with_noise,
		const float* param, size_t pstride, float* state, size_t sstride,
		unsigned fstim[], RNG rng[],
		const float currentEPSP[], const float currentIPSP[], float currentExternal[],
		uint64_t recentFiring[], unsigned fired[])
{
	const float* p_v_rest = param + IF_V_REST * pstride;
	const float* p_c_m = param + IF_C_M * pstride;
	const float* p_tau_m = param + IF_TAU_M * pstride;
	const float* p_tau_refrac = param + IF_TAU_REFRAC * pstride;
	const float* p_tau_syn_E = param + IF_TAU_SYN_E * pstride;
	const float* p_tau_syn_I = param + IF_TAU_SYN_I * pstride;
	const float* p_I_offset = param + IF_I_OFFSET * pstride;
	const float* p_v_reset = param + IF_V_RESET * pstride;
	const float* p_v_thresh = param + IF_V_THRESH * pstride;
	const float* p_sigma = with_noise ? param + IF_SIGMA * pstride : 0;
	float* vv = state + IF_V * sstride;
	float* Ie0 = state + IF_IE * sstride;
	float* Ii0 = state + IF_II * sstride;
	float* lastfired = state + IF_LASTFIRED * sstride;

	int nn = int(end - start);
#pragma omp parallel for default(shared)
	for(int nl = 0; nl < nn; nl++) {
		unsigned ng = start + nl;
		float Ie = ((1.0f - 1.0f / p_tau_syn_E[nl]) * Ie0[nl]) + currentEPSP[ng];
		float Ii = ((1.0f - 1.0f / p_tau_syn_I[nl]) * Ii0[nl]) + currentIPSP[ng];
		float I = Ie + Ii + currentExternal[ng] + p_I_offset[nl];
		if(with_noise && p_sigma[nl] != 0.0f) {
			I += p_sigma[nl] * nrand(&rng[nl]);
		}
		Ie0[nl] = Ie;
		Ii0[nl] = Ii;
		currentExternal[ng] = 0.0f;
		float v = vv[nl];
		bool refractory = lastfired[nl] <= p_tau_refrac[nl];
		if(!refractory) {
			float c_m = p_c_m[nl];
			float v_rest = p_v_rest[nl];
			float tau_m = p_tau_m[nl];
			v += I / c_m + (v_rest - v) / tau_m;
		}
		fired[ng] = (!refractory && v > p_v_thresh[nl]) || fstim[ng];
		fstim[ng] = 0;
		recentFiring[ng] = (recentFiring[ng] << 1) | (uint64_t) fired[ng];
		if(fired[ng]) {
			lastfired[nl] = 1;
			vv[nl] = p_v_reset[nl];
		} else {
			lastfired[nl] += 1;
			vv[nl] = v;
		}
	}
}

/* ----------------------------------------------------------------- STDP */

/* src/nemo/StdpFunction.cpp:12-60 (validation), :73-108 (bit masks) and
 * src/nemo/StdpProcess.cpp:20-138 (window arithmetic). */
struct StdpProcess {
	std::vector<fix_t> fnPre, fnPost;
	int postPreWindow;
	uint64_t potentiationBits, depressionBits, prePostBits, postPreBits, postFireMask;
	fix_t minE, maxE, minI, maxI;

	static uint64_t getBits(const std::vector<float>& pre, const std::vector<float>& post,
			bool potentiation)
	{
		uint64_t bits = 0;
		int n = 0;
		for(std::vector<float>::const_reverse_iterator f = post.rbegin(); f != post.rend(); ++f, ++n) {
			if(potentiation ? (*f > 0.0f) : (*f < 0.0f)) bits |= uint64_t(1) << n;
		}
		for(std::vector<float>::const_iterator f = pre.begin(); f != pre.end(); ++f, ++n) {
			if(potentiation ? (*f > 0.0f) : (*f < 0.0f)) bits |= uint64_t(1) << n;
		}
		return bits;
	}

	StdpProcess(const std::vector<float>& pre, const std::vector<float>& post,
			float fminE, float fmaxE, float fminI, float fmaxI, unsigned fb)
	{
		if(fmaxI > 0.0f || fminI > 0.0f)
			throw exception(NEMO_INVALID_INPUT, "STDP function should not have positive limits for inhibitory synapses");
		if(fmaxE < 0.0f || fminE < 0.0f)
			throw exception(NEMO_INVALID_INPUT, "STDP function should not have negative limits for inhibitory synapses");
		if(fmaxE < fminE)
			throw exception(NEMO_INVALID_INPUT, "STDP should have maximum excitatory weight >= minimum excitatory weight");
		if(fmaxI > fminI)
			throw exception(NEMO_INVALID_INPUT, "STDP should have abs(maximum inhibitory weight) >= abs(minimum inhibitory weight)");
		if(pre.size() + post.size() > 64)
			throw exception(NEMO_INVALID_INPUT, "size of STDP window too large");
		minE = fx_toFix(fminE, fb);
		maxE = fx_toFix(fmaxE, fb);
		minI = fx_toFix(fminI, fb);
		maxI = fx_toFix(fmaxI, fb);
		for(size_t i = 0; i < pre.size(); ++i) fnPre.push_back(fx_toFix(pre[i], fb));
		for(size_t i = 0; i < post.size(); ++i) fnPost.push_back(fx_toFix(post[i], fb));
		potentiationBits = getBits(pre, post, true);
		depressionBits = getBits(pre, post, false);
		int prePostWindow = int(pre.size());
		postPreWindow = int(post.size());
		prePostBits = (~(~uint64_t(0) << uint64_t(prePostWindow))) << uint64_t(postPreWindow);
		postPreBits = ~(~uint64_t(0) << uint64_t(postPreWindow));
		postFireMask = uint64_t(1) << postPreWindow;
	}

	static const unsigned NO_APPLICATION = ~0u;

	/* src/nemo/StdpProcess.cpp:68-74 */
	unsigned closestPreFire(uint64_t arrivals) const {
		uint64_t valid = arrivals & prePostBits;
		if(!valid) return NO_APPLICATION;
		return unsigned(__builtin_ctzll(valid >> postPreWindow));
	}

	/* src/nemo/StdpProcess.cpp:77-83 */
	unsigned closestPostFire(uint64_t arrivals) const {
		uint64_t valid = arrivals & postPreBits;
		if(!valid) return NO_APPLICATION;
		return unsigned(__builtin_clzll(valid << uint64_t(64 - postPreWindow)));
	}

	/* src/nemo/StdpProcess.cpp:87-113 */
	fix_t updateRegion(uint64_t arrivals) const {
		fix_t w_diff = 0;
		if(arrivals) {
			unsigned dt_pre = closestPreFire(arrivals);
			unsigned dt_post = closestPostFire(arrivals);
			if(dt_pre < dt_post) {
				w_diff = fnPre.at(dt_pre);
			} else if(dt_post < dt_pre) {
				w_diff = fnPost.at(dt_post);
			}
		}
		return w_diff;
	}

	/* src/nemo/StdpProcess.cpp:116-124 */
	fix_t weightChange(uint64_t preFiring) const {
		return updateRegion(preFiring & potentiationBits) + updateRegion(preFiring & depressionBits);
	}

	/* src/nemo/StdpProcess.cpp:128-138: sign sticks, zero sticks */
	fix_t updatedWeight(fix_t w_old, fix_t w_diff) const {
		fix_t w_new = 0;
		if(w_old > 0) {
			w_new = std::min(maxE, std::max(fix_t(w_old + w_diff), minE));
		} else if(w_old < 0) {
			w_new = std::min(minI, std::max(fix_t(w_old + w_diff), maxI));
		}
		return w_new;
	}
};

/* Kuramoto oscillators, src/nemo/cpu/plugins/Kuramoto.cpp. Parameter 0 is
 * the natural frequency (radians per cycle), state 0 the phase, kept for
 * KURAMOTO_HISTORY cycles: plane c mod KURAMOTO_HISTORY holds the phase of
 * cycle c (src/nemo/plugins/Kuramoto.h:4). The incoming couplings of each
 * oscillator are its reverse-matrix entries in the reference's order
 * (kStart/kSource/kDelay/kWeight, sim-local indices, built in finalize). */
static const unsigned KURAMOTO_HISTORY = 32;

/* Kuramoto.cpp:37-42 */
static inline float*
kuramoto_phase(float* base, size_t stride, unsigned cycle)
{
	return base + size_t(cycle % KURAMOTO_HISTORY) * stride;
}

/* Kuramoto.cpp:121 and :150 -- the remainder is float, the constant
 * 2.0f*M_PI of the conditional is a double, so for a negative phase the sum
 * is formed in double and rounded to float on assignment */
static inline float
kuramoto_wrap(float phase)
{
	return fmodf(phase, 2.0f * M_PI) + (phase < 0.0f ? 2.0f * M_PI : 0.0f);
}

/* Kuramoto.cpp:18-29: sum of w * sin(source - target) in reverse-matrix order */
static inline float
kuramoto_sum(const std::vector<float>& weight, const std::vector<float>& sourcePhase, float targetPhase)
{
	float sum = 0.0f;
	for(size_t i = 0; i < weight.size(); ++i) {
		sum += weight[i] * sinf(sourcePhase[i] - targetPhase);
	}
	return sum;
}

/* Kuramoto.cpp:85-123: fourth-order Runge-Kutta step of one cycle; the
 * source phases are those of cycle `cycle - (delay - 1)` */
static void
update_kuramoto(unsigned start, unsigned end, unsigned cycle, const float* frequency,
		float* stateBase, size_t historyStride, const std::vector<size_t>& kStart,
		const std::vector<unsigned>& kSource, const std::vector<unsigned>& kDelay,
		const std::vector<float>& kWeight)
{
	const float* phase0 = kuramoto_phase(stateBase, historyStride, cycle);
	float* phase1 = kuramoto_phase(stateBase, historyStride, cycle + 1u);
	/* the oscillators are independent within a cycle (delays <= 31 never
	 * read the plane being written), so they may be split over threads */
#pragma omp parallel for schedule(static)
	for(unsigned n = start; n < end; ++n) {
		std::vector<float> weight, sourcePhase;
		const float f = frequency[n];
		float targetPhase = phase0[n];
		/* Kuramoto.cpp:48-80 (loadIncoming) */
		weight.clear();
		sourcePhase.clear();
		for(size_t r = kStart[n]; r < kStart[n + 1]; ++r) {
			weight.push_back(kWeight[r]);
			sourcePhase.push_back(kuramoto_phase(stateBase, historyStride, cycle - int(kDelay[r] - 1))[kSource[r]]);
		}
		float k0 = f + kuramoto_sum(weight, sourcePhase, targetPhase);
		float k1 = f + kuramoto_sum(weight, sourcePhase, targetPhase + k0 * 0.5f);
		float k2 = f + kuramoto_sum(weight, sourcePhase, targetPhase + k1 * 0.5f);
		float k3 = f + kuramoto_sum(weight, sourcePhase, targetPhase + k2);
		targetPhase += (k0 + 2 * k1 + 2 * k2 + k3) / 6.0f;
		phase1[n] = kuramoto_wrap(targetPhase);
	}
}

/* Kuramoto.cpp:133-152: the history before cycle 0, by running the
 * uncoupled model backwards from the initial phase in plane 0 */
static void
init_kuramoto(unsigned start, unsigned end, const float* frequency, float* stateBase, size_t historyStride)
{
	for(unsigned t = 0; t < KURAMOTO_HISTORY - 1u; ++t) {
		const float* phase0 = kuramoto_phase(stateBase, historyStride, 0u - t);
		float* phase1 = kuramoto_phase(stateBase, historyStride, 0u - (t + 1u));
		for(unsigned n = start; n < end; ++n) {
			const float phase = phase0[n] - frequency[n];
			phase1[n] = kuramoto_wrap(phase);
		}
	}
}

/* -------------------------------------------------------------- network */

struct Synapse {
	unsigned target;
	unsigned delay;
	float weight;
	unsigned char plastic;
};

struct NeuronGroup {
	unsigned model;
	std::vector<std::vector<float> > param;  // [nparam][n]
	std::vector<std::vector<float> > state;  // [nstate][n]
	std::vector<unsigned> gidx;
};

/* Forward row of one (source, delay): src/nemo/ConnectivityMatrix.hpp:38-63 */
struct FAxonTerminal { unsigned target; fix_t weight; };

struct Oracle {
	/* ---- network (src/nemo/NetworkImpl.cpp) */
	std::vector<NeuronGroup> groups;
	int typeOf[NMODELS];
	std::unordered_map<unsigned, std::pair<unsigned, unsigned> > netMapper;  // gidx -> (group, index)
	/* synapses in insertion order (the per-source Axon vectors of
	 * src/nemo/Axon.cpp:9-20, flattened); ids are (source << 32) | ordinal */
	std::vector<unsigned> synSrc, synTgt;
	std::vector<unsigned char> synDelay, synPlastic;
	std::vector<float> synWeight;
	std::unordered_map<unsigned, uint32_t> outCount;
	unsigned maxDelayNet;

	bool hasStdp;
	std::vector<float> stdpPre, stdpPost;
	float sMinE, sMaxE, sMinI, sMaxI;

	/* ---- simulation (src/nemo/cpu/Simulation.cpp) */
	bool finalized;
	size_t N;
	std::vector<unsigned> groupBase;                 // sim-local base per group
	std::vector<unsigned> localToGlobal;
	std::unordered_map<unsigned, unsigned> globalToLocal;
	std::vector<unsigned> localGroup;
	std::vector<std::vector<float> > simParam;       // per group: [nparam*n]
	std::vector<std::vector<float> > simState;       // per group: [history][nstate][n]
	std::vector<unsigned> stateCurrent;              // per group: plane of the current state

	/* reverse matrix of every synapse, for the Kuramoto model
	 * (src/nemo/construction/RCM.cpp:150-171 with STDP off; per target in
	 * the order of the network's synapse iterator) */
	std::vector<size_t> kStart;
	std::vector<unsigned> kSource, kDelay;
	std::vector<float> kWeight;
	std::vector<std::vector<RNG> > simRng;

	std::vector<unsigned> fired;
	std::vector<uint64_t> recentFiring;
	std::vector<uint64_t> delays;
	std::vector<wfix_t> fxE, fxI;
	std::vector<float> curE, curI, curExt;
	std::vector<unsigned> fstim;

	unsigned maxDelay;
	std::vector<size_t> rowStart;                    // CSR over (source, delay)
	std::vector<FAxonTerminal> terminals;

	/* synapse queries: per local source, the flat indices of its synapses in
	 * ordinal order, and each flat synapse's position in terminals[] */
	std::vector<size_t> qStart;
	std::vector<size_t> qFlat;
	std::vector<size_t> flatTerm;

	/* reverse matrix of plastic synapses (src/nemo/construction/RCM.cpp:150-171) */
	std::vector<StdpProcess> stdp;                   // 0 or 1 element
	std::vector<size_t> rcmStart;                    // CSR over local target
	std::vector<unsigned> rcmSource, rcmDelay;
	std::vector<size_t> rcmForward;                  // index into terminals
	std::vector<fix_t> rcmAcc;

	unsigned long long cycle;
	std::vector<std::vector<unsigned> > firingFifo;
	std::vector<unsigned> lastFired;
	double wallMs;

	Oracle() : maxDelayNet(0), hasStdp(false), finalized(false), N(0), maxDelay(0),
		cycle(0), wallMs(0) {
		for(unsigned m = 0; m < NMODELS; ++m) typeOf[m] = -1;
	}

	unsigned groupFor(unsigned model) {
		if(model >= NMODELS) throw exception(NEMO_INVALID_INPUT, "unknown neuron model");
		if(typeOf[model] < 0) {
			typeOf[model] = int(groups.size());
			NeuronGroup g;
			g.model = model;
			g.param.resize(TYPES[model].nparam);
			g.state.resize(TYPES[model].nstate);
			groups.push_back(g);
		}
		return unsigned(typeOf[model]);
	}

	/* src/nemo/NetworkImpl.cpp:84-92, src/nemo/Neurons.cpp:29-47 */
	void addNeuron(unsigned model, unsigned g, unsigned nargs, const float* args) {
		unsigned gi = groupFor(model);
		NeuronGroup& grp = groups[gi];
		const TypeInfo& ti = TYPES[model];
		if(nargs != ti.nparam + ti.nstate) {
			std::ostringstream m;
			m << "Unexpected number of parameters/state variables when adding neuron. Expected "
				<< ti.nparam + ti.nstate << ", found " << nargs;
			throw exception(NEMO_INVALID_INPUT, m.str());
		}
		if(netMapper.count(g)) {
			std::ostringstream m;
			m << "Duplicate neuron index " << g;
			throw exception(NEMO_INVALID_INPUT, m.str());
		}
		for(unsigned i = 0; i < ti.nparam; ++i) grp.param[i].push_back(args[i]);
		for(unsigned i = 0; i < ti.nstate; ++i) grp.state[i].push_back(args[ti.nparam + i]);
		netMapper[g] = std::make_pair(gi, unsigned(grp.gidx.size()));
		grp.gidx.push_back(g);
	}

	/* src/nemo/NetworkImpl.cpp:103-125 */
	uint64_t addSynapse(unsigned s, unsigned t, unsigned d, float w, unsigned char plastic) {
		if(d < 1) {
			std::ostringstream m;
			m << "Invalid delay (" << d << ") for synapse between " << s << " and " << t;
			throw exception(NEMO_INVALID_INPUT, m.str());
		}
		appendSynapse(s, t, d, w, plastic);
		uint32_t& c = outCount[s];
		uint64_t id = (uint64_t(s) << 32) | uint64_t(c);
		c += 1;
		return id;
	}

	void appendSynapse(unsigned s, unsigned t, unsigned d, float w, unsigned char plastic) {
		if(d > 255) {
			throw exception(NEMO_INVALID_INPUT, "delay out of range");
		}
		synSrc.push_back(s);
		synTgt.push_back(t);
		synDelay.push_back((unsigned char) d);
		synWeight.push_back(w);
		synPlastic.push_back(plastic ? 1 : 0);
		maxDelayNet = std::max(maxDelayNet, d);
	}

	unsigned localIdx(unsigned g) const {
		std::unordered_map<unsigned, unsigned>::const_iterator i = globalToLocal.find(g);
		if(i == globalToLocal.end()) {
			std::ostringstream m;
			m << "Non-existing neuron index " << g;
			throw exception(NEMO_INVALID_INPUT, m.str());
		}
		return i->second;
	}

	/* src/nemo/cpu/Simulation.cpp:36-82, src/nemo/cpu/Neurons.cpp:7-48,
	 * src/nemo/ConnectivityMatrix.cpp:65-162 */
	void finalize() {
		if(finalized) throw exception(NEMO_LOGIC_ERROR, "simulation already created");
		if(maxDelayNet > 64) {
			std::ostringstream m;
			m << "The network has synapses with delay " << maxDelayNet
				<< "ms. The CPU backend supports a maximum of 64 ms";
			throw exception(NEMO_INVALID_INPUT, m.str());
		}
		if(hasStdp) {
			stdp.push_back(StdpProcess(stdpPre, stdpPost, sMinE, sMaxE, sMinI, sMaxI, FBITS));
			/* src/nemo/StdpFunction.cpp:52-60 */
			if(maxDelayNet + stdpPost.size() > 64) {
				throw exception(NEMO_INVALID_INPUT, "Max delay + postfire part of STDP function must be <= 64");
			}
		}
		/* Kuramoto oscillators. src/nemo/construction/RCM.cpp:90-97 refuses
		 * STDP together with a type that reads the reverse matrix. The two
		 * further conditions are this restatement's: the reference's CPU
		 * plugin indexes its arrays with simulation-wide indices
		 * (Kuramoto.cpp:105-110), which stays inside them only when the
		 * oscillators are the network's only neurons (the CUDA plugin asserts
		 * the same, src/nemo/cuda/plugins/Kuramoto.cu:270), and a delay of
		 * KURAMOTO_HISTORY or more would read the plane being written this
		 * cycle (Kuramoto.cu:116 asserts delay - 1 < KURAMOTO_HISTORY - 1). */
		bool rcmType = false, otherType = false;
		for(size_t gi = 0; gi < groups.size(); ++gi) {
			if(groups[gi].gidx.empty()) continue;
			if(TYPES[groups[gi].model].rcm) rcmType = true; else otherType = true;
		}
		if(rcmType && hasStdp) {
			throw exception(NEMO_API_UNSUPPORTED,
					"The current version does not support a mix of STDP with neuron types which require reverse connectivity for other purposes");
		}
		if(rcmType && otherType) {
			throw exception(NEMO_API_UNSUPPORTED, "Kuramoto oscillators cannot share a network with other neuron types");
		}
		if(rcmType && maxDelayNet >= KURAMOTO_HISTORY) {
			std::ostringstream m;
			m << "Kuramoto couplings support delays up to " << KURAMOTO_HISTORY - 1 << " cycles; found " << maxDelayNet;
			throw exception(NEMO_INVALID_INPUT, m.str());
		}
		/* neuron groups in type order, sim-local indices contiguous */
		unsigned base = 0;
		groupBase.resize(groups.size());
		simParam.resize(groups.size());
		simState.resize(groups.size());
		stateCurrent.assign(groups.size(), 0u);
		simRng.resize(groups.size());
		for(size_t gi = 0; gi < groups.size(); ++gi) {
			NeuronGroup& grp = groups[gi];
			const TypeInfo& ti = TYPES[grp.model];
			size_t n = grp.gidx.size();
			groupBase[gi] = base;
			simParam[gi].resize(ti.nparam * n);
			simState[gi].assign(size_t(ti.history) * ti.nstate * n, 0.0f);
			for(unsigned p = 0; p < ti.nparam; ++p)
				std::copy(grp.param[p].begin(), grp.param[p].end(), simParam[gi].begin() + p * n);
			for(unsigned p = 0; p < ti.nstate; ++p)
				std::copy(grp.state[p].begin(), grp.state[p].end(), simState[gi].begin() + p * n);
			for(size_t i = 0; i < n; ++i) {
				globalToLocal[grp.gidx[i]] = unsigned(base + i);
				localToGlobal.push_back(grp.gidx[i]);
				localGroup.push_back(unsigned(gi));
			}
			simRng[gi].resize(n);
			if(n > 0) initialiseRng(base, unsigned(base + n - 1), simRng[gi]);
			base += unsigned(n);
		}
		N = base;

		fired.assign(N, 0);
		recentFiring.assign(N, 0);
		delays.assign(N, 0);
		fxE.assign(N, 0);
		fxI.assign(N, 0);
		curE.assign(N, 0.0f);
		curI.assign(N, 0.0f);
		curExt.assign(N, 0.0f);
		fstim.assign(N, 0);

		/* forward matrix rows keyed (source, delay); every referenced source and
		 * target must exist (verifySynapseTerminals, ConnectivityMatrix.cpp:166-189) */
		maxDelay = maxDelayNet;
		const size_t S = synSrc.size();
		std::vector<unsigned> ls(S), lt(S);
		std::vector<size_t> rowCount(N * size_t(std::max(maxDelay, 1u)), 0);
		for(size_t i = 0; i < S; ++i) {
			std::unordered_map<unsigned, unsigned>::const_iterator si = globalToLocal.find(synSrc[i]);
			if(si == globalToLocal.end()) {
				std::ostringstream m;
				m << "Invalid synapse source neuron " << synSrc[i];
				throw exception(NEMO_INVALID_INPUT, m.str());
			}
			std::unordered_map<unsigned, unsigned>::const_iterator ti = globalToLocal.find(synTgt[i]);
			if(ti == globalToLocal.end()) {
				std::ostringstream m;
				m << "Invalid synapse target neuron " << synTgt[i] << " (source: " << synSrc[i] << ")";
				throw exception(NEMO_INVALID_INPUT, m.str());
			}
			ls[i] = si->second;
			lt[i] = ti->second;
			rowCount[size_t(ls[i]) * maxDelay + synDelay[i] - 1] += 1;
		}
		rowStart.assign(rowCount.size() + 1, 0);
		for(size_t r = 0; r < rowCount.size(); ++r) rowStart[r + 1] = rowStart[r] + rowCount[r];
		terminals.resize(rowStart.back());
		std::vector<size_t> fill(rowStart.begin(), rowStart.end() - 1);

		/* src/nemo/ConnectivityMatrix.cpp:97-119: weights to Q11.20, rows in
		 * insertion order, per-source delay bits (OutgoingDelays.cpp:70-80) */
		flatTerm.resize(S);
		std::vector<size_t> rcmCount(N, 0);
		for(size_t i = 0; i < S; ++i) {
			size_t row = size_t(ls[i]) * maxDelay + synDelay[i] - 1;
			FAxonTerminal term = { lt[i], fx_toFix(synWeight[i], FBITS) };
			size_t pos = fill[row]++;
			terminals[pos] = term;
			flatTerm[i] = pos;
			delays[ls[i]] |= uint64_t(1) << (synDelay[i] - 1);
			if(hasStdp && synPlastic[i]) rcmCount[lt[i]] += 1;
		}

		/* ordinal order per source for id queries */
		qStart.assign(N + 1, 0);
		for(size_t i = 0; i < S; ++i) qStart[ls[i] + 1] += 1;
		for(size_t n = 0; n < N; ++n) qStart[n + 1] += qStart[n];
		qFlat.resize(S);
		{
			std::vector<size_t> qc(qStart.begin(), qStart.end() - 1);
			for(size_t i = 0; i < S; ++i) qFlat[qc[ls[i]]++] = i;
		}

		if(hasStdp) {
			rcmStart.assign(N + 1, 0);
			for(size_t n = 0; n < N; ++n) rcmStart[n + 1] = rcmStart[n] + rcmCount[n];
			size_t nr = rcmStart.back();
			rcmSource.resize(nr);
			rcmDelay.resize(nr);
			rcmForward.resize(nr);
			rcmAcc.assign(nr, 0);
			std::vector<size_t> rf(rcmStart.begin(), rcmStart.end() - 1);
			for(size_t i = 0; i < S; ++i) {
				if(!synPlastic[i]) continue;
				size_t r = rf[lt[i]]++;
				rcmSource[r] = ls[i];
				rcmDelay[r] = synDelay[i];
				rcmForward[r] = flatTerm[i];
			}
		}
		if(rcmType) {
			/* every synapse, per target in the order of the network's synapse
			 * iterator: sources ascending by global index, each source's
			 * synapses in insertion order (src/nemo/NetworkImpl.cpp:311-325
			 * over the std::map of axons; src/nemo/ConnectivityMatrix.cpp:79-87) */
			std::vector<size_t> order(S);
			for(size_t i = 0; i < S; ++i) order[i] = i;
			std::stable_sort(order.begin(), order.end(),
					[&](size_t a, size_t b) { return synSrc[a] < synSrc[b]; });
			kStart.assign(N + 1, 0);
			for(size_t i = 0; i < S; ++i) kStart[lt[i] + 1] += 1;
			for(size_t n = 0; n < N; ++n) kStart[n + 1] += kStart[n];
			kSource.resize(S);
			kDelay.resize(S);
			kWeight.resize(S);
			std::vector<size_t> kf(kStart.begin(), kStart.end() - 1);
			for(size_t k = 0; k < S; ++k) {
				size_t i = order[k];
				size_t r = kf[lt[i]]++;
				kSource[r] = ls[i];
				kDelay[r] = synDelay[i];
				kWeight[r] = synWeight[i];
			}
			/* src/nemo/cpu/Neurons.cpp:42-46: the type's init after its RNG */
			for(size_t gi = 0; gi < groups.size(); ++gi) {
				if(groups[gi].model != MODEL_KURAMOTO || groups[gi].gidx.empty()) continue;
				size_t n = groups[gi].gidx.size();
				init_kuramoto(groupBase[gi], unsigned(groupBase[gi] + n), &simParam[gi][0], &simState[gi][0], n);
			}
		}
		firingFifo.assign(1, std::vector<unsigned>());  // dummy entry (FiringBuffer.cpp:7-12)
		finalized = true;
		cycle = 0;
	}

	/* src/nemo/cpu/Simulation.cpp:258-307. Spikes of the last maxDelay cycles
	 * are delivered along the rows of the delays they have synapses at. */
	void deliverSpikes() {
		uint64_t validSpikes = maxDelay >= 64 ? ~uint64_t(0) : ~((~uint64_t(0)) << maxDelay);
		for(size_t source = 0; source < N; ++source) {
			uint64_t f = recentFiring[source] & validSpikes & delays[source];
			while(f) {
				unsigned p = unsigned(__builtin_ctzll(f));
				f &= f - 1;
				unsigned delay = p + 1;
				size_t row = source * maxDelay + delay - 1;
				for(size_t s = rowStart[row]; s < rowStart[row + 1]; ++s) {
					const FAxonTerminal& t = terminals[s];
					if(t.weight >= 0) fxE[t.target] += t.weight;
					else fxI[t.target] += t.weight;
				}
			}
		}
		int ncount = int(N);
#pragma omp parallel for default(shared)
		for(int n = 0; n < ncount; n++) {
			curE[n] = wfx_toFloat(fxE[n], FBITS);
			fxE[n] = 0;
			curI[n] = wfx_toFloat(fxI[n], FBITS);
			fxI[n] = 0;
		}
	}

	/* src/nemo/ConnectivityMatrix.cpp:193-227 */
	void accumulateStdp() {
		if(stdp.empty()) return;
		const StdpProcess& sp = stdp[0];
		for(size_t target = 0; target < N; ++target) {
			if(!(recentFiring[target] & sp.postFireMask)) continue;
			for(size_t r = rcmStart[target]; r < rcmStart[target + 1]; ++r) {
				uint64_t preFiring = recentFiring[rcmSource[r]] >> rcmDelay[r];
				fix_t w_diff = sp.weightChange(preFiring);
				if(w_diff != 0) {
					rcmAcc[r] = fix_t(uint32_t(rcmAcc[r]) + uint32_t(w_diff));
				}
			}
		}
	}

	/* src/nemo/ConnectivityMatrix.cpp:241-294 */
	void applyStdp(float reward) {
		if(stdp.empty()) {
			throw exception(NEMO_LOGIC_ERROR, "applyStdp called, but no STDP model specified");
		}
		const StdpProcess& sp = stdp[0];
		fix_t fx_reward = fx_toFix(reward, FBITS);
		if(fx_reward == 0 && reward != 0) {
			std::ostringstream m;
			m << "STDP reward rounded down to zero. The smallest valid reward is "
				<< fx_toFloat(1, FBITS);
			throw exception(NEMO_INVALID_INPUT, m.str());
		}
		if(fx_reward == 0) {
			std::fill(rcmAcc.begin(), rcmAcc.end(), 0);
		}
		for(size_t r = 0; r < rcmAcc.size(); ++r) {
			fix_t& w_old = terminals[rcmForward[r]].weight;
			fix_t w_new = sp.updatedWeight(w_old, fx_mul(fx_reward, rcmAcc[r], FBITS));
			if(w_old != w_new) w_old = w_new;
			rcmAcc[r] = 0;
		}
	}

	/* src/nemo/nemo_c.cpp:440-461 + src/nemo/cpu/Simulation.cpp:94-111,140-216 */
	void step(const unsigned* fs, size_t nf, const unsigned* ii, const float* iv, size_t ni) {
		if(!finalized) throw exception(NEMO_LOGIC_ERROR, "simulation not created");
		std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
		for(size_t i = 0; i < nf; ++i) fstim.at(localIdx(fs[i])) = 1;
		for(size_t i = 0; i < ni; ++i) curExt[localIdx(ii[i])] = iv[i];
		deliverSpikes();
		for(size_t gi = 0; gi < groups.size(); ++gi) {
			size_t n = groups[gi].gidx.size();
			if(n == 0) continue;
			unsigned start = groupBase[gi], end = unsigned(groupBase[gi] + n);
			unsigned model = groups[gi].model;
			if(model == MODEL_IZHIKEVICH) {
				update_izhikevich(start, end, &simParam[gi][0], n, &simState[gi][0], n,
						&fstim[0], &simRng[gi][0], &curE[0], &curI[0], &curExt[0],
						&recentFiring[0], &fired[0]);
			} else if(model == MODEL_POISSON_SOURCE) {
				update_poisson_source(start, end, &simParam[gi][0], &fstim[0], &simRng[gi][0],
						&recentFiring[0], &fired[0]);
			} else if(model == MODEL_INPUT) {
				update_input(start, end, &fstim[0], &recentFiring[0], &fired[0]);
			} else if(model == MODEL_KURAMOTO) {
				/* src/nemo/cpu/Neurons.cpp:121-139: the current plane moves on
				 * before the update; the cycle is the timer's (unsigned) */
				stateCurrent[gi] = unsigned((unsigned(cycle) + 1u) % KURAMOTO_HISTORY);
				update_kuramoto(start, end, unsigned(cycle), &simParam[gi][0], &simState[gi][0], n,
						kStart, kSource, kDelay, kWeight);
			} else {
				update_if_curr_exp(start, end, model == MODEL_IF_CURR_EXP_NOISE,
						&simParam[gi][0], n, &simState[gi][0], n,
						&fstim[0], &simRng[gi][0], &curE[0], &curI[0], &curExt[0],
						&recentFiring[0], &fired[0]);
			}
		}
		accumulateStdp();
		/* setFiring (Simulation.cpp:199-208): ascending local index */
		std::vector<unsigned> out;
		for(size_t n = 0; n < N; ++n) {
			if(fired[n]) out.push_back(localToGlobal[n]);
		}
		firingFifo.push_back(out);
		cycle += 1;
		/* readFiring (FiringBuffer.cpp:33-41) */
		firingFifo.erase(firingFifo.begin());
		lastFired = firingFifo.front();
		std::chrono::steady_clock::time_point t1 = std::chrono::steady_clock::now();
		wallMs += std::chrono::duration<double, std::milli>(t1 - t0).count();
	}

	std::pair<unsigned, unsigned> groupIndex(unsigned g) const {
		unsigned l = localIdx(g);
		unsigned gi = localGroup[l];
		return std::make_pair(gi, l - groupBase[gi]);
	}

	float getState(unsigned g, unsigned var) const {
		std::pair<unsigned, unsigned> a = groupIndex(g);
		const TypeInfo& ti = TYPES[groups[a.first].model];
		if(var >= ti.nstate) {
			std::ostringstream m;
			m << "Invalid state variable index " << var;
			throw exception(NEMO_INVALID_INPUT, m.str());
		}
		size_t n = groups[a.first].gidx.size();
		return simState[a.first][(size_t(stateCurrent[a.first]) * ti.nstate + var) * n + a.second];
	}

	float getParam(unsigned g, unsigned p) const {
		std::pair<unsigned, unsigned> a = groupIndex(g);
		const TypeInfo& ti = TYPES[groups[a.first].model];
		if(p >= ti.nparam) {
			std::ostringstream m;
			m << "Invalid parameter index " << p;
			throw exception(NEMO_INVALID_INPUT, m.str());
		}
		return simParam[a.first][p * groups[a.first].gidx.size() + a.second];
	}

	void setState(unsigned g, unsigned var, float v) {
		std::pair<unsigned, unsigned> a = groupIndex(g);
		const TypeInfo& ti = TYPES[groups[a.first].model];
		if(var >= ti.nstate) {
			std::ostringstream m;
			m << "Invalid state variable index " << var;
			throw exception(NEMO_INVALID_INPUT, m.str());
		}
		size_t n = groups[a.first].gidx.size();
		simState[a.first][(size_t(stateCurrent[a.first]) * ti.nstate + var) * n + a.second] = v;
	}

	void setParam(unsigned g, unsigned p, float v) {
		std::pair<unsigned, unsigned> a = groupIndex(g);
		if(p >= TYPES[groups[a.first].model].nparam) {
			std::ostringstream m;
			m << "Invalid parameter index " << p;
			throw exception(NEMO_INVALID_INPUT, m.str());
		}
		simParam[a.first][p * groups[a.first].gidx.size() + a.second] = v;
	}

	void setNeuron(unsigned g, unsigned nargs, const float* args) {
		std::pair<unsigned, unsigned> a = groupIndex(g);
		const TypeInfo& ti = TYPES[groups[a.first].model];
		if(nargs != ti.nparam + ti.nstate) {
			std::ostringstream m;
			m << "Unexpected number of parameters/state variables when modifying neuron. Expected "
				<< ti.nparam + ti.nstate << ", found " << nargs;
			throw exception(NEMO_INVALID_INPUT, m.str());
		}
		/* src/nemo/cpu/Neurons.cpp:66-81: the state goes to plane 0, whichever
		 * plane is current */
		size_t n = groups[a.first].gidx.size();
		for(unsigned i = 0; i < ti.nparam; ++i) simParam[a.first][i * n + a.second] = args[i];
		for(unsigned i = 0; i < ti.nstate; ++i) simState[a.first][i * n + a.second] = args[ti.nparam + i];
	}

	float membranePotential(unsigned g) const {
		std::pair<unsigned, unsigned> a = groupIndex(g);
		return getState(g, TYPES[groups[a.first].model].membranePotential);
	}

	/* flat index of synapse id (src/nemo/ConnectivityMatrix.cpp:343-360) */
	size_t flatOf(uint64_t id) const {
		unsigned src = unsigned(id >> 32);
		std::unordered_map<unsigned, unsigned>::const_iterator i = globalToLocal.find(src);
		if(i == globalToLocal.end() || qStart[i->second + 1] == qStart[i->second]) {
			std::ostringstream m;
			m << "Non-existing neuron id (" << src << ") in synapse query";
			throw exception(NEMO_INVALID_INPUT, m.str());
		}
		size_t k = size_t(id & 0xffffffffu);
		if(k >= qStart[i->second + 1] - qStart[i->second]) {
			throw exception(NEMO_INVALID_INPUT, "synapse id out of range");
		}
		return qFlat[qStart[i->second] + k];
	}

	float synapseWeight(uint64_t id) const {
		return fx_toFloat(terminals[flatTerm[flatOf(id)]].weight, FBITS);
	}
	unsigned synapseTarget(uint64_t id) const {
		return localToGlobal[terminals[flatTerm[flatOf(id)]].target];
	}
	unsigned synapseDelay(uint64_t id) const {
		return synDelay[flatOf(id)];
	}
	unsigned char synapsePlastic(uint64_t id) const {
		return synPlastic[flatOf(id)];
	}
};

} // namespace oracle


/* ============================================================ C boundary */

using namespace oracle;

static std::string g_err;

#define ORACLE_CALL(...) \
	try { __VA_ARGS__; return NEMO_OK; } \
	catch(oracle::exception& e) { g_err = e.what(); return e.code; } \
	catch(std::exception& e) { g_err = e.what(); return NEMO_UNKNOWN_ERROR; } \
	catch(...) { g_err = "unknown exception"; return NEMO_UNKNOWN_ERROR; }

extern "C" {

const char* oracle_strerror() { return g_err.c_str(); }

void* oracle_new() { return new Oracle(); }

void oracle_delete(void* o) { delete static_cast<Oracle*>(o); }

int oracle_num_threads()
{
#ifdef _OPENMP
	return omp_get_max_threads();
#else
	return 1;
#endif
}

/* Threads of the calling host thread's parallel loops (one oracle per host
 * thread can then run many small networks side by side). */
void oracle_set_threads(int n)
{
#ifdef _OPENMP
	omp_set_num_threads(n > 0 ? n : 1);
#else
	(void) n;
#endif
}

int oracle_add_neurons(void* o, unsigned model, const unsigned* gidx, size_t n,
		unsigned nargs, const float* args)
{
	ORACLE_CALL(
		Oracle* s = static_cast<Oracle*>(o);
		for(size_t i = 0; i < n; ++i) s->addNeuron(model, gidx[i], nargs, args + i * nargs);
	)
}

int oracle_add_synapses(void* o, const unsigned* src, const unsigned* tgt, const unsigned* delay,
		const float* w, const unsigned char* plastic, size_t n, uint64_t* ids)
{
	ORACLE_CALL(
		Oracle* s = static_cast<Oracle*>(o);
		for(size_t i = 0; i < n; ++i) {
			uint64_t id = s->addSynapse(src[i], tgt[i], delay[i], w[i], plastic ? plastic[i] : 0);
			if(ids) ids[i] = id;
		}
	)
}

/* Bulk form for generated networks: sources given as a CSR over sources
 * 0..nsrc-1 (neuron k owns synapses [rowptr[k], rowptr[k+1])). */
int oracle_add_synapses_csr(void* o, size_t nsrc, const uint64_t* rowptr, const unsigned* tgt,
		const unsigned char* delay, const float* w, const unsigned char* plastic)
{
	ORACLE_CALL(
		Oracle* s = static_cast<Oracle*>(o);
		size_t total = size_t(rowptr[nsrc] - rowptr[0]);
		s->synSrc.reserve(s->synSrc.size() + total);
		s->synTgt.reserve(s->synTgt.size() + total);
		s->synDelay.reserve(s->synDelay.size() + total);
		s->synWeight.reserve(s->synWeight.size() + total);
		s->synPlastic.reserve(s->synPlastic.size() + total);
		for(size_t k = 0; k < nsrc; ++k) {
			for(uint64_t i = rowptr[k]; i < rowptr[k + 1]; ++i) {
				if(delay[i] < 1) throw exception(NEMO_INVALID_INPUT, "Invalid delay (0)");
				s->appendSynapse(unsigned(k), tgt[i], delay[i], w[i], plastic ? plastic[i] : (unsigned char) 0);
			}
			if(rowptr[k + 1] > rowptr[k]) s->outCount[unsigned(k)] += uint32_t(rowptr[k + 1] - rowptr[k]);
		}
	)
}

int oracle_set_stdp(void* o, const float* pre, size_t npre, const float* post, size_t npost,
		float minE, float maxE, float minI, float maxI)
{
	ORACLE_CALL(
		Oracle* s = static_cast<Oracle*>(o);
		s->hasStdp = true;
		s->stdpPre.assign(pre, pre + npre);
		s->stdpPost.assign(post, post + npost);
		s->sMinE = minE; s->sMaxE = maxE; s->sMinI = minI; s->sMaxI = maxI;
		/* validate now, as StdpFunction's constructor does */
		StdpProcess check(s->stdpPre, s->stdpPost, minE, maxE, minI, maxI, FBITS);
		(void) check;
	)
}

int oracle_finalize(void* o) { ORACLE_CALL(static_cast<Oracle*>(o)->finalize()) }

size_t oracle_neuron_count(void* o) { return static_cast<Oracle*>(o)->N; }

int oracle_step(void* o, const unsigned* fs, size_t nf, const unsigned* ii, const float* iv, size_t ni,
		unsigned* fired_out, size_t* nfired)
{
	ORACLE_CALL(
		Oracle* s = static_cast<Oracle*>(o);
		s->step(fs, nf, ii, iv, ni);
		if(fired_out) std::copy(s->lastFired.begin(), s->lastFired.end(), fired_out);
		if(nfired) *nfired = s->lastFired.size();
	)
}

/* Runs nsteps without stimulus; writes the fired count of every step to
 * counts[] and the fired ids back to back to ids[] (capacity cap). Returns
 * NEMO_INVALID_INPUT if the ids do not fit. */
int oracle_run(void* o, unsigned nsteps, unsigned* counts, unsigned* ids, size_t cap, size_t* total)
{
	ORACLE_CALL(
		Oracle* s = static_cast<Oracle*>(o);
		size_t k = 0;
		for(unsigned t = 0; t < nsteps; ++t) {
			s->step(0, 0, 0, 0, 0);
			counts[t] = unsigned(s->lastFired.size());
			if(ids) {
				if(k + s->lastFired.size() > cap) throw exception(NEMO_INVALID_INPUT, "raster capacity exceeded");
				std::copy(s->lastFired.begin(), s->lastFired.end(), ids + k);
			}
			k += s->lastFired.size();
		}
		if(total) *total = k;
	)
}

int oracle_apply_stdp(void* o, float reward) { ORACLE_CALL(static_cast<Oracle*>(o)->applyStdp(reward)) }

int oracle_get_state(void* o, unsigned g, unsigned var, float* out)
{ ORACLE_CALL(*out = static_cast<Oracle*>(o)->getState(g, var)) }

int oracle_get_param(void* o, unsigned g, unsigned p, float* out)
{ ORACLE_CALL(*out = static_cast<Oracle*>(o)->getParam(g, p)) }

int oracle_set_state(void* o, unsigned g, unsigned var, float v)
{ ORACLE_CALL(static_cast<Oracle*>(o)->setState(g, var, v)) }

int oracle_set_param(void* o, unsigned g, unsigned p, float v)
{ ORACLE_CALL(static_cast<Oracle*>(o)->setParam(g, p, v)) }

int oracle_set_neuron(void* o, unsigned g, unsigned nargs, const float* args)
{ ORACLE_CALL(static_cast<Oracle*>(o)->setNeuron(g, nargs, args)) }

int oracle_get_membrane_potential(void* o, unsigned g, float* out)
{ ORACLE_CALL(*out = static_cast<Oracle*>(o)->membranePotential(g)) }

/* All state variable `var` of every neuron, in sim-local order. */
int oracle_get_state_all(void* o, unsigned var, float* out)
{
	ORACLE_CALL(
		Oracle* s = static_cast<Oracle*>(o);
		/* NaN for a neuron whose type has no state variable var */
		for(size_t l = 0; l < s->N; ++l) {
			out[l] = var < TYPES[s->groups[s->localGroup[l]].model].nstate
				? s->getState(s->localToGlobal[l], var) : std::nanf("");
		}
	)
}

/* RNG state (4 words) of every neuron, in sim-local order. */
int oracle_get_rng_all(void* o, uint32_t* out)
{
	ORACLE_CALL(
		Oracle* s = static_cast<Oracle*>(o);
		for(size_t l = 0; l < s->N; ++l) {
			unsigned gi = s->localGroup[l];
			const RNG& r = s->simRng[gi][l - s->groupBase[gi]];
			for(int k = 0; k < 4; ++k) out[4 * l + k] = r.state[k];
		}
	)
}

int oracle_get_synapse_weight(void* o, uint64_t id, float* w)
{ ORACLE_CALL(*w = static_cast<Oracle*>(o)->synapseWeight(id)) }

int oracle_get_synapse_target(void* o, uint64_t id, unsigned* t)
{ ORACLE_CALL(*t = static_cast<Oracle*>(o)->synapseTarget(id)) }

int oracle_get_synapse_delay(void* o, uint64_t id, unsigned* d)
{ ORACLE_CALL(*d = static_cast<Oracle*>(o)->synapseDelay(id)) }

int oracle_get_synapse_plastic(void* o, uint64_t id, unsigned char* p)
{ ORACLE_CALL(*p = static_cast<Oracle*>(o)->synapsePlastic(id)) }

/* Fixed-point weights of all synapses of sources 0..nsrc-1 in insertion
 * order (the order of oracle_add_synapses_csr), for bulk weight comparison. */
int oracle_get_weights_fx(void* o, size_t nsrc, int32_t* out)
{
	ORACLE_CALL(
		Oracle* s = static_cast<Oracle*>(o);
		(void) nsrc;
		for(size_t i = 0; i < s->flatTerm.size(); ++i) {
			out[i] = s->terminals[s->flatTerm[i]].weight;
		}
	)
}

unsigned long oracle_elapsed_simulation(void* o) { return (unsigned long) static_cast<Oracle*>(o)->cycle; }
double oracle_elapsed_wall_ms(void* o) { return static_cast<Oracle*>(o)->wallMs; }
void oracle_reset_timer(void* o) { static_cast<Oracle*>(o)->cycle = 0; static_cast<Oracle*>(o)->wallMs = 0; }

/* ---- scalar pieces, exposed for known-answer tests */

int oracle_fx_toFix(float f, int32_t* out) { ORACLE_CALL(*out = fx_toFix(f, FBITS)) }
float oracle_wfx_toFloat(int64_t v) { return wfx_toFloat(v, FBITS); }
int32_t oracle_fx_mul(int32_t a, int32_t b) { return fx_mul(a, b, FBITS); }

void oracle_urand_seq(const uint32_t* state4, size_t n, uint32_t* out)
{
	RNG r;
	std::memcpy(r.state, state4, sizeof(r.state));
	for(size_t i = 0; i < n; ++i) out[i] = urand(&r);
}

void oracle_nrand_seq(const uint32_t* state4, size_t n, float* out)
{
	RNG r;
	std::memcpy(r.state, state4, sizeof(r.state));
	for(size_t i = 0; i < n; ++i) out[i] = nrand(&r);
}

/* Host libm on the two nrand sub-domains: log input x = m * 2^-24 for
 * m = first..first+n-1, and sin input a = F(i) * 1.46291808e-9f where F(i)
 * enumerates the distinct floats of a 32-bit unsigned (see nemosim_amd's
 * nrand domain enumeration, mirrored here). */
void oracle_libm_logf_grid(uint32_t first, size_t n, float* out)
{
	for(size_t i = 0; i < n; ++i) out[i] = logf(float(first + i) * 0x1p-24f);
}

void oracle_libm_sinf(const float* a, size_t n, float* out)
{
	for(size_t i = 0; i < n; ++i) out[i] = sinf(a[i]);
}

/* The Kuramoto phase wrap with the host libm (Kuramoto.cpp:121) */
void oracle_kuramoto_wrap(const float* a, size_t n, float* out)
{
	for(size_t i = 0; i < n; ++i) out[i] = kuramoto_wrap(a[i]);
}

void oracle_init_rng(unsigned minIdx, unsigned maxIdx, uint32_t* out)
{
	std::vector<RNG> r(maxIdx - minIdx + 1);
	initialiseRng(minIdx, maxIdx, r);
	for(size_t i = 0; i < r.size(); ++i)
		for(int k = 0; k < 4; ++k) out[4 * i + k] = r[i].state[k];
}

uint32_t oracle_mt19937_nth(unsigned n)
{
	std::mt19937 mt;
	uint32_t v = 0;
	for(unsigned i = 0; i < n; ++i) v = uint32_t(mt());
	return v;
}

}  // extern "C"
