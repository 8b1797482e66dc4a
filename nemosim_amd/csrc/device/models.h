/*
 * models.h -- per-neuron update arithmetic of the built-in neuron models and
 * the per-neuron Gaussian RNG, shared by every HIP kernel that updates
 * neurons (the fused step kernel and the plugin kernels) and by the host
 * build of the library.
 *
 * Every expression keeps the operation order of NeMo's CPU plugins, and the
 * library is compiled with -ffp-contract=off, so each value is the same IEEE
 * result the reference CPU backend computes.
 */
#ifndef NEMOSIM_MODELS_H
#define NEMOSIM_MODELS_H

#include <stdint.h>
#include "libm_f32.h"

namespace nemo_dev {

/* Neuron model ids (the order of nemosim's built-in model table) */
enum { MODEL_IZHIKEVICH = 0, MODEL_IF_CURR_EXP = 1, MODEL_IF_CURR_EXP_NOISE = 2,
	MODEL_POISSON_SOURCE = 3, MODEL_INPUT = 4 };

/* Largest parameter / state counts over the built-in models; device arrays
 * are [MAX_PARAM][n] and [MAX_STATE][n]. */
enum { MAX_PARAM = 10, MAX_STATE = 4 };

/* Izhikevich indices: src/nemo/plugins/Izhikevich.h:6-13 */
enum { IZ_A = 0, IZ_B = 1, IZ_C = 2, IZ_D = 3, IZ_SIGMA = 4 };
enum { IZ_U = 0, IZ_V = 1 };

/* IF_curr_exp indices: src/nemo/plugins/IF_curr_exp.h:4-16; IF_SIGMA is the
 * 10th parameter of IF_curr_exp_noise. */
enum { IF_V_REST = 0, IF_V_RESET = 1, IF_C_M = 2, IF_TAU_M = 3, IF_TAU_REFRAC = 4,
	IF_TAU_SYN_E = 5, IF_TAU_SYN_I = 6, IF_V_THRESH = 7, IF_I_OFFSET = 8, IF_SIGMA = 9 };
enum { IF_V = 0, IF_IE = 1, IF_II = 2, IF_LASTFIRED = 3 };

/* Parameter / state counts and RNG use per model (the model descriptors,
 * src/nemo/plugins/Izhikevich.ini:1-15, src/nemo/plugins/IF_curr_exp.ini:14-36) */
struct ModelInfoDev { int nparam; int nstate; bool normalRng; };

NEMO_HD inline ModelInfoDev model_info_dev(int m)
{
	ModelInfoDev r;
	if(m == MODEL_IZHIKEVICH) { r.nparam = 5; r.nstate = 2; r.normalRng = true; }
	else if(m == MODEL_IF_CURR_EXP) { r.nparam = 9; r.nstate = 4; r.normalRng = false; }
	else if(m == MODEL_POISSON_SOURCE) { r.nparam = 1; r.nstate = 0; r.normalRng = true; }
	else if(m == MODEL_INPUT) { r.nparam = 0; r.nstate = 0; r.normalRng = false; }
	else { r.nparam = 10; r.nstate = 4; r.normalRng = true; }
	return r;
}

struct Rng { uint32_t s[4]; };

/* xorshift128, src/nemo/RNG.cpp:9-18 */
NEMO_HD inline uint32_t urand(Rng& r)
{
	uint32_t t = (r.s[0] ^ (r.s[0] << 11));
	r.s[0] = r.s[1];
	r.s[1] = r.s[2];
	r.s[2] = r.s[3];
	r.s[3] = (r.s[3] ^ (r.s[3] >> 19)) ^ (t ^ (t >> 8));
	return r.s[3];
}

/* Correctly rounded single-precision square root (the double square root of
 * a float, rounded once, is the correctly rounded float square root). */
NEMO_HD inline float sqrt_f32(float x)
{
	return float(__builtin_sqrt(double(x)));
}

/* Correctly rounded single-precision division, same argument. */
NEMO_HD inline float div_f32(float a, float b)
{
	return float(double(a) / double(b));
}

/* Box-Muller sine branch, src/nemo/RNG.cpp:25-33 */
NEMO_HD inline float nrand(Rng& r)
{
	float a = float(urand(r)) * 1.4629180792671596810513378043098e-9f;
	float b = float(urand(r)) * 0.00000000023283064365386962890625f;
	float l = nemo_libm::logf_glibc(1.0f - b);
	float rr = sqrt_f32(-2.0f * l);
	return nemo_libm::sinf_rng(a) * rr;
}

/* Q11.20 accumulator to float with saturation, src/nemo/fixedpoint.cpp:41-49.
 * The accumulators reaching here are already clamped to the int32 range. */
NEMO_HD inline float fx20_to_float(int32_t v)
{
	return float(v) / 1048576.0f;
}

/* Izhikevich update of one neuron, src/nemo/cpu/plugins/Izhikevich.cpp:53-94,
 * with the Gaussian input sigma * nrand() already drawn (`noise`, added only
 * when sigma != 0, after E + I + external as the reference adds it).
 * p = {a, b, c, d, sigma}; s = {u, v}. Returns the fired flag. */
NEMO_HD inline bool izhikevich_drawn(const float* p, float* s, float noise,
		float currentE, float currentI, float currentExt, bool fstim)
{
	float I = currentE + currentI + currentExt;
	if(p[IZ_SIGMA] != 0.0f) {
		I += noise;
	}
	bool fired = false;
	float u = s[IZ_U];
	float v = s[IZ_V];
	const float a = p[IZ_A];
	const float b = p[IZ_B];
	for(int t = 0; t < 4; ++t) {
		if(!fired) {
			/* the right-hand side and the add to v are double, rounded once */
			v = float(double(v) + double(0.25f) * ((0.04 * double(v) + 5.0) * double(v) + 140.0 - double(u) + double(I)));
			u += 0.25f * (a * (b * v - u));
			fired = double(v) >= 30.0;
		}
	}
	fired = fired || fstim;
	if(fired) {
		v = p[IZ_C];
		u += p[IZ_D];
	}
	s[IZ_U] = u;
	s[IZ_V] = v;
	return fired;
}

/* Izhikevich update drawing its Gaussian input from the neuron's stream */
NEMO_HD inline bool izhikevich(const float* p, float* s, Rng& rng,
		float currentE, float currentI, float currentExt, bool fstim)
{
	const float noise = (p[IZ_SIGMA] != 0.0f) ? p[IZ_SIGMA] * nrand(rng) : 0.0f;
	return izhikevich_drawn(p, s, noise, currentE, currentI, currentExt, fstim);
}

/* IF_curr_exp update of one neuron, src/nemo/cpu/plugins/IF_curr_exp.cpp:177-222;
 * with_noise adds sigma * nrand() to the total current (IF_curr_exp_noise). */
NEMO_HD inline bool if_curr_exp(const float* p, float* s, Rng& rng, bool with_noise,
		float currentE, float currentI, float currentExt, bool fstim)
{
	float Ie = ((1.0f - div_f32(1.0f, p[IF_TAU_SYN_E])) * s[IF_IE]) + currentE;
	float Ii = ((1.0f - div_f32(1.0f, p[IF_TAU_SYN_I])) * s[IF_II]) + currentI;
	float I = Ie + Ii + currentExt + p[IF_I_OFFSET];
	if(with_noise && p[IF_SIGMA] != 0.0f) {
		I += p[IF_SIGMA] * nrand(rng);
	}
	s[IF_IE] = Ie;
	s[IF_II] = Ii;
	float v = s[IF_V];
	bool refractory = s[IF_LASTFIRED] <= p[IF_TAU_REFRAC];
	if(!refractory) {
		v += div_f32(I, p[IF_C_M]) + div_f32(p[IF_V_REST] - v, p[IF_TAU_M]);
	}
	bool fired = (!refractory && v > p[IF_V_THRESH]) || fstim;
	if(fired) {
		s[IF_LASTFIRED] = 1.0f;
		s[IF_V] = p[IF_V_RESET];
	} else {
		s[IF_LASTFIRED] += 1.0f;
		s[IF_V] = v;
	}
	return fired;
}

/* PoissonSource, src/nemo/cpu/plugins/PoissonSource.cpp:22-29: parameter 0
 * is the firing probability per cycle; fires when the low 16 bits of the
 * neuron's next urand are below rate * 2^16, or when forced; input is
 * ignored. The float -> unsigned conversion goes through int64, which is
 * what the host compiler's conversion does (so a negative or NaN rate gives
 * the host's threshold, not a clamped one). */
NEMO_HD inline bool poisson_source(const float* p, Rng& rng, bool fstim)
{
	const uint32_t threshold = uint32_t(int64_t(p[0] * 65536.0f));
	const uint32_t draw = urand(rng) & 0xffffu;
	return draw < threshold || fstim;
}

NEMO_HD inline bool update_neuron(int model, const float* p, float* s, Rng& rng,
		float currentE, float currentI, float currentExt, bool fstim)
{
	if(model == MODEL_IZHIKEVICH) {
		return izhikevich(p, s, rng, currentE, currentI, currentExt, fstim);
	}
	if(model == MODEL_POISSON_SOURCE) {
		return poisson_source(p, rng, fstim);
	}
	if(model == MODEL_INPUT) {
		/* Input, src/nemo/cpu/plugins/Input.cpp:22-27: fires exactly when forced */
		return fstim;
	}
	return if_curr_exp(p, s, rng, model == MODEL_IF_CURR_EXP_NOISE, currentE, currentI, currentExt, fstim);
}

/* ------------------------------------------------------------------ STDP */

/* STDP window tables, src/nemo/StdpProcess.cpp:20-46 */
struct StdpParams {
	uint64_t potentiation;
	uint64_t depression;
	uint64_t prePost;
	uint64_t postPre;
	int32_t postLen;
	int32_t preLen;
	int32_t fnPre[64];
	int32_t fnPost[64];
	int32_t minE, maxE, minI, maxI;
};

/* src/nemo/StdpProcess.cpp:87-113 */
NEMO_HD inline int32_t stdp_region(const StdpParams& sp, uint64_t arrivals)
{
	int32_t w = 0;
	if(arrivals) {
		const unsigned NONE = ~0u;
		uint64_t pre = arrivals & sp.prePost;
		uint64_t post = arrivals & sp.postPre;
		unsigned dtPre = pre ? unsigned(__builtin_ctzll(pre >> sp.postLen)) : NONE;
		unsigned dtPost = post ? unsigned(__builtin_clzll(post << (64 - sp.postLen))) : NONE;
		if(dtPre < dtPost) {
			w = sp.fnPre[dtPre];
		} else if(dtPost < dtPre) {
			w = sp.fnPost[dtPost];
		}
	}
	return w;
}

/* src/nemo/StdpProcess.cpp:116-124 */
NEMO_HD inline int32_t stdp_weight_change(const StdpParams& sp, uint64_t preFiring)
{
	return int32_t(uint32_t(stdp_region(sp, preFiring & sp.potentiation))
			+ uint32_t(stdp_region(sp, preFiring & sp.depression)));
}

/* src/nemo/StdpProcess.cpp:128-138 */
NEMO_HD inline int32_t stdp_updated_weight(const StdpParams& sp, int32_t w_old, int32_t w_diff)
{
	int32_t w_new = 0;
	int32_t sum = int32_t(uint32_t(w_old) + uint32_t(w_diff));
	if(w_old > 0) {
		w_new = sum > sp.minE ? sum : sp.minE;
		w_new = sp.maxE < w_new ? sp.maxE : w_new;
	} else if(w_old < 0) {
		w_new = sum > sp.maxI ? sum : sp.maxI;
		w_new = sp.minI < w_new ? sp.minI : w_new;
	}
	return w_new;
}

/* src/nemo/fixedpoint.hpp:42-46 */
NEMO_HD inline int32_t fx_mul20(int32_t a, int32_t b)
{
	int64_t r = int64_t(a) * int64_t(b);
	return int32_t(r >> 20);
}

} // namespace nemo_dev

#endif
