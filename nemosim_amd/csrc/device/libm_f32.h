/*
 * libm_f32.h -- single-precision logf / sinf that reproduce the host C
 * library bit for bit on the inputs of NeMo's Gaussian RNG.
 *
 * NeMo's nrand() (src/nemo/RNG.cpp:25-33) calls the host's logf and sinf.
 * glibc >= 2.28 implements both with the algorithms published in ARM's
 * "optimized-routines" project (math/logf.c, math/sinf.c, math/sincosf.h,
 * math/logf_data.c, math/sincosf_data.c): a 16-entry table plus a degree-3
 * polynomial in double for logf, and a quadrant reduction plus even/odd
 * double polynomials for sinf. This file restates those algorithms in
 * double precision so they run on gfx950 FP64 and on the host alike.
 *
 * On FMA-capable x86-64 hosts glibc runs an FMA build of these sources, in
 * which the compiler fuses every product whose only use is a sum. On the
 * RNG's arguments the fused and unfused forms give the same results: logf
 * at all x = m * 2^-24 (m = 0..2^24) and sinf at all
 * a = float(u) * 1.4629180792671596810513378043098e-9f, u a 32-bit unsigned,
 * were compared with the host libm with zero differences
 * (tests/test_libm_f32.py repeats that check). The RNG uses the unfused form
 * (sinf_rng), which needs fewer registers in the step kernel. Beyond those
 * arguments the unfused form differs from the host in about 1 of 10^7
 * (near multiples of pi/2 of magnitude ~100, where the reduction cancels),
 * so sinf_glibc, used by the Kuramoto model on arbitrary phase differences,
 * is restated in the fused form: fma() in the same places as the host
 * build (the reduction x - n*pi/2 and each multiply-add of the polynomials).
 *
 * logf is supported for x in [0, 1] only (no NaN / negative paths). sinf
 * takes every float: besides the RNG's arguments, the Kuramoto model feeds
 * it phase differences of either sign and whatever magnitude the coupling
 * produces, so the large-argument reduction (|a| >= 120, by integer
 * multiplication with 2/pi) is restated too. The wider domain is checked
 * against the host C library by tests/test_libm_f32.py and, on the device,
 * by the GPU libm tests.
 */
#ifndef NEMOSIM_LIBM_F32_H
#define NEMOSIM_LIBM_F32_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define NEMO_HD __host__ __device__
#else
#define NEMO_HD
#endif

namespace nemo_libm {

NEMO_HD inline uint32_t f2u(float f) { union { float f; uint32_t u; } x; x.f = f; return x.u; }
NEMO_HD inline float u2f(uint32_t u) { union { float f; uint32_t u; } x; x.u = u; return x.f; }

/* 1/c and log(c) for the 16 sub-intervals of [0x3f330000, 2*0x3f330000) */
struct LogfEntry { double invc, logc; };

/* constexpr, so that clang makes the table available to device code too */
constexpr LogfEntry LOGF_TABLE[16] = {
	{ 0x1.661ec79f8f3bep+0, -0x1.57bf7808caadep-2 },
	{ 0x1.571ed4aaf883dp+0, -0x1.2bef0a7c06ddbp-2 },
	{ 0x1.49539f0f010bp+0, -0x1.01eae7f513a67p-2 },
	{ 0x1.3c995b0b80385p+0, -0x1.b31d8a68224e9p-3 },
	{ 0x1.30d190c8864a5p+0, -0x1.6574f0ac07758p-3 },
	{ 0x1.25e227b0b8eap+0, -0x1.1aa2bc79c81p-3 },
	{ 0x1.1bb4a4a1a343fp+0, -0x1.a4e76ce8c0e5ep-4 },
	{ 0x1.12358f08ae5bap+0, -0x1.1973c5a611cccp-4 },
	{ 0x1.0953f419900a7p+0, -0x1.252f438e10c1ep-5 },
	{ 0x1p+0, 0x0p+0 },
	{ 0x1.e608cfd9a47acp-1, 0x1.aa5aa5df25984p-5 },
	{ 0x1.ca4b31f026aap-1, 0x1.c5e53aa362eb4p-4 },
	{ 0x1.b2036576afce6p-1, 0x1.526e57720db08p-3 },
	{ 0x1.9c2d163a1aa2dp-1, 0x1.bc2860d22477p-3 },
	{ 0x1.886e6037841edp-1, 0x1.1058bc8a07ee1p-2 },
	{ 0x1.767dcf5534862p-1, 0x1.4043057b6ee09p-2 },
};

/* logf for x in [0, 1]: x = 2^k z with z in [OFF, 2 OFF); log(x) =
 * log1p(z/c - 1) + log(c) + k ln2, log1p by a degree-3 polynomial. */
NEMO_HD inline float logf_glibc(float x)
{
	const double A0 = -0x1.00ea348b88334p-2;
	const double A1 = 0x1.5575b0be00b6ap-2;
	const double A2 = -0x1.ffffef20a4123p-2;
	const double LN2 = 0x1.62e42fefa39efp-1;
	const uint32_t OFF = 0x3f330000u;

	uint32_t ix = f2u(x);
	if(ix == 0x3f800000u) {
		return 0.0f;
	}
	if(ix == 0u) {
		return -__builtin_inff();
	}
	uint32_t tmp = ix - OFF;
	int i = int((tmp >> 19) & 15u);
	int k = int32_t(tmp) >> 23;
	uint32_t iz = ix - (tmp & 0xff800000u);
	const LogfEntry& e = LOGF_TABLE[i];
	double z = double(u2f(iz));
	double r = z * e.invc - 1.0;
	double y0 = e.logc + double(k) * LN2;
	double r2 = r * r;
	double y = A1 * r + A2;
	y = A0 * r2 + y;
	y = y * r2 + (y0 + r);
	return float(y);
}

/* Sine / cosine polynomial coefficients */
struct SincosfSet {
	double c0, c1, c2, c3, c4, s1, s2, s3;
};

NEMO_HD inline float sinf_poly(double x, double x2, const SincosfSet& p, int n)
{
	if((n & 1) == 0) {
		double x3 = x * x2;
		double s1 = fma(x2, p.s3, p.s2);
		double x7 = x3 * x2;
		double s = fma(x3, p.s1, x);
		return float(fma(x7, s1, s));
	} else {
		double x4 = x2 * x2;
		double c2 = fma(x2, p.c4, p.c3);
		double c1 = fma(x2, p.c1, p.c0);
		double x6 = x4 * x2;
		double c = fma(x4, p.c2, c1);
		return float(fma(x6, c2, c));
	}
}

/* The polynomials without fused multiply-adds */
NEMO_HD inline float sinf_poly_unfused(double x, double x2, const SincosfSet& p, int n)
{
	if((n & 1) == 0) {
		double x3 = x * x2;
		double s1 = p.s2 + x2 * p.s3;
		double x7 = x3 * x2;
		double s = x + x3 * p.s1;
		return float(s + x7 * s1);
	} else {
		double x4 = x2 * x2;
		double c2 = p.c3 + x2 * p.c4;
		double c1 = p.c0 + x2 * p.c1;
		double x6 = x4 * x2;
		double c = c1 + x4 * p.c2;
		return float(c + x6 * c2);
	}
}

/* sinf for the RNG's arguments a in [0, 2*pi + 1ulp] (unfused form):
 * |a| < pi/4 directly, otherwise one reduction step by pi/2 (n rounded
 * through a 2^24-scaled integer). In quadrants 2 and 3 glibc evaluates the
 * cosine polynomial with every coefficient negated, which equals the
 * negated result (negation is exact in each product and sum). */
NEMO_HD inline float sinf_rng(float y)
{
	const SincosfSet P0 = { 0x1p0, -0x1.ffffffd0c621cp-2, 0x1.55553e1068f19p-5,
		-0x1.6c087e89a359dp-10, 0x1.99343027bf8c3p-16,
		-0x1.555545995a603p-3, 0x1.1107605230bc4p-7, -0x1.994eb3774cf24p-13 };
	const double HPI_INV = 0x1.45F306DC9C883p+23;
	const double HPI = 0x1.921FB54442D18p0;

	uint32_t top = (f2u(y) >> 20) & 0x7ffu;
	double x = y;
	if(top < ((f2u(0x1.921FB6p-1f) >> 20) & 0x7ffu)) {
		if(top < ((f2u(0x1p-12f) >> 20) & 0x7ffu)) {
			return y;
		}
		return sinf_poly_unfused(x, x * x, P0, 0);
	}
	double r = x * HPI_INV;
	int n = (int32_t(r) + 0x800000) >> 24;
	x = x - double(n) * HPI;
	double s = ((n & 3) == 1 || (n & 3) == 2) ? -1.0 : 1.0;
	const float v = sinf_poly_unfused(x * s, x * x, P0, n);
	return ((n & 3) == 3) ? -v : v;
}

/* Bits of 2/pi: entry i is floor(2/pi * 2^(8(i+1))) mod 2^32, i.e. the
 * 32-bit window ending at byte i+1 of the binary expansion. The table is a
 * property of the number (tests/test_libm_f32.py recomputes it from pi with
 * integer arithmetic). */
constexpr uint32_t INV_PIO4_BITS[24] = {
	0xa2u, 0xa2f9u, 0xa2f983u, 0xa2f9836eu, 0xf9836e4eu, 0x836e4e44u, 0x6e4e4415u, 0x4e441529u,
	0x441529fcu, 0x1529fc27u, 0x29fc2757u, 0xfc2757d1u, 0x2757d1f5u, 0x57d1f534u, 0xd1f534ddu, 0xf534ddc0u,
	0x34ddc0dbu, 0xddc0db62u, 0xc0db6295u, 0xdb629599u, 0x6295993cu, 0x95993c43u, 0x993c4390u, 0x3c439041u,
};

/* Reduction of |a| >= 120 (the float's bits in xi, sign ignored) to
 * r in about [-pi/4, pi/4] and the quadrant n: the 24-bit significand times
 * a 96-bit window of 2/pi (chosen by the exponent) gives the product modulo
 * 4 as a 2.62 fixed-point number, exactly; n is its rounded integer part,
 * and the remainder, converted to double, is scaled by pi/2 * 2^-62. */
NEMO_HD inline double sinf_reduce_large(uint32_t xi, int* np)
{
	const uint32_t* w = &INV_PIO4_BITS[(xi >> 26) & 15u];
	const int shift = int((xi >> 23) & 7u);
	uint32_t m = ((xi & 0xffffffu) | 0x800000u) << shift;
	uint64_t hi = uint64_t(uint32_t(m * w[0]));           /* only its low 32 bits matter */
	uint64_t mid = uint64_t(m) * w[4];
	uint64_t lo = uint64_t(m) * w[8];
	uint64_t acc = ((lo >> 32) | (hi << 32)) + mid;
	uint64_t n = (acc + (uint64_t(1) << 61)) >> 62;
	acc -= n << 62;
	*np = int(n);
	return double(int64_t(acc)) * 0x1.921FB54442D18p-62;
}

/* The polynomial of quadrant n on the reduced argument x, with sign index
 * m (n, or n plus the argument's sign bit in the large-argument path):
 * x is negated for m = 1, 2; glibc evaluates the cosine polynomial with
 * every coefficient negated when m & 2, which equals the negated result
 * (negation is exact in each product and sum), so one coefficient set
 * serves every quadrant. */
NEMO_HD inline float sinf_quadrant(double x, int n, int m)
{
	const SincosfSet P0 = { 0x1p0, -0x1.ffffffd0c621cp-2, 0x1.55553e1068f19p-5,
		-0x1.6c087e89a359dp-10, 0x1.99343027bf8c3p-16,
		-0x1.555545995a603p-3, 0x1.1107605230bc4p-7, -0x1.994eb3774cf24p-13 };
	const double s = ((m & 3) == 1 || (m & 3) == 2) ? -1.0 : 1.0;
	const float v = sinf_poly(x * s, x * x, P0, n);
	return ((n & 1) && (m & 2)) ? -v : v;
}

/* sinf of any float: |a| < pi/4 directly (a itself below 2^-12); |a| < 120
 * by one reduction step by pi/2, n rounded through a 2^24-scaled integer;
 * larger finite |a| by sinf_reduce_large; NaN for infinities and NaN. */
NEMO_HD inline float sinf_glibc(float y)
{
	const SincosfSet P0 = { 0x1p0, -0x1.ffffffd0c621cp-2, 0x1.55553e1068f19p-5,
		-0x1.6c087e89a359dp-10, 0x1.99343027bf8c3p-16,
		-0x1.555545995a603p-3, 0x1.1107605230bc4p-7, -0x1.994eb3774cf24p-13 };
	const double HPI_INV = 0x1.45F306DC9C883p+23;
	const double HPI = 0x1.921FB54442D18p0;

	const uint32_t top = (f2u(y) >> 20) & 0x7ffu;
	double x = y;
	if(top < ((f2u(0x1.921FB6p-1f) >> 20) & 0x7ffu)) {
		if(top < ((f2u(0x1p-12f) >> 20) & 0x7ffu)) {
			return y;
		}
		return sinf_poly(x, x * x, P0, 0);
	}
	if(top < ((f2u(120.0f) >> 20) & 0x7ffu)) {
		const double r = x * HPI_INV;
		const int n = (int32_t(r) + 0x800000) >> 24;
		x = fma(-double(n), HPI, x);
		return sinf_quadrant(x, n, n);
	}
	if(top < ((f2u(__builtin_inff()) >> 20) & 0x7ffu)) {
		const uint32_t xi = f2u(y);
		int n = 0;
		x = sinf_reduce_large(xi, &n);
		return sinf_quadrant(x, n, n + int(xi >> 31));
	}
	return (y - y) / (y - y);
}

/* The phase wrap of the Kuramoto model (src/nemo/cpu/plugins/Kuramoto.cpp:
 * 121, 150): fmodf by 2*pi rounded to float, then, for a negative phase,
 * 2*pi added in double -- the reference's expression mixes the float
 * remainder with the double constant 2.0f*M_PI -- and the sum rounded to
 * float. fmodf is exact, so any correct implementation gives the same
 * remainder. */
NEMO_HD inline float kuramoto_wrap(float phase)
{
	const double TWO_PI = 6.283185307179586;
	const float r = fmodf(phase, float(TWO_PI));
	return float(double(r) + (phase < 0.0f ? TWO_PI : 0.0));
}

} // namespace nemo_libm

#endif
