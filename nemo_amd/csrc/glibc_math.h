/*
 * Bit-exact single-precision logf, sinf and expf for HIP device code (and for
 * host code, where tests compare them against the C library).
 *
 * NeMo's per-neuron Gaussian noise calls the host C library's logf and sinf
 * (src/nemo/RNG.cpp:28-32) and its HH model calls expf (src/nemo/cpu/plugins/
 * HH.cpp:92-98).  On the x86-64 hosts this backend is checked against, those
 * are glibc 2.35's implementations, which come from Szabolcs Nagy's
 * "optimized-routines" (Arm Ltd., MIT/Apache-2.0; in glibc since 2.28).  They
 * are not correctly rounded (logf 0.82 ulp, sinf 0.56 ulp, expf 0.51 ulp), so
 * a bit-exact device port must evaluate the same published algorithm with the
 * same tables, polynomials and operation order.  On hosts with FMA, glibc
 * selects its FMA build of these routines, in which the compiler fused every
 * a*b+c; the fma() calls below reproduce that build.
 *
 * Algorithms (restated from the published method, not from any binary):
 *   logf: x = 2^k z, z in [0x3f330000, 2*0x3f330000); 16-entry table of
 *         (1/c, log c); log x = k ln2 + log c + log1p(z/c - 1), log1p by a
 *         degree-4 polynomial evaluated in double.
 *   sinf: |x| < 0.75 (top-12-bit test against pi/4): odd polynomial in double;
 *         |x| < 120: quadrant n = round(x 2/pi) by a 2^24-scaled product,
 *         r = x - n pi/2, then the sine or cosine polynomial by quadrant.
 *   expf: x/ln2 = (k + r)/32 with k rounded via the 1.5*2^52 shift; 2^(k/32)
 *         from a 32-entry table of 2^(i/32) bit patterns, 2^(r/32) by a cubic.
 *
 * Domains: logf all finite x > 0 (+ special cases); sinf |x| < 120 (larger
 * arguments need the Payne-Hanek path and are rejected with NaN -- no caller
 * in this library produces them); expf all finite x.
 * The tests check these against the host library exhaustively over every
 * float the callers can produce.
 */
#ifndef NEMO_AMD_GLIBC_MATH_H
#define NEMO_AMD_GLIBC_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define NEMO_HD __host__ __device__ inline
#else
#define NEMO_HD inline
#include <math.h>
#include <string.h>
#endif

namespace nemo_math {

NEMO_HD uint32_t asuint(float f) { return __builtin_bit_cast(uint32_t, f); }
NEMO_HD float asfloat(uint32_t u) { return __builtin_bit_cast(float, u); }
NEMO_HD uint64_t asuint64(double f) { return __builtin_bit_cast(uint64_t, f); }
NEMO_HD double asdouble(uint64_t u) { return __builtin_bit_cast(double, u); }

NEMO_HD double fma_d(double a, double b, double c) { return __builtin_fma(a, b, c); }

/* ---------------------------------------------------------------- logf */

NEMO_HD float logf_exact(float x)
{
	/* (1/c, log c) for the 16 subintervals of [0x3f330000, 2*0x3f330000) */
	const double invc[16] = {
		0x1.661ec79f8f3bep+0, 0x1.571ed4aaf883dp+0, 0x1.49539f0f010bp+0, 0x1.3c995b0b80385p+0,
		0x1.30d190c8864a5p+0, 0x1.25e227b0b8eap+0, 0x1.1bb4a4a1a343fp+0, 0x1.12358f08ae5bap+0,
		0x1.0953f419900a7p+0, 0x1p+0, 0x1.e608cfd9a47acp-1, 0x1.ca4b31f026aap-1,
		0x1.b2036576afce6p-1, 0x1.9c2d163a1aa2dp-1, 0x1.886e6037841edp-1, 0x1.767dcf5534862p-1,
	};
	const double logc[16] = {
		-0x1.57bf7808caadep-2, -0x1.2bef0a7c06ddbp-2, -0x1.01eae7f513a67p-2, -0x1.b31d8a68224e9p-3,
		-0x1.6574f0ac07758p-3, -0x1.1aa2bc79c81p-3, -0x1.a4e76ce8c0e5ep-4, -0x1.1973c5a611cccp-4,
		-0x1.252f438e10c1ep-5, 0x0p+0, 0x1.aa5aa5df25984p-5, 0x1.c5e53aa362eb4p-4,
		0x1.526e57720db08p-3, 0x1.bc2860d22477p-3, 0x1.1058bc8a07ee1p-2, 0x1.4043057b6ee09p-2,
	};
	const double A0 = -0x1.00ea348b88334p-2, A1 = 0x1.5575b0be00b6ap-2, A2 = -0x1.ffffef20a4123p-2;
	const double Ln2 = 0x1.62e42fefa39efp-1;

	uint32_t ix = asuint(x);
	if(ix == 0x3f800000u) {
		return 0.0f;
	}
	if(ix - 0x00800000u >= 0x7f800000u - 0x00800000u) {
		if(ix * 2 == 0) return -__builtin_inff();
		if(ix == 0x7f800000u) return x;
		if((ix & 0x80000000u) || ix * 2 >= 0xff000000u) return __builtin_nanf("");
		/* subnormal: normalise */
		ix = asuint(x * 0x1p23f);
		ix -= 23u << 23;
	}
	uint32_t tmp = ix - 0x3f330000u;
	int i = int((tmp >> (23 - 4)) % 16);
	int k = int32_t(tmp) >> 23;
	uint32_t iz = ix - (tmp & 0xff800000u);
	double z = double(asfloat(iz));
	double r = fma_d(z, invc[i], -1.0);
	double y0 = fma_d(double(k), Ln2, logc[i]);
	double r2 = r * r;
	double y = fma_d(A1, r, A2);
	y = fma_d(A0, r2, y);
	y = fma_d(y, r2, y0 + r);
	return float(y);
}

/* ---------------------------------------------------------------- sinf */

NEMO_HD uint32_t abstop12(float x) { return (asuint(x) >> 20) & 0x7ff; }

/* sine (n even) or cosine (n odd) polynomial; q selects the sign table half */
NEMO_HD float sinf_poly(double x, double x2, int q, int n)
{
	const double S1 = -0x1.555545995a603p-3, S2 = 0x1.1107605230bc4p-7, S3 = -0x1.994eb3774cf24p-13;
	if((n & 1) == 0) {
		double x3 = x * x2;
		double s1 = fma_d(x2, S3, S2);
		double x7 = x3 * x2;
		double s = fma_d(x3, S1, x);
		return float(fma_d(x7, s1, s));
	} else {
		const double sg = q ? -1.0 : 1.0;
		const double C0 = sg * 0x1p0, C1 = sg * -0x1.ffffffd0c621cp-2, C2 = sg * 0x1.55553e1068f19p-5,
				C3 = sg * -0x1.6c087e89a359dp-10, C4 = sg * 0x1.99343027bf8c3p-16;
		double x4 = x2 * x2;
		double c2 = fma_d(x2, C4, C3);
		double c1 = fma_d(x2, C1, C0);
		double x6 = x4 * x2;
		double c = fma_d(x4, C2, c1);
		return float(fma_d(x6, c2, c));
	}
}

NEMO_HD float sinf_exact(float y)
{
	const double hpi_inv = 0x1.45F306DC9C883p+23; /* 2/pi * 2^24 */
	const double hpi = 0x1.921FB54442D18p0;       /* pi/2 */
	double x = y;
	if(abstop12(y) < abstop12(0x1.921fb6p-1f)) {
		double s = x * x;
		if(abstop12(y) < abstop12(0x1p-12f)) {
			return y;
		}
		return sinf_poly(x, s, 0, 0);
	} else if(abstop12(y) < abstop12(120.0f)) {
		double r = x * hpi_inv;
		int n = ((int32_t)r + 0x800000) >> 24;
		x = fma_d(-double(n), hpi, x);
		double s = ((n & 3) == 1 || (n & 3) == 2) ? -1.0 : 1.0; /* sign of {1,-1,-1,1} */
		return sinf_poly(x * s, x * x, (n & 2) ? 1 : 0, n);
	}
	return __builtin_nanf("");
}

/* ---------------------------------------------------------------- expf */

/* The published expf's table: the bit patterns of 2^(i/32) with i << 47
 * subtracted, i < 32. */
#define NEMO_EXPF_TABLE { \
		0x3ff0000000000000ULL, 0x3fefd9b0d3158574ULL, 0x3fefb5586cf9890fULL, 0x3fef9301d0125b51ULL, \
		0x3fef72b83c7d517bULL, 0x3fef54873168b9aaULL, 0x3fef387a6e756238ULL, 0x3fef1e9df51fdee1ULL, \
		0x3fef06fe0a31b715ULL, 0x3feef1a7373aa9cbULL, 0x3feedea64c123422ULL, 0x3feece086061892dULL, \
		0x3feebfdad5362a27ULL, 0x3feeb42b569d4f82ULL, 0x3feeab07dd485429ULL, 0x3feea47eb03a5585ULL, \
		0x3feea09e667f3bcdULL, 0x3fee9f75e8ec5f74ULL, 0x3feea11473eb0187ULL, 0x3feea589994cce13ULL, \
		0x3feeace5422aa0dbULL, 0x3feeb737b0cdc5e5ULL, 0x3feec49182a3f090ULL, 0x3feed503b23e255dULL, \
		0x3feee89f995ad3adULL, 0x3feeff76f2fb5e47ULL, 0x3fef199bdd85529cULL, 0x3fef3720dcef9069ULL, \
		0x3fef5818dcfba487ULL, 0x3fef7c97337b9b5fULL, 0x3fefa4afa2a490daULL, 0x3fefd0765b6e4540ULL, \
	}

/* Entry i of the table (i < 32). */
NEMO_HD uint64_t expf_table(unsigned i)
{
	const uint64_t T[32] = NEMO_EXPF_TABLE;
	return T[i % 32];
}

/* expf with the table at tab (tab[i] == expf_table(i); a caller may keep it
 * in LDS).  The special cases of the published routine (|x| >= 88: -inf
 * gives 0, NaN and +inf give x + x, overflow inf, underflow 0) are selected
 * after the main path instead of branched around it: the main path is
 * harmless for those inputs (any table index is in range), and a wave whose
 * lanes take different cases stays on one instruction stream. */
NEMO_HD float expf_exact_tab(float x, const uint64_t* tab)
{
	const double InvLn2N = 0x1.71547652b82fep+0 * 32;
	const double SHIFT = 0x1.8p+52;
	const double C0 = 0x1.c6af84b912394p-5 / 32 / 32 / 32;
	const double C1 = 0x1.ebfce50fac4f3p-3 / 32 / 32;
	const double C2 = 0x1.62e42ff0c52d6p-1 / 32;

	double xd = double(x);
	double kd = fma_d(InvLn2N, xd, SHIFT);
	uint64_t ki = asuint64(kd);
	kd -= SHIFT;
	double r = fma_d(InvLn2N, xd, -kd);
	uint64_t t = tab[ki % 32];
	t += ki << (52 - 5);
	double s = asdouble(t);
	double z = fma_d(C0, r, C1);
	double r2 = r * r;
	double y = fma_d(C2, r, 1.0);
	y = fma_d(z, r2, y);
	y = y * s;
	const float main = float(y);

	const uint32_t abstop = (asuint(x) >> 20) & 0x7ff;
	float special = main;
	special = x < -0x1.9fe368p6f ? 0.0f : special;          /* underflow, and -inf */
	special = x > 0x1.62e42ep6f ? __builtin_inff() : special; /* overflow, and +inf (= x + x) */
	special = x != x ? x + x : special;                      /* NaN */
	return abstop >= (asuint(88.0f) >> 20) ? special : main;
}

NEMO_HD float expf_exact(float x)
{
	const uint64_t T[32] = NEMO_EXPF_TABLE;
	return expf_exact_tab(x, T);
}

/* ---------------------------------------------------------------- fmodf */

/* x - n y for the integer n = trunc(x / y), which IEEE 754 defines exactly
 * (the result is always representable), so any correct implementation --
 * the C library's, this one -- gives the same bits.  Shift-and-subtract
 * long division of the significands: the remainder of the aligned
 * significand of x by that of y, one exponent step at a time. */
NEMO_HD float fmodf_exact(float x, float y)
{
	uint32_t ux = asuint(x), uy = asuint(y);
	const uint32_t sign = ux & 0x80000000u;
	int ex = int((ux >> 23) & 0xff), ey = int((uy >> 23) & 0xff);
	if((uy << 1) == 0 || (uy << 1) > 0xff000000u || ex == 0xff) return (x * y) / (x * y); /* NaN */
	if((ux << 1) <= (uy << 1)) return (ux << 1) == (uy << 1) ? 0.0f * x : x;
	/* significands with the implicit bit; subnormals normalised, their
	 * exponent going below 1 */
	if(ex == 0) {
		ux <<= 9;
		while(!(ux & 0x80000000u)) {
			--ex;
			ux <<= 1;
		}
		ux >>= 8;
	} else {
		ux = (ux & 0x007fffffu) | 0x00800000u;
	}
	if(ey == 0) {
		uy <<= 9;
		while(!(uy & 0x80000000u)) {
			--ey;
			uy <<= 1;
		}
		uy >>= 8;
	} else {
		uy = (uy & 0x007fffffu) | 0x00800000u;
	}
	for(; ex > ey; --ex) {
		if(ux >= uy) {
			ux -= uy;
			if(ux == 0) return 0.0f * x;
		}
		ux <<= 1;
	}
	if(ux >= uy) {
		ux -= uy;
		if(ux == 0) return 0.0f * x;
	}
	while(!(ux & 0x00800000u)) {
		ux <<= 1;
		--ex;
	}
	if(ex > 0) {
		ux = (ux & 0x007fffffu) | (uint32_t(ex) << 23);
	} else {
		ux >>= 1 - ex;
	}
	return asfloat(ux | sign);
}

/* An argument of expf divided by a constant c > 0 (HH's rate exponents
 * divide by 20, 80 and 18, or by 1), with inv = 1/c rounded to float: the
 * product x inv corrected once by its exact FMA residual x - q c.  For every
 * float x of magnitude at least 2^-122 this is the correctly rounded x / c,
 * bit for bit; below that both quotients are under 2^-126 in magnitude, so
 * expf of either is 1.  Infinite x keep the product (the residual would be
 * NaN).  The tests check both statements over every float for the four
 * divisors.  It replaces a correctly rounded division, the longest link in
 * HH's substep chain. */
NEMO_HD float div_by_const(float x, float c, float inv)
{
	const float q0 = x * inv;
	const float r = __builtin_fmaf(-q0, c, x);
	const float q1 = __builtin_fmaf(r, inv, q0);
	return __builtin_isfinite(q0) ? q1 : q0;
}

} // namespace nemo_math

#endif
