/*
 * zoracle.c — test-only CPU restatement of zimg's per-plane arithmetic.
 * See zoracle.h for the scope and what pins each part. This file is the
 * checker, never the product: the shipped library does not link it.
 */
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "zoracle.h"

/* ======================================================================
 * libm restatement. Method: ARM optimized-routines single-precision
 * powf / expf / logf (Szabolcs Nagy, 2017-2018, MIT), the code vendored by
 * the reference for its unit tests (test/extra/musl-libm/{powf,expf,logf}.c)
 * and swapped in by test/main.cpp:10-18. Evaluation is in double, no FMA.
 * The table constants must be identical for bit-exact results.
 * ====================================================================== */

static uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static float u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }
static uint64_t d2u(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static double u2d(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

/* 2^(i/32) as double bits, minus i<<47 (exp2f_data.c:16-27). */
static const uint64_t EXP2_TAB[32] = {
	0x3ff0000000000000ULL, 0x3fefd9b0d3158574ULL, 0x3fefb5586cf9890fULL, 0x3fef9301d0125b51ULL,
	0x3fef72b83c7d517bULL, 0x3fef54873168b9aaULL, 0x3fef387a6e756238ULL, 0x3fef1e9df51fdee1ULL,
	0x3fef06fe0a31b715ULL, 0x3feef1a7373aa9cbULL, 0x3feedea64c123422ULL, 0x3feece086061892dULL,
	0x3feebfdad5362a27ULL, 0x3feeb42b569d4f82ULL, 0x3feeab07dd485429ULL, 0x3feea47eb03a5585ULL,
	0x3feea09e667f3bcdULL, 0x3fee9f75e8ec5f74ULL, 0x3feea11473eb0187ULL, 0x3feea589994cce13ULL,
	0x3feeace5422aa0dbULL, 0x3feeb737b0cdc5e5ULL, 0x3feec49182a3f090ULL, 0x3feed503b23e255dULL,
	0x3feee89f995ad3adULL, 0x3feeff76f2fb5e47ULL, 0x3fef199bdd85529cULL, 0x3fef3720dcef9069ULL,
	0x3fef5818dcfba487ULL, 0x3fef7c97337b9b5fULL, 0x3fefa4afa2a490daULL, 0x3fefd0765b6e4540ULL,
};
static const double EXP2_POLY[3] = { 0x1.c6af84b912394p-5, 0x1.ebfce50fac4f3p-3, 0x1.62e42ff0c52d6p-1 };
static const double EXP2_SHIFT_SCALED = 0x1.8p+52 / 32;
static const double EXP2_SHIFT = 0x1.8p+52;
static const double EXP_INVLN2_SCALED = 0x1.71547652b82fep+0 * 32;
static const double EXP_POLY_SCALED[3] = {
	0x1.c6af84b912394p-5 / 32 / 32 / 32, 0x1.ebfce50fac4f3p-3 / 32 / 32, 0x1.62e42ff0c52d6p-1 / 32,
};

/* Reciprocal centres and log2 of 16 subintervals of [0x3f330000, 2*that] (powf_data.c). */
static const double POWF_INVC[16] = {
	0x1.661ec79f8f3bep+0, 0x1.571ed4aaf883dp+0, 0x1.49539f0f010bp+0, 0x1.3c995b0b80385p+0,
	0x1.30d190c8864a5p+0, 0x1.25e227b0b8eap+0, 0x1.1bb4a4a1a343fp+0, 0x1.12358f08ae5bap+0,
	0x1.0953f419900a7p+0, 0x1p+0, 0x1.e608cfd9a47acp-1, 0x1.ca4b31f026aap-1,
	0x1.b2036576afce6p-1, 0x1.9c2d163a1aa2dp-1, 0x1.886e6037841edp-1, 0x1.767dcf5534862p-1,
};
static const double POWF_LOG2C[16] = {
	-0x1.efec65b963019p-2, -0x1.b0b6832d4fca4p-2, -0x1.7418b0a1fb77bp-2, -0x1.39de91a6dcf7bp-2,
	-0x1.01d9bf3f2b631p-2, -0x1.97c1d1b3b7afp-3, -0x1.2f9e393af3c9fp-3, -0x1.960cbbf788d5cp-4,
	-0x1.a6f9db6475fcep-5, 0x0p+0, 0x1.338ca9f24f53dp-4, 0x1.476a9543891bap-3,
	0x1.e840b4ac4e4d2p-3, 0x1.40645f0c6651cp-2, 0x1.88e9c2c1b9ff8p-2, 0x1.ce0a44eb17bccp-2,
};
static const double POWF_POLY[5] = {
	0x1.27616c9496e0bp-2, -0x1.71969a075c67ap-2, 0x1.ec70a6ca7baddp-2, -0x1.7154748bef6c8p-1,
	0x1.71547652ab82bp0,
};
/* natural-log table for logf (logf_data.c): same invc, logc = ln(c). */
static const double LOGF_LOGC[16] = {
	-0x1.57bf7808caadep-2, -0x1.2bef0a7c06ddbp-2, -0x1.01eae7f513a67p-2, -0x1.b31d8a68224e9p-3,
	-0x1.6574f0ac07758p-3, -0x1.1aa2bc79c81p-3, -0x1.a4e76ce8c0e5ep-4, -0x1.1973c5a611cccp-4,
	-0x1.252f438e10c1ep-5, 0x0p+0, 0x1.aa5aa5df25984p-5, 0x1.c5e53aa362eb4p-4,
	0x1.526e57720db08p-3, 0x1.bc2860d22477p-3, 0x1.1058bc8a07ee1p-2, 0x1.4043057b6ee09p-2,
};
static const double LOGF_LN2 = 0x1.62e42fefa39efp-1;
static const double LOGF_POLY[3] = { -0x1.00ea348b88334p-2, 0x1.5575b0be00b6ap-2, -0x1.ffffef20a4123p-2 };

#define LOG_OFF 0x3f330000u

static double orc_log2_core(uint32_t ix)
{
	uint32_t tmp = ix - LOG_OFF;
	int i = (int)((tmp >> 19) % 16);
	uint32_t top = tmp & 0xff800000u;
	uint32_t iz = ix - top;
	int k = (int32_t)top >> 23;
	double z = (double)u2f(iz);
	double r = z * POWF_INVC[i] - 1.0;
	double y0 = POWF_LOG2C[i] + (double)k;
	double r2 = r * r;
	double y = POWF_POLY[0] * r + POWF_POLY[1];
	double p = POWF_POLY[2] * r + POWF_POLY[3];
	double r4 = r2 * r2;
	double q = POWF_POLY[4] * r + y0;
	q = p * r2 + q;
	y = y * r4 + q;
	return y;
}

static float orc_exp2_core(double xd, uint32_t sign_bias)
{
	double kd = xd + EXP2_SHIFT_SCALED;
	uint64_t ki = d2u(kd);
	kd -= EXP2_SHIFT_SCALED;
	double r = xd - kd;
	uint64_t t = EXP2_TAB[ki % 32];
	uint64_t ski = ki + sign_bias;
	t += ski << 47;
	double s = u2d(t);
	double z = EXP2_POLY[0] * r + EXP2_POLY[1];
	double r2 = r * r;
	double y = EXP2_POLY[2] * r + 1.0;
	y = z * r2 + y;
	y = y * s;
	return (float)y;
}

static int orc_checkint(uint32_t iy)
{
	int e = iy >> 23 & 0xff;
	if (e < 0x7f)
		return 0;
	if (e > 0x7f + 23)
		return 2;
	if (iy & ((1u << (0x7f + 23 - e)) - 1))
		return 0;
	if (iy & (1u << (0x7f + 23 - e)))
		return 1;
	return 2;
}

static int orc_zeroinfnan(uint32_t ix) { return 2 * ix - 1 >= 2u * 0x7f800000u - 1; }

float orc_powf(float x, float y)
{
	uint32_t sign_bias = 0;
	uint32_t ix = f2u(x), iy = f2u(y);

	if (ix - 0x00800000u >= 0x7f800000u - 0x00800000u || orc_zeroinfnan(iy)) {
		if (orc_zeroinfnan(iy)) {
			if (2 * iy == 0)
				return 1.0f;
			if (ix == 0x3f800000u)
				return 1.0f;
			if (2 * ix > 2u * 0x7f800000u || 2 * iy > 2u * 0x7f800000u)
				return x + y;
			if (2 * ix == 2 * 0x3f800000u)
				return 1.0f;
			if ((2 * ix < 2 * 0x3f800000u) == !(iy & 0x80000000u))
				return 0.0f;
			return y * y;
		}
		if (orc_zeroinfnan(ix)) {
			float x2 = x * x;
			if ((ix & 0x80000000u) && orc_checkint(iy) == 1)
				x2 = -x2;
			return (iy & 0x80000000u) ? 1.0f / x2 : x2;
		}
		if (ix & 0x80000000u) {
			int yint = orc_checkint(iy);
			if (yint == 0)
				return NAN;
			if (yint == 1)
				sign_bias = 1u << (5 + 11);
			ix &= 0x7fffffffu;
		}
		if (ix < 0x00800000u) {
			ix = f2u(x * 0x1p23f);
			ix &= 0x7fffffffu;
			ix -= 23u << 23;
		}
	}
	double logx = orc_log2_core(ix);
	double ylogx = (double)y * logx;
	if (((d2u(ylogx) >> 47) & 0xffff) >= (d2u(126.0) >> 47)) {
		if (ylogx > 0x1.fffffffd1d571p+6)
			return sign_bias ? -INFINITY : INFINITY;
		if (ylogx <= -150.0)
			return sign_bias ? -0.0f : 0.0f;
	}
	return orc_exp2_core(ylogx, sign_bias);
}

float orc_expf(float x)
{
	uint32_t abstop = (f2u(x) >> 20) & 0x7ff;
	double xd = (double)x;

	if (abstop >= (f2u(88.0f) >> 20)) {
		if (f2u(x) == f2u(-INFINITY))
			return 0.0f;
		if (abstop >= (f2u(INFINITY) >> 20))
			return x + x;
		if (x > 0x1.62e42ep6f)
			return INFINITY;
		if (x < -0x1.9fe368p6f)
			return 0.0f;
	}
	double z = EXP_INVLN2_SCALED * xd;
	double kd = z + EXP2_SHIFT;
	uint64_t ki = d2u(kd);
	kd -= EXP2_SHIFT;
	double r = z - kd;
	uint64_t t = EXP2_TAB[ki % 32];
	t += ki << 47;
	double s = u2d(t);
	z = EXP_POLY_SCALED[0] * r + EXP_POLY_SCALED[1];
	double r2 = r * r;
	double y = EXP_POLY_SCALED[2] * r + 1.0;
	y = z * r2 + y;
	y = y * s;
	return (float)y;
}

float orc_logf(float x)
{
	uint32_t ix = f2u(x);
	if (ix == 0x3f800000u)
		return 0.0f;
	if (ix - 0x00800000u >= 0x7f800000u - 0x00800000u) {
		if (ix * 2 == 0)
			return -INFINITY;
		if (ix == 0x7f800000u)
			return x;
		if ((ix & 0x80000000u) || ix * 2 >= 0xff000000u)
			return NAN;
		ix = f2u(x * 0x1p23f);
		ix -= 23u << 23;
	}
	uint32_t tmp = ix - LOG_OFF;
	int i = (int)((tmp >> 19) % 16);
	int k = (int32_t)tmp >> 23;
	uint32_t iz = ix - (tmp & (0x1ffu << 23));
	double z = (double)u2f(iz);
	double r = z * POWF_INVC[i] - 1.0;
	double y0 = LOGF_LOGC[i] + (double)k * LOGF_LN2;
	double r2 = r * r;
	double y = LOGF_POLY[1] * r + LOGF_POLY[2];
	y = LOGF_POLY[0] * r2 + y;
	y = y * r2 + (y0 + r);
	return (float)y;
}

/* log10f: the fdlibm single-precision method (Sun Microsystems, as vendored
 * in test/extra/musl-libm/log10f.c): reduce to [sqrt(2)/2, sqrt(2)], evaluate
 * log(1+f) with the Lg1..Lg4 polynomial in float, recombine with split
 * log10(2) and 1/ln(10) constants. Float arithmetic throughout. */
float orc_log10f(float x)
{
	static const float ivln10hi = 4.3432617188e-01f, ivln10lo = -3.1689971365e-05f;
	static const float log10_2hi = 3.0102920532e-01f, log10_2lo = 7.9034151668e-07f;
	static const float Lg1 = 0xaaaaaa.0p-24f, Lg2 = 0xccce13.0p-25f, Lg3 = 0x91e9ee.0p-25f, Lg4 = 0xf89e26.0p-26f;
	uint32_t ix = f2u(x);
	int k = 0;
	if (ix < 0x00800000u || ix >> 31) {
		if (ix << 1 == 0)
			return -1 / (x * x);
		if (ix >> 31)
			return (x - x) / 0.0f;
		k -= 25;
		x *= 0x1p25f;
		ix = f2u(x);
	} else if (ix >= 0x7f800000u) {
		return x;
	} else if (ix == 0x3f800000u) {
		return 0;
	}
	ix += 0x3f800000u - 0x3f3504f3u;
	k += (int)(ix >> 23) - 0x7f;
	ix = (ix & 0x007fffffu) + 0x3f3504f3u;
	x = u2f(ix);
	float f = x - 1.0f;
	float s = f / (2.0f + f);
	float z = s * s;
	float w = z * z;
	float t1 = w * (Lg2 + w * Lg4);
	float t2 = z * (Lg1 + w * Lg3);
	float R = t2 + t1;
	float hfsq = 0.5f * f * f;
	float hi = f - hfsq;
	hi = u2f(f2u(hi) & 0xfffff000u);
	float lo = f - hi - hfsq + s * (hfsq + R);
	float dk = (float)k;
	return dk * log10_2lo + (lo + hi) * ivln10lo + lo * ivln10hi + hi * ivln10hi + dk * log10_2hi;
}

/* ======================================================================
 * Resampling filters — resize/filter.cpp:19-246.
 * ====================================================================== */

static const double ORC_PI = 3.14159265358979323846;

static double orc_sinc(double x)
{
	/* filter.cpp:21-25 (zimg_x_sin defaults to libm sin) */
	return x == 0.0 ? 1.0 : sin(x * ORC_PI) / (x * ORC_PI);
}

static double orc_poly3(double x, double c0, double c1, double c2, double c3)
{
	return c0 + x * (c1 + x * (c2 + x * c3)); /* filter.cpp:27-30 */
}

static double orc_round_halfup(double x)
{
	/* filter.cpp:32-40 */
	return x < 0 ? floor(x + 0.5) : floor(x + 0.49999999999999994);
}

unsigned orc_filter_support(int type, double param_a)
{
	switch (type) {
	case ORC_FILTER_POINT: return 0;
	case ORC_FILTER_BILINEAR: return 1;
	case ORC_FILTER_BICUBIC: return 2;
	case ORC_FILTER_SPLINE16: return 2;
	case ORC_FILTER_SPLINE36: return 3;
	case ORC_FILTER_SPLINE64: return 4;
	case ORC_FILTER_LANCZOS: return (unsigned)param_a;
	default: return 0;
	}
}

double orc_filter_eval(int type, double param_a, double param_b, double x)
{
	switch (type) {
	case ORC_FILTER_POINT:
		return 1.0; /* filter.cpp:136 */
	case ORC_FILTER_BILINEAR:
		return fmax(1.0 - fabs(x), 0.0); /* :141-144 */
	case ORC_FILTER_BICUBIC: {
		/* :147-169; coefficients computed in the constructor order. */
		double b = param_a, c = param_b;
		double p0 = (6.0 - 2.0 * b) / 6.0;
		double p2 = (-18.0 + 12.0 * b + 6.0 * c) / 6.0;
		double p3 = (12.0 - 9.0 * b - 6.0 * c) / 6.0;
		double q0 = (8.0 * b + 24.0 * c) / 6.0;
		double q1 = (-12.0 * b - 48.0 * c) / 6.0;
		double q2 = (6.0 * b + 30.0 * c) / 6.0;
		double q3 = (-b - 6.0 * c) / 6.0;
		x = fabs(x);
		if (x < 1.0)
			return orc_poly3(x, p0, 0.0, p2, p3);
		else if (x < 2.0)
			return orc_poly3(x, q0, q1, q2, q3);
		return 0.0;
	}
	case ORC_FILTER_SPLINE16:
		/* :174-186 */
		x = fabs(x);
		if (x < 1.0)
			return orc_poly3(x, 1.0, -1.0 / 5.0, -9.0 / 5.0, 1.0);
		if (x < 2.0) {
			x -= 1.0;
			return orc_poly3(x, 0.0, -7.0 / 15.0, 4.0 / 5.0, -1.0 / 3.0);
		}
		return 0.0;
	case ORC_FILTER_SPLINE36:
		/* :191-206 */
		x = fabs(x);
		if (x < 1.0)
			return orc_poly3(x, 1.0, -3.0 / 209.0, -453.0 / 209.0, 13.0 / 11.0);
		if (x < 2.0) {
			x -= 1.0;
			return orc_poly3(x, 0.0, -156.0 / 209.0, 270.0 / 209.0, -6.0 / 11.0);
		}
		if (x < 3.0) {
			x -= 2.0;
			return orc_poly3(x, 0.0, 26.0 / 209.0, -45.0 / 209.0, 1.0 / 11.0);
		}
		return 0.0;
	case ORC_FILTER_SPLINE64:
		/* :211-229 */
		x = fabs(x);
		if (x < 1.0)
			return orc_poly3(x, 1.0, -3.0 / 2911.0, -6387.0 / 2911.0, 49.0 / 41.0);
		if (x < 2.0) {
			x -= 1.0;
			return orc_poly3(x, 0.0, -2328.0 / 2911.0, 4032.0 / 2911.0, -24.0 / 41.0);
		}
		if (x < 3.0) {
			x -= 2.0;
			return orc_poly3(x, 0.0, 582.0 / 2911.0, -1008.0 / 2911.0, 6.0 / 41.0);
		}
		if (x < 4.0) {
			x -= 3.0;
			return orc_poly3(x, 0.0, -97.0 / 2911.0, 168.0 / 2911.0, -1.0 / 41.0);
		}
		return 0.0;
	case ORC_FILTER_LANCZOS: {
		/* :242-246 */
		unsigned taps = (unsigned)param_a;
		x = fabs(x);
		return x < taps ? orc_sinc(x) * orc_sinc(x / taps) : 0.0;
	}
	default:
		return 0.0;
	}
}

/* One row of the sparse resampling matrix. Restates the allocation rule of
 * RowMatrix (common/matrix.cpp:14-21,104-129): an element becomes part of the
 * stored range only when an assignment changes its value. */
typedef struct {
	size_t lo, hi; /* stored range [lo, hi); lo == hi means empty */
	double *v;     /* dense scratch of src_dim entries, zero outside range */
} orc_row;

static void orc_row_assign(orc_row *r, size_t j, double val)
{
	double cur = (j >= r->lo && j < r->hi) ? r->v[j] : 0.0;
	if (cur != val) {
		if (r->lo == r->hi) {
			r->lo = j;
			r->hi = j + 1;
		} else {
			if (j < r->lo) r->lo = j;
			if (j >= r->hi) r->hi = j + 1;
		}
		r->v[j] = val;
	}
}

static double orc_row_get(const orc_row *r, size_t j)
{
	return (j >= r->lo && j < r->hi) ? r->v[j] : 0.0;
}

int orc_compute_filter64(int type, double param_a, double param_b,
                         unsigned src_dim, unsigned dst_dim, double shift, double width,
                         unsigned *filter_width_out,
                         float *data, int16_t *data_i16, double *data64, unsigned *left_out);

int orc_compute_filter(int type, double param_a, double param_b,
                       unsigned src_dim, unsigned dst_dim, double shift, double width,
                       unsigned *filter_width_out,
                       float *data, int16_t *data_i16, unsigned *left_out)
{
	return orc_compute_filter64(type, param_a, param_b, src_dim, dst_dim, shift, width, filter_width_out, data, data_i16, 0, left_out);
}

int orc_compute_filter64(int type, double param_a, double param_b,
                         unsigned src_dim, unsigned dst_dim, double shift, double width,
                         unsigned *filter_width_out,
                         float *data, int16_t *data_i16, double *data64, unsigned *left_out)
{
	/* filter.cpp:249-306 */
	double scale = (double)dst_dim / width;
	double step = scale < 1.0 ? scale : 1.0;
	double support = (double)orc_filter_support(type, param_a) / step;
	unsigned filter_size = (unsigned)ceil(support) * 2u;
	if (filter_size < 1)
		filter_size = 1;
	if (support > (double)(UINT_MAX / 2))
		return 1;

	size_t *row_lo = (size_t *)malloc(sizeof(size_t) * dst_dim);
	size_t *row_hi = (size_t *)malloc(sizeof(size_t) * dst_dim);
	double **row_vals = (double **)calloc(dst_dim, sizeof(double *));
	double *dense = (double *)calloc(src_dim, sizeof(double));
	if (!row_lo || !row_hi || !row_vals || !dense)
		return 2;

	for (unsigned i = 0; i < dst_dim; ++i) {
		orc_row r = { 0, 0, dense };
		double pos = (i + 0.5) / scale + shift;
		double begin_pos = orc_round_halfup(pos - filter_size / 2.0) + 0.5;

		double total = 0.0;
		for (unsigned j = 0; j < filter_size; ++j) {
			double xpos = begin_pos + j;
			total += orc_filter_eval(type, param_a, param_b, (xpos - pos) * step);
		}

		size_t left = SIZE_MAX;
		for (unsigned j = 0; j < filter_size; ++j) {
			double xpos = begin_pos + j;
			double real_pos;

			if (xpos < 0.0)
				real_pos = -xpos;
			else if (xpos >= src_dim)
				real_pos = 2.0 * src_dim - xpos;
			else
				real_pos = xpos;

			double hi = nextafter((double)src_dim, -INFINITY);
			if (real_pos < 0.0) real_pos = 0.0;
			if (real_pos > hi) real_pos = hi;

			size_t idx = (size_t)floor(real_pos);
			double contrib = orc_filter_eval(type, param_a, param_b, (xpos - pos) * step) / total;
			orc_row_assign(&r, idx, orc_row_get(&r, idx) + contrib);
			if (idx < left)
				left = idx;
		}

		if (orc_row_get(&r, left) == 0.0) {
			orc_row_assign(&r, left, DBL_EPSILON);
			orc_row_assign(&r, left, 0.0);
		}

		row_lo[i] = r.lo;
		row_hi[i] = r.hi;
		row_vals[i] = (double *)malloc(sizeof(double) * (r.hi - r.lo));
		memcpy(row_vals[i], dense + r.lo, sizeof(double) * (r.hi - r.lo));
		memset(dense + r.lo, 0, sizeof(double) * (r.hi - r.lo));
	}

	/* matrix_to_filter, filter.cpp:43-127 */
	size_t width_fw = 0;
	for (unsigned i = 0; i < dst_dim; ++i) {
		if (row_hi[i] - row_lo[i] > width_fw)
			width_fw = row_hi[i] - row_lo[i];
	}
	*filter_width_out = (unsigned)width_fw;

	if (data) {
		for (unsigned i = 0; i < dst_dim; ++i) {
			size_t left = row_lo[i] < src_dim - width_fw ? row_lo[i] : src_dim - width_fw;
			double f32_err = 0.0;
			double i16_err = 0.0;
			int16_t i16_sum = 0;
			int16_t i16_greatest = 0;
			size_t i16_greatest_idx = 0;

			for (size_t j = 0; j < width_fw; ++j) {
				size_t col = left + j;
				double coeff = (col >= row_lo[i] && col < row_hi[i]) ? row_vals[i][col - row_lo[i]] : 0.0;
				double expected_f32 = coeff - f32_err;
				double expected_i16 = coeff * (1 << 14) - i16_err;
				float c_f32 = (float)expected_f32;
				int16_t c_i16 = (int16_t)lrint(expected_i16);

				f32_err = (double)c_f32 - expected_f32;
				i16_err = (double)c_i16 - expected_i16;

				if (abs(c_i16) > abs(i16_greatest)) {
					i16_greatest = c_i16;
					i16_greatest_idx = j;
				}
				i16_sum = (int16_t)(i16_sum + c_i16);

				data[i * width_fw + j] = c_f32;
				data_i16[i * width_fw + j] = c_i16;
				if (data64)
					data64[i * width_fw + j] = coeff;
			}
			data_i16[i * width_fw + i16_greatest_idx] = (int16_t)(data_i16[i * width_fw + i16_greatest_idx] + ((1 << 14) - i16_sum));
			left_out[i] = (unsigned)left;
		}
	}

	for (unsigned i = 0; i < dst_dim; ++i)
		free(row_vals[i]);
	free(row_vals);
	free(row_lo);
	free(row_hi);
	free(dense);
	return 0;
}

/* ======================================================================
 * Resize kernels — resize/resize_impl.cpp:31-114.
 * ====================================================================== */

static uint16_t orc_pack_u16(int32_t x, int32_t pixel_max)
{
	x = ((x + (1 << 13)) >> 14) + 32768; /* :36-42 (−INT16_MIN) */
	if (x > pixel_max) x = pixel_max;
	if (x < 0) x = 0;
	return (uint16_t)x;
}

#define ROWP(T, base, stride, i) ((T *)((char *)(base) + (ptrdiff_t)(i) * (stride)))
#define CROWP(T, base, stride, i) ((const T *)((const char *)(base) + (ptrdiff_t)(i) * (stride)))

void orc_resize_h_u16(const int16_t *coeffs, const unsigned *left, unsigned fw, unsigned dst_w,
                      const uint16_t *src, ptrdiff_t src_stride, uint16_t *dst, ptrdiff_t dst_stride,
                      unsigned height, unsigned depth)
{
	int32_t pixel_max = (int32_t)((1UL << depth) - 1);
	for (unsigned i = 0; i < height; ++i) {
		const uint16_t *s = CROWP(uint16_t, src, src_stride, i);
		uint16_t *d = ROWP(uint16_t, dst, dst_stride, i);
		for (unsigned j = 0; j < dst_w; ++j) {
			int32_t accum = 0;
			for (unsigned k = 0; k < fw; ++k)
				accum += (int32_t)coeffs[(size_t)j * fw + k] * ((int32_t)s[left[j] + k] - 32768);
			d[j] = orc_pack_u16(accum, pixel_max);
		}
	}
}

void orc_resize_v_u16(const int16_t *coeffs, const unsigned *left, unsigned fw, unsigned dst_h,
                      const uint16_t *src, ptrdiff_t src_stride, uint16_t *dst, ptrdiff_t dst_stride,
                      unsigned width, unsigned depth)
{
	int32_t pixel_max = (int32_t)((1UL << depth) - 1);
	for (unsigned i = 0; i < dst_h; ++i) {
		uint16_t *d = ROWP(uint16_t, dst, dst_stride, i);
		for (unsigned j = 0; j < width; ++j) {
			int32_t accum = 0;
			for (unsigned k = 0; k < fw; ++k)
				accum += (int32_t)coeffs[(size_t)i * fw + k] * ((int32_t)CROWP(uint16_t, src, src_stride, left[i] + k)[j] - 32768);
			d[j] = orc_pack_u16(accum, pixel_max);
		}
	}
}

void orc_resize_h_f32(const float *coeffs, const unsigned *left, unsigned fw, unsigned dst_w,
                      const float *src, ptrdiff_t src_stride, float *dst, ptrdiff_t dst_stride,
                      unsigned height)
{
	for (unsigned i = 0; i < height; ++i) {
		const float *s = CROWP(float, src, src_stride, i);
		float *d = ROWP(float, dst, dst_stride, i);
		for (unsigned j = 0; j < dst_w; ++j) {
			float accum = 0;
			for (unsigned k = 0; k < fw; ++k)
				accum += coeffs[(size_t)j * fw + k] * s[left[j] + k];
			d[j] = accum;
		}
	}
}

void orc_resize_v_f32(const float *coeffs, const unsigned *left, unsigned fw, unsigned dst_h,
                      const float *src, ptrdiff_t src_stride, float *dst, ptrdiff_t dst_stride,
                      unsigned width)
{
	for (unsigned i = 0; i < dst_h; ++i) {
		float *d = ROWP(float, dst, dst_stride, i);
		for (unsigned j = 0; j < width; ++j) {
			float accum = 0;
			for (unsigned k = 0; k < fw; ++k)
				accum += coeffs[(size_t)i * fw + k] * CROWP(float, src, src_stride, left[i] + k)[j];
			d[j] = accum;
		}
	}
}

/* ======================================================================
 * Depth — depth/quantize.h:12-44, quantize.cpp:25-144, depth_convert.cpp,
 * dither.cpp.
 * ====================================================================== */

static int32_t orc_integer_offset(int type, unsigned depth, int full, int chroma)
{
	if (type >= 2) return 0;
	if (chroma) return 1L << (depth - 1);
	if (!full) return 16L << (depth - 8);
	return 0;
}

static int32_t orc_integer_range(int type, unsigned depth, int full, int chroma, int ycgco)
{
	if (type >= 2) return 1;
	if (full) return (1L << depth) - 1;
	if (chroma && !ycgco) return 224L << (depth - 8);
	return 219L << (depth - 8);
}

void orc_scale_offset(int type_in, unsigned depth_in, int full_in, int chroma_in, int ycgco_in,
                      int type_out, unsigned depth_out, int full_out, int chroma_out, int ycgco_out,
                      float *scale, float *offset)
{
	double range_in = orc_integer_range(type_in, depth_in, full_in, chroma_in, ycgco_in);
	double offset_in = orc_integer_offset(type_in, depth_in, full_in, chroma_in);
	double range_out = orc_integer_range(type_out, depth_out, full_out, chroma_out, ycgco_out);
	double offset_out = orc_integer_offset(type_out, depth_out, full_out, chroma_out);
	*scale = (float)(range_out / range_in);
	*offset = (float)(-offset_in * range_out / range_in + offset_out);
}

float orc_half_to_float(uint16_t h)
{
	/* quantize.cpp:25-80, restated as: scale the 11-bit significand by the exponent. */
	uint32_t sign = (uint32_t)(h >> 15) << 31;
	uint32_t e = (h >> 10) & 0x1f;
	uint32_t m = h & 0x3ff;
	uint32_t bits;
	if (e == 0x1f) {
		bits = sign | 0x7f800000u | (m ? ((m << 13) | 0x00400000u) : 0);
	} else if (e == 0) {
		if (m == 0) {
			bits = sign;
		} else {
			int sh = 0;
			while (!(m & 0x400)) { m <<= 1; ++sh; }
			m &= 0x3ff;
			bits = sign | ((uint32_t)(127 - 15 - sh + 1) << 23) | (m << 13);
		}
	} else {
		bits = sign | ((e + 112) << 23) | (m << 13);
	}
	return u2f(bits);
}

uint16_t orc_float_to_half(float f)
{
	/* quantize.cpp:84-144 */
	uint32_t w = f2u(f);
	uint32_t sign = w >> 31;
	uint32_t e = (w >> 23) & 0xff;
	uint32_t m = w & 0x7fffff;
	uint32_t eh, mh;
	if (e == 0xff) {
		eh = 0x1f;
		mh = m ? ((m >> 13) | 0x200) : 0;
	} else {
		uint32_t sh, ma;
		if (e <= 112) {
			sh = 13 + 112 - e + 1;
			if (sh > 31) sh = 31;
			ma = m | (1u << 23);
			eh = 0;
		} else {
			sh = 13;
			ma = m;
			eh = e - 112;
		}
		uint32_t half = 1u << (sh - 1);
		mh = (ma + half - 1 + ((ma >> sh) & 1)) >> sh;
		if (mh > 0x3ff) {
			mh &= 0x3ff;
			eh += 1;
		}
		if (eh >= 0x1f) {
			eh = 0x1f;
			mh = 0;
		}
	}
	return (uint16_t)((sign << 15) | (eh << 10) | mh);
}

static float orc_load(int type, const void *row, unsigned j)
{
	switch (type) {
	case 0: return (float)((const uint8_t *)row)[j];
	case 1: return (float)((const uint16_t *)row)[j];
	case 2: return orc_half_to_float(((const uint16_t *)row)[j]);
	default: return ((const float *)row)[j];
	}
}

void orc_convert_to_float(int type_in, const void *src, ptrdiff_t src_stride,
                          float *dst, ptrdiff_t dst_stride, unsigned width, unsigned height,
                          float scale, float offset)
{
	/* depth_convert.cpp:41-64: integers get x*scale+offset (two roundings);
	 * half input is widened exactly (half_to_float_n ignores scale/offset). */
	for (unsigned i = 0; i < height; ++i) {
		const void *s = CROWP(char, src, src_stride, i);
		float *d = ROWP(float, dst, dst_stride, i);
		for (unsigned j = 0; j < width; ++j) {
			if (type_in == 2)
				d[j] = orc_load(2, s, j);
			else if (type_in == 3)
				d[j] = orc_load(3, s, j);
			else
				d[j] = orc_load(type_in, s, j) * scale + offset;
		}
	}
}

void orc_convert_to_half(int type_in, const void *src, ptrdiff_t src_stride,
                         uint16_t *dst, ptrdiff_t dst_stride, unsigned width, unsigned height,
                         float scale, float offset)
{
	/* depth_convert.cpp:32-38 (integer_to_half) and 58-63 (float_to_half_n) */
	for (unsigned i = 0; i < height; ++i) {
		const void *s = CROWP(char, src, src_stride, i);
		uint16_t *d = ROWP(uint16_t, dst, dst_stride, i);
		for (unsigned j = 0; j < width; ++j) {
			if (type_in == 3)
				d[j] = orc_float_to_half(orc_load(3, s, j));
			else
				d[j] = orc_float_to_half(orc_load(type_in, s, j) * scale + offset);
		}
	}
}

void orc_left_shift(int type_in, const void *src, ptrdiff_t src_stride,
                    int type_out, void *dst, ptrdiff_t dst_stride,
                    unsigned width, unsigned height, unsigned shift)
{
	/* depth_convert.cpp:23-30 */
	for (unsigned i = 0; i < height; ++i) {
		const void *s = CROWP(char, src, src_stride, i);
		void *d = ROWP(char, dst, dst_stride, i);
		for (unsigned j = 0; j < width; ++j) {
			unsigned x = type_in == 0 ? ((const uint8_t *)s)[j] : ((const uint16_t *)s)[j];
			x <<= shift;
			if (type_out == 0)
				((uint8_t *)d)[j] = (uint8_t)x;
			else
				((uint16_t *)d)[j] = (uint16_t)x;
		}
	}
}

void orc_bayer_table(uint8_t out[256])
{
	/* The 16x16 Bayer index matrix of dither.cpp:140-157, regenerated:
	 * bit pairs of (row xor col, col) interleaved from the most significant end. */
	for (unsigned i = 0; i < 16; ++i) {
		for (unsigned j = 0; j < 16; ++j) {
			unsigned x = i ^ j, v = 0;
			for (unsigned b = 0; b < 4; ++b) {
				v |= ((x >> b) & 1u) << (7 - 2 * b);
				v |= ((j >> b) & 1u) << (6 - 2 * b);
			}
			out[i * 16 + j] = (uint8_t)v;
		}
	}
}

static void orc_store_int(int type_out, void *row, unsigned j, long q)
{
	if (type_out == 0)
		((uint8_t *)row)[j] = (uint8_t)q;
	else
		((uint16_t *)row)[j] = (uint16_t)q;
}

void orc_dither_ordered(int dither_type, const uint8_t *blue_noise, unsigned plane,
                        int type_in, const void *src, ptrdiff_t src_stride,
                        int type_out, void *dst, ptrdiff_t dst_stride,
                        unsigned width, unsigned height, float scale, float offset, unsigned bits)
{
	/* dither.cpp:37-55 with the tables of :159-234. */
	float bayer[4][256];
	float blue[64 * 64];
	if (dither_type == 1) {
		uint8_t raw[256];
		orc_bayer_table(raw);
		for (unsigned t = 0; t < 256; ++t)
			bayer[0][t] = (float)(raw[t] + 1) / (float)(255 + 2) - 0.5f;
		for (unsigned i = 0; i < 16; ++i) {
			for (unsigned j = 0; j < 16; ++j) {
				bayer[1][i * 16 + j] = bayer[0][i * 16 + (15 - j)];
				bayer[2][i * 16 + j] = bayer[0][(15 - i) * 16 + j];
				bayer[3][i * 16 + j] = bayer[0][j * 16 + i];
			}
		}
	} else if (dither_type == 2) {
		for (unsigned t = 0; t < 64 * 64; ++t)
			blue[t] = (float)(blue_noise[t] + 1) / (float)(255 + 2) - 0.5f;
	}

	float maxval = (float)(1UL << bits) - 1;
	for (unsigned i = 0; i < height; ++i) {
		const void *s = CROWP(char, src, src_stride, i);
		void *d = ROWP(char, dst, dst_stride, i);
		const float *dither;
		unsigned doff, dmask;
		static const float zeros[16] = { 0 };

		if (dither_type == 1) {
			dither = &bayer[plane % 4][(i % 16) * 16];
			doff = 0;
			dmask = 15;
		} else if (dither_type == 2) {
			static const unsigned offs[4] = { (0 << 8) | 0, (32 << 8) | 12, (16 << 8) | 55, (48 << 8) | 26 };
			unsigned hoff = offs[plane % 4] >> 8, voff = offs[plane % 4] & 0xff;
			dither = &blue[((i + voff) % 64) * 64];
			doff = hoff;
			dmask = 63;
		} else {
			dither = zeros;
			doff = 0;
			dmask = 15;
		}

		for (unsigned j = 0; j < width; ++j) {
			float x = orc_load(type_in, s, j) * scale + offset;
			float dv = dither[(doff + j) & dmask];
			x += dv;
			if (x < 0.0f) x = 0.0f;
			if (x > maxval) x = maxval;
			orc_store_int(type_out, d, j, lrintf(x));
		}
	}
}

void orc_dither_ed(int type_in, const void *src, ptrdiff_t src_stride,
                   int type_out, void *dst, ptrdiff_t dst_stride,
                   unsigned width, unsigned height, float scale, float offset, unsigned bits)
{
	/* dither.cpp:57-89 driven as in ErrorDiffusion::process (:330-348). */
	float *ctx = (float *)calloc((size_t)(width + 2) * 2, sizeof(float));
	float *err_a = ctx, *err_b = ctx + (width + 2);
	float maxval = (float)(1UL << bits) - 1;

	for (unsigned i = 0; i < height; ++i) {
		const void *s = CROWP(char, src, src_stride, i);
		void *d = ROWP(char, dst, dst_stride, i);
		const float *top = (i % 2) ? err_a : err_b;
		float *cur = (i % 2) ? err_b : err_a;

		for (unsigned j = 0; j < width; ++j) {
			unsigned je = j + 1;
			float x = orc_load(type_in, s, j) * scale + offset;
			float err = 0;
			err += cur[je - 1] * (7.0f / 16.0f);
			err += top[je + 1] * (3.0f / 16.0f);
			err += top[je + 0] * (5.0f / 16.0f);
			err += top[je - 1] * (1.0f / 16.0f);
			x += err;
			if (x < 0.0f) x = 0.0f;
			if (x > maxval) x = maxval;
			long q = lrintf(x);
			orc_store_int(type_out, d, j, q);
			cur[je] = x - (float)q;
		}
	}
	free(ctx);
}

/* ======================================================================
 * Colourspace — colorspace/gamma.cpp, colorspace_param.cpp, matrix3.cpp,
 * operation.cpp, operation_impl.cpp, colorspace_graph.cpp.
 * Internal enumerations (colorspace.h:18-63).
 * ====================================================================== */

enum { M_UNSPEC, M_RGB, M_601, M_709, M_FCC, M_240M, M_YCGCO, M_2020_NCL, M_2020_CL, M_CD_NCL, M_CD_CL, M_LMS, M_ICTCP, M_COUNT };
enum { T_UNSPEC, T_LINEAR, T_LOG100, T_LOG316, T_709, T_470M, T_470BG, T_240M, T_XVYCC, T_1361, T_SRGB, T_2084, T_428, T_B67, T_COUNT };
enum { P_UNSPEC, P_470M, P_470BG, P_SMPTE_C, P_709, P_FILM, P_2020, P_XYZ, P_DCI_P3, P_DCI_P3_D65, P_EBU3213, P_COUNT };

static const float REC709_ALPHA = 1.09929682680944f;
static const float REC709_BETA = 0.018053968510807f;
static const float SMPTE_240M_ALPHA = 1.111572195921731f;
static const float SMPTE_240M_BETA = 0.022821585529445f;
static const float SRGB_ALPHA = 1.055010718947587f;
static const float SRGB_BETA = 0.003041282560128f;
static const float ST2084_M1 = 0.1593017578125f;
static const float ST2084_M2 = 78.84375f;
static const float ST2084_C1 = 0.8359375f;
static const float ST2084_C2 = 18.8515625f;
static const float ST2084_C3 = 18.6875f;
static const float ARIB_B67_A = 0.17883277f;
static const float ARIB_B67_B = 0.28466892f;
static const float ARIB_B67_C = 0.55991073f;
static const float ST2084_OOTF_SCALE = 59.49080238715383f;

/* Transfer function ids (one per function of gamma.cpp:63-338). */
enum {
	F_709_OETF, F_709_INV_OETF, F_LOG100_OETF, F_LOG100_INV_OETF, F_LOG316_OETF, F_LOG316_INV_OETF,
	F_470M_OETF, F_470M_INV_OETF, F_470BG_OETF, F_470BG_INV_OETF, F_240M_OETF, F_240M_INV_OETF,
	F_XVYCC_OETF, F_XVYCC_INV_OETF, F_1361_OETF, F_1361_INV_OETF, F_B67_OETF, F_B67_INV_OETF,
	F_1886_EOTF, F_1886_INV_EOTF, F_1361_EOTF, F_1361_INV_EOTF, F_XVYCC_EOTF, F_XVYCC_INV_EOTF,
	F_SRGB_EOTF, F_SRGB_INV_EOTF, F_2084_EOTF, F_2084_INV_EOTF, F_428_EOTF, F_428_INV_EOTF,
	F_B67_EOTF, F_B67_INV_EOTF, F_2084_OETF, F_2084_INV_OETF,
};

static float fmaxf_(float a, float b) { return a < b ? b : a; } /* std::max(a, b) */

static float tf_709_oetf(float x)
{
	x = fmaxf_(x, 0.0f);
	if (x < REC709_BETA)
		x = x * 4.5f;
	else
		x = REC709_ALPHA * orc_powf(x, 0.45f) - (REC709_ALPHA - 1.0f);
	return x;
}

static float tf_709_inv_oetf(float x)
{
	x = fmaxf_(x, 0.0f);
	if (x < 4.5f * REC709_BETA)
		x = x / 4.5f;
	else
		x = orc_powf((x + (REC709_ALPHA - 1.0f)) / REC709_ALPHA, 1.0f / 0.45f);
	return x;
}

static float tf_1886_eotf(float x) { return x < 0.0f ? 0.0f : orc_powf(x, 2.4f); }
static float tf_1886_inv_eotf(float x) { return x < 0.0f ? 0.0f : orc_powf(x, 1.0f / 2.4f); }

static float tf_b67_oetf(float x)
{
	x = fmaxf_(x, 0.0f);
	if (x <= (1.0f / 12.0f))
		x = sqrtf(3.0f * x);
	else
		x = ARIB_B67_A * orc_logf(12.0f * x - ARIB_B67_B) + ARIB_B67_C;
	return x;
}

static float tf_b67_inv_oetf(float x)
{
	x = fmaxf_(x, 0.0f);
	if (x <= 0.5f)
		x = (x * x) * (1.0f / 3.0f);
	else
		x = (orc_expf((x - ARIB_B67_C) / ARIB_B67_A) + ARIB_B67_B) / 12.0f;
	return x;
}

static float tf_2084_eotf(float x)
{
	if (x > 0.0f) {
		float xpow = orc_powf(x, 1.0f / ST2084_M2);
		float num = fmaxf_(xpow - ST2084_C1, 0.0f);
		float den = fmaxf_(ST2084_C2 - ST2084_C3 * xpow, FLT_MIN);
		x = orc_powf(num / den, 1.0f / ST2084_M1);
	} else {
		x = 0.0f;
	}
	return x;
}

static float tf_2084_inv_eotf(float x)
{
	if (x > 0.0f) {
		float xpow = orc_powf(x, ST2084_M1);
		float num = (ST2084_C1 - 1.0f) + (ST2084_C2 - ST2084_C3) * xpow;
		float den = 1.0f + ST2084_C3 * xpow;
		x = orc_powf(1.0f + num / den, ST2084_M2);
	} else {
		x = 0.0f;
	}
	return x;
}

static float tf_eval(int f, float x);

static float tf_eval(int f, float x)
{
	switch (f) {
	case F_709_OETF: return tf_709_oetf(x);
	case F_709_INV_OETF: return tf_709_inv_oetf(x);
	case F_LOG100_OETF: return x <= 0.01f ? 0.0f : 1.0f + orc_log10f(x) / 2.0f;
	case F_LOG100_INV_OETF: return x <= 0.0f ? 0.01f : orc_powf(10.0f, 2 * (x - 1.0f));
	case F_LOG316_OETF: return x <= 0.00316227766f ? 0.0f : 1.0f + orc_log10f(x) / 2.5f;
	case F_LOG316_INV_OETF: return x <= 0.0f ? 0.00316227766f : orc_powf(10.0f, 2.5f * (x - 1.0f));
	case F_470M_OETF: return x < 0.0f ? 0.0f : orc_powf(x, 2.2f);
	case F_470M_INV_OETF: return x < 0.0f ? 0.0f : orc_powf(x, 1.0f / 2.2f);
	case F_470BG_OETF: return x < 0.0f ? 0.0f : orc_powf(x, 2.8f);
	case F_470BG_INV_OETF: return x < 0.0f ? 0.0f : orc_powf(x, 1.0f / 2.8f);
	case F_240M_OETF:
		x = fmaxf_(x, 0.0f);
		if (x < 4.0f * SMPTE_240M_BETA)
			x = x / 4.0f;
		else
			x = orc_powf((x + (SMPTE_240M_ALPHA - 1.0f)) / SMPTE_240M_ALPHA, 1.0f / 0.45f);
		return x;
	case F_240M_INV_OETF:
		x = fmaxf_(x, 0.0f);
		if (x < SMPTE_240M_BETA)
			x = x * 4.0f;
		else
			x = SMPTE_240M_ALPHA * orc_powf(x, 0.45f) - (SMPTE_240M_ALPHA - 1.0f);
		return x;
	case F_XVYCC_OETF: return copysignf(tf_709_oetf(fabsf(x)), x);
	case F_XVYCC_INV_OETF: return copysignf(tf_709_inv_oetf(fabsf(x)), x);
	case F_1361_OETF: return x >= 0.0f ? tf_709_oetf(x) : -tf_709_oetf(-x * 4.0f) / 4.0f;
	case F_1361_INV_OETF: return x >= 0.0f ? tf_709_inv_oetf(x) : -tf_709_inv_oetf(-x * 4.0f) / 4.0f;
	case F_B67_OETF: return tf_b67_oetf(x);
	case F_B67_INV_OETF: return tf_b67_inv_oetf(x);
	case F_1886_EOTF: return tf_1886_eotf(x);
	case F_1886_INV_EOTF: return tf_1886_inv_eotf(x);
	case F_1361_EOTF: return (x < 0.0f || x > 1.0f) ? tf_eval(F_1361_INV_OETF, x) : tf_1886_eotf(x);
	case F_1361_INV_EOTF: return (x < 0.0f || x > 1.0f) ? tf_eval(F_1361_OETF, x) : tf_1886_inv_eotf(x);
	case F_XVYCC_EOTF: return (x < 0.0f || x > 1.0f) ? tf_eval(F_XVYCC_INV_OETF, x) : tf_1886_eotf(x);
	case F_XVYCC_INV_EOTF: return (x < 0.0f || x > 1.0f) ? tf_eval(F_XVYCC_OETF, x) : tf_1886_inv_eotf(x);
	case F_SRGB_EOTF:
		x = fmaxf_(x, 0.0f);
		if (x < 12.92f * SRGB_BETA)
			x = x / 12.92f;
		else
			x = orc_powf((x + (SRGB_ALPHA - 1.0f)) / SRGB_ALPHA, 2.4f);
		return x;
	case F_SRGB_INV_EOTF:
		x = fmaxf_(x, 0.0f);
		if (x < SRGB_BETA)
			x = x * 12.92f;
		else
			x = SRGB_ALPHA * orc_powf(x, 1.0f / 2.4f) - (SRGB_ALPHA - 1.0f);
		return x;
	case F_2084_EOTF: return tf_2084_eotf(x);
	case F_2084_INV_EOTF: return tf_2084_inv_eotf(x);
	case F_428_EOTF: return x < 0.0f ? 0.0f : 52.37f / 48.0f * orc_powf(x, 2.6f);
	case F_428_INV_EOTF: return x < 0.0f ? 0.0f : orc_powf(x * (48.0f / 52.37f), 1.0f / 2.6f);
	case F_B67_EOTF: {
		float y = tf_b67_inv_oetf(x);
		return y < 0.0f ? y : orc_powf(y, 1.2f);
	}
	case F_B67_INV_EOTF: {
		float y = x < 0.0f ? x : orc_powf(x, 1.0f / 1.2f);
		return tf_b67_oetf(y);
	}
	case F_2084_OETF: {
		/* ootf_st2084 (gamma.cpp:50-53) then inverse EOTF */
		float y = tf_1886_eotf(tf_709_oetf(x * ST2084_OOTF_SCALE)) / 100.0f;
		return tf_2084_inv_eotf(y);
	}
	case F_2084_INV_OETF: {
		float y = tf_2084_eotf(x);
		return tf_709_inv_oetf(tf_1886_inv_eotf(y * 100.0f)) / ST2084_OOTF_SCALE;
	}
	default:
		return NAN;
	}
}

/* select_transfer_function (gamma.cpp:341-409): function ids. Returns 0 on success. */
static int select_tf(int transfer, int scene, int *to_linear, int *to_gamma)
{
	switch (transfer) {
	case T_LOG100: *to_linear = F_LOG100_INV_OETF; *to_gamma = F_LOG100_OETF; return 0;
	case T_LOG316: *to_linear = F_LOG316_INV_OETF; *to_gamma = F_LOG316_OETF; return 0;
	case T_709:
		*to_linear = scene ? F_709_INV_OETF : F_1886_EOTF;
		*to_gamma = scene ? F_709_OETF : F_1886_INV_EOTF;
		return 0;
	case T_470M: *to_linear = F_470M_OETF; *to_gamma = F_470M_INV_OETF; return 0;
	case T_470BG: *to_linear = F_470BG_OETF; *to_gamma = F_470BG_INV_OETF; return 0;
	case T_240M:
		*to_linear = scene ? F_240M_INV_OETF : F_1886_EOTF;
		*to_gamma = scene ? F_240M_OETF : F_1886_INV_EOTF;
		return 0;
	case T_1361:
		*to_linear = scene ? F_1361_INV_OETF : F_1361_EOTF;
		*to_gamma = scene ? F_1361_OETF : F_1361_INV_EOTF;
		return 0;
	case T_XVYCC:
		*to_linear = scene ? F_XVYCC_INV_OETF : F_XVYCC_EOTF;
		*to_gamma = scene ? F_XVYCC_OETF : F_XVYCC_INV_EOTF;
		return 0;
	case T_SRGB: *to_linear = F_SRGB_EOTF; *to_gamma = F_SRGB_INV_EOTF; return 0;
	case T_2084:
		*to_linear = scene ? F_2084_INV_OETF : F_2084_EOTF;
		*to_gamma = scene ? F_2084_OETF : F_2084_INV_EOTF;
		return 0;
	case T_428: *to_linear = F_428_EOTF; *to_gamma = F_428_INV_EOTF; return 0;
	case T_B67:
		*to_linear = scene ? F_B67_INV_OETF : F_B67_EOTF;
		*to_gamma = scene ? F_B67_OETF : F_B67_INV_EOTF;
		return 0;
	default:
		return 1;
	}
}

void orc_transfer_scales(int transfer, double peak_luminance, int scene_referred,
                         float *to_linear_scale, float *to_gamma_scale)
{
	*to_linear_scale = 1.0f;
	*to_gamma_scale = 1.0f;
	if (transfer == T_2084) {
		*to_linear_scale = (float)(10000.0f / peak_luminance);
		*to_gamma_scale = (float)(peak_luminance / 10000.0f);
	} else if (transfer == T_B67) {
		*to_linear_scale = scene_referred ? 12.0f : (float)(1000.0 / peak_luminance);
		*to_gamma_scale = scene_referred ? 1.0f / 12.0f : (float)(peak_luminance / 1000.0);
	}
}

int orc_transfer_eval(int transfer, int to_linear, int scene_referred, float x, float *y)
{
	int fl, fg;
	if (select_tf(transfer, scene_referred, &fl, &fg))
		return 1;
	*y = tf_eval(to_linear ? fl : fg, x);
	return 0;
}

/* ---- 3x3 matrices (matrix3.cpp, colorspace_param.cpp) ---- */

typedef struct { double m[3][3]; } mat3;
typedef struct { double v[3]; } vec3;

static double det2(double a00, double a01, double a10, double a11) { return a00 * a11 - a01 * a10; }

static mat3 mat_mul(mat3 a, mat3 b)
{
	mat3 r;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j) {
			double acc = 0;
			for (int k = 0; k < 3; ++k)
				acc += a.m[i][k] * b.m[k][j];
			r.m[i][j] = acc;
		}
	return r;
}

static vec3 mat_vec(mat3 a, vec3 v)
{
	vec3 r;
	for (int i = 0; i < 3; ++i) {
		double acc = 0;
		for (int k = 0; k < 3; ++k)
			acc += a.m[i][k] * v.v[k];
		r.v[i] = acc;
	}
	return r;
}

static mat3 mat_inv(mat3 a)
{
	double (*m)[3] = a.m;
	double det = 0;
	det += m[0][0] * det2(m[1][1], m[1][2], m[2][1], m[2][2]);
	det -= m[0][1] * det2(m[1][0], m[1][2], m[2][0], m[2][2]);
	det += m[0][2] * det2(m[1][0], m[1][1], m[2][0], m[2][1]);
	mat3 r;
	r.m[0][0] = det2(m[1][1], m[1][2], m[2][1], m[2][2]) / det;
	r.m[0][1] = det2(m[0][2], m[0][1], m[2][2], m[2][1]) / det;
	r.m[0][2] = det2(m[0][1], m[0][2], m[1][1], m[1][2]) / det;
	r.m[1][0] = det2(m[1][2], m[1][0], m[2][2], m[2][0]) / det;
	r.m[1][1] = det2(m[0][0], m[0][2], m[2][0], m[2][2]) / det;
	r.m[1][2] = det2(m[0][2], m[0][0], m[1][2], m[1][0]) / det;
	r.m[2][0] = det2(m[1][0], m[1][1], m[2][0], m[2][1]) / det;
	r.m[2][1] = det2(m[0][1], m[0][0], m[2][1], m[2][0]) / det;
	r.m[2][2] = det2(m[0][0], m[0][1], m[1][0], m[1][1]) / det;
	return r;
}

static mat3 mat_transpose(mat3 a)
{
	mat3 r;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			r.m[i][j] = a.m[j][i];
	return r;
}

static mat3 mat_identity(void)
{
	mat3 r = { { { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } } };
	return r;
}

static vec3 xy_to_xyz(double x, double y)
{
	vec3 r = { { x / y, 1.0, (1.0 - x - y) / y } };
	return r;
}

static vec3 vcross(vec3 a, vec3 b)
{
	vec3 r = { { a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2], a.v[0] * b.v[1] - a.v[1] * b.v[0] } };
	return r;
}

static double vdot(vec3 a, vec3 b) { return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2]; }

/* Set when a derivation hits a case the reference rejects with InternalError
 * (unknown Kr/Kb matrix or primaries without xy coordinates). */
static int g_csp_err;

static int get_yuv_constants(double *kr, double *kb, int matrix)
{
	switch (matrix) {
	case M_RGB: *kr = 0; *kb = 0; return 0;
	case M_FCC: *kr = 0.3; *kb = 0.11; return 0;
	case M_240M: *kr = 0.212; *kb = 0.087; return 0;
	case M_601: *kr = 0.299; *kb = 0.114; return 0;
	case M_709: *kr = 0.2126; *kb = 0.0722; return 0;
	case M_2020_NCL:
	case M_2020_CL: *kr = 0.2627; *kb = 0.0593; return 0;
	default: g_csp_err = 1; *kr = 0; *kb = 0; return 1;
	}
}

static vec3 white_point(int p)
{
	switch (p) {
	case P_470M:
	case P_FILM: return xy_to_xyz(0.31, 0.316);
	case P_XYZ: return xy_to_xyz(1.0 / 3.0, 1.0 / 3.0);
	case P_DCI_P3: return xy_to_xyz(0.314, 0.351);
	default: return xy_to_xyz(0.3127, 0.3290);
	}
}

static void primaries_xy(double out[3][2], int p)
{
	static const double T470M[3][2] = { { 0.670, 0.330 }, { 0.210, 0.710 }, { 0.140, 0.080 } };
	static const double T470BG[3][2] = { { 0.640, 0.330 }, { 0.290, 0.600 }, { 0.150, 0.060 } };
	static const double TSMPTEC[3][2] = { { 0.630, 0.340 }, { 0.310, 0.595 }, { 0.155, 0.070 } };
	static const double T709[3][2] = { { 0.640, 0.330 }, { 0.300, 0.600 }, { 0.150, 0.060 } };
	static const double TFILM[3][2] = { { 0.681, 0.319 }, { 0.243, 0.692 }, { 0.145, 0.049 } };
	static const double T2020[3][2] = { { 0.708, 0.292 }, { 0.170, 0.797 }, { 0.131, 0.046 } };
	static const double TP3[3][2] = { { 0.680, 0.320 }, { 0.265, 0.690 }, { 0.150, 0.060 } };
	static const double TEBU[3][2] = { { 0.630, 0.340 }, { 0.295, 0.605 }, { 0.155, 0.077 } };
	const double (*src)[2] = T709;
	switch (p) {
	case P_470M: src = T470M; break;
	case P_470BG: src = T470BG; break;
	case P_SMPTE_C: src = TSMPTEC; break;
	case P_709: src = T709; break;
	case P_FILM: src = TFILM; break;
	case P_2020: src = T2020; break;
	case P_DCI_P3:
	case P_DCI_P3_D65: src = TP3; break;
	case P_EBU3213: src = TEBU; break;
	default: g_csp_err = 1; break;
	}
	memcpy(out, src, sizeof(double) * 6);
}

static mat3 rgb_to_yuv_from_kr_kb(double kr, double kb)
{
	mat3 r;
	double kg = 1.0 - kr - kb;
	double uscale = 1.0 / (2.0 - 2.0 * kb);
	double vscale = 1.0 / (2.0 - 2.0 * kr);
	r.m[0][0] = kr; r.m[0][1] = kg; r.m[0][2] = kb;
	r.m[1][0] = -kr * uscale; r.m[1][1] = -kg * uscale; r.m[1][2] = (1.0 - kb) * uscale;
	r.m[2][0] = (1.0 - kr) * vscale; r.m[2][1] = -kg * vscale; r.m[2][2] = -kb * vscale;
	return r;
}

static mat3 ncl_rgb_to_yuv(int matrix)
{
	double kr = 0, kb = 0;
	if (matrix == M_YCGCO) {
		mat3 r = { { { 0.25, 0.5, 0.25 }, { -0.25, 0.5, -0.25 }, { 0.5, 0, -0.5 } } };
		return r;
	}
	if (matrix == M_LMS) {
		mat3 r = { { { 1688.0 / 4096.0, 2146.0 / 4096.0, 262.0 / 4096.0 },
		             { 683.0 / 4096.0, 2951.0 / 4096.0, 462.0 / 4096.0 },
		             { 99.0 / 4096.0, 309.0 / 4096.0, 3688.0 / 4096.0 } } };
		return r;
	}
	get_yuv_constants(&kr, &kb, matrix);
	return rgb_to_yuv_from_kr_kb(kr, kb);
}

static void yuv_constants_from_primaries(double *kr, double *kb, int p)
{
	double xy[3][2];
	primaries_xy(xy, p);
	vec3 r = xy_to_xyz(xy[0][0], xy[0][1]);
	vec3 g = xy_to_xyz(xy[1][0], xy[1][1]);
	vec3 b = xy_to_xyz(xy[2][0], xy[2][1]);
	vec3 w = white_point(p);
	vec3 xr = { { r.v[0], g.v[0], b.v[0] } };
	vec3 yr = { { r.v[1], g.v[1], b.v[1] } };
	vec3 zr = { { r.v[2], g.v[2], b.v[2] } };
	*kr = vdot(w, vcross(g, b)) / vdot(xr, vcross(yr, zr));
	*kb = vdot(w, vcross(r, g)) / vdot(xr, vcross(yr, zr));
}

static mat3 ncl_rgb_to_yuv_from_primaries(int p)
{
	double kr, kb;
	if (p == P_709) return ncl_rgb_to_yuv(M_709);
	if (p == P_2020) return ncl_rgb_to_yuv(M_2020_NCL);
	yuv_constants_from_primaries(&kr, &kb, p);
	return rgb_to_yuv_from_kr_kb(kr, kb);
}

static mat3 ncl_yuv_to_rgb_from_primaries(int p)
{
	double kr, kb;
	if (p == P_709) return mat_inv(ncl_rgb_to_yuv(M_709));
	if (p == P_2020) return mat_inv(ncl_rgb_to_yuv(M_2020_NCL));
	yuv_constants_from_primaries(&kr, &kb, p);
	return mat_inv(rgb_to_yuv_from_kr_kb(kr, kb));
}

static mat3 lms_to_ictcp(int transfer)
{
	if (transfer == T_B67) {
		mat3 r = { { { 0.5, 0.5, 0.0 }, { 3625.0 / 4096.0, -7465.0 / 4096.0, 3840.0 / 4096.0 },
		             { 9500.0 / 4096.0, -9212.0 / 4096.0, -288.0 / 4096.0 } } };
		return r;
	} else {
		mat3 r = { { { 0.5, 0.5, 0.0 }, { 6610.0 / 4096.0, -13613.0 / 4096.0, 7003.0 / 4096.0 },
		             { 17933.0 / 4096.0, -17390.0 / 4096.0, -543.0 / 4096.0 } } };
		return r;
	}
}

static mat3 gamut_rgb_to_xyz(int p)
{
	if (p == P_XYZ) return mat_identity();
	double xy[3][2];
	primaries_xy(xy, p);
	mat3 x;
	for (int i = 0; i < 3; ++i) {
		vec3 c = xy_to_xyz(xy[i][0], xy[i][1]);
		for (int k = 0; k < 3; ++k) x.m[i][k] = c.v[k];
	}
	x = mat_transpose(x);
	vec3 w = white_point(p);
	vec3 s = mat_vec(mat_inv(x), w);
	mat3 m;
	for (int i = 0; i < 3; ++i)
		for (int k = 0; k < 3; ++k)
			m.m[i][k] = x.m[i][k] * s.v[k];
	return m;
}

static mat3 gamut_xyz_to_rgb(int p)
{
	if (p == P_XYZ) return mat_identity();
	return mat_inv(gamut_rgb_to_xyz(p));
}

static mat3 white_adapt(int pin, int pout)
{
	mat3 bradford = { { { 0.8951, 0.2664, -0.1614 }, { -0.7502, 1.7135, 0.0367 }, { 0.0389, -0.0685, 1.0296 } } };
	vec3 wi = white_point(pin), wo = white_point(pout);
	if (wi.v[0] == wo.v[0] && wi.v[1] == wo.v[1] && wi.v[2] == wo.v[2])
		return mat_identity();
	vec3 ri = mat_vec(bradford, wi), ro = mat_vec(bradford, wo);
	mat3 m = { { { 0 } } };
	m.m[0][0] = ro.v[0] / ri.v[0];
	m.m[1][1] = ro.v[1] / ri.v[1];
	m.m[2][2] = ro.v[2] / ri.v[2];
	return mat_mul(mat_mul(mat_inv(bradford), m), bradford);
}

void orc_ncl_rgb_to_yuv(int matrix, double out[9])
{
	mat3 m = ncl_rgb_to_yuv(matrix);
	memcpy(out, m.m, sizeof(double) * 9);
}

void orc_gamut_matrix(int p_in, int p_out, int adapt, double out[9])
{
	mat3 m = adapt ? mat_mul(mat_mul(gamut_xyz_to_rgb(p_out), white_adapt(p_in, p_out)), gamut_rgb_to_xyz(p_in))
	               : mat_mul(gamut_xyz_to_rgb(p_out), gamut_rgb_to_xyz(p_in));
	memcpy(out, m.m, sizeof(double) * 9);
}

/* ---- operations (operation_impl.cpp) and path search (colorspace_graph.cpp) ---- */

enum { OP_MATRIX, OP_GAMMA, OP_B67, OP_B67_INV, OP_CL_TO_RGB, OP_CL_TO_YUV, OP_MATRIX_FMA, OP_LUT_LINEAR, OP_LUT_GAMMA };

/* Nonzero while building the chain the reference runs with
 * allow_approximate_gamma on an x86 CPU with AVX2, FMA and F16C (cpu AUTO,
 * operation_impl_x86.cpp:12-77): matrices with FMA
 * (operation_impl_avx2.cpp:43-68), gamma and inverse gamma through lookup
 * tables (:258-382), and HLG taken scene-referred (operation.cpp:14-17). */
static int g_csp_approx;
enum { E_NCL_RGB_TO_YUV, E_NCL_YUV_TO_RGB, E_LIN_TO_GAMMA, E_GAMMA_TO_LIN, E_CL_RGB_TO_YUV, E_CL_YUV_TO_RGB,
       E_GAMUT, E_LMS_TO_ICTCP, E_ICTCP_TO_LMS };

typedef struct { int m, t, p; } csp_t;

static int csp_index(csp_t c) { return (c.m * T_COUNT + c.t) * P_COUNT + c.p; }
static csp_t csp_from_index(int i) { csp_t c = { i / (T_COUNT * P_COUNT), (i / P_COUNT) % T_COUNT, i % P_COUNT }; return c; }

static int is_valid_csp(csp_t c)
{
	int valid_2020cl = c.m == M_2020_CL && c.t == T_709;
	int valid_ictcp = c.m == M_ICTCP && (c.t == T_2084 || c.t == T_B67) && c.p == P_2020;
	int valid_lms = c.m == M_LMS && (c.t == T_LINEAR || c.t == T_2084 || c.t == T_B67) && c.p == P_2020;
	return !(c.m == M_2020_CL && !valid_2020cl) &&
	       !(c.m == M_CD_NCL && c.p == P_UNSPEC) &&
	       !(c.m == M_CD_CL && c.p == P_UNSPEC) &&
	       !(c.m == M_ICTCP && !valid_ictcp) &&
	       !(c.m == M_LMS && !valid_lms);
}

typedef struct { csp_t to; int kind; } edge_t;

static int neighbours(csp_t c, edge_t *e)
{
	int n = 0;
#define ADD(M, T, P, K) do { e[n].to.m = (M); e[n].to.t = (T); e[n].to.p = (P); e[n].kind = (K); ++n; } while (0)
	if (c.m == M_RGB) {
		for (int m = 0; m < M_COUNT; ++m) {
			if (m == M_UNSPEC || m == M_RGB || m == M_2020_CL || m == M_CD_NCL || m == M_CD_CL || m == M_LMS || m == M_ICTCP)
				continue;
			ADD(m, c.t, c.p, E_NCL_RGB_TO_YUV);
		}
		if (c.p != P_UNSPEC)
			ADD(M_CD_NCL, c.t, c.p, E_NCL_RGB_TO_YUV);
		if (c.t == T_LINEAR) {
			for (int t = 0; t < T_COUNT; ++t) {
				if (t != c.t && t != T_UNSPEC) {
					ADD(c.m, t, c.p, E_LIN_TO_GAMMA);
					if (c.p != P_UNSPEC)
						ADD(M_CD_CL, t, c.p, E_CL_RGB_TO_YUV);
				}
			}
			if (c.p != P_UNSPEC) {
				for (int p = 0; p < P_COUNT; ++p)
					if (p != c.p && p != P_UNSPEC)
						ADD(c.m, c.t, p, E_GAMUT);
			}
			ADD(M_2020_CL, T_709, c.p, E_CL_RGB_TO_YUV);
			if (c.p == P_2020)
				ADD(M_LMS, c.t, c.p, E_NCL_RGB_TO_YUV);
		} else if (c.t != T_UNSPEC) {
			ADD(c.m, T_LINEAR, c.p, E_GAMMA_TO_LIN);
		}
	} else if (c.m == M_2020_CL || c.m == M_CD_CL) {
		ADD(M_RGB, T_LINEAR, c.p, E_CL_YUV_TO_RGB);
	} else if (c.m == M_LMS) {
		if (c.t == T_2084 || c.t == T_B67) {
			ADD(M_ICTCP, c.t, c.p, E_LMS_TO_ICTCP);
			ADD(c.m, T_LINEAR, c.p, E_GAMMA_TO_LIN);
		}
		if (c.t == T_LINEAR) {
			ADD(M_RGB, c.t, c.p, E_NCL_YUV_TO_RGB);
			ADD(c.m, T_2084, c.p, E_LIN_TO_GAMMA);
			ADD(c.m, T_B67, c.p, E_LIN_TO_GAMMA);
		}
	} else if (c.m == M_ICTCP) {
		if (c.t == T_2084 || c.t == T_B67)
			ADD(M_LMS, c.t, c.p, E_ICTCP_TO_LMS);
	} else if (c.m != M_UNSPEC) {
		ADD(M_RGB, c.t, c.p, E_NCL_YUV_TO_RGB);
	}
#undef ADD
	return n;
}

static void set_matrix_op(orc_csp_ops *ops, int k, mat3 m)
{
	ops->kind[k] = g_csp_approx ? OP_MATRIX_FMA : OP_MATRIX;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			ops->m[k][i][j] = (float)m.m[i][j];
}

static int make_op_inner(orc_csp_ops *ops, int k, int kind, csp_t in, csp_t out, double peak, int scene, int adapt);

static int make_op(orc_csp_ops *ops, int k, int kind, csp_t in, csp_t out, double peak, int scene, int adapt)
{
	g_csp_err = 0;
	int rc = make_op_inner(ops, k, kind, in, out, peak, scene, adapt);
	return rc || g_csp_err;
}

static int make_op_inner(orc_csp_ops *ops, int k, int kind, csp_t in, csp_t out, double peak, int scene, int adapt)
{
	int fl, fg;
	float sl, sg;
	int display_b67;
	mat3 m;

	switch (kind) {
	case E_NCL_YUV_TO_RGB:
		m = in.m == M_CD_NCL ? ncl_yuv_to_rgb_from_primaries(in.p) : mat_inv(ncl_rgb_to_yuv(in.m));
		set_matrix_op(ops, k, m);
		return 0;
	case E_NCL_RGB_TO_YUV:
		m = out.m == M_CD_NCL ? ncl_rgb_to_yuv_from_primaries(out.p) : ncl_rgb_to_yuv(out.m);
		set_matrix_op(ops, k, m);
		return 0;
	case E_ICTCP_TO_LMS:
		set_matrix_op(ops, k, mat_inv(lms_to_ictcp(in.t)));
		return 0;
	case E_LMS_TO_ICTCP:
		set_matrix_op(ops, k, lms_to_ictcp(in.t));
		return 0;
	case E_GAMMA_TO_LIN:
		display_b67 = in.p != P_UNSPEC && !scene && !g_csp_approx;
		if (in.t == T_B67 && display_b67) {
			m = ncl_rgb_to_yuv_from_primaries(in.p);
			orc_transfer_scales(T_B67, peak, 0, &sl, &sg);
			ops->kind[k] = OP_B67_INV;
			ops->kr[k] = (float)m.m[0][0]; ops->kg[k] = (float)m.m[0][1]; ops->kb[k] = (float)m.m[0][2];
			ops->post[k] = sl;
			return 0;
		}
		if (select_tf(in.t, scene, &fl, &fg))
			return 1;
		orc_transfer_scales(in.t, peak, scene, &sl, &sg);
		ops->kind[k] = g_csp_approx ? OP_LUT_LINEAR : OP_GAMMA;
		ops->func[k] = fl;
		ops->pre[k] = 1.0f;
		ops->post[k] = sl;
		return 0;
	case E_LIN_TO_GAMMA:
		display_b67 = out.p != P_UNSPEC && !scene && !g_csp_approx;
		if (out.t == T_B67 && display_b67) {
			m = ncl_rgb_to_yuv_from_primaries(out.p);
			orc_transfer_scales(T_B67, peak, 0, &sl, &sg);
			ops->kind[k] = OP_B67;
			ops->kr[k] = (float)m.m[0][0]; ops->kg[k] = (float)m.m[0][1]; ops->kb[k] = (float)m.m[0][2];
			ops->pre[k] = sg;
			return 0;
		}
		if (select_tf(out.t, scene, &fl, &fg))
			return 1;
		orc_transfer_scales(out.t, peak, scene, &sl, &sg);
		ops->kind[k] = g_csp_approx ? OP_LUT_GAMMA : OP_GAMMA;
		ops->func[k] = fg;
		ops->pre[k] = sg;
		ops->post[k] = 1.0f;
		return 0;
	case E_GAMUT: {
		double mm[9];
		orc_gamut_matrix(in.p, out.p, adapt, mm);
		memcpy(m.m, mm, sizeof(mm));
		set_matrix_op(ops, k, m);
		return 0;
	}
	case E_CL_YUV_TO_RGB:
	case E_CL_RGB_TO_YUV: {
		/* The factories assert transfer == REC_709 (operation_impl.cpp:364,377);
		 * a release build of the reference assumes it (zassert.h ASSUME_CONDITION),
		 * so the CL operations always use the REC_709 scene-referred pair. */
		csp_t c = kind == E_CL_YUV_TO_RGB ? in : out;
		c.t = T_709;
		if (select_tf(c.t, 1, &fl, &fg))
			return 1;
		orc_transfer_scales(c.t, peak, 1, &sl, &sg);
		m = c.m == M_CD_CL ? ncl_rgb_to_yuv_from_primaries(c.p) : ncl_rgb_to_yuv(c.m);
		float kr = (float)m.m[0][0], kg = (float)m.m[0][1], kb = (float)m.m[0][2];
		ops->kind[k] = kind == E_CL_YUV_TO_RGB ? OP_CL_TO_RGB : OP_CL_TO_YUV;
		ops->func[k] = kind == E_CL_YUV_TO_RGB ? fl : fg;
		ops->aux_func[k] = fg;
		ops->kr[k] = kr; ops->kg[k] = kg; ops->kb[k] = kb;
		/* constructor constants (operation_impl.cpp:184-187, 249-252) from to_gamma */
		ops->nb[k] = tf_eval(fg, 1.0f - (float)m.m[0][2]);
		ops->pb[k] = 1.0f - tf_eval(fg, (float)m.m[0][2]);
		ops->nr[k] = tf_eval(fg, 1.0f - (float)m.m[0][0]);
		ops->pr[k] = 1.0f - tf_eval(fg, (float)m.m[0][0]);
		ops->post[k] = sl;
		ops->pre[k] = sg;
		return 0;
	}
	default:
		return 1;
	}
}

int orc_colorspace_build(int m_in, int t_in, int p_in, int m_out, int t_out, int p_out,
                         double peak_luminance, int scene_referred, int chromatic_adaptation,
                         orc_csp_ops *ops, char *trace, size_t trace_len)
{
	/* colorspace.cpp:95-122: SMPTE 240M transfer is remapped unless scene-referred. */
	csp_t in = { m_in, t_in, p_in }, out = { m_out, t_out, p_out };
	if (!scene_referred) {
		if (in.t == T_240M) in.t = T_709;
		if (out.t == T_240M) out.t = T_709;
	}
	memset(ops, 0, sizeof(*ops));
	if (trace && trace_len)
		trace[0] = '\0';
	if (in.m == out.m && in.t == out.t && in.p == out.p)
		return 0;
	if (!is_valid_csp(in) || !is_valid_csp(out))
		return -1;

	enum { NSTATE = M_COUNT * T_COUNT * P_COUNT };
	static int visited[NSTATE], parent[NSTATE], pkind[NSTATE], queue[NSTATE];
	memset(visited, 0, sizeof(visited));
	int qh = 0, qt = 0;
	int src = csp_index(in), dst = csp_index(out), v = -1;
	visited[src] = 1;
	queue[qt++] = src;
	while (qh < qt) {
		v = queue[qh++];
		if (v == dst)
			break;
		edge_t e[64];
		int n = neighbours(csp_from_index(v), e);
		for (int i = 0; i < n; ++i) {
			int w = csp_index(e[i].to);
			if (visited[w])
				continue;
			visited[w] = 1;
			queue[qt++] = w;
			parent[w] = v;
			pkind[w] = e[i].kind;
		}
	}
	if (v != dst)
		return -1;

	int path[16], nk = 0;
	for (int x = dst; x != src; x = parent[x])
		path[nk++] = x;
	if (nk > 6)
		return -1;

	for (int i = 0; i < nk; ++i) {
		int to = path[nk - 1 - i];
		int from = parent[to];
		if (make_op(ops, i, pkind[to], csp_from_index(from), csp_from_index(to), peak_luminance, scene_referred, chromatic_adaptation))
			return -1;
		if (trace) {
			size_t len = strlen(trace);
			snprintf(trace + len, trace_len - len, "%d:%d->%d;", pkind[to], from, to);
		}
	}
	ops->n = nk;
	return nk;
}

int orc_colorspace_build_approx(int m_in, int t_in, int p_in, int m_out, int t_out, int p_out,
                                double peak_luminance, int scene_referred, int chromatic_adaptation,
                                orc_csp_ops *ops, char *trace, size_t trace_len)
{
	g_csp_approx = 1;
	int rc = orc_colorspace_build(m_in, t_in, p_in, m_out, t_out, p_out, peak_luminance, scene_referred, chromatic_adaptation,
	                              ops, trace, trace_len);
	g_csp_approx = 0;
	return rc;
}

/* Lookup-table gamma of the AVX2 path. Entries are computed where they are
 * used, by the table's own formula. To linear (operation_impl_avx2.cpp:
 * 122-156, 258-274): 2^16 + 1 entries, entry i = f(i / 65536 * 2 - 0.5) * post,
 * index = cvtps2dq(fma(x, 32768, 16384)) (round to nearest even; NaN and
 * values outside int32 give INT32_MIN) clamped to [0, 65536]. To gamma
 * (:194-240, 304-320): 2^16 entries indexed by the binary16 bits of x
 * (round to nearest even), entry h = f(half(h) * pre). */
static float lut_to_linear(int func, float post, float x)
{
	float v = fmaf(x, 32768.0f, 16384.0f);
	int32_t idx = (v >= -2147483648.0f && v < 2147483648.0f) ? (int32_t)rintf(v) : INT32_MIN;
	idx = idx < 0 ? 0 : idx > 65536 ? 65536 : idx;
	float g = (float)idx / 65536.0f * 2.0f - 0.5f;
	return tf_eval(func, g) * post;
}

static float lut_to_gamma(int func, float pre, float x)
{
	return tf_eval(func, orc_half_to_float(orc_float_to_half(x)) * pre);
}

void orc_colorspace_apply(const orc_csp_ops *ops, float *p0, float *p1, float *p2, size_t n)
{
	for (int k = 0; k < ops->n; ++k) {
		switch (ops->kind[k]) {
		case OP_MATRIX: {
			/* operation_impl.cpp:27-45 */
			const float (*m)[3] = ops->m[k];
			for (size_t i = 0; i < n; ++i) {
				float a = p0[i], b = p1[i], c = p2[i];
				float x = m[0][0] * a + m[0][1] * b + m[0][2] * c;
				float y = m[1][0] * a + m[1][1] * b + m[1][2] * c;
				float z = m[2][0] * a + m[2][1] * b + m[2][2] * c;
				p0[i] = x; p1[i] = y; p2[i] = z;
			}
			break;
		}
		case OP_MATRIX_FMA: {
			/* operation_impl_avx2.cpp:43-68: a product, then two fused multiply-adds */
			const float (*m)[3] = ops->m[k];
			for (size_t i = 0; i < n; ++i) {
				float a = p0[i], b = p1[i], c = p2[i];
				float x = fmaf(m[0][2], c, fmaf(m[0][1], b, m[0][0] * a));
				float y = fmaf(m[1][2], c, fmaf(m[1][1], b, m[1][0] * a));
				float z = fmaf(m[2][2], c, fmaf(m[2][1], b, m[2][0] * a));
				p0[i] = x; p1[i] = y; p2[i] = z;
			}
			break;
		}
		case OP_LUT_LINEAR:
		case OP_LUT_GAMMA: {
			float *pl[3] = { p0, p1, p2 };
			for (int p = 0; p < 3; ++p)
				for (size_t i = 0; i < n; ++i)
					pl[p][i] = ops->kind[k] == OP_LUT_LINEAR ? lut_to_linear(ops->func[k], ops->post[k], pl[p][i])
					                                         : lut_to_gamma(ops->func[k], ops->pre[k], pl[p][i]);
			break;
		}
		case OP_GAMMA: {
			/* :64-76 */
			float *pl[3] = { p0, p1, p2 };
			for (int p = 0; p < 3; ++p)
				for (size_t i = 0; i < n; ++i)
					pl[p][i] = ops->post[k] * tf_eval(ops->func[k], pl[p][i] * ops->pre[k]);
			break;
		}
		case OP_B67: {
			/* :92-114 */
			const float gamma = 1.2f;
			for (size_t i = 0; i < n; ++i) {
				float r = p0[i] * ops->pre[k], g = p1[i] * ops->pre[k], b = p2[i] * ops->pre[k];
				float yd = fmaxf_(ops->kr[k] * r + ops->kg[k] * g + ops->kb[k] * b, FLT_MIN);
				float ys_inv = orc_powf(yd, (1.0f - gamma) / gamma);
				p0[i] = tf_b67_oetf(r * ys_inv);
				p1[i] = tf_b67_oetf(g * ys_inv);
				p2[i] = tf_b67_oetf(b * ys_inv);
			}
			break;
		}
		case OP_B67_INV: {
			/* :130-157 */
			const float gamma = 1.2f;
			for (size_t i = 0; i < n; ++i) {
				float r = tf_b67_inv_oetf(p0[i]), g = tf_b67_inv_oetf(p1[i]), b = tf_b67_inv_oetf(p2[i]);
				float ys = fmaxf_(ops->kr[k] * r + ops->kg[k] * g + ops->kb[k] * b, FLT_MIN);
				ys = orc_powf(ys, gamma - 1.0f);
				r = r * ys; g = g * ys; b = b * ys;
				p0[i] = r * ops->post[k]; p1[i] = g * ops->post[k]; p2[i] = b * ops->post[k];
			}
			break;
		}
		case OP_CL_TO_RGB: {
			/* :190-222 */
			for (size_t i = 0; i < n; ++i) {
				float y = p0[i], u = p1[i], v = p2[i];
				float bmy = u < 0 ? u * 2.0f * ops->nb[k] : u * 2.0f * ops->pb[k];
				float rmy = v < 0 ? v * 2.0f * ops->nr[k] : v * 2.0f * ops->pr[k];
				float b = tf_eval(ops->func[k], bmy + y);
				float r = tf_eval(ops->func[k], rmy + y);
				y = tf_eval(ops->func[k], y);
				float g = (y - ops->kr[k] * r - ops->kb[k] * b) / ops->kg[k];
				p0[i] = r * ops->post[k]; p1[i] = g * ops->post[k]; p2[i] = b * ops->post[k];
			}
			break;
		}
		case OP_CL_TO_YUV: {
			/* :255-284 */
			for (size_t i = 0; i < n; ++i) {
				float r = p0[i] * ops->pre[k], g = p1[i] * ops->pre[k], b = p2[i] * ops->pre[k];
				float y = tf_eval(ops->func[k], ops->kr[k] * r + ops->kg[k] * g + ops->kb[k] * b);
				float u, v;
				b = tf_eval(ops->func[k], b);
				r = tf_eval(ops->func[k], r);
				u = (b - y < 0.0f) ? (b - y) / (2.0f * ops->nb[k]) : (b - y) / (2.0f * ops->pb[k]);
				v = (r - y < 0.0f) ? (r - y) / (2.0f * ops->nr[k]) : (r - y) / (2.0f * ops->pr[k]);
				p0[i] = y; p1[i] = u; p2[i] = v;
			}
			break;
		}
		}
	}
}

/* ======================================================================
 * Unresize — unresize/bilinear.cpp:31-172 (context), unresize_impl.cpp:27-85
 * (line kernels). The least-squares inverse of a bilinear upscale.
 * ====================================================================== */

/* Entry j of row i of the bilinear matrix A (out x in): rows hold one or two
 * consecutive stored entries from *left (bilinear.cpp:64-108). */
typedef struct { unsigned left, n; double v[2]; } orc_brow;

static double brow_at(const orc_brow *r, size_t j)
{
	return (j < r->left || j >= r->left + r->n) ? 0.0 : r->v[j - r->left];
}

int orc_unresize_context(unsigned up_dim, unsigned orig_dim, double shift, unsigned *row_size,
                         float *coef, unsigned *offset, float *lu_c, float *lu_l, float *lu_u)
{
	const unsigned in = orig_dim, out = up_dim;
	if (in > out || !in)
		return -1;
	const double sh = -shift * in / out;
	orc_brow *a = calloc(out, sizeof(orc_brow));
	size_t *tl = calloc(in, sizeof(size_t)), *tr = calloc(in, sizeof(size_t));
	double *c = calloc(in, sizeof(double)), *l = calloc(in, sizeof(double)), *u = calloc(in, sizeof(double));
	double leftmost = 0.5 + sh, rightmost = in - 0.5 + sh;
	double lf = floor(leftmost), rf = floor(rightmost);
	unsigned li = (unsigned)(lf < 0 ? 0 : lf > in - 1.0 ? in - 1.0 : lf);
	unsigned ri = (unsigned)(rf < 0 ? 0 : rf > in - 1.0 ? in - 1.0 : rf);
	for (unsigned i = 0; i < out; ++i) {
		double position = (i + 0.5) * in / out;
		if (position <= leftmost) {
			a[i].left = li; a[i].n = 1; a[i].v[0] = 1.0;
		} else if (position >= rightmost) {
			a[i].left = ri; a[i].n = 1; a[i].v[0] = 1.0;
		} else {
			unsigned left_idx = (unsigned)floor(position - leftmost);
			double distance = position - left_idx - leftmost;
			a[i].left = left_idx; a[i].n = 2;
			a[i].v[0] = 1.0 - distance; a[i].v[1] = distance;
		}
	}
	/* rows of A' (transpose) trimmed to their nonzero ends (matrix.cpp:154-200) */
	for (unsigned col = 0; col < in; ++col) {
		size_t first = out, last = 0;
		for (unsigned i = 0; i < out; ++i)
			if (brow_at(&a[i], col) != 0) { if (first == out) first = i; last = i; }
		if (first < out) { tl[col] = first; tr[col] = last + 1; } else { tl[col] = tr[col] = 0; }
	}
	/* tridiagonal of A'A, each entry summed over the A' row's range in order (matrix.cpp:202-219) */
#define ORC_PROD(i, j) ({ double acc_ = 0; for (size_t k_ = tl[i]; k_ < tr[i]; ++k_) acc_ += brow_at(&a[k_], (i)) * brow_at(&a[k_], (j)); acc_; })
	const double eps = DBL_EPSILON;
	if (in == 1) {
		l[0] = ORC_PROD(0, 0);
	} else {
		c[0] = 0;
		l[0] = ORC_PROD(0, 0);
		u[0] = ORC_PROD(0, 1) / (l[0] + eps);
		for (size_t i = 1; i + 1 < in; ++i) {
			c[i] = ORC_PROD(i, i - 1);
			l[i] = ORC_PROD(i, i) - c[i] * u[i - 1];
			u[i] = ORC_PROD(i, i + 1) / (l[i] + eps);
		}
		c[in - 1] = ORC_PROD(in - 1, in - 2);
		l[in - 1] = ORC_PROD(in - 1, in - 1) - c[in - 1] * u[in - 2];
		u[in - 1] = 0;
	}
#undef ORC_PROD
	size_t rs = 0;
	for (unsigned i = 0; i < in; ++i)
		if (tr[i] - tl[i] > rs) rs = tr[i] - tl[i];
	*row_size = (unsigned)rs;
	if (coef) {
		for (unsigned i = 0; i < in; ++i) {
			size_t left = tl[i] < out - rs ? tl[i] : out - rs;
			for (size_t j = 0; j < rs; ++j)
				coef[i * rs + j] = j < tr[i] - left ? (float)brow_at(&a[left + j], i) : 0.0f;
			offset[i] = (unsigned)left;
			lu_c[i] = (float)c[i];
			lu_l[i] = (float)(1.0 / (l[i] + eps));
			lu_u[i] = (float)u[i];
		}
	}
	free(a); free(tl); free(tr); free(c); free(l); free(u);
	return 0;
}

void orc_unresize_h_f32(const float *coef, const unsigned *offset, unsigned row_size, const float *lu_c, const float *lu_l,
                        const float *lu_u, unsigned out_w, const float *src, ptrdiff_t src_stride, float *dst,
                        ptrdiff_t dst_stride, unsigned height)
{
	/* unresize_impl.cpp:27-56 */
	for (unsigned r = 0; r < height; ++r) {
		const float *s = CROWP(float, src, src_stride, r);
		float *d = ROWP(float, dst, dst_stride, r);
		float z = 0.0f, w = 0.0f;
		for (unsigned j = 0; j < out_w; ++j) {
			float accum = 0.0f;
			for (unsigned k = 0; k < row_size; ++k)
				accum += coef[(size_t)j * row_size + k] * s[offset[j] + k];
			z = (accum - lu_c[j] * z) * lu_l[j];
			d[j] = z;
		}
		for (unsigned j = out_w; j != 0; --j) {
			w = d[j - 1] - lu_u[j - 1] * w;
			d[j - 1] = w;
		}
	}
}

void orc_unresize_v_f32(const float *coef, const unsigned *offset, unsigned row_size, const float *lu_c, const float *lu_l,
                        const float *lu_u, unsigned out_h, const float *src, ptrdiff_t src_stride, float *dst,
                        ptrdiff_t dst_stride, unsigned width)
{
	/* unresize_impl.cpp:58-85, 150-156: forward over every row, then back */
	for (unsigned i = 0; i < out_h; ++i) {
		float *d = ROWP(float, dst, dst_stride, i);
		for (unsigned j = 0; j < width; ++j) {
			float z = i ? CROWP(float, dst, dst_stride, i - 1)[j] : 0.0f;
			float accum = 0.0f;
			for (unsigned k = 0; k < row_size; ++k)
				accum += coef[(size_t)i * row_size + k] * CROWP(float, src, src_stride, offset[i] + k)[j];
			z = (accum - lu_c[i] * z) * lu_l[i];
			d[j] = z;
		}
	}
	for (unsigned i = out_h; i != 0; --i) {
		for (unsigned j = 0; j < width; ++j) {
			float w = i < out_h ? CROWP(float, dst, dst_stride, i)[j] : 0.0f;
			w = CROWP(float, dst, dst_stride, i - 1)[j] - lu_u[i - 1] * w;
			ROWP(float, dst, dst_stride, i - 1)[j] = w;
		}
	}
}
