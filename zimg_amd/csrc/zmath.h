// Scalar math shared by the host planner and the gfx950 kernels:
//  * powf / expf / logf: the ARM optimized-routines single-precision method
//    (Szabolcs Nagy 2017-2018, MIT; the code the reference vendors for its unit
//    tests, test/extra/musl-libm/{powf,expf,logf}.c). Evaluated in FP64 with no
//    contraction, so results are bit-identical to the reference's libm on
//    x86-64 (glibc and musl both implement this method).
//  * log10f: the fdlibm float method (test/extra/musl-libm/log10f.c).
//  * every transfer function of src/zimg/colorspace/gamma.cpp:40-338.
// The tables below are the method's constants; bit-exactness requires them to
// be identical to the published ones.
#pragma once

#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
// Always inlined: the transfer dispatch is on a wave-uniform id, so inlining
// turns it into one scalar branch instead of a call with spilled registers.
#define ZHD __host__ __device__ __attribute__((always_inline)) inline
#define ZCONST __constant__
#else
#define ZHD inline
#define ZCONST
#endif

namespace zamd {
namespace zm {

// Transfer function identifiers, one per function of gamma.cpp.
enum TransferFunc : int {
	F_709_OETF, F_709_INV_OETF, F_LOG100_OETF, F_LOG100_INV_OETF, F_LOG316_OETF, F_LOG316_INV_OETF,
	F_470M_OETF, F_470M_INV_OETF, F_470BG_OETF, F_470BG_INV_OETF, F_240M_OETF, F_240M_INV_OETF,
	F_XVYCC_OETF, F_XVYCC_INV_OETF, F_1361_OETF, F_1361_INV_OETF, F_B67_OETF, F_B67_INV_OETF,
	F_1886_EOTF, F_1886_INV_EOTF, F_1361_EOTF, F_1361_INV_EOTF, F_XVYCC_EOTF, F_XVYCC_INV_EOTF,
	F_SRGB_EOTF, F_SRGB_INV_EOTF, F_2084_EOTF, F_2084_INV_EOTF, F_428_EOTF, F_428_INV_EOTF,
	F_B67_EOTF, F_B67_INV_EOTF, F_2084_OETF, F_2084_INV_OETF, F_COUNT
};

ZHD uint32_t fbits(float x) { uint32_t u; std::memcpy(&u, &x, 4); return u; }
ZHD float bitsf(uint32_t u) { float x; std::memcpy(&x, &u, 4); return x; }
ZHD uint64_t dbits(double x) { uint64_t u; std::memcpy(&u, &x, 8); return u; }
ZHD double bitsd(uint64_t u) { double x; std::memcpy(&x, &u, 8); return x; }

// IEEE binary16 bits to binary32, exactly (a NaN is made quiet and keeps its
// payload), as quantize.cpp:24-77 widens a half.
ZHD float half_to_float(uint16_t h)
{
	const uint32_t sign = static_cast<uint32_t>(h & 0x8000u) << 16;
	const uint32_t e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
	if (e == 0x1F)
		return bitsf(sign | 0x7F800000u | (m ? 0x00400000u | (m << 13) : 0u));
	if (e == 0) {
		const float f = static_cast<float>(m) * 0x1p-24f; // exact: m < 2^10
		return sign ? -f : f;
	}
	return bitsf(sign | ((e + 112u) << 23) | (m << 13));
}

// 2^(i/32) (double bits) minus i<<47.
ZCONST static const uint64_t kExp2Tab[32] = {
	0x3ff0000000000000ULL, 0x3fefd9b0d3158574ULL, 0x3fefb5586cf9890fULL, 0x3fef9301d0125b51ULL,
	0x3fef72b83c7d517bULL, 0x3fef54873168b9aaULL, 0x3fef387a6e756238ULL, 0x3fef1e9df51fdee1ULL,
	0x3fef06fe0a31b715ULL, 0x3feef1a7373aa9cbULL, 0x3feedea64c123422ULL, 0x3feece086061892dULL,
	0x3feebfdad5362a27ULL, 0x3feeb42b569d4f82ULL, 0x3feeab07dd485429ULL, 0x3feea47eb03a5585ULL,
	0x3feea09e667f3bcdULL, 0x3fee9f75e8ec5f74ULL, 0x3feea11473eb0187ULL, 0x3feea589994cce13ULL,
	0x3feeace5422aa0dbULL, 0x3feeb737b0cdc5e5ULL, 0x3feec49182a3f090ULL, 0x3feed503b23e255dULL,
	0x3feee89f995ad3adULL, 0x3feeff76f2fb5e47ULL, 0x3fef199bdd85529cULL, 0x3fef3720dcef9069ULL,
	0x3fef5818dcfba487ULL, 0x3fef7c97337b9b5fULL, 0x3fefa4afa2a490daULL, 0x3fefd0765b6e4540ULL,
};
// Reciprocal centres of 16 subintervals of [0x3f330000, 2*0x3f330000) and their log2 / ln.
ZCONST static const double kInvC[16] = {
	0x1.661ec79f8f3bep+0, 0x1.571ed4aaf883dp+0, 0x1.49539f0f010bp+0, 0x1.3c995b0b80385p+0,
	0x1.30d190c8864a5p+0, 0x1.25e227b0b8eap+0, 0x1.1bb4a4a1a343fp+0, 0x1.12358f08ae5bap+0,
	0x1.0953f419900a7p+0, 0x1p+0, 0x1.e608cfd9a47acp-1, 0x1.ca4b31f026aap-1,
	0x1.b2036576afce6p-1, 0x1.9c2d163a1aa2dp-1, 0x1.886e6037841edp-1, 0x1.767dcf5534862p-1,
};
ZCONST static const double kLog2C[16] = {
	-0x1.efec65b963019p-2, -0x1.b0b6832d4fca4p-2, -0x1.7418b0a1fb77bp-2, -0x1.39de91a6dcf7bp-2,
	-0x1.01d9bf3f2b631p-2, -0x1.97c1d1b3b7afp-3, -0x1.2f9e393af3c9fp-3, -0x1.960cbbf788d5cp-4,
	-0x1.a6f9db6475fcep-5, 0x0p+0, 0x1.338ca9f24f53dp-4, 0x1.476a9543891bap-3,
	0x1.e840b4ac4e4d2p-3, 0x1.40645f0c6651cp-2, 0x1.88e9c2c1b9ff8p-2, 0x1.ce0a44eb17bccp-2,
};
ZCONST static const double kLnC[16] = {
	-0x1.57bf7808caadep-2, -0x1.2bef0a7c06ddbp-2, -0x1.01eae7f513a67p-2, -0x1.b31d8a68224e9p-3,
	-0x1.6574f0ac07758p-3, -0x1.1aa2bc79c81p-3, -0x1.a4e76ce8c0e5ep-4, -0x1.1973c5a611cccp-4,
	-0x1.252f438e10c1ep-5, 0x0p+0, 0x1.aa5aa5df25984p-5, 0x1.c5e53aa362eb4p-4,
	0x1.526e57720db08p-3, 0x1.bc2860d22477p-3, 0x1.1058bc8a07ee1p-2, 0x1.4043057b6ee09p-2,
};

// Addends of the three multiply-adds of the powf polynomials whose other
// factor is also a constant (log2: the r and r^2 terms' pairs, exp2: the r^2
// pair). In the fused form a VALU instruction reads at most one scalar
// operand, so the device reads these from LDS into registers instead of
// rebuilding a 64-bit literal with two moves at every use.
constexpr double kPolyAdd[3] = { -0x1.71969a075c67ap-2, -0x1.7154748bef6c8p-1, 0x1.ebfce50fac4f3p-3 };

#if defined(__HIP_DEVICE_COMPILE__)
// Device code reads per-workgroup LDS copies of the tables (the lookups are
// per lane, so the copies turn vector memory loads into LDS reads). Every
// kernel that evaluates these functions calls lds_tables_init() first.
static __shared__ uint64_t s_exp2_tab[32];
static __shared__ double s_inv_c[16], s_log2_c[16], s_ln_c[16], s_poly_add[4];
#define ZM_EXP2TAB(i) s_exp2_tab[i]
#define ZM_INVC(i) s_inv_c[i]
#define ZM_LOG2C(i) s_log2_c[i]
#define ZM_LNC(i) s_ln_c[i]
#define ZM_POLYADD(i) s_poly_add[i]

__device__ __forceinline__ void lds_tables_init()
{
	for (unsigned k = threadIdx.x; k < 32; k += blockDim.x) {
		s_exp2_tab[k] = kExp2Tab[k];
		if (k < 16) {
			s_inv_c[k] = kInvC[k];
			s_log2_c[k] = kLog2C[k];
			s_ln_c[k] = kLnC[k];
		}
		if (k < 3)
			s_poly_add[k] = kPolyAdd[k];
	}
	__syncthreads();
}
#else
#define ZM_EXP2TAB(i) kExp2Tab[i]
#define ZM_INVC(i) kInvC[i]
#define ZM_LOG2C(i) kLog2C[i]
#define ZM_LNC(i) kLnC[i]
#define ZM_POLYADD(i) kPolyAdd[i]
#if defined(__HIPCC__)
// Host pass of a HIP translation unit: kernels are parsed, never run here.
__device__ inline void lds_tables_init() {}
#endif
#endif

// a * b + c: two roundings (the reference's libm, built without
// contraction) or, with FMA, one. The fused form is used only where an
// exhaustive comparison over every float input showed the same float results
// (see powf_fused below).
template <bool FMA>
ZHD double mul_add(double a, double b, double c)
{
	if constexpr (FMA)
		return __builtin_fma(a, b, c);
	else
		return a * b + c;
}

// bitsd(tab + (k << 47)), modulo 2^64: the shifted term's low word is zero,
// so only the high word changes (no carry). The device adds on the high word
// of a two-word vector (a plain 64-bit form is compiled into a 64-bit add).
ZHD double exp2_scale(uint64_t tab, uint32_t k)
{
#if defined(__HIP_DEVICE_COMPILE__)
	typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
	u32x2 w = __builtin_bit_cast(u32x2, tab);
	w.y += k << 15;
	return __builtin_bit_cast(double, w);
#else
	return bitsd(tab + (static_cast<uint64_t>(k) << 47));
#endif
}

// kPolyAdd[i]: from the device's LDS copy in the fused form, else the constant.
template <bool FMA>
ZHD double poly_add(int i)
{
	if constexpr (FMA)
		return ZM_POLYADD(i);
	else
		return kPolyAdd[i];
}

template <bool FMA = false>
ZHD float exp2_from_log(double xd, uint32_t sign_bias)
{
	// 2^xd = 2^(k/32) * 2^r, |r| <= 1/64; cubic for 2^r.
	const double shift = 0x1.8p+52 / 32;
	double kd = xd + shift;
	uint64_t ki = dbits(kd);
	kd -= shift;
	double r = xd - kd;
	double s = exp2_scale(ZM_EXP2TAB(ki % 32), static_cast<uint32_t>(ki + sign_bias));
	double z = mul_add<FMA>(0x1.c6af84b912394p-5, r, poly_add<FMA>(2));
	double r2 = r * r;
	double y = mul_add<FMA>(0x1.62e42ff0c52d6p-1, r, 1.0);
	y = mul_add<FMA>(z, r2, y);
	return static_cast<float>(y * s);
}

template <bool FMA = false>
ZHD double log2_reduced(uint32_t ix)
{
	// x = 2^k z, z in [0x3f330000, 2*0x3f330000); log2(x) = log1p(z/c - 1)/ln2 + log2(c) + k.
	uint32_t tmp = ix - 0x3f330000u;
	int i = static_cast<int>((tmp >> 19) % 16);
	uint32_t top = tmp & 0xff800000u;
	int k = static_cast<int32_t>(top) >> 23;
	double z = static_cast<double>(bitsf(ix - top));
	double r = mul_add<FMA>(z, ZM_INVC(i), -1.0);
	double y0 = ZM_LOG2C(i) + static_cast<double>(k);
	double r2 = r * r;
	double y = mul_add<FMA>(0x1.27616c9496e0bp-2, r, poly_add<FMA>(0));
	double p = mul_add<FMA>(0x1.ec70a6ca7baddp-2, r, poly_add<FMA>(1));
	double r4 = r2 * r2;
	double q = mul_add<FMA>(0x1.71547652ab82bp0, r, y0);
	q = mul_add<FMA>(p, r2, q);
	return mul_add<FMA>(y, r4, q);
}

ZHD int int_class(uint32_t iy)
{
	// 0: not an integer, 1: odd integer, 2: even integer.
	int e = iy >> 23 & 0xff;
	if (e < 0x7f) return 0;
	if (e > 0x7f + 23) return 2;
	if (iy & ((1u << (0x7f + 23 - e)) - 1)) return 0;
	if (iy & (1u << (0x7f + 23 - e))) return 1;
	return 2;
}

ZHD bool zero_inf_nan(uint32_t ix) { return 2 * ix - 1 >= 2u * 0x7f800000u - 1; }

// (w >> 15) & 0xffff; one bit-field extract on the device (the generic form
// is widened back into a 64-bit mask and compare).
ZHD uint32_t hi_bits_15_16(uint32_t w)
{
#if defined(__HIP_DEVICE_COMPILE__)
	return __builtin_amdgcn_ubfe(w, 15, 16);
#else
	return (w >> 15) & 0xffffu;
#endif
}

template <bool FMA = false>
ZHD float powf_t(float x, float y)
{
	uint32_t sign_bias = 0;
	uint32_t ix = fbits(x), iy = fbits(y);
	if (ix - 0x00800000u >= 0x7f800000u - 0x00800000u || zero_inf_nan(iy)) {
		if (zero_inf_nan(iy)) {
			if (2 * iy == 0) return 1.0f;
			if (ix == 0x3f800000u) return 1.0f;
			if (2 * ix > 2u * 0x7f800000u || 2 * iy > 2u * 0x7f800000u) return x + y;
			if (2 * ix == 2 * 0x3f800000u) return 1.0f;
			if ((2 * ix < 2 * 0x3f800000u) == !(iy & 0x80000000u)) return 0.0f;
			return y * y;
		}
		if (zero_inf_nan(ix)) {
			float x2 = x * x;
			if ((ix & 0x80000000u) && int_class(iy) == 1) x2 = -x2;
			return (iy & 0x80000000u) ? 1.0f / x2 : x2;
		}
		if (ix & 0x80000000u) {
			int yint = int_class(iy);
			if (yint == 0) return bitsf(0x7fc00000u);
			if (yint == 1) sign_bias = 1u << 16;
			ix &= 0x7fffffffu;
		}
		if (ix < 0x00800000u) {
			ix = fbits(x * 0x1p23f) & 0x7fffffffu;
			ix -= 23u << 23;
		}
	}
	double ylogx = static_cast<double>(y) * log2_reduced<FMA>(ix);
	// Exponent bits of |ylogx| >= 126 (the test on bits 47-62, on the high word).
	if (hi_bits_15_16(static_cast<uint32_t>(dbits(ylogx) >> 32)) >= static_cast<uint32_t>(dbits(126.0) >> 47)) {
		if (ylogx > 0x1.fffffffd1d571p+6) return sign_bias ? -bitsf(0x7f800000u) : bitsf(0x7f800000u);
		if (ylogx <= -150.0) return sign_bias ? -0.0f : 0.0f;
	}
	return exp2_from_log<FMA>(ylogx, sign_bias);
}

// The reference's powf (no contraction).
ZHD float powf_(float x, float y) { return powf_t<false>(x, y); }

ZHD float expf_(float x)
{
	uint32_t abstop = (fbits(x) >> 20) & 0x7ff;
	if (abstop >= (fbits(88.0f) >> 20)) {
		if (fbits(x) == 0xff800000u) return 0.0f;
		if (abstop >= 0x7f8) return x + x;
		if (x > 0x1.62e42ep6f) return bitsf(0x7f800000u);
		if (x < -0x1.9fe368p6f) return 0.0f;
	}
	const double inv_ln2_32 = 0x1.71547652b82fep+0 * 32;
	const double shift = 0x1.8p+52;
	double z = inv_ln2_32 * static_cast<double>(x);
	double kd = z + shift;
	uint64_t ki = dbits(kd);
	kd -= shift;
	double r = z - kd;
	uint64_t t = ZM_EXP2TAB(ki % 32) + (ki << 47);
	double s = bitsd(t);
	const double c0 = 0x1.c6af84b912394p-5 / 32 / 32 / 32, c1 = 0x1.ebfce50fac4f3p-3 / 32 / 32, c2 = 0x1.62e42ff0c52d6p-1 / 32;
	z = c0 * r + c1;
	double r2 = r * r;
	double y = c2 * r + 1.0;
	y = z * r2 + y;
	return static_cast<float>(y * s);
}

ZHD float logf_(float x)
{
	uint32_t ix = fbits(x);
	if (ix == 0x3f800000u) return 0.0f;
	if (ix - 0x00800000u >= 0x7f800000u - 0x00800000u) {
		if (ix * 2 == 0) return -bitsf(0x7f800000u);
		if (ix == 0x7f800000u) return x;
		if ((ix & 0x80000000u) || ix * 2 >= 0xff000000u) return bitsf(0x7fc00000u);
		ix = fbits(x * 0x1p23f);
		ix -= 23u << 23;
	}
	uint32_t tmp = ix - 0x3f330000u;
	int i = static_cast<int>((tmp >> 19) % 16);
	int k = static_cast<int32_t>(tmp) >> 23;
	double z = static_cast<double>(bitsf(ix - (tmp & (0x1ffu << 23))));
	double r = z * ZM_INVC(i) - 1.0;
	double y0 = ZM_LNC(i) + static_cast<double>(k) * 0x1.62e42fefa39efp-1;
	double r2 = r * r;
	double y = 0x1.5575b0be00b6ap-2 * r + -0x1.ffffef20a4123p-2;
	y = -0x1.00ea348b88334p-2 * r2 + y;
	y = y * r2 + (y0 + r);
	return static_cast<float>(y);
}

ZHD float log10f_(float x)
{
	const float ivln10hi = 4.3432617188e-01f, ivln10lo = -3.1689971365e-05f;
	const float log10_2hi = 3.0102920532e-01f, log10_2lo = 7.9034151668e-07f;
	const float Lg1 = 0xaaaaaa.0p-24f, Lg2 = 0xccce13.0p-25f, Lg3 = 0x91e9ee.0p-25f, Lg4 = 0xf89e26.0p-26f;
	uint32_t ix = fbits(x);
	int k = 0;
	if (ix < 0x00800000u || ix >> 31) {
		if (ix << 1 == 0) return -1 / (x * x);
		if (ix >> 31) return (x - x) / 0.0f;
		k -= 25;
		x *= 0x1p25f;
		ix = fbits(x);
	} else if (ix >= 0x7f800000u) {
		return x;
	} else if (ix == 0x3f800000u) {
		return 0;
	}
	ix += 0x3f800000u - 0x3f3504f3u;
	k += static_cast<int>(ix >> 23) - 0x7f;
	ix = (ix & 0x007fffffu) + 0x3f3504f3u;
	x = bitsf(ix);
	float f = x - 1.0f;
	float s = f / (2.0f + f);
	float z = s * s;
	float w = z * z;
	float t1 = w * (Lg2 + w * Lg4);
	float t2 = z * (Lg1 + w * Lg3);
	float R = t2 + t1;
	float hfsq = 0.5f * f * f;
	float hi = bitsf(fbits(f - hfsq) & 0xfffff000u);
	float lo = f - hi - hfsq + s * (hfsq + R);
	float dk = static_cast<float>(k);
	return dk * log10_2lo + (lo + hi) * ivln10lo + lo * ivln10hi + hi * ivln10hi + dk * log10_2hi;
}

ZHD float sqrtf_(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
	return __builtin_sqrtf(x);
#else
	return __builtin_sqrtf(x);
#endif
}

// ---- transfer functions (gamma.cpp:14-338) ----

constexpr float REC709_ALPHA = 1.09929682680944f;
constexpr float REC709_BETA = 0.018053968510807f;
constexpr float SMPTE_240M_ALPHA = 1.111572195921731f;
constexpr float SMPTE_240M_BETA = 0.022821585529445f;
constexpr float SRGB_ALPHA = 1.055010718947587f;
constexpr float SRGB_BETA = 0.003041282560128f;
constexpr float ST2084_M1 = 0.1593017578125f;
constexpr float ST2084_M2 = 78.84375f;
constexpr float ST2084_C1 = 0.8359375f;
constexpr float ST2084_C2 = 18.8515625f;
constexpr float ST2084_C3 = 18.6875f;
constexpr float ARIB_B67_A = 0.17883277f;
constexpr float ARIB_B67_B = 0.28466892f;
constexpr float ARIB_B67_C = 0.55991073f;
constexpr float ST2084_OOTF_SCALE = 59.49080238715383f;
constexpr float FLT_MIN_ = 1.17549435082228750797e-38f;

ZHD float maxf(float a, float b) { return a < b ? b : a; } // std::max order

// Float bits of a constant expression (template arguments).
#define ZM_FB(x) __builtin_bit_cast(uint32_t, (x))

// powf(x, y) for an exponent fixed at compile time (float bits YB), with
// fused multiply-adds in the FP64 steps. For each exponent accepted here the
// float result equals the reference's unfused powf for all 2^32 bit patterns
// of x, the inputs listed below included (they return the reference's value);
// tests/test_powf_fma.py repeats that comparison over every input.
template <uint32_t YB>
ZHD float powf_fused(float x)
{
	float r = powf_t<true>(x, bitsf(YB));
	const uint32_t ix = fbits(x);
	if constexpr (YB == ZM_FB(2.4f)) {
		r = ix == 0x3b9445eeu ? bitsf(0x361e949fu) : r;
		r = ix == 0x3d109d26u ? bitsf(0x39ab8eb9u) : r;
	} else if constexpr (YB == ZM_FB(ST2084_M1)) {
		r = ix == 0x1371da3eu ? bitsf(0x387a7341u) : r;
	} else if constexpr (YB == ZM_FB(ST2084_M2)) {
		r = ix == 0x3f727c4eu ? bitsf(0x3c63ae8cu) : r;
		r = ix == 0x3f5233b4u ? bitsf(0x343f4271u) : r;
	} else {
		static_assert(YB == ZM_FB(1.0f / 2.4f) || YB == ZM_FB(0.45f) || YB == ZM_FB(1.0f / 0.45f) || YB == ZM_FB(1.0f / ST2084_M2) ||
		                  YB == ZM_FB(1.0f / ST2084_M1) || YB == ZM_FB((1.0f - 1.2f) / 1.2f) || YB == ZM_FB(1.2f - 1.0f),
		              "powf_fused: exponent without an exhaustive comparison");
	}
	return r;
}

// powf(x, y) with y fixed: fused (FU) or the reference's.
template <bool FU, uint32_t YB>
ZHD float powf_k(float x)
{
	if constexpr (FU)
		return powf_fused<YB>(x);
	else
		return powf_(x, bitsf(YB));
}

template <bool FU = false>
ZHD float rec709_oetf(float x)
{
	x = maxf(x, 0.0f);
	return x < REC709_BETA ? x * 4.5f : REC709_ALPHA * powf_k<FU, ZM_FB(0.45f)>(x) - (REC709_ALPHA - 1.0f);
}

template <bool FU = false>
ZHD float rec709_inverse_oetf(float x)
{
	x = maxf(x, 0.0f);
	return x < 4.5f * REC709_BETA ? x / 4.5f : powf_k<FU, ZM_FB(1.0f / 0.45f)>((x + (REC709_ALPHA - 1.0f)) / REC709_ALPHA);
}

template <bool FU = false>
ZHD float rec1886_eotf(float x) { return x < 0.0f ? 0.0f : powf_k<FU, ZM_FB(2.4f)>(x); }
template <bool FU = false>
ZHD float rec1886_inverse_eotf(float x) { return x < 0.0f ? 0.0f : powf_k<FU, ZM_FB(1.0f / 2.4f)>(x); }

ZHD float arib_b67_oetf(float x)
{
	x = maxf(x, 0.0f);
	return x <= (1.0f / 12.0f) ? sqrtf_(3.0f * x) : ARIB_B67_A * logf_(12.0f * x - ARIB_B67_B) + ARIB_B67_C;
}

ZHD float arib_b67_inverse_oetf(float x)
{
	x = maxf(x, 0.0f);
	return x <= 0.5f ? (x * x) * (1.0f / 3.0f) : (expf_((x - ARIB_B67_C) / ARIB_B67_A) + ARIB_B67_B) / 12.0f;
}

template <bool FU = false>
ZHD float st2084_eotf(float x)
{
	if (!(x > 0.0f)) return 0.0f;
	float xpow = powf_k<FU, ZM_FB(1.0f / ST2084_M2)>(x);
	float num = maxf(xpow - ST2084_C1, 0.0f);
	float den = maxf(ST2084_C2 - ST2084_C3 * xpow, FLT_MIN_);
	return powf_k<FU, ZM_FB(1.0f / ST2084_M1)>(num / den);
}

template <bool FU = false>
ZHD float st2084_inverse_eotf(float x)
{
	if (!(x > 0.0f)) return 0.0f;
	float xpow = powf_k<FU, ZM_FB(ST2084_M1)>(x);
	float num = (ST2084_C1 - 1.0f) + (ST2084_C2 - ST2084_C3) * xpow;
	float den = 1.0f + ST2084_C3 * xpow;
	return powf_k<FU, ZM_FB(ST2084_M2)>(1.0f + num / den);
}

ZHD float copysign_(float mag, float sgn) { return bitsf((fbits(mag) & 0x7fffffffu) | (fbits(sgn) & 0x80000000u)); }
ZHD float fabs_(float x) { return bitsf(fbits(x) & 0x7fffffffu); }

// One function per TransferFunc id (gamma.cpp:63-338 via select_transfer_function).
ZHD float t_log100_oetf(float x) { return x <= 0.01f ? 0.0f : 1.0f + log10f_(x) / 2.0f; }
ZHD float t_log100_inv_oetf(float x) { return x <= 0.0f ? 0.01f : powf_(10.0f, 2 * (x - 1.0f)); }
ZHD float t_log316_oetf(float x) { return x <= 0.00316227766f ? 0.0f : 1.0f + log10f_(x) / 2.5f; }
ZHD float t_log316_inv_oetf(float x) { return x <= 0.0f ? 0.00316227766f : powf_(10.0f, 2.5f * (x - 1.0f)); }
ZHD float t_470m_oetf(float x) { return x < 0.0f ? 0.0f : powf_(x, 2.2f); }
ZHD float t_470m_inv_oetf(float x) { return x < 0.0f ? 0.0f : powf_(x, 1.0f / 2.2f); }
ZHD float t_470bg_oetf(float x) { return x < 0.0f ? 0.0f : powf_(x, 2.8f); }
ZHD float t_470bg_inv_oetf(float x) { return x < 0.0f ? 0.0f : powf_(x, 1.0f / 2.8f); }
ZHD float t_240m_oetf(float x)
{
	x = maxf(x, 0.0f);
	return x < 4.0f * SMPTE_240M_BETA ? x / 4.0f : powf_((x + (SMPTE_240M_ALPHA - 1.0f)) / SMPTE_240M_ALPHA, 1.0f / 0.45f);
}
ZHD float t_240m_inv_oetf(float x)
{
	x = maxf(x, 0.0f);
	return x < SMPTE_240M_BETA ? x * 4.0f : SMPTE_240M_ALPHA * powf_(x, 0.45f) - (SMPTE_240M_ALPHA - 1.0f);
}
ZHD float t_xvycc_oetf(float x) { return copysign_(rec709_oetf(fabs_(x)), x); }
ZHD float t_xvycc_inv_oetf(float x) { return copysign_(rec709_inverse_oetf(fabs_(x)), x); }
ZHD float t_1361_oetf(float x) { return x >= 0.0f ? rec709_oetf(x) : -rec709_oetf(-x * 4.0f) / 4.0f; }
ZHD float t_1361_inv_oetf(float x) { return x >= 0.0f ? rec709_inverse_oetf(x) : -rec709_inverse_oetf(-x * 4.0f) / 4.0f; }
ZHD float t_1361_eotf(float x)
{
	if (x < 0.0f || x > 1.0f) return x >= 0.0f ? rec709_inverse_oetf(x) : -rec709_inverse_oetf(-x * 4.0f) / 4.0f;
	return rec1886_eotf(x);
}
ZHD float t_1361_inv_eotf(float x)
{
	if (x < 0.0f || x > 1.0f) return x >= 0.0f ? rec709_oetf(x) : -rec709_oetf(-x * 4.0f) / 4.0f;
	return rec1886_inverse_eotf(x);
}
ZHD float t_xvycc_eotf(float x)
{
	if (x < 0.0f || x > 1.0f) return copysign_(rec709_inverse_oetf(fabs_(x)), x);
	return rec1886_eotf(x);
}
ZHD float t_xvycc_inv_eotf(float x)
{
	if (x < 0.0f || x > 1.0f) return copysign_(rec709_oetf(fabs_(x)), x);
	return rec1886_inverse_eotf(x);
}
template <bool FU = false>
ZHD float t_srgb_eotf(float x)
{
	x = maxf(x, 0.0f);
	return x < 12.92f * SRGB_BETA ? x / 12.92f : powf_k<FU, ZM_FB(2.4f)>((x + (SRGB_ALPHA - 1.0f)) / SRGB_ALPHA);
}
template <bool FU = false>
ZHD float t_srgb_inv_eotf(float x)
{
	x = maxf(x, 0.0f);
	return x < SRGB_BETA ? x * 12.92f : SRGB_ALPHA * powf_k<FU, ZM_FB(1.0f / 2.4f)>(x) - (SRGB_ALPHA - 1.0f);
}
ZHD float t_428_eotf(float x) { return x < 0.0f ? 0.0f : 52.37f / 48.0f * powf_(x, 2.6f); }
ZHD float t_428_inv_eotf(float x) { return x < 0.0f ? 0.0f : powf_(x * (48.0f / 52.37f), 1.0f / 2.6f); }
ZHD float t_b67_eotf(float x)
{
	float y = arib_b67_inverse_oetf(x);
	return y < 0.0f ? y : powf_(y, 1.2f);
}
ZHD float t_b67_inv_eotf(float x)
{
	float y = x < 0.0f ? x : powf_(x, 1.0f / 1.2f);
	return arib_b67_oetf(y);
}
ZHD float t_2084_oetf(float x) { return st2084_inverse_eotf(rec1886_eotf(rec709_oetf(x * ST2084_OOTF_SCALE)) / 100.0f); }
ZHD float t_2084_inv_oetf(float x) { return rec709_inverse_oetf(rec1886_inverse_eotf(st2084_eotf(x) * 100.0f)) / ST2084_OOTF_SCALE; }

#define ZAMD_TRANSFERS(X) \
	X(F_709_OETF, rec709_oetf) X(F_709_INV_OETF, rec709_inverse_oetf) \
	X(F_LOG100_OETF, t_log100_oetf) X(F_LOG100_INV_OETF, t_log100_inv_oetf) \
	X(F_LOG316_OETF, t_log316_oetf) X(F_LOG316_INV_OETF, t_log316_inv_oetf) \
	X(F_470M_OETF, t_470m_oetf) X(F_470M_INV_OETF, t_470m_inv_oetf) \
	X(F_470BG_OETF, t_470bg_oetf) X(F_470BG_INV_OETF, t_470bg_inv_oetf) \
	X(F_240M_OETF, t_240m_oetf) X(F_240M_INV_OETF, t_240m_inv_oetf) \
	X(F_XVYCC_OETF, t_xvycc_oetf) X(F_XVYCC_INV_OETF, t_xvycc_inv_oetf) \
	X(F_1361_OETF, t_1361_oetf) X(F_1361_INV_OETF, t_1361_inv_oetf) \
	X(F_B67_OETF, arib_b67_oetf) X(F_B67_INV_OETF, arib_b67_inverse_oetf) \
	X(F_1886_EOTF, rec1886_eotf) X(F_1886_INV_EOTF, rec1886_inverse_eotf) \
	X(F_1361_EOTF, t_1361_eotf) X(F_1361_INV_EOTF, t_1361_inv_eotf) \
	X(F_XVYCC_EOTF, t_xvycc_eotf) X(F_XVYCC_INV_EOTF, t_xvycc_inv_eotf) \
	X(F_SRGB_EOTF, t_srgb_eotf) X(F_SRGB_INV_EOTF, t_srgb_inv_eotf) \
	X(F_2084_EOTF, st2084_eotf) X(F_2084_INV_EOTF, st2084_inverse_eotf) \
	X(F_428_EOTF, t_428_eotf) X(F_428_INV_EOTF, t_428_inv_eotf) \
	X(F_B67_EOTF, t_b67_eotf) X(F_B67_INV_EOTF, t_b67_inv_eotf) \
	X(F_2084_OETF, t_2084_oetf) X(F_2084_INV_OETF, t_2084_inv_oetf)

ZHD float transfer(int f, float x)
{
	switch (f) {
#define ZAMD_CASE(id, fn) case id: return fn(x);
	ZAMD_TRANSFERS(ZAMD_CASE)
#undef ZAMD_CASE
	default:
		return bitsf(0x7fc00000u);
	}
}

// Transfers whose powf steps have a fused form (powf_fused): the same float
// results as the reference's for every input, with fewer instructions.
#define ZAMD_FUSED_TRANSFERS(X) \
	X(F_709_OETF, rec709_oetf) X(F_709_INV_OETF, rec709_inverse_oetf) \
	X(F_1886_EOTF, rec1886_eotf) X(F_1886_INV_EOTF, rec1886_inverse_eotf) \
	X(F_SRGB_EOTF, t_srgb_eotf) X(F_SRGB_INV_EOTF, t_srgb_inv_eotf) \
	X(F_2084_EOTF, st2084_eotf) X(F_2084_INV_EOTF, st2084_inverse_eotf)

// transfer() with the function fixed at compile time (the fused form where one exists).
template <int F>
ZHD float transfer_c(float x)
{
#define ZAMD_CASE(id, fn) if constexpr (F == id) return fn<true>(x); else
	ZAMD_FUSED_TRANSFERS(ZAMD_CASE)
#undef ZAMD_CASE
#define ZAMD_CASE(id, fn) if constexpr (F == id) return fn(x); else
	ZAMD_TRANSFERS(ZAMD_CASE)
#undef ZAMD_CASE
	return bitsf(0x7fc00000u);
}

// transfer() over n values with the function selected once.
template <int N>
ZHD void transfer_n(int f, float *x)
{
	switch (f) {
#define ZAMD_CASE(id, fn) case id: for (int i = 0; i < N; ++i) x[i] = fn(x[i]); return;
	ZAMD_TRANSFERS(ZAMD_CASE)
#undef ZAMD_CASE
	default:
		for (int i = 0; i < N; ++i) x[i] = bitsf(0x7fc00000u);
	}
}

} // namespace zm
} // namespace zamd
