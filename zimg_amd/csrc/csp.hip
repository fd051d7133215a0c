// The fused colourspace stage: ColorspaceConversionImpl::process
// (colorspace.cpp:45-77) with the depth conversions before and after it, up
// to six in-place operations on three float channels (operation_impl.cpp:
// 21-285, operation_impl_avx2.cpp:43-382 for the approximate-gamma forms).
//
// Two kernels run the same per-operation code:
//  * k_colorspace reads the operation list at run time (one uniform branch
//    per operation and per transfer function);
//  * k_colorspace_chain<...> has the operation kinds and transfer functions
//    as template parameters, for the chains that conversions between the
//    common video systems produce (matrix / transfer / primaries / transfer /
//    matrix, and PQ -> HLG). Its code is straight-line: no dispatch, every
//    operation constant in scalar registers, no scalar spills, and a short
//    instruction stream for the instruction cache. Its powf steps use the
//    fused form where one exists (zm::powf_fused: fused multiply-adds in the
//    FP64 polynomials, with float results equal to the reference's for every
//    input). Both kernels give identical outputs.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "devutil.h"
#include "kernels.h"
#include "zmath.h"

namespace zamd {
namespace dev {

namespace {

// Pixels per thread of the colourspace kernels (consecutive blocks of kThreads
// columns): the run-time operation list, and the compile-time chains (their
// straight-line code keeps more pixels' FP64 chains in flight in fewer
// registers: 8 pixels measured 5 % faster than 4 on C2, 2 slower).
constexpr int kCspPix = 4;
constexpr int kCspChainPix = 8;

template <int NP>
__device__ __forceinline__ void op_matrix(const CspOp &op, float *a, float *b, float *c)
{
#pragma unroll
	for (int p = 0; p < NP; ++p) {
		const float x = op.m[0] * a[p] + op.m[1] * b[p] + op.m[2] * c[p];
		const float y = op.m[3] * a[p] + op.m[4] * b[p] + op.m[5] * c[p];
		const float z = op.m[6] * a[p] + op.m[7] * b[p] + op.m[8] * c[p];
		a[p] = x; b[p] = y; c[p] = z;
	}
}

template <int NP>
__device__ __forceinline__ void op_matrix_fma(const CspOp &op, float *a, float *b, float *c)
{
#pragma unroll
	for (int p = 0; p < NP; ++p) {
		const float x = __builtin_fmaf(op.m[2], c[p], __builtin_fmaf(op.m[1], b[p], op.m[0] * a[p]));
		const float y = __builtin_fmaf(op.m[5], c[p], __builtin_fmaf(op.m[4], b[p], op.m[3] * a[p]));
		const float z = __builtin_fmaf(op.m[8], c[p], __builtin_fmaf(op.m[7], b[p], op.m[6] * a[p]));
		a[p] = x; b[p] = y; c[p] = z;
	}
}

template <int NP>
__device__ __forceinline__ void op_lut_linear(const CspOp &op, float *a, float *b, float *c)
{
	// index = cvtps2dq(fma(x, 2^15, 2^14)): round to nearest even, NaN and
	// values outside int32 to INT32_MIN; clamped to the table
	auto idx = [](float x) {
		const float v = __builtin_fmaf(x, 32768.0f, 16384.0f);
		const int i = (v >= -2147483648.0f && v < 2147483648.0f) ? static_cast<int>(__builtin_rintf(v)) : INT_MIN;
		return static_cast<unsigned>(min(max(i, 0), static_cast<int>(kLutLinearSize) - 1));
	};
#pragma unroll
	for (int p = 0; p < NP; ++p) {
		a[p] = op.lut[idx(a[p])];
		b[p] = op.lut[idx(b[p])];
		c[p] = op.lut[idx(c[p])];
	}
}

template <int NP>
__device__ __forceinline__ void op_lut_gamma(const CspOp &op, float *a, float *b, float *c)
{
#pragma unroll
	for (int p = 0; p < NP; ++p) {
		a[p] = op.lut[float_to_half(a[p])];
		b[p] = op.lut[float_to_half(b[p])];
		c[p] = op.lut[float_to_half(c[p])];
	}
}

// GammaOperationC (operation_impl.cpp:79-100): post * f(pre * x). F >= 0
// fixes the transfer function at compile time; F < 0 selects op.func once
// for all 3 NP values.
template <int NP, int F>
__device__ __forceinline__ void op_gamma(const CspOp &op, float *a, float *b, float *c)
{
	float x[3 * NP];
#pragma unroll
	for (int p = 0; p < NP; ++p) {
		x[p] = a[p] * op.pre;
		x[NP + p] = b[p] * op.pre;
		x[2 * NP + p] = c[p] * op.pre;
	}
	if constexpr (F >= 0) {
#pragma unroll
		for (int k = 0; k < 3 * NP; ++k)
			x[k] = zm::transfer_c<F>(x[k]);
	} else {
		zm::transfer_n<3 * NP>(op.func, x);
	}
#pragma unroll
	for (int p = 0; p < NP; ++p) {
		a[p] = op.post * x[p];
		b[p] = op.post * x[NP + p];
		c[p] = op.post * x[2 * NP + p];
	}
}

// AribB67OperationC (operation_impl.cpp:102-125): inverse OOTF scaling by
// Ys^((1 - gamma) / gamma), then (OETF true) the HLG OETF. FU: the fused
// powf (the same results, fewer instructions).
template <int NP, bool OETF, bool FU = false>
__device__ __forceinline__ void op_b67(const CspOp &op, float *a, float *b, float *c)
{
#pragma unroll
	for (int p = 0; p < NP; ++p) {
		constexpr float gamma = 1.2f;
		const float r = a[p] * op.pre, g = b[p] * op.pre, bl = c[p] * op.pre;
		const float yd = zm::maxf(op.kr * r + op.kg * g + op.kb * bl, zm::FLT_MIN_);
		const float ys_inv = zm::powf_k<FU, ZM_FB((1.0f - gamma) / gamma)>(yd);
		if constexpr (OETF) {
			a[p] = zm::arib_b67_oetf(r * ys_inv);
			b[p] = zm::arib_b67_oetf(g * ys_inv);
			c[p] = zm::arib_b67_oetf(bl * ys_inv);
		} else {
			a[p] = r * ys_inv;
			b[p] = g * ys_inv;
			c[p] = bl * ys_inv;
		}
	}
}

// AribB67InverseOperationC (operation_impl.cpp:127-150).
template <int NP, bool FU = false>
__device__ __forceinline__ void op_b67_inv(const CspOp &op, float *a, float *b, float *c)
{
#pragma unroll
	for (int p = 0; p < NP; ++p) {
		constexpr float gamma = 1.2f;
		float r = zm::arib_b67_inverse_oetf(a[p]), g = zm::arib_b67_inverse_oetf(b[p]), bl = zm::arib_b67_inverse_oetf(c[p]);
		float ys = zm::maxf(op.kr * r + op.kg * g + op.kb * bl, zm::FLT_MIN_);
		ys = zm::powf_k<FU, ZM_FB(gamma - 1.0f)>(ys);
		r = r * ys; g = g * ys; bl = bl * ys;
		a[p] = r * op.post; b[p] = g * op.post; c[p] = bl * op.post;
	}
}

// CLToRGBOperationC (operation_impl.cpp:152-188).
template <int NP>
__device__ __forceinline__ void op_cl_to_rgb(const CspOp &op, float *a, float *b, float *c)
{
	float x[3 * NP];
#pragma unroll
	for (int p = 0; p < NP; ++p) {
		const float y = a[p], u = b[p], v = c[p];
		const float bmy = u < 0 ? u * 2.0f * op.nb : u * 2.0f * op.pb;
		const float rmy = v < 0 ? v * 2.0f * op.nr : v * 2.0f * op.pr;
		x[p] = bmy + y;
		x[NP + p] = rmy + y;
		x[2 * NP + p] = y;
	}
	zm::transfer_n<3 * NP>(op.func, x);
#pragma unroll
	for (int p = 0; p < NP; ++p) {
		const float bl = x[p], r = x[NP + p], yl = x[2 * NP + p];
		const float g = (yl - op.kr * r - op.kb * bl) / op.kg;
		a[p] = r * op.post; b[p] = g * op.post; c[p] = bl * op.post;
	}
}

// CLToYUVOperationC (operation_impl.cpp:190-227).
template <int NP>
__device__ __forceinline__ void op_cl_to_yuv(const CspOp &op, float *a, float *b, float *c)
{
	float x[3 * NP];
#pragma unroll
	for (int p = 0; p < NP; ++p) {
		const float r = a[p] * op.pre, g = b[p] * op.pre, bl = c[p] * op.pre;
		x[p] = op.kr * r + op.kg * g + op.kb * bl;
		x[NP + p] = bl;
		x[2 * NP + p] = r;
	}
	zm::transfer_n<3 * NP>(op.func, x);
#pragma unroll
	for (int p = 0; p < NP; ++p) {
		const float y = x[p], bl = x[NP + p], r = x[2 * NP + p];
		const float u = (bl - y < 0.0f) ? (bl - y) / (2.0f * op.nb) : (bl - y) / (2.0f * op.pb);
		const float v = (r - y < 0.0f) ? (r - y) / (2.0f * op.nr) : (r - y) / (2.0f * op.pr);
		a[p] = y; b[p] = u; c[p] = v;
	}
}

// One operation of a run-time list: the operation (and a transfer
// function's selection) is decided once for the NP pixels of three channels.
template <int NP>
__device__ __forceinline__ void apply_op(const CspOp &op, float *a, float *b, float *c)
{
	switch (op.kind) {
	case CSP_MATRIX_FMA: op_matrix_fma<NP>(op, a, b, c); break;
	case CSP_LUT_LINEAR: op_lut_linear<NP>(op, a, b, c); break;
	case CSP_LUT_GAMMA: op_lut_gamma<NP>(op, a, b, c); break;
	case CSP_MATRIX: op_matrix<NP>(op, a, b, c); break;
	case CSP_GAMMA: op_gamma<NP, -1>(op, a, b, c); break;
	case CSP_B67: op_b67<NP, true>(op, a, b, c); break;
	case CSP_B67_SCALE: op_b67<NP, false>(op, a, b, c); break;
	case CSP_B67_INV: op_b67_inv<NP>(op, a, b, c); break;
	case CSP_CL_TO_RGB: op_cl_to_rgb<NP>(op, a, b, c); break;
	case CSP_CL_TO_YUV: op_cl_to_yuv<NP>(op, a, b, c); break;
	default: break;
	}
}

// One slot of a compile-time chain: kind K (-1: no operation), transfer F.
template <int NP, int K, int F>
__device__ __forceinline__ void apply_slot(const CspOp &op, float *a, float *b, float *c)
{
	if constexpr (K == CSP_MATRIX)
		op_matrix<NP>(op, a, b, c);
	else if constexpr (K == CSP_GAMMA)
		op_gamma<NP, F>(op, a, b, c);
	else if constexpr (K == CSP_B67)
		op_b67<NP, true, true>(op, a, b, c);
	else if constexpr (K == CSP_B67_INV)
		op_b67_inv<NP, true>(op, a, b, c);
	else
		static_assert(K < 0, "operation kind without a chain form");
}

// The loads (with their depth conversion) of a thread's NPX pixels.
template <int NPX>
__device__ __forceinline__ void csp_load(const CspArg &a, unsigned f, unsigned i, unsigned jb, float (*v)[NPX])
{
#pragma unroll
	for (int p = 0; p < 3; ++p) {
		const char *srow = row_ptr(frame_ptr(a.src[p], f), a.src[p].stride, i);
#pragma unroll
		for (int k = 0; k < NPX; ++k) {
			const unsigned j = min(jb + k * kThreads, a.width - 1);
			float x = load_as_float(srow, j, a.in[p].type);
			if (a.in[p].op == 2)
				x = x * a.in[p].scale + a.in[p].offset;
			v[p][k] = x;
		}
	}
}

template <int NPX>
__device__ __forceinline__ void csp_store(const CspArg &a, unsigned f, unsigned i, unsigned jb, float (*v)[NPX])
{
#pragma unroll
	for (int p = 0; p < 3; ++p) {
		char *drow = row_ptr(frame_ptr_w(a.dst[p], f), a.dst[p].stride, i);
#pragma unroll
		for (int k = 0; k < NPX; ++k) {
			const unsigned j = jb + k * kThreads;
			if (j >= a.width)
				break;
			if (a.out[p].op == 1)
				store_uint(drow, j, a.out[p].type, quantize(v[p][k], a.out[p], i, j));
			else
				store_float(drow, j, a.out[p].type, v[p][k]);
		}
	}
}

__global__ void __launch_bounds__(kThreads) k_colorspace(CspArg a)
{
	zm::lds_tables_init();
	const unsigned f = blockIdx.z;
	const unsigned i = blockIdx.y;
	const unsigned jb = blockIdx.x * (kThreads * kCspPix) + threadIdx.x;
	if (i >= a.height || jb >= a.width)
		return;
	float v[3][kCspPix];
	csp_load<kCspPix>(a, f, i, jb, v);
	for (int k = 0; k < a.prog.n; ++k)
		apply_op<kCspPix>(a.prog.op[k], v[0], v[1], v[2]);
	csp_store<kCspPix>(a, f, i, jb, v);
}

template <int K0, int F0, int K1, int F1, int K2, int F2, int K3, int F3, int K4, int F4>
__global__ void __launch_bounds__(kThreads) k_colorspace_chain(CspArg a)
{
	zm::lds_tables_init();
	const unsigned f = blockIdx.z;
	const unsigned i = blockIdx.y;
	const unsigned jb = blockIdx.x * (kThreads * kCspChainPix) + threadIdx.x;
	if (i >= a.height || jb >= a.width)
		return;
	float v[3][kCspChainPix];
	csp_load<kCspChainPix>(a, f, i, jb, v);
	if constexpr (K0 >= 0) apply_slot<kCspChainPix, K0, F0>(a.prog.op[0], v[0], v[1], v[2]);
	if constexpr (K1 >= 0) apply_slot<kCspChainPix, K1, F1>(a.prog.op[1], v[0], v[1], v[2]);
	if constexpr (K2 >= 0) apply_slot<kCspChainPix, K2, F2>(a.prog.op[2], v[0], v[1], v[2]);
	if constexpr (K3 >= 0) apply_slot<kCspChainPix, K3, F3>(a.prog.op[3], v[0], v[1], v[2]);
	if constexpr (K4 >= 0) apply_slot<kCspChainPix, K4, F4>(a.prog.op[4], v[0], v[1], v[2]);
	csp_store<kCspChainPix>(a, f, i, jb, v);
}

// Chains with a compile-time kernel: (kind, transfer) per slot, -1 past the end.
#define M_ CSP_MATRIX, 0
#define G_(f) CSP_GAMMA, zm::f
#define B67_ CSP_B67, 0
#define END_ -1, 0
#define ZAMD_CSP_CHAINS(X) \
	/* YUV -> RGB, display-referred (BT.1886) linearisation, primaries, back (BT.709/601/2020 SDR) */ \
	X(M_, G_(F_1886_EOTF), M_, G_(F_1886_INV_EOTF), M_) \
	/* the same, scene-referred (BT.709 OETF) */ \
	X(M_, G_(F_709_INV_OETF), M_, G_(F_709_OETF), M_) \
	/* RGB in: linearise, primaries, transfer, YUV out */ \
	X(G_(F_1886_EOTF), M_, G_(F_1886_INV_EOTF), M_, END_) \
	X(M_, G_(F_1886_EOTF), M_, G_(F_1886_INV_EOTF), END_) \
	/* SDR <-> PQ (BT.2100) */ \
	X(M_, G_(F_1886_EOTF), M_, G_(F_2084_INV_EOTF), M_) \
	X(M_, G_(F_2084_EOTF), M_, G_(F_1886_INV_EOTF), M_) \
	/* PQ -> HLG (ST 2084 EOTF, inverse OOTF and HLG OETF), RGB */ \
	X(G_(F_2084_EOTF), B67_, END_, END_, END_) \
	/* sRGB <-> linear */ \
	X(G_(F_SRGB_EOTF), END_, END_, END_, END_) \
	X(G_(F_SRGB_INV_EOTF), END_, END_, END_, END_)

struct ChainKey {
	int kind[5], func[5];
};

using CspKernel = void (*)(CspArg);

struct ChainEntry {
	ChainKey key;
	CspKernel fn;
};

#define ZAMD_CHAIN_ENTRY(k0, f0, k1, f1, k2, f2, k3, f3, k4, f4) \
	{ { { k0, k1, k2, k3, k4 }, { f0, f1, f2, f3, f4 } }, k_colorspace_chain<k0, f0, k1, f1, k2, f2, k3, f3, k4, f4> },
#define ZAMD_CHAIN_X(...) ZAMD_CHAIN_ENTRY(__VA_ARGS__)

const ChainEntry kChains[] = { ZAMD_CSP_CHAINS(ZAMD_CHAIN_X) };

#undef ZAMD_CHAIN_X
#undef ZAMD_CHAIN_ENTRY
#undef M_
#undef G_
#undef B67_
#undef END_

// The compile-time kernel of a program, or null.
CspKernel chain_kernel(const CspProgram &prog)
{
	if (prog.n < 1 || prog.n > 5)
		return nullptr;
	for (const ChainEntry &e : kChains) {
		bool ok = true;
		for (int s = 0; s < 5 && ok; ++s) {
			if (s >= prog.n) {
				ok = e.key.kind[s] < 0;
				continue;
			}
			const CspOp &op = prog.op[s];
			ok = op.kind == e.key.kind[s] && (op.kind != CSP_GAMMA || op.func == e.key.func[s]);
		}
		if (ok)
			return e.fn;
	}
	return nullptr;
}

} // namespace

int launch_colorspace(const CspArg &a, void *stream)
{
	CspKernel k = chain_kernel(a.prog);
	const unsigned px = k ? kCspChainPix : kCspPix;
	if (!k)
		k = k_colorspace;
	dim3 grid((a.width + kThreads * px - 1) / (kThreads * px), a.height, a.frames);
	CspArg arg = a;
	void *args[] = { &arg };
	return static_cast<int>(hipLaunchKernel(reinterpret_cast<const void *>(k), grid, dim3(kThreads), args, 0, static_cast<hipStream_t>(stream)));
}

bool colorspace_has_chain_kernel(const CspProgram &prog) { return chain_kernel(prog) != nullptr; }

} // namespace dev
} // namespace zamd
