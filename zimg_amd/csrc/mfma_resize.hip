// Precision form of the FLOAT resize passes: each pass is a banded dense
// contraction on gfx950 matrix cores (v_mfma_f32_16x16x32_{f16,bf16}) with
// FP32 accumulation. Samples and coefficients are split into a 16-bit high
// part and a 16-bit remainder, x = hi(x) + lo(x), and each 16x16 output tile
// accumulates hi*hi + hi*lo + lo*hi, which recovers about twice the 16-bit
// format's precision (the lo*lo term is below FP32 rounding). Coefficients
// are the FP64 entries of the resampling matrix (filter.cpp:249-306) split on
// the host; the data split happens in registers.
//
// Lane layout of the 16x16x32 operations (wave64): lane l holds, for A
// (16 x 32, row-major M x K) row l % 16, k = 8 (l / 16) .. + 7; for B
// (32 x 16, K x N) column l % 16 with the same k; the 16 x 16 FP32 result
// holds row 4 (l / 16) + r, r < 4, column l % 16.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.h"

namespace zamd {
namespace dev {

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr unsigned kWaves = 4;

__device__ __forceinline__ float load_sample(const char *row, unsigned j, const InXform &in)
{
	float x;
	if (in.type == 0) x = static_cast<float>(reinterpret_cast<const uint8_t *>(row)[j]);
	else if (in.type == 1) x = static_cast<float>(reinterpret_cast<const uint16_t *>(row)[j]);
	else if (in.type == 2) x = static_cast<float>(__builtin_bit_cast(_Float16, reinterpret_cast<const uint16_t *>(row)[j]));
	else x = reinterpret_cast<const float *>(row)[j];
	if (in.op == 2) x = x * in.scale + in.offset;
	return x;
}

__device__ __forceinline__ void store_sample(char *row, unsigned j, float v, const OutXform &o)
{
	if (o.op == 1) {
		float x = v * o.scale + o.offset;
		const float maxval = static_cast<float>(1UL << o.bits) - 1;
		x = x < 0.0f ? 0.0f : (maxval < x ? maxval : x);
		const unsigned q = static_cast<unsigned>(static_cast<int>(__builtin_rintf(x)));
		if (o.type == 0) reinterpret_cast<uint8_t *>(row)[j] = static_cast<uint8_t>(q);
		else reinterpret_cast<uint16_t *>(row)[j] = static_cast<uint16_t>(q);
	} else if (o.type == 2) {
		reinterpret_cast<uint16_t *>(row)[j] = __builtin_bit_cast(uint16_t, static_cast<_Float16>(v));
	} else {
		reinterpret_cast<float *>(row)[j] = v;
	}
}

// Split 8 FP32 samples into high and low 16-bit operands.
template <bool BF16>
struct Split;

template <>
struct Split<false> {
	using V = h8;
	__device__ __forceinline__ static void run(const float *x, V &hi, V &lo)
	{
#pragma unroll
		for (int i = 0; i < 8; ++i) {
			const _Float16 h = static_cast<_Float16>(x[i]);
			hi[i] = h;
			lo[i] = static_cast<_Float16>(x[i] - static_cast<float>(h));
		}
	}
	__device__ __forceinline__ static f4 mma(V a, V b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};

template <>
struct Split<true> {
	using V = b8;
	__device__ __forceinline__ static void run(const float *x, V &hi, V &lo)
	{
#pragma unroll
		for (int i = 0; i < 8; ++i) {
			const __bf16 h = static_cast<__bf16>(x[i]);
			hi[i] = h;
			lo[i] = static_cast<__bf16>(x[i] - static_cast<float>(h));
		}
	}
	__device__ __forceinline__ static f4 mma(V a, V b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};

// Coefficient operand of tile t, k-step s for this lane: 8 high then 8 low
// 16-bit values; each (tile, step) block is 64 lanes x 16 values.
template <bool BF16>
__device__ __forceinline__ void load_coef(const MfmaTab &tab, unsigned t, unsigned s, unsigned lane, typename Split<BF16>::V &hi,
                                          typename Split<BF16>::V &lo)
{
	const uint4 *blk = reinterpret_cast<const uint4 *>(tab.coef) + (static_cast<size_t>(tab.off[t]) + s) * 128;
	const uint4 h = blk[lane], l = blk[64 + lane];
	hi = __builtin_bit_cast(typename Split<BF16>::V, h);
	lo = __builtin_bit_cast(typename Split<BF16>::V, l);
}

// Horizontal pass: wave w of the block computes rows r0 + 16 w .. + 15 of
// output columns 16 blockIdx.x .. + 15, as (16 rows x K) x (K x 16 columns).
template <bool BF16>
__global__ void __launch_bounds__(64 * kWaves) k_mfma_h(MfmaArg a)
{
	using S = Split<BF16>;
	const unsigned lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	const unsigned t = blockIdx.x;
	const unsigned j0 = 16 * t;
	const unsigned r0 = (blockIdx.y * kWaves + w) * 16;
	const unsigned f = blockIdx.z;
	if (r0 >= a.rows)
		return;
	const char *src = a.src.ptr + static_cast<long long>(f) * a.src.fstride;
	char *dst = a.dst.ptr + static_cast<long long>(f) * a.dst.fstride;
	const unsigned row = min(r0 + (lane & 15), a.rows - 1);
	const char *srow = src + static_cast<long long>(row) * a.src.stride;
	const int kb = a.tab.kb[t];
	const unsigned nks = a.tab.nks[t];
	const unsigned kl = 8 * (lane >> 4);

	f4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
	for (unsigned s = 0; s < nks; ++s) {
		float x[8];
		const int c0 = kb + static_cast<int>(32 * s + kl);
		if (a.vec16 && a.in.type == 3 && a.in.op == 0 && c0 + 8 <= static_cast<int>(a.src_len)) {
			const float4 *p = reinterpret_cast<const float4 *>(srow + 4 * static_cast<size_t>(c0));
			const float4 u = p[0], v = p[1];
			x[0] = u.x; x[1] = u.y; x[2] = u.z; x[3] = u.w;
			x[4] = v.x; x[5] = v.y; x[6] = v.z; x[7] = v.w;
		} else {
#pragma unroll
			for (int i = 0; i < 8; ++i) {
				const int c = c0 + i;
				x[i] = (c >= 0 && c < static_cast<int>(a.src_len)) ? load_sample(srow, static_cast<unsigned>(c), a.in) : 0.0f;
			}
		}
		typename S::V xh, xl, ch, cl;
		S::run(x, xh, xl);
		load_coef<BF16>(a.tab, t, s, lane, ch, cl);
		acc = S::mma(xh, ch, acc);
		acc = S::mma(xh, cl, acc);
		acc = S::mma(xl, ch, acc);
	}
	const unsigned j = j0 + (lane & 15);
	if (j >= a.dst_len)
		return;
#pragma unroll
	for (int r = 0; r < 4; ++r) {
		const unsigned orow = r0 + 4 * (lane >> 4) + r;
		if (orow < a.rows)
			store_sample(dst + static_cast<long long>(orow) * a.dst.stride, j, acc[r], a.out);
	}
}

// Vertical pass: wave w of the block computes columns c0 + 16 w .. + 15 of
// output rows 16 blockIdx.y .. + 15, as (16 rows x K) x (K x 16 columns).
template <bool BF16>
__global__ void __launch_bounds__(64 * kWaves) k_mfma_v(MfmaArg a)
{
	using S = Split<BF16>;
	const unsigned lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	const unsigned t = blockIdx.y;
	const unsigned i0 = 16 * t;
	const unsigned c0 = (blockIdx.x * kWaves + w) * 16;
	const unsigned f = blockIdx.z;
	if (c0 >= a.cols)
		return;
	const char *src = a.src.ptr + static_cast<long long>(f) * a.src.fstride;
	char *dst = a.dst.ptr + static_cast<long long>(f) * a.dst.fstride;
	const unsigned col = min(c0 + (lane & 15), a.cols - 1);
	const int kb = a.tab.kb[t];
	const unsigned nks = a.tab.nks[t];
	const unsigned kl = 8 * (lane >> 4);

	f4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
	for (unsigned s = 0; s < nks; ++s) {
		float x[8];
		const int rbase = kb + static_cast<int>(32 * s + kl);
#pragma unroll
		for (int i = 0; i < 8; ++i) {
			const int r = rbase + i;
			x[i] = (r >= 0 && r < static_cast<int>(a.src_len)) ? load_sample(src + static_cast<long long>(r) * a.src.stride, col, a.in) : 0.0f;
		}
		typename S::V xh, xl, ch, cl;
		S::run(x, xh, xl);
		load_coef<BF16>(a.tab, t, s, lane, ch, cl);
		acc = S::mma(ch, xh, acc);
		acc = S::mma(cl, xh, acc);
		acc = S::mma(ch, xl, acc);
	}
	const unsigned c = c0 + (lane & 15);
	if (c >= a.cols)
		return;
#pragma unroll
	for (int r = 0; r < 4; ++r) {
		const unsigned orow = i0 + 4 * (lane >> 4) + r;
		if (orow < a.dst_len)
			store_sample(dst + static_cast<long long>(orow) * a.dst.stride, c, acc[r], a.out);
	}
}

} // namespace

int launch_mfma_resize(const MfmaArg &a, void *stream)
{
	hipStream_t s = static_cast<hipStream_t>(stream);
	if (a.horizontal) {
		dim3 grid((a.dst_len + 15) / 16, (a.rows + 16 * kWaves - 1) / (16 * kWaves), a.frames);
		if (a.bf16) hipLaunchKernelGGL(k_mfma_h<true>, grid, dim3(64 * kWaves), 0, s, a);
		else hipLaunchKernelGGL(k_mfma_h<false>, grid, dim3(64 * kWaves), 0, s, a);
	} else {
		dim3 grid((a.cols + 16 * kWaves - 1) / (16 * kWaves), (a.dst_len + 15) / 16, a.frames);
		if (a.bf16) hipLaunchKernelGGL(k_mfma_v<true>, grid, dim3(64 * kWaves), 0, s, a);
		else hipLaunchKernelGGL(k_mfma_v<false>, grid, dim3(64 * kWaves), 0, s, a);
	}
	return static_cast<int>(hipGetLastError());
}

} // namespace dev
} // namespace zamd
