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
#include <type_traits>

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
	const uint4 *blk = reinterpret_cast<const uint4 *>(tab.coef) + (static_cast<size_t>(tab.tile[t].off) + s) * 128;
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
	const MfmaTile tt = a.tab.tile[t];
	const int kb = tt.kb;
	const unsigned nks = tt.nks;
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
	const MfmaTile tt = a.tab.tile[t];
	const int kb = tt.kb;
	const unsigned nks = tt.nks;
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

// Split 4 FP32 values into packed high and low 16-bit parts.
template <bool BF16>
__device__ __forceinline__ void split4(const float *x, uint2 &hi, uint2 &lo)
{
	using T = typename std::conditional<BF16, __bf16, _Float16>::type;
	typedef T t4 __attribute__((ext_vector_type(4)));
	t4 h, l;
#pragma unroll
	for (int i = 0; i < 4; ++i) {
		const T v = static_cast<T>(x[i]);
		h[i] = v;
		l[i] = static_cast<T>(x[i] - static_cast<float>(v));
	}
	hi = __builtin_bit_cast(uint2, h);
	lo = __builtin_bit_cast(uint2, l);
}

// Four consecutive samples of a source row from column col (a multiple of
// 4); columns at or beyond the row length read as zero.
template <bool FAST>
__device__ __forceinline__ void fused_load(const MfmaFusedArg &a, const char *row, int col, float *x)
{
	if (FAST && col + 4 <= static_cast<int>(a.src_w)) {
		const float4 v = *reinterpret_cast<const float4 *>(row + 4 * static_cast<size_t>(col));
		x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
		return;
	}
#pragma unroll
	for (int i = 0; i < 4; ++i) {
		const int c = col + i;
		if (FAST)
			x[i] = c < static_cast<int>(a.src_w) ? reinterpret_cast<const float *>(row)[c] : 0.0f;
		else
			x[i] = c < static_cast<int>(a.src_w) ? load_sample(row, static_cast<unsigned>(c), a.in) : 0.0f;
	}
}

// 16-byte LDS operand: the 8 consecutive 16-bit values at base + run, or zero
// when the run starts past the window (then base itself is read, in bounds,
// and discarded: no branch).
__device__ __forceinline__ uint4 lds_run(const uint16_t *base, unsigned run, bool inside)
{
	const uint4 v = *reinterpret_cast<const uint4 *>(base + (inside ? run : 0u));
	return inside ? v : make_uint4(0, 0, 0, 0);
}

// Two-pass resize, horizontal pass first, intermediate in LDS.
//
// Pass 1 (wave w, output column tile t1 = kFusedWaves blockIdx.x + w): for
// each 16-row group of the intermediate window, A = 16 source rows x 32
// columns read from the slab, B = the tile's coefficient block; the 16 x 16
// result (lane: rows 4 (l / 16) + r, column l % 16) is split into hi/lo and
// stored column-major, so 4 rows of one column are one 8-byte LDS write.
// Pass 2 (same wave, same 16 columns, each 16-row output tile t2 of the
// workgroup): A = the transposed intermediate (16 columns x 32 rows, 16-byte
// LDS runs along a column), B = the vertical coefficient block (copied to
// LDS at the start); the result lane holds columns 4 (l / 16) + r of output
// row l % 16. All output tiles accumulate in registers; the stores come last
// (FLOAT output goes through an LDS copy of the workgroup's tile, so a store
// instruction writes whole 256-byte row segments).
template <bool BF16, bool FAST>
__global__ void __launch_bounds__(64 * kFusedWaves) __attribute__((amdgpu_waves_per_eu(3, 8))) k_mfma_fused(MfmaFusedArg a)
{
	using S = Split<BF16>;
	using V = typename S::V;
	extern __shared__ uint4 fused_lds[];
	uint16_t *lds = reinterpret_cast<uint16_t *>(fused_lds);
	const unsigned tid = threadIdx.x, lane = tid & 63;
	const unsigned w = __builtin_amdgcn_readfirstlane(tid >> 6);
	const unsigned f = blockIdx.z;
	const unsigned nt1 = (a.dst_w + 15) / 16, nt2 = (a.dst_h + 15) / 16;
	const unsigned nb2 = a.tile_h / 16;
	const unsigned t1first = blockIdx.x * kFusedWaves;
	const unsigned t1last = min(t1first + kFusedWaves, nt1) - 1;
	const unsigned t2first = blockIdx.y * nb2;
	const unsigned t2last = min(t2first + nb2, nt2) - 1;

	// Tile records (clamped indices, so every load is unconditional).
	MfmaTile h1[kFusedWaves], v2[kFusedMaxRowTiles];
#pragma unroll
	for (unsigned i = 0; i < kFusedWaves; ++i)
		h1[i] = a.th.tile[min(t1first + i, t1last)];
#pragma unroll
	for (unsigned i = 0; i < kFusedMaxRowTiles; ++i)
		v2[i] = a.tv.tile[min(t2first + i, t2last)];

	// Source column window [k0, k0 + wc) and intermediate row window [r0, r0 + wr).
	int k0 = h1[0].kb, kend = h1[0].kend;
#pragma unroll
	for (unsigned i = 1; i < kFusedWaves; ++i) {
		k0 = min(k0, h1[i].kb);
		kend = max(kend, h1[i].kend);
	}
	int r0 = v2[0].kb, rend = v2[0].kend;
#pragma unroll
	for (unsigned i = 1; i < kFusedMaxRowTiles; ++i) {
		r0 = min(r0, v2[i].kb);
		rend = max(rend, v2[i].kend);
	}
	const unsigned wc = static_cast<unsigned>(kend - k0), wr = static_cast<unsigned>(rend - r0);
	const unsigned nrg = (wr + 15) / 16;

	const unsigned sc = a.slab_stride, sr = a.mid_stride;
	uint16_t *const mid_hi = lds + 4 * 16 * sc;
	uint16_t *const mid_lo = mid_hi + 16 * kFusedWaves * sr;
	uint4 *const coef2 = reinterpret_cast<uint4 *>(mid_lo + 16 * kFusedWaves * sr);

	const char *src = a.src.ptr + static_cast<long long>(f) * a.src.fstride;
	char *dst = a.dst.ptr + static_cast<long long>(f) * a.dst.fstride;

	const unsigned t1 = t1first + w;
	const bool active = t1 <= t1last;
	const MfmaTile my1 = a.th.tile[min(t1, t1last)];
	const unsigned nks1 = my1.nks;
	const unsigned off1 = static_cast<unsigned>(my1.kb - k0);
	const unsigned kl = 8 * (lane >> 4);

	// Global loads below are unconditional (clamped to valid addresses), so
	// the compiler's waits before each use count only the loads after it.
	V ch1[kFusedMaxSteps], cl1[kFusedMaxSteps];
#pragma unroll
	for (unsigned s = 0; s < kFusedMaxSteps; ++s)
		load_coef<BF16>(a.th, min(t1, t1last), min(s, nks1 - 1), lane, ch1[s], cl1[s]);

	// The workgroup's vertical coefficient blocks (consecutive in memory,
	// at most kFusedMaxCoefBlocks): 4 uint4 per thread, loaded first and
	// written to LDS once the first two slabs are in flight.
	const unsigned cb0 = v2[0].off;
	unsigned cbn = 0;
#pragma unroll
	for (unsigned i = 0; i < kFusedMaxRowTiles; ++i)
		cbn = max(cbn, v2[i].off + v2[i].nks);
	cbn = (cbn - cb0) * 128;
	const uint4 *cg = reinterpret_cast<const uint4 *>(a.tv.coef) + static_cast<size_t>(cb0) * 128;
	const uint4 cp0 = cg[min(tid, cbn - 1)], cp1 = cg[min(tid + 256, cbn - 1)];
	const uint4 cp2 = cg[min(tid + 512, cbn - 1)], cp3 = cg[min(tid + 768, cbn - 1)];

	// Slab staging: thread (rr = tid / 16) loads row rr, 4-column groups
	// (tid % 16) + 16 j of the window. Loads run two row groups ahead of the
	// matrix-core work (two register sets); the LDS slab is double-buffered.
	// Every lane loads 4 float4 per slab (groups past the window repeat its
	// last group, rows past the window repeat its last row group); at the
	// right edge of the plane a load takes the aligned float4 holding the
	// last column (in the same page as valid data) and zeroes the columns past
	// the edge.
	const unsigned rr = tid >> 4, cq = tid & 15;
	const unsigned nq = wc / 4;
	const int last4 = (static_cast<int>(a.src_w) - 1) & ~3;
	float4 pa[4], pb[4];
	auto fetch = [&](unsigned g, float4 (&pre)[4]) {
		const int row = min(r0 + static_cast<int>(16 * min(g, nrg - 1) + rr), static_cast<int>(a.src_h) - 1);
		const char *srow = src + static_cast<long long>(row) * a.src.stride;
#pragma unroll
		for (unsigned j = 0; j < 4; ++j) {
			const int col = k0 + static_cast<int>(4 * min(cq + 16 * j, nq - 1));
			if (FAST) {
				const int cl = min(col, last4);
				const float4 v = *reinterpret_cast<const float4 *>(srow + 4 * static_cast<size_t>(cl));
				const int n = col == cl ? static_cast<int>(a.src_w) - col : 0; // valid columns from col
				pre[j] = make_float4(n > 0 ? v.x : 0.0f, n > 1 ? v.y : 0.0f, n > 2 ? v.z : 0.0f, n > 3 ? v.w : 0.0f);
			} else {
				float x[4];
				fused_load<FAST>(a, srow, col, x);
				pre[j] = make_float4(x[0], x[1], x[2], x[3]);
			}
		}
	};
	// Lanes whose group is past the window store the window's last group
	// again (same values, same place): no lane conditions here either.
	auto deposit = [&](unsigned g, const float4 (&pre)[4]) {
		uint16_t *hi = lds + (g & 1) * 2 * 16 * sc + rr * sc;
		uint16_t *lo = hi + 16 * sc;
#pragma unroll
		for (unsigned j = 0; j < 4; ++j) {
			const unsigned q = min(cq + 16 * j, nq - 1);
			uint2 h, l;
			const float x[4] = { pre[j].x, pre[j].y, pre[j].z, pre[j].w };
			split4<BF16>(x, h, l);
			*reinterpret_cast<uint2 *>(hi + 4 * q) = h;
			*reinterpret_cast<uint2 *>(lo + 4 * q) = l;
		}
	};
	auto horizontal = [&](unsigned g) {
		if (!active)
			return;
		const uint16_t *hi = lds + (g & 1) * 2 * 16 * sc + (lane & 15) * sc;
		const uint16_t *lo = hi + 16 * sc;
		f4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
		for (unsigned s = 0; s < kFusedMaxSteps; ++s) {
			if (s < nks1) {
				const unsigned run = off1 + 32 * s + kl;
				const V xh = __builtin_bit_cast(V, lds_run(hi, run, run < wc));
				const V xl = __builtin_bit_cast(V, lds_run(lo, run, run < wc));
				acc = S::mma(xh, ch1[s], acc);
				acc = S::mma(xh, cl1[s], acc);
				acc = S::mma(xl, ch1[s], acc);
			}
		}
		const float v[4] = { acc[0], acc[1], acc[2], acc[3] };
		uint2 h, l;
		split4<BF16>(v, h, l);
		const unsigned at = (16 * w + (lane & 15)) * sr + 16 * g + 4 * (lane >> 4);
		*reinterpret_cast<uint2 *>(mid_hi + at) = h;
		*reinterpret_cast<uint2 *>(mid_lo + at) = l;
	};

	fetch(0, pa);
	fetch(1, pb);
	// Lanes past the region rewrite its last entry with the same value.
	coef2[min(tid, cbn - 1)] = cp0;
	coef2[min(tid + 256, cbn - 1)] = cp1;
	coef2[min(tid + 512, cbn - 1)] = cp2;
	coef2[min(tid + 768, cbn - 1)] = cp3;
	// Row groups as nested blocks (nrg <= kFusedMaxGroups): each group is
	// reached only from the one before it, so the wait for a slab's loads
	// counts exactly the loads issued after them.
	auto group = [&](auto gc) {
		constexpr unsigned g = decltype(gc)::value;
		if constexpr ((g & 1) == 0) {
			deposit(g, pa);
			__syncthreads();
			fetch(g + 2, pa);
		} else {
			deposit(g, pb);
			__syncthreads();
			fetch(g + 2, pb);
		}
		horizontal(g);
	};
	using std::integral_constant;
	group(integral_constant<unsigned, 0>{});
	if (nrg > 1) { group(integral_constant<unsigned, 1>{});
	if (nrg > 2) { group(integral_constant<unsigned, 2>{});
	if (nrg > 3) { group(integral_constant<unsigned, 3>{});
	if (nrg > 4) { group(integral_constant<unsigned, 4>{});
	if (nrg > 5) { group(integral_constant<unsigned, 5>{});
	if (nrg > 6) { group(integral_constant<unsigned, 6>{});
	if (nrg > 7) { group(integral_constant<unsigned, 7>{});
	if (nrg > 8) { group(integral_constant<unsigned, 8>{});
	if (nrg > 9) { group(integral_constant<unsigned, 9>{});
	if (nrg > 10) { group(integral_constant<unsigned, 10>{});
	if (nrg > 11) { group(integral_constant<unsigned, 11>{});
	} } } } } } } } } } }
	static_assert(kFusedMaxGroups == 12, "one nested block per row group");
	__syncthreads();
	if (!active && !(FAST && a.stage_out))
		return;

	const uint16_t *mh = mid_hi + (16 * w + (lane & 15)) * sr;
	const uint16_t *ml = mid_lo + (16 * w + (lane & 15)) * sr;
	f4 acc[kFusedMaxRowTiles];
#pragma unroll
	for (unsigned i = 0; i < kFusedMaxRowTiles; ++i) {
		acc[i] = f4{ 0.0f, 0.0f, 0.0f, 0.0f };
		if (!active || t2first + i > t2last)
			continue;
		const unsigned off2 = static_cast<unsigned>(v2[i].kb - r0);
		const uint4 *cb = coef2 + (v2[i].off - cb0) * 128;
		for (unsigned s = 0; s < v2[i].nks; ++s) {
			const V ch = __builtin_bit_cast(V, cb[128 * s + lane]);
			const V cl = __builtin_bit_cast(V, cb[128 * s + 64 + lane]);
			const unsigned run = off2 + 32 * s + kl;
			const V xh = __builtin_bit_cast(V, lds_run(mh, run, run < wr));
			const V xl = __builtin_bit_cast(V, lds_run(ml, run, run < wr));
			acc[i] = S::mma(xh, ch, acc[i]);
			acc[i] = S::mma(xh, cl, acc[i]);
			acc[i] = S::mma(xl, ch, acc[i]);
		}
	}
	if (FAST && a.stage_out) {
		// The workgroup's output tile goes through LDS (rows of 64 floats at a
		// stride of 68) so that each store instruction writes whole 256-byte
		// row segments instead of 16 rows x 64 bytes.
		constexpr unsigned kRow = 16 * kFusedWaves + 4;
		float *ob = reinterpret_cast<float *>(mid_hi);
		__syncthreads();
		if (active) {
#pragma unroll
			for (unsigned i = 0; i < kFusedMaxRowTiles; ++i) {
				if (t2first + i <= t2last)
					*reinterpret_cast<f4 *>(ob + (16 * i + (lane & 15)) * kRow + 16 * w + 4 * (lane >> 4)) = acc[i];
			}
		}
		__syncthreads();
		const unsigned row0 = 16 * t2first;
		const unsigned nrows = min(a.tile_h, a.dst_h - row0);
		const unsigned col = 16 * t1first + 4 * (tid & 15);
		for (unsigned r = tid >> 4; r < nrows; r += 64 * kFusedWaves / 16) {
			const f4 v = *reinterpret_cast<const f4 *>(ob + r * kRow + 4 * (tid & 15));
			char *drow = dst + static_cast<long long>(row0 + r) * a.dst.stride;
			if (col + 4 <= a.dst_w) {
				*reinterpret_cast<float4 *>(drow + 4 * static_cast<size_t>(col)) = make_float4(v[0], v[1], v[2], v[3]);
			} else {
#pragma unroll
				for (int k = 0; k < 4; ++k)
					if (col + k < a.dst_w)
						reinterpret_cast<float *>(drow)[col + k] = v[k];
			}
		}
		return;
	}
	const unsigned c0 = 16 * t1 + 4 * (lane >> 4);
#pragma unroll
	for (unsigned i = 0; i < kFusedMaxRowTiles; ++i) {
		const unsigned o = 16 * (t2first + i) + (lane & 15);
		if (t2first + i > t2last || o >= a.dst_h)
			continue;
		char *drow = dst + static_cast<long long>(o) * a.dst.stride;
		if (FAST && c0 + 4 <= a.dst_w) {
			*reinterpret_cast<float4 *>(drow + 4 * static_cast<size_t>(c0)) = make_float4(acc[i][0], acc[i][1], acc[i][2], acc[i][3]);
		} else {
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				if (c0 + r < a.dst_w) {
					if (FAST)
						reinterpret_cast<float *>(drow)[c0 + r] = acc[i][r];
					else
						store_sample(drow, c0 + r, acc[i][r], a.out);
				}
			}
		}
	}
}

} // namespace

int launch_mfma_fused(const MfmaFusedArg &a, void *stream)
{
	hipStream_t s = static_cast<hipStream_t>(stream);
	dim3 grid((a.dst_w + 16 * kFusedWaves - 1) / (16 * kFusedWaves), (a.dst_h + a.tile_h - 1) / a.tile_h, a.frames);
	const bool fast = a.vec_in && a.vec_out;
	if (a.bf16 && fast) hipLaunchKernelGGL((k_mfma_fused<true, true>), grid, dim3(64 * kFusedWaves), a.lds_bytes, s, a);
	else if (a.bf16) hipLaunchKernelGGL((k_mfma_fused<true, false>), grid, dim3(64 * kFusedWaves), a.lds_bytes, s, a);
	else if (fast) hipLaunchKernelGGL((k_mfma_fused<false, true>), grid, dim3(64 * kFusedWaves), a.lds_bytes, s, a);
	else hipLaunchKernelGGL((k_mfma_fused<false, false>), grid, dim3(64 * kFusedWaves), a.lds_bytes, s, a);
	return static_cast<int>(hipGetLastError());
}

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
