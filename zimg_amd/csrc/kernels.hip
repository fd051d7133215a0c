// gfx950 kernels of the per-plane path: fused two-pass resize, point depth
// conversions, the fused colourspace stage and the error-diffusion wavefront.
//
// Arithmetic is the reference's portable C path (resize_impl.cpp:31-114,
// depth_convert.cpp:23-64, dither.cpp:37-89, operation_impl.cpp:21-285),
// evaluated in the same order with FP contraction disabled (-ffp-contract=off),
// IEEE division/sqrt and denormals kept, so outputs are bit-identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "bluenoise.h"
#include "kernels.h"
#include "zmath.h"

namespace zamd {
namespace dev {

namespace {

constexpr unsigned kThreads = 256;

__device__ __forceinline__ const char *frame_ptr(const PlaneArg &p, unsigned f) { return p.ptr + static_cast<long long>(f) * p.fstride; }
__device__ __forceinline__ char *frame_ptr_w(const PlaneArg &p, unsigned f) { return p.ptr + static_cast<long long>(f) * p.fstride; }

__device__ __forceinline__ const char *row_ptr(const char *base, long long stride, unsigned r) { return base + static_cast<long long>(r) * stride; }
__device__ __forceinline__ char *row_ptr(char *base, long long stride, unsigned r) { return base + static_cast<long long>(r) * stride; }

// half_to_float (quantize.cpp:24-77): exact widening; a NaN keeps its payload
// and is made quiet.
__device__ __forceinline__ float half_to_float(uint16_t h)
{
	if ((h & 0x7C00u) == 0x7C00u && (h & 0x03FFu))
		return __uint_as_float((static_cast<uint32_t>(h & 0x8000u) << 16) | 0x7FC00000u | (static_cast<uint32_t>(h & 0x03FFu) << 13));
	return static_cast<float>(__builtin_bit_cast(_Float16, h));
}

// float_to_half (quantize.cpp:79-144): IEEE binary16 round-to-nearest-even,
// overflow to infinity; a NaN keeps the top of its payload and is made quiet.
__device__ __forceinline__ uint16_t float_to_half(float x)
{
	const uint32_t u = __float_as_uint(x);
	if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu))
		return static_cast<uint16_t>(((u >> 16) & 0x8000u) | 0x7E00u | ((u & 0x007FFFFFu) >> 13));
	return __builtin_bit_cast(uint16_t, static_cast<_Float16>(x));
}

// Storage types: 0 BYTE, 1 WORD, 2 HALF, 3 FLOAT.
__device__ __forceinline__ float load_as_float(const char *row, unsigned j, int type)
{
	if (type == 0) return static_cast<float>(reinterpret_cast<const uint8_t *>(row)[j]);
	if (type == 1) return static_cast<float>(reinterpret_cast<const uint16_t *>(row)[j]);
	if (type == 2) return half_to_float(reinterpret_cast<const uint16_t *>(row)[j]);
	return reinterpret_cast<const float *>(row)[j];
}

// Raw storage bits (for copies and integer shifts).
__device__ __forceinline__ unsigned load_as_uint(const char *row, unsigned j, int type)
{
	if (type == 0) return reinterpret_cast<const uint8_t *>(row)[j];
	if (type == 1 || type == 2) return reinterpret_cast<const uint16_t *>(row)[j];
	return reinterpret_cast<const uint32_t *>(row)[j];
}

__device__ __forceinline__ void store_uint(char *row, unsigned j, int type, unsigned v)
{
	if (type == 0) reinterpret_cast<uint8_t *>(row)[j] = static_cast<uint8_t>(v);
	else if (type == 1 || type == 2) reinterpret_cast<uint16_t *>(row)[j] = static_cast<uint16_t>(v);
	else reinterpret_cast<uint32_t *>(row)[j] = v;
}

// An unquantised value into a floating-point plane (HALF or FLOAT).
__device__ __forceinline__ void store_float(char *row, unsigned j, int type, float v)
{
	if (type == 2) reinterpret_cast<uint16_t *>(row)[j] = float_to_half(v);
	else reinterpret_cast<float *>(row)[j] = v;
}

// Bayer 16x16 ordered-dither offset for output pixel (i, j) of plane variant p
// (dither.cpp:137-217): index bits interleave (row xor col, col) from the top;
// variants flip horizontally, vertically, or transpose; value (t+1)/257 - 1/2.
__device__ __forceinline__ float bayer_offset(int plane, unsigned i, unsigned j)
{
	unsigned r = i & 15, c = j & 15;
	switch (plane & 3) {
	case 1: c = 15 - c; break;
	case 2: r = 15 - r; break;
	case 3: { unsigned t = r; r = c; c = t; } break;
	default: break;
	}
	const unsigned x = r ^ c;
	unsigned v = 0;
#pragma unroll
	for (unsigned b = 0; b < 4; ++b) {
		v |= ((x >> b) & 1u) << (7 - 2 * b);
		v |= ((c >> b) & 1u) << (6 - 2 * b);
	}
	return static_cast<float>(static_cast<int>(v) + 1) / static_cast<float>(255u + 2u) - 0.5f;
}

__constant__ uint8_t c_blue_noise[ZAMD_BLUE_NOISE_LEN * ZAMD_BLUE_NOISE_LEN] = { ZAMD_BLUE_NOISE_BYTES };

// Blue-noise offset for output pixel (i, j) of plane variant p (dither.cpp:219-232):
// each variant starts the 64x64 map at its own (column, row) offset.
__device__ __forceinline__ float blue_noise_offset(int plane, unsigned i, unsigned j)
{
	const unsigned variant = static_cast<unsigned>(plane) & 3;
	const unsigned hoff = variant == 0 ? 0 : variant == 1 ? 32 : variant == 2 ? 16 : 48;
	const unsigned voff = variant == 0 ? 0 : variant == 1 ? 12 : variant == 2 ? 55 : 26;
	const unsigned t = c_blue_noise[((i + voff) % ZAMD_BLUE_NOISE_LEN) * ZAMD_BLUE_NOISE_LEN + ((hoff + j) & (ZAMD_BLUE_NOISE_LEN - 1))];
	return static_cast<float>(static_cast<int>(t) + 1) / static_cast<float>(255u + 2u) - 0.5f;
}

// dither_ordered (dither.cpp:37-55) for one value already converted to float.
__device__ __forceinline__ unsigned quantize(float x, const OutXform &o, unsigned i, unsigned j)
{
	x = x * o.scale + o.offset;
	const float d = o.dither == 1 ? bayer_offset(o.plane, i, j) : o.dither == 2 ? blue_noise_offset(o.plane, i, j) : 0.0f;
	x += d;
	const float maxval = static_cast<float>(1UL << o.bits) - 1;
	x = x < 0.0f ? 0.0f : (maxval < x ? maxval : x);
	return static_cast<unsigned>(static_cast<int>(__builtin_rintf(x)));
}

// pack_pixel_u16 (resize_impl.cpp:36-42)
__device__ __forceinline__ uint16_t pack_word(int32_t acc, int32_t pixel_max)
{
	int32_t x = ((acc + (1 << 13)) >> 14) + 32768;
	x = x > pixel_max ? pixel_max : x;
	x = x < 0 ? 0 : x;
	return static_cast<uint16_t>(x);
}

template <bool WORD>
using lds_t = std::conditional_t<WORD, uint16_t, float>;

template <bool WORD>
__device__ __forceinline__ lds_t<WORD> load_in(const char *row, unsigned j, const InXform &in)
{
	if constexpr (WORD) {
		unsigned v = load_as_uint(row, j, in.type);
		if (in.op == 1) v <<= in.shift;
		return static_cast<uint16_t>(v);
	} else {
		float v = load_as_float(row, j, in.type);
		if (in.op == 2) v = v * in.scale + in.offset;
		return v;
	}
}

template <bool WORD>
__device__ __forceinline__ void store_out(char *row, unsigned j, lds_t<WORD> v, const OutXform &o, unsigned i)
{
	if (o.op == 1) {
		store_uint(row, j, o.type, quantize(static_cast<float>(v), o, i, j));
	} else if constexpr (WORD) {
		reinterpret_cast<uint16_t *>(row)[j] = v;
	} else {
		store_float(row, j, o.type, v);
	}
}

// One output of a resize pass: taps k < fw over consecutive inputs.
template <bool WORD>
__device__ __forceinline__ lds_t<WORD> taps_contig(const lds_t<WORD> *x, const PassArg &p, unsigned row, int pixel_max)
{
	if constexpr (WORD) {
		const int16_t *c = p.ci + static_cast<size_t>(row) * p.fw;
		int32_t acc = 0;
		for (int k = 0; k < p.fw; ++k)
			acc += static_cast<int32_t>(c[k]) * (static_cast<int32_t>(x[k]) - 32768);
		return pack_word(acc, pixel_max);
	} else {
		const float *c = p.cf + static_cast<size_t>(row) * p.fw;
		float acc = 0;
		for (int k = 0; k < p.fw; ++k)
			acc += c[k] * x[k];
		return acc;
	}
}

// Same over a strided column (vertical taps).
template <bool WORD>
__device__ __forceinline__ lds_t<WORD> taps_strided(const lds_t<WORD> *x, unsigned xstride, const PassArg &p, unsigned row, int pixel_max)
{
	if constexpr (WORD) {
		const int16_t *c = p.ci + static_cast<size_t>(row) * p.fw;
		int32_t acc = 0;
		for (int k = 0; k < p.fw; ++k)
			acc += static_cast<int32_t>(c[k]) * (static_cast<int32_t>(x[static_cast<size_t>(k) * xstride]) - 32768);
		return pack_word(acc, pixel_max);
	} else {
		const float *c = p.cf + static_cast<size_t>(row) * p.fw;
		float acc = 0;
		for (int k = 0; k < p.fw; ++k)
			acc += c[k] * x[static_cast<size_t>(k) * xstride];
		return acc;
	}
}

// Fused resize of one output tile (tile_h rows x tile_w columns) of one frame.
//  H first: stage source rows [r0, r1) x cols [c0, c1) in LDS, horizontal pass
//           into the intermediate tile, vertical pass into the output.
//  V first: vertical pass from HBM into the intermediate tile (columns the
//           horizontal pass needs), horizontal pass into the output.
// The intermediate keeps the reference's storage type (WORD rounded and
// clamped, or FLOAT), so fusing does not change a single bit.
template <bool WORD>
__global__ void __launch_bounds__(kThreads) k_resize(ResizeArg a)
{
	using T = lds_t<WORD>;
	extern __shared__ __align__(16) unsigned char smem[];
	T *in_t = reinterpret_cast<T *>(smem);
	T *mid_t = in_t + static_cast<size_t>(a.lds_in_rows) * a.lds_in_cols;

	const unsigned tid = threadIdx.x;
	const unsigned f = blockIdx.z;
	const unsigned j0 = blockIdx.x * a.tile_w, i0 = blockIdx.y * a.tile_h;
	if (j0 >= a.dst_w || i0 >= a.dst_h)
		return;
	const unsigned ncols = min(a.tile_w, a.dst_w - j0), nrows = min(a.tile_h, a.dst_h - i0);
	const char *src = frame_ptr(a.src, f);
	char *dst = frame_ptr_w(a.dst, f);
	const unsigned mstride = a.lds_mid_cols;

	if (a.h_first && a.lds_in_rows == 0) {
		// Single horizontal pass with unsorted offsets: taps straight from HBM.
		const PassArg &ph = a.p1;
		for (unsigned idx = tid; idx < nrows * ncols; idx += kThreads) {
			const unsigned i = idx / ncols, c = idx - i * ncols;
			const unsigned j = j0 + c;
			const char *srow = row_ptr(src, a.src.stride, i0 + i);
			T x[64];
			const int fw = ph.fw < 64 ? ph.fw : 64;
			for (int k = 0; k < fw; ++k)
				x[k] = load_in<WORD>(srow, ph.left[j] + k, a.in);
			const T v = taps_contig<WORD>(x, ph, j, a.pixel_max);
			store_out<WORD>(row_ptr(dst, a.dst.stride, i0 + i), j, v, a.out, i0 + i);
		}
	} else if (a.h_first) {
		const PassArg &ph = a.p1;
		unsigned r0, r1;
		if (a.passes == 2) {
			r0 = a.p2.left[i0];
			r1 = a.p2.left[i0 + nrows - 1] + a.p2.fw;
		} else {
			r0 = i0;
			r1 = i0 + nrows;
		}
		const unsigned c0 = ph.left[j0], c1 = ph.left[j0 + ncols - 1] + ph.fw;
		const unsigned w_in = c1 - c0, h_in = r1 - r0;
		const unsigned istride = a.lds_in_cols;

		for (unsigned idx = tid; idx < w_in * h_in; idx += kThreads) {
			const unsigned r = idx / w_in, c = idx - r * w_in;
			in_t[r * istride + c] = load_in<WORD>(row_ptr(src, a.src.stride, r0 + r), c0 + c, a.in);
		}
		__syncthreads();

		for (unsigned idx = tid; idx < h_in * ncols; idx += kThreads) {
			const unsigned r = idx / ncols, c = idx - r * ncols;
			const unsigned j = j0 + c;
			const T v = taps_contig<WORD>(in_t + r * istride + (ph.left[j] - c0), ph, j, a.pixel_max);
			if (a.passes == 2)
				mid_t[r * mstride + c] = v;
			else
				store_out<WORD>(row_ptr(dst, a.dst.stride, r0 + r), j, v, a.out, r0 + r);
		}
		if (a.passes == 2) {
			__syncthreads();
			const PassArg &pv = a.p2;
			for (unsigned idx = tid; idx < nrows * ncols; idx += kThreads) {
				const unsigned i = idx / ncols, c = idx - i * ncols;
				const unsigned top = pv.left[i0 + i] - r0;
				const T v = taps_strided<WORD>(mid_t + top * mstride + c, mstride, pv, i0 + i, a.pixel_max);
				store_out<WORD>(row_ptr(dst, a.dst.stride, i0 + i), j0 + c, v, a.out, i0 + i);
			}
		}
	} else {
		const PassArg &pv = a.p1;
		unsigned c0, c1;
		if (a.passes == 2) {
			c0 = a.p2.left[j0];
			c1 = a.p2.left[j0 + ncols - 1] + a.p2.fw;
		} else {
			c0 = j0;
			c1 = j0 + ncols;
		}
		const unsigned w_mid = c1 - c0;

		for (unsigned idx = tid; idx < nrows * w_mid; idx += kThreads) {
			const unsigned i = idx / w_mid, c = idx - i * w_mid;
			const unsigned top = pv.left[i0 + i];
			T v;
			if constexpr (WORD) {
				const int16_t *cf = pv.ci + static_cast<size_t>(i0 + i) * pv.fw;
				int32_t acc = 0;
				for (int k = 0; k < pv.fw; ++k)
					acc += static_cast<int32_t>(cf[k]) *
					       (static_cast<int32_t>(load_in<true>(row_ptr(src, a.src.stride, top + k), c0 + c, a.in)) - 32768);
				v = pack_word(acc, a.pixel_max);
			} else {
				const float *cf = pv.cf + static_cast<size_t>(i0 + i) * pv.fw;
				float acc = 0;
				for (int k = 0; k < pv.fw; ++k)
					acc += cf[k] * load_in<false>(row_ptr(src, a.src.stride, top + k), c0 + c, a.in);
				v = acc;
			}
			if (a.passes == 2)
				mid_t[i * mstride + c] = v;
			else
				store_out<WORD>(row_ptr(dst, a.dst.stride, i0 + i), c0 + c, v, a.out, i0 + i);
		}
		if (a.passes == 2) {
			__syncthreads();
			const PassArg &ph = a.p2;
			for (unsigned idx = tid; idx < nrows * ncols; idx += kThreads) {
				const unsigned i = idx / ncols, c = idx - i * ncols;
				const unsigned j = j0 + c;
				const T v = taps_contig<WORD>(mid_t + i * mstride + (ph.left[j] - c0), ph, j, a.pixel_max);
				store_out<WORD>(row_ptr(dst, a.dst.stride, i0 + i), j, v, a.out, i0 + i);
			}
		}
	}
}

// ---------------------------------------------------------------------------
// Fast fused resize. Same arithmetic as k_resize, organised for throughput:
// each thread owns two adjacent output columns for the whole tile, so the
// horizontal taps of both columns live in registers and are reused on every
// row; vertical taps are uniform per output row (scalar loads); LDS and HBM
// are touched in column pairs (4 or 8 bytes per lane). Loops carry no integer
// division. Requires sorted left offsets, <= FWH horizontal taps and planes
// whose base pointers and row strides are multiples of 16 bytes.
// ---------------------------------------------------------------------------

template <bool WORD>
struct Pair;
template <>
struct Pair<true> {
	uint16_t a, b;
};
template <>
struct Pair<false> {
	float a, b;
};

template <bool WORD>
__device__ __forceinline__ Pair<WORD> load_pair(const char *row, unsigned c, unsigned limit, const InXform &in)
{
	// Columns c (even) and c + 1; the second is read only when c + 1 < limit.
	Pair<WORD> p;
	const bool two = c + 1 < limit;
	if constexpr (WORD) {
		unsigned x0, x1 = 0;
		if (in.type == 0) {
			if (two) {
				const uint16_t v = *reinterpret_cast<const uint16_t *>(row + c);
				x0 = v & 0xFF;
				x1 = v >> 8;
			} else {
				x0 = reinterpret_cast<const uint8_t *>(row)[c];
			}
		} else {
			if (two) {
				const uint32_t v = *reinterpret_cast<const uint32_t *>(row + 2 * static_cast<size_t>(c));
				x0 = v & 0xFFFF;
				x1 = v >> 16;
			} else {
				x0 = reinterpret_cast<const uint16_t *>(row)[c];
			}
		}
		if (in.op == 1) {
			x0 <<= in.shift;
			x1 <<= in.shift;
		}
		p.a = static_cast<uint16_t>(x0);
		p.b = static_cast<uint16_t>(x1);
	} else {
		float x0, x1 = 0.0f;
		if (in.type == 3) {
			if (two) {
				const float2 v = *reinterpret_cast<const float2 *>(row + 4 * static_cast<size_t>(c));
				x0 = v.x;
				x1 = v.y;
			} else {
				x0 = reinterpret_cast<const float *>(row)[c];
			}
		} else {
			x0 = load_as_float(row, c, in.type);
			if (two)
				x1 = load_as_float(row, c + 1, in.type);
		}
		if (in.op == 2) {
			x0 = x0 * in.scale + in.offset;
			x1 = x1 * in.scale + in.offset;
		}
		p.a = x0;
		p.b = x1;
	}
	return p;
}

template <bool WORD>
__device__ __forceinline__ void store_pair(char *row, unsigned j, lds_t<WORD> v0, lds_t<WORD> v1, bool two, const OutXform &o, unsigned i)
{
	if (o.op == 1) {
		const unsigned q0 = quantize(static_cast<float>(v0), o, i, j);
		if (!two) {
			store_uint(row, j, o.type, q0);
			return;
		}
		const unsigned q1 = quantize(static_cast<float>(v1), o, i, j + 1);
		if (o.type == 0)
			*reinterpret_cast<uint16_t *>(row + j) = static_cast<uint16_t>(q0 | (q1 << 8));
		else
			*reinterpret_cast<uint32_t *>(row + 2 * static_cast<size_t>(j)) = q0 | (q1 << 16);
		return;
	}
	if constexpr (WORD) {
		if (two)
			*reinterpret_cast<uint32_t *>(row + 2 * static_cast<size_t>(j)) = static_cast<uint32_t>(v0) | (static_cast<uint32_t>(v1) << 16);
		else
			reinterpret_cast<uint16_t *>(row)[j] = v0;
	} else if (o.type == 2) {
		if (two)
			*reinterpret_cast<uint32_t *>(row + 2 * static_cast<size_t>(j)) = static_cast<uint32_t>(float_to_half(v0)) | (static_cast<uint32_t>(float_to_half(v1)) << 16);
		else
			reinterpret_cast<uint16_t *>(row)[j] = float_to_half(v0);
	} else {
		if (two)
			*reinterpret_cast<float2 *>(row + 4 * static_cast<size_t>(j)) = make_float2(v0, v1);
		else
			reinterpret_cast<float *>(row)[j] = v0;
	}
}

// Per-thread horizontal taps of two columns held in registers.
template <bool WORD, int FWH>
struct HTaps {
	using C = std::conditional_t<WORD, int32_t, float>;
	C c0[FWH], c1[FWH];
	int l0, l1;

	__device__ __forceinline__ void load(const PassArg &p, unsigned j, bool two, int base)
	{
		const unsigned j1 = two ? j + 1 : j;
		l0 = static_cast<int>(p.left[j]) - base;
		l1 = static_cast<int>(p.left[j1]) - base;
#pragma unroll
		for (int k = 0; k < FWH; ++k) {
			if (k < p.fw) {
				if constexpr (WORD) {
					c0[k] = p.ci[static_cast<size_t>(j) * p.fw + k];
					c1[k] = p.ci[static_cast<size_t>(j1) * p.fw + k];
				} else {
					c0[k] = p.cf[static_cast<size_t>(j) * p.fw + k];
					c1[k] = p.cf[static_cast<size_t>(j1) * p.fw + k];
				}
			}
		}
	}

	__device__ __forceinline__ void apply(const lds_t<WORD> *x, int fw, int pixel_max, lds_t<WORD> &v0, lds_t<WORD> &v1) const
	{
		if constexpr (WORD) {
			int32_t a0 = 0, a1 = 0;
#pragma unroll
			for (int k = 0; k < FWH; ++k) {
				if (k < fw) {
					a0 += c0[k] * (static_cast<int32_t>(x[l0 + k]) - 32768);
					a1 += c1[k] * (static_cast<int32_t>(x[l1 + k]) - 32768);
				}
			}
			v0 = pack_word(a0, pixel_max);
			v1 = pack_word(a1, pixel_max);
		} else {
			float a0 = 0, a1 = 0;
#pragma unroll
			for (int k = 0; k < FWH; ++k) {
				if (k < fw) {
					a0 += c0[k] * x[l0 + k];
					a1 += c1[k] * x[l1 + k];
				}
			}
			v0 = a0;
			v1 = a1;
		}
	}
};

template <bool WORD>
__device__ __forceinline__ void v_taps_pair(const lds_t<WORD> *x, unsigned pitch, const PassArg &p, unsigned row, int pixel_max,
                                            lds_t<WORD> &v0, lds_t<WORD> &v1)
{
	if constexpr (WORD) {
		const int16_t *c = p.ci + static_cast<size_t>(row) * p.fw;
		int32_t a0 = 0, a1 = 0;
		for (int k = 0; k < p.fw; ++k) {
			const uint32_t pr = *reinterpret_cast<const uint32_t *>(x + static_cast<size_t>(k) * pitch);
			const int32_t ck = c[k];
			a0 += ck * (static_cast<int32_t>(pr & 0xFFFF) - 32768);
			a1 += ck * (static_cast<int32_t>(pr >> 16) - 32768);
		}
		v0 = pack_word(a0, pixel_max);
		v1 = pack_word(a1, pixel_max);
	} else {
		const float *c = p.cf + static_cast<size_t>(row) * p.fw;
		float a0 = 0, a1 = 0;
		for (int k = 0; k < p.fw; ++k) {
			const float2 pr = *reinterpret_cast<const float2 *>(x + static_cast<size_t>(k) * pitch);
			a0 += c[k] * pr.x;
			a1 += c[k] * pr.y;
		}
		v0 = a0;
		v1 = a1;
	}
}

template <bool WORD, int FWH>
__global__ void __launch_bounds__(kThreads) k_resize_fast(ResizeArg a)
{
	using T = lds_t<WORD>;
	extern __shared__ __align__(16) unsigned char smem[];
	T *xin = reinterpret_cast<T *>(smem);
	T *mid = xin + static_cast<size_t>(a.lds_in_rows) * a.lds_in_cols;

	const unsigned t = threadIdx.x;
	const unsigned f = blockIdx.z;
	const unsigned j0 = blockIdx.x * a.tile_w, i0 = blockIdx.y * a.tile_h;
	const unsigned nrows = min(a.tile_h, a.dst_h - i0);
	const unsigned ncols = min(a.tile_w, a.dst_w - j0);
	const char *src = frame_ptr(a.src, f);
	char *dst = frame_ptr_w(a.dst, f);
	const unsigned jc = 2 * t;            // this thread's column pair inside the tile
	const bool active = jc < ncols;
	const bool two = jc + 1 < ncols;
	const unsigned j = j0 + (active ? jc : 0);

	if (a.h_first) {
		const PassArg &ph = a.p1;
		const unsigned r0 = a.passes == 2 ? a.p2.left[i0] : i0;
		const unsigned r1 = a.passes == 2 ? a.p2.left[i0 + nrows - 1] + a.p2.fw : i0 + nrows;
		const unsigned c0 = ph.left[j0] & ~1u;
		const unsigned c1 = ph.left[j0 + ncols - 1] + ph.fw;
		const unsigned pin = a.lds_in_cols;
		const unsigned npairs = (c1 - c0 + 1) / 2;

		// Stage source rows [r0, r1) x cols [c0, c1) (c0 even) in LDS, in column pairs.
		for (unsigned r = 0; r < r1 - r0; ++r) {
			const char *srow = row_ptr(src, a.src.stride, r0 + r);
			for (unsigned q = t; q < npairs; q += kThreads) {
				const Pair<WORD> p = load_pair<WORD>(srow, c0 + 2 * q, a.src_w, a.in);
				T *d = xin + r * pin + 2 * q;
				d[0] = p.a;
				d[1] = p.b;
			}
		}
		__syncthreads();

		HTaps<WORD, FWH> taps;
		taps.load(ph, j, two, static_cast<int>(c0));
		const unsigned span = r1 - r0;
		const unsigned pm = a.lds_mid_cols;
		for (unsigned r = 0; r < span; ++r) {
			T v0, v1;
			taps.apply(xin + r * pin, ph.fw, a.pixel_max, v0, v1);
			if (!active)
				continue;
			if (a.passes == 2) {
				mid[r * pm + jc] = v0;
				mid[r * pm + jc + 1] = v1;
			} else {
				store_pair<WORD>(row_ptr(dst, a.dst.stride, r0 + r), j, v0, v1, two, a.out, r0 + r);
			}
		}
		if (a.passes == 2) {
			__syncthreads();
			const PassArg &pv = a.p2;
			for (unsigned i = 0; i < nrows; ++i) {
				const unsigned top = pv.left[i0 + i] - r0;
				T v0, v1;
				v_taps_pair<WORD>(mid + top * pm + jc, pm, pv, i0 + i, a.pixel_max, v0, v1);
				if (active)
					store_pair<WORD>(row_ptr(dst, a.dst.stride, i0 + i), j, v0, v1, two, a.out, i0 + i);
			}
		}
	} else if (a.passes == 1) {
		// Vertical pass only: column pairs straight from HBM to HBM.
		const PassArg &pv = a.p1;
		for (unsigned i = 0; i < nrows; ++i) {
			const unsigned top = pv.left[i0 + i];
			T v0, v1;
			if constexpr (WORD) {
				const int16_t *cv = pv.ci + static_cast<size_t>(i0 + i) * pv.fw;
				int32_t a0 = 0, a1 = 0;
				for (int k = 0; k < pv.fw; ++k) {
					const Pair<true> p = load_pair<true>(row_ptr(src, a.src.stride, top + k), j, j0 + ncols, a.in);
					a0 += static_cast<int32_t>(cv[k]) * (static_cast<int32_t>(p.a) - 32768);
					a1 += static_cast<int32_t>(cv[k]) * (static_cast<int32_t>(p.b) - 32768);
				}
				v0 = pack_word(a0, a.pixel_max);
				v1 = pack_word(a1, a.pixel_max);
			} else {
				const float *cv = pv.cf + static_cast<size_t>(i0 + i) * pv.fw;
				float a0 = 0, a1 = 0;
				for (int k = 0; k < pv.fw; ++k) {
					const Pair<false> p = load_pair<false>(row_ptr(src, a.src.stride, top + k), j, j0 + ncols, a.in);
					a0 += cv[k] * p.a;
					a1 += cv[k] * p.b;
				}
				v0 = a0;
				v1 = a1;
			}
			if (active)
				store_pair<WORD>(row_ptr(dst, a.dst.stride, i0 + i), j, v0, v1, two, a.out, i0 + i);
		}
	} else {
		// V first: vertical pass from HBM into the intermediate tile (the source
		// columns the horizontal pass needs), then horizontal pass to HBM.
		const PassArg &pv = a.p1;
		const PassArg &ph = a.p2;
		const unsigned c0 = ph.left[j0] & ~1u;
		const unsigned c1 = ph.left[j0 + ncols - 1] + ph.fw;
		const unsigned pm = a.lds_mid_cols;
		const unsigned npairs = (c1 - c0 + 1) / 2;
		for (unsigned q = t; q < npairs; q += kThreads) {
			const unsigned c = c0 + 2 * q;
			for (unsigned i = 0; i < nrows; ++i) {
				const unsigned top = pv.left[i0 + i];
				T v0, v1;
				if constexpr (WORD) {
					const int16_t *cv = pv.ci + static_cast<size_t>(i0 + i) * pv.fw;
					int32_t a0 = 0, a1 = 0;
					for (int k = 0; k < pv.fw; ++k) {
						const Pair<true> p = load_pair<true>(row_ptr(src, a.src.stride, top + k), c, a.src_w, a.in);
						a0 += static_cast<int32_t>(cv[k]) * (static_cast<int32_t>(p.a) - 32768);
						a1 += static_cast<int32_t>(cv[k]) * (static_cast<int32_t>(p.b) - 32768);
					}
					v0 = pack_word(a0, a.pixel_max);
					v1 = pack_word(a1, a.pixel_max);
				} else {
					const float *cv = pv.cf + static_cast<size_t>(i0 + i) * pv.fw;
					float a0 = 0, a1 = 0;
					for (int k = 0; k < pv.fw; ++k) {
						const Pair<false> p = load_pair<false>(row_ptr(src, a.src.stride, top + k), c, a.src_w, a.in);
						a0 += cv[k] * p.a;
						a1 += cv[k] * p.b;
					}
					v0 = a0;
					v1 = a1;
				}
				mid[i * pm + 2 * q] = v0;
				mid[i * pm + 2 * q + 1] = v1;
			}
		}
		__syncthreads();
		HTaps<WORD, FWH> taps;
		taps.load(ph, j, two, static_cast<int>(c0));
		for (unsigned i = 0; i < nrows; ++i) {
			T v0, v1;
			taps.apply(mid + i * pm, ph.fw, a.pixel_max, v0, v1);
			if (active)
				store_pair<WORD>(row_ptr(dst, a.dst.stride, i0 + i), j, v0, v1, two, a.out, i0 + i);
		}
	}
}

// Point stages: depth conversion, ordered dither, sub-rectangle copy, fill.
__global__ void __launch_bounds__(kThreads) k_point(PointArg a)
{
	const unsigned f = blockIdx.z;
	const unsigned i = blockIdx.y;
	const unsigned j = blockIdx.x * kThreads + threadIdx.x;
	if (j >= a.width || i >= a.height)
		return;
	char *drow = row_ptr(frame_ptr_w(a.dst, f), a.dst.stride, i);
	if (a.fill) {
		store_uint(drow, j, a.out.type, a.fill_bits);
		return;
	}
	const char *srow = row_ptr(frame_ptr(a.src, f), a.src.stride, i + a.crop_top);
	const unsigned js = j + a.crop_left;
	if (a.in.op == 1) {
		store_uint(drow, j, a.out.type, load_as_uint(srow, js, a.in.type) << a.in.shift);
		return;
	}
	if (a.in.op == 0 && a.out.op == 0) {
		store_uint(drow, j, a.out.type, load_as_uint(srow, js, a.in.type));
		return;
	}
	float x = load_as_float(srow, js, a.in.type);
	if (a.in.op == 2)
		x = x * a.in.scale + a.in.offset;
	if (a.out.op == 1)
		store_uint(drow, j, a.out.type, quantize(x, a.out, i, j));
	else
		store_float(drow, j, a.out.type, x);
}

// ColorspaceConversionImpl::process (colorspace.cpp:45-77): up to six in-place
// operations on three float channels, fused with the depth conversions before
// and after it.
__device__ __forceinline__ void apply_op(const CspOp &op, float &a, float &b, float &c)
{
	switch (op.kind) {
	case CSP_MATRIX: {
		const float x = op.m[0] * a + op.m[1] * b + op.m[2] * c;
		const float y = op.m[3] * a + op.m[4] * b + op.m[5] * c;
		const float z = op.m[6] * a + op.m[7] * b + op.m[8] * c;
		a = x; b = y; c = z;
		break;
	}
	case CSP_GAMMA:
		a = op.post * zm::transfer(op.func, a * op.pre);
		b = op.post * zm::transfer(op.func, b * op.pre);
		c = op.post * zm::transfer(op.func, c * op.pre);
		break;
	case CSP_B67: {
		const float gamma = 1.2f;
		const float r = a * op.pre, g = b * op.pre, bl = c * op.pre;
		const float yd = zm::maxf(op.kr * r + op.kg * g + op.kb * bl, zm::FLT_MIN_);
		const float ys_inv = zm::powf_(yd, (1.0f - gamma) / gamma);
		a = zm::arib_b67_oetf(r * ys_inv);
		b = zm::arib_b67_oetf(g * ys_inv);
		c = zm::arib_b67_oetf(bl * ys_inv);
		break;
	}
	case CSP_B67_INV: {
		const float gamma = 1.2f;
		float r = zm::arib_b67_inverse_oetf(a), g = zm::arib_b67_inverse_oetf(b), bl = zm::arib_b67_inverse_oetf(c);
		float ys = zm::maxf(op.kr * r + op.kg * g + op.kb * bl, zm::FLT_MIN_);
		ys = zm::powf_(ys, gamma - 1.0f);
		r = r * ys; g = g * ys; bl = bl * ys;
		a = r * op.post; b = g * op.post; c = bl * op.post;
		break;
	}
	case CSP_CL_TO_RGB: {
		const float y = a, u = b, v = c;
		const float bmy = u < 0 ? u * 2.0f * op.nb : u * 2.0f * op.pb;
		const float rmy = v < 0 ? v * 2.0f * op.nr : v * 2.0f * op.pr;
		const float bl = zm::transfer(op.func, bmy + y);
		const float r = zm::transfer(op.func, rmy + y);
		const float yl = zm::transfer(op.func, y);
		const float g = (yl - op.kr * r - op.kb * bl) / op.kg;
		a = r * op.post; b = g * op.post; c = bl * op.post;
		break;
	}
	case CSP_CL_TO_YUV: {
		float r = a * op.pre, g = b * op.pre, bl = c * op.pre;
		const float y = zm::transfer(op.func, op.kr * r + op.kg * g + op.kb * bl);
		bl = zm::transfer(op.func, bl);
		r = zm::transfer(op.func, r);
		const float u = (bl - y < 0.0f) ? (bl - y) / (2.0f * op.nb) : (bl - y) / (2.0f * op.pb);
		const float v = (r - y < 0.0f) ? (r - y) / (2.0f * op.nr) : (r - y) / (2.0f * op.pr);
		a = y; b = u; c = v;
		break;
	}
	default:
		break;
	}
}

__global__ void __launch_bounds__(kThreads) k_colorspace(CspArg a)
{
	const unsigned f = blockIdx.z;
	const unsigned i = blockIdx.y;
	const unsigned j = blockIdx.x * kThreads + threadIdx.x;
	if (j >= a.width || i >= a.height)
		return;
	float v[3];
#pragma unroll
	for (int p = 0; p < 3; ++p) {
		const char *srow = row_ptr(frame_ptr(a.src[p], f), a.src[p].stride, i);
		float x = load_as_float(srow, j, a.in[p].type);
		if (a.in[p].op == 2)
			x = x * a.in[p].scale + a.in[p].offset;
		v[p] = x;
	}
	for (int k = 0; k < a.prog.n; ++k)
		apply_op(a.prog.op[k], v[0], v[1], v[2]);
#pragma unroll
	for (int p = 0; p < 3; ++p) {
		char *drow = row_ptr(frame_ptr_w(a.dst[p], f), a.dst[p].stride, i);
		if (a.out[p].op == 1)
			store_uint(drow, j, a.out[p].type, quantize(v[p], a.out[p], i, j));
		else
			store_float(drow, j, a.out[p].type, v[p]);
	}
}

// Floyd-Steinberg error diffusion (dither.cpp:57-89, 286-349) as a row-skewed
// wavefront. Lane l of wave w owns row band_base + 64 w + l and handles column
// j = T - 2 l - w D at step T, so the three errors it needs from the row above
// were produced by lane l-1 one, two and three steps earlier and arrive by a
// one-lane shuffle. Across waves the boundary row goes through an LDS ring;
// waves are offset by D = 2*63 + 2 + kEdChunk columns and synchronise every
// kEdChunk steps, so every ring entry is written a full chunk before it is
// read. Across bands the last row is kept in an LDS row buffer. The float
// operation order equals the sequential reference, hence bit-exact output.
constexpr int kEdWaves = 16;
constexpr int kEdChunk = 16;
constexpr int kEdDelay = 2 * 63 + 2 + kEdChunk;
constexpr int kEdRing = 256;

__device__ __forceinline__ float ed_load(const char *srow, int j, int W, int type, bool row_ok)
{
	return (row_ok && j >= 0 && j < W) ? load_as_float(srow, static_cast<unsigned>(j), type) : 0.0f;
}

__global__ void __launch_bounds__(kEdWaves * 64) k_error_diffusion(EdArg a)
{
	extern __shared__ __align__(16) float ed_smem[];
	float *ring = ed_smem;                      // [kEdWaves][kEdRing]
	float *prev_row = ring + kEdWaves * kEdRing; // [width + 2], padded row of the previous band's last row

	// blockIdx.x: frame, blockIdx.y: plane of this launch.
	const unsigned f = blockIdx.x;
	const unsigned pl = blockIdx.y;
	const PlaneArg &sp = a.src[pl];
	const PlaneArg &dp = a.dst[pl];
	const int lane = threadIdx.x & 63;
	const int w = threadIdx.x >> 6;
	const int W = static_cast<int>(a.width);
	const char *src = frame_ptr(sp, f);
	char *dst = frame_ptr_w(dp, f);
	const float maxval = static_cast<float>(1UL << a.out.bits) - 1;
	const float scale = a.scale[pl], offset = a.offset[pl];

	for (int t = threadIdx.x; t < W + 2; t += kEdWaves * 64)
		prev_row[t] = 0.0f;
	__syncthreads();

	const int band_rows = kEdWaves * 64;
	const int steps = W + 2 * 63 + (kEdWaves - 1) * kEdDelay + 1;
	const int nchunks = (steps + kEdChunk - 1) / kEdChunk;

	for (int base = 0; base < static_cast<int>(a.height); base += band_rows) {
		const int r = base + w * 64 + lane;
		const bool row_ok = r < static_cast<int>(a.height);
		const char *srow = row_ptr(src, sp.stride, row_ok ? r : 0);
		char *drow = row_ptr(dst, dp.stride, row_ok ? r : 0);
		float e_left = 0.0f, t_m1 = 0.0f, t_0 = 0.0f, t_p1 = 0.0f, prod = 0.0f;

		// Inputs of a chunk are fetched one chunk ahead, so the HBM latency of
		// this lane's row is off the step-to-step critical path.
		const int jbase = -2 * lane - w * kEdDelay;
		float xin[kEdChunk], xnext[kEdChunk];
#pragma unroll
		for (int s = 0; s < kEdChunk; ++s)
			xnext[s] = ed_load(srow, jbase + s, W, a.in.type, row_ok);

		for (int ch = 0; ch < nchunks; ++ch) {
#pragma unroll
			for (int s = 0; s < kEdChunk; ++s)
				xin[s] = xnext[s];
#pragma unroll
			for (int s = 0; s < kEdChunk; ++s)
				xnext[s] = ed_load(srow, jbase + (ch + 1) * kEdChunk + s, W, a.in.type, row_ok);
#pragma unroll
			for (int s = 0; s < kEdChunk; ++s) {
				const int T = ch * kEdChunk + s;
				const int j = T - 2 * lane - w * kEdDelay;
				const float up = __shfl_up(prod, 1);
				float recv;
				if (lane == 0) {
					const int jn = j + 1;
					if (jn < 0 || jn >= W)
						recv = 0.0f;
					else if (w == 0)
						recv = prev_row[jn + 1];
					else
						recv = ring[(w - 1) * kEdRing + (jn & (kEdRing - 1))];
				} else {
					recv = up;
				}
				t_m1 = t_0;
				t_0 = t_p1;
				t_p1 = recv;
				prod = 0.0f;
				if (row_ok && j >= 0 && j < W) {
					float x = xin[s] * scale + offset;
					float err = 0;
					err += e_left * (7.0f / 16.0f);
					err += t_p1 * (3.0f / 16.0f);
					err += t_0 * (5.0f / 16.0f);
					err += t_m1 * (1.0f / 16.0f);
					x += err;
					x = x < 0.0f ? 0.0f : (maxval < x ? maxval : x);
					const float q = __builtin_rintf(x);
					store_uint(drow, static_cast<unsigned>(j), a.out.type, static_cast<unsigned>(static_cast<int>(q)));
					const float e = x - q;
					e_left = e;
					prod = e;
					if (lane == 63) {
						if (w < kEdWaves - 1)
							ring[w * kEdRing + (j & (kEdRing - 1))] = e;
						else
							prev_row[j + 1] = e;
					}
				}
			}
			__syncthreads();
		}
		// A band whose last rows are past the image leaves zeros for nobody; the
		// next band (if any) always starts with a complete previous row.
		__syncthreads();
	}
}

template <bool WORD>
int launch_resize_t(const ResizeArg &a, void *stream)
{
	dim3 grid((a.dst_w + a.tile_w - 1) / a.tile_w, (a.dst_h + a.tile_h - 1) / a.tile_h, a.frames);
	const size_t lds = resize_lds_bytes(a);
	hipStream_t s = static_cast<hipStream_t>(stream);
	switch (a.fwh) {
	case 0: hipLaunchKernelGGL(k_resize<WORD>, grid, dim3(kThreads), lds, s, a); break;
	case 4: hipLaunchKernelGGL((k_resize_fast<WORD, 4>), grid, dim3(kThreads), lds, s, a); break;
	case 8: hipLaunchKernelGGL((k_resize_fast<WORD, 8>), grid, dim3(kThreads), lds, s, a); break;
	case 12: hipLaunchKernelGGL((k_resize_fast<WORD, 12>), grid, dim3(kThreads), lds, s, a); break;
	case 16: hipLaunchKernelGGL((k_resize_fast<WORD, 16>), grid, dim3(kThreads), lds, s, a); break;
	case 24: hipLaunchKernelGGL((k_resize_fast<WORD, 24>), grid, dim3(kThreads), lds, s, a); break;
	case 32: hipLaunchKernelGGL((k_resize_fast<WORD, 32>), grid, dim3(kThreads), lds, s, a); break;
	default: return static_cast<int>(hipErrorInvalidValue);
	}
	return static_cast<int>(hipGetLastError());
}

} // namespace

size_t resize_lds_bytes(const ResizeArg &a)
{
	const size_t elem = a.word ? 2 : 4;
	return elem * (static_cast<size_t>(a.lds_in_rows) * a.lds_in_cols + static_cast<size_t>(a.lds_mid_rows) * a.lds_mid_cols);
}

int launch_resize(const ResizeArg &a, void *stream)
{
	return a.word ? launch_resize_t<true>(a, stream) : launch_resize_t<false>(a, stream);
}

int launch_point(const PointArg &a, void *stream)
{
	dim3 grid((a.width + kThreads - 1) / kThreads, a.height, a.frames);
	hipLaunchKernelGGL(k_point, grid, dim3(kThreads), 0, static_cast<hipStream_t>(stream), a);
	return static_cast<int>(hipGetLastError());
}

int launch_colorspace(const CspArg &a, void *stream)
{
	dim3 grid((a.width + kThreads - 1) / kThreads, a.height, a.frames);
	hipLaunchKernelGGL(k_colorspace, grid, dim3(kThreads), 0, static_cast<hipStream_t>(stream), a);
	return static_cast<int>(hipGetLastError());
}

int launch_error_diffusion(const EdArg &a, void *stream)
{
	const size_t lds = sizeof(float) * (static_cast<size_t>(kEdWaves) * kEdRing + a.width + 2);
	hipLaunchKernelGGL(k_error_diffusion, dim3(a.frames, a.planes), dim3(kEdWaves * 64), lds, static_cast<hipStream_t>(stream), a);
	return static_cast<int>(hipGetLastError());
}

} // namespace dev
} // namespace zamd
