// gfx950 kernels of the per-plane path: fused two-pass resize, point depth
// conversions, unresize and the error-diffusion wavefront (the colourspace
// stage is in csp.hip).
//
// Arithmetic is the reference's portable C path (resize_impl.cpp:31-114,
// depth_convert.cpp:23-64, dither.cpp:37-89, operation_impl.cpp:21-285),
// evaluated in the same order with FP contraction disabled (-ffp-contract=off),
// IEEE division/sqrt and denormals kept, so outputs are bit-identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <type_traits>

#include "devutil.h"
#include "kernels.h"

namespace zamd {
namespace dev {

namespace {

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

// ---------------------------------------------------------------------------
// Tiled resize. Same arithmetic as k_resize, organised for throughput:
//  * the source rectangle a tile needs is staged in LDS with 16-byte loads,
//    several in flight per thread, with the load transform applied once;
//  * WORD samples are kept as x - 32768 in int16 and the Q1.14 taps are
//    applied in pairs with v_dot2_i32_i16 (exact: the int32 sum of the
//    products is the reference's accumulator, resize_impl.cpp:44-95), using
//    tables whose rows start on an even sample and carry a zero tap where
//    the real window starts on an odd one;
//  * horizontal taps of a thread's two columns live in registers for the
//    whole tile; vertical taps are uniform per row (scalar loads);
//  * between two passes the intermediate stays in LDS; for WORD the
//    horizontal-first intermediate interleaves row pairs so that the
//    vertical pass also runs on dot2.
// Requires sorted left offsets and planes whose base pointers and row
// strides are multiples of 16 bytes.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int32_t dot2(uint32_t x, int32_t c, int32_t acc)
{
	typedef short v2i16 __attribute__((ext_vector_type(2)));
	return __builtin_amdgcn_sdot2(__builtin_bit_cast(v2i16, x), __builtin_bit_cast(v2i16, c), acc, false);
}

// pack_word minus 32768: the biased int16 the next pass consumes (upper
// half of the result unspecified).
__device__ __forceinline__ uint32_t pack_biased(int32_t acc, int32_t pixel_max)
{
	return static_cast<uint32_t>(med3_i32((acc + (1 << 13)) >> 14, -32768, pixel_max - 32768));
}

// Accumulators of the tiled kernels start at the rounding terms instead of
// 0 (the int32 sum is the same, so are the results): kBiasedInit for
// pack_biased, kWordInit for pack_word; the *_pre forms then skip the add.
constexpr int32_t kBiasedInit = 1 << 13;
constexpr int32_t kWordInit = (32768 << 14) + (1 << 13);
__device__ __forceinline__ uint32_t pack_biased_pre(int32_t acc, int32_t pixel_max)
{
	return static_cast<uint32_t>(med3_i32(acc >> 14, -32768, pixel_max - 32768));
}
__device__ __forceinline__ uint16_t pack_word_pre(int32_t acc, int32_t pixel_max)
{
	return static_cast<uint16_t>(med3_i32(acc >> 14, 0, pixel_max));
}

// Shift every 16-bit half of w left by s (s < 16), then bias to int16 (x ^
// 0x8000 = x - 32768; the matrix-core kernel passes 0x80808080, which also
// turns each low byte into the signed byte low - 128).
__device__ __forceinline__ uint32_t shift_bias2(uint32_t w, unsigned s, uint32_t bias = 0x80008000u)
{
	const uint32_t m = (0xFFFFu << s) & 0xFFFFu;
	return ((w << s) & (m | (m << 16))) ^ bias;
}

// Stage source rows [r0, r0 + nr) x columns [c0, c0 + nc) into LDS with row
// pitch `pin` samples. c0 and nc are multiples of 16, pin of 16. Rows at or
// past `rows_valid` and columns at or past `src_w` are stored as zero
// samples; every tap that could reach them has a zero coefficient.
//
// All 16-byte loads of a batch are issued before any is used, and none is
// conditional: a thread past the end of the tile repeats the last chunk
// (same value to the same place), and a chunk at or past the edge of the
// plane is first read from a clamped address inside the row (the aligned
// 16 bytes at the row's start or at its last full chunk, in the same page as
// valid data) and then rebuilt sample by sample.
template <bool WORD, int U>
__device__ __forceinline__ void stage_tile(lds_t<WORD> *xin, unsigned pin, const char *src, long long stride, unsigned r0,
                                           unsigned nr, unsigned rows_valid, unsigned c0, unsigned nc, unsigned src_w,
                                           const InXform &in)
{
	const unsigned esz = in.type == 0 ? 1 : in.type == 3 ? 4 : 2;
	const unsigned ce = 16 / esz; // samples per 16-byte chunk
	const unsigned cpr = nc / ce; // chunks per row
	const unsigned total = nr * cpr;
	const unsigned last_full = src_w >= ce ? (src_w - ce) / ce * ce : 0; // last chunk start inside the row
	// Chunk e = base + u kThreads + thread as (row, column), advanced by
	// kThreads chunks at a time without a division; chunks past the end are
	// the last chunk (row nr - 1, column cpr - 1).
	const unsigned step_rows = kThreads / cpr, step_cols = kThreads % cpr;
	unsigned row = threadIdx.x / cpr, col = threadIdx.x - (threadIdx.x / cpr) * cpr;
	for (unsigned base = 0; base < total; base += U * kThreads) {
		uint4 raw[U];
		unsigned rr[U], cc[U];
		bool full[U];
#pragma unroll
		for (int u = 0; u < U; ++u) {
			const bool past = row >= nr;
			rr[u] = past ? nr - 1 : row;
			cc[u] = c0 + (past ? cpr - 1 : col) * ce;
			col += step_cols;
			row += step_rows;
			if (col >= cpr) {
				col -= cpr;
				++row;
			}
			const unsigned r = min(r0 + rr[u], rows_valid - 1);
			full[u] = r0 + rr[u] < rows_valid && cc[u] + ce <= src_w;
			const unsigned cl = min(cc[u], last_full);
			raw[u] = *reinterpret_cast<const uint4 *>(row_ptr(src, stride, r) + static_cast<size_t>(cl) * esz);
		}
#pragma unroll
		for (int u = 0; u < U; ++u) {
			const unsigned r = r0 + rr[u];
			if (!full[u]) {
				// Chunk below the plane (zero) or straddling its right edge:
				// per-sample loads inside the row.
				uint32_t w[4] = { 0, 0, 0, 0 };
				if (r < rows_valid) {
					const char *srow = row_ptr(src, stride, r);
#pragma unroll
					for (unsigned k = 0; k < 16; ++k) {
						const unsigned c = cc[u] + k;
						if (k < ce && c < src_w) {
							const uint32_t v = load_as_uint(srow, c, in.type);
							if (esz == 4)
								w[k & 3] = v;
							else if (esz == 2)
								w[(k >> 1) & 3] |= v << (16 * (k & 1));
							else
								w[k >> 2] |= v << (8 * (k & 3));
						}
					}
				}
				raw[u] = make_uint4(w[0], w[1], w[2], w[3]);
			}
			lds_t<WORD> *d = xin + static_cast<size_t>(rr[u]) * pin + (cc[u] - c0);
			const uint32_t wv[4] = { raw[u].x, raw[u].y, raw[u].z, raw[u].w };
			if constexpr (WORD) {
				const unsigned s = in.op == 1 ? in.shift : 0;
				if (esz == 2) {
					*reinterpret_cast<uint4 *>(d) = make_uint4(shift_bias2(wv[0], s), shift_bias2(wv[1], s), shift_bias2(wv[2], s), shift_bias2(wv[3], s));
				} else {
					uint32_t o[8];
#pragma unroll
					for (int k = 0; k < 4; ++k) {
						o[2 * k] = shift_bias2((wv[k] & 0xFFu) | ((wv[k] & 0xFF00u) << 8), s);
						o[2 * k + 1] = shift_bias2(((wv[k] >> 16) & 0xFFu) | ((wv[k] >> 8) & 0xFF0000u), s);
					}
					reinterpret_cast<uint4 *>(d)[0] = make_uint4(o[0], o[1], o[2], o[3]);
					reinterpret_cast<uint4 *>(d)[1] = make_uint4(o[4], o[5], o[6], o[7]);
				}
			} else {
				if (esz == 4) {
					float4 v = make_float4(__uint_as_float(wv[0]), __uint_as_float(wv[1]), __uint_as_float(wv[2]), __uint_as_float(wv[3]));
					if (in.op == 2) {
						v.x = v.x * in.scale + in.offset;
						v.y = v.y * in.scale + in.offset;
						v.z = v.z * in.scale + in.offset;
						v.w = v.w * in.scale + in.offset;
					}
					*reinterpret_cast<float4 *>(d) = v;
				} else {
					// 16 BYTE or 8 WORD / HALF samples of the chunk, constant word indices.
					auto conv = [&](uint32_t bits) {
						float x = in.type == 2 ? half_to_float(static_cast<uint16_t>(bits)) : static_cast<float>(bits);
						return in.op == 2 ? x * in.scale + in.offset : x;
					};
					if (esz == 2 && in.type != 2) {
						// WORD: the conversion chosen once per chunk (uniform), not per sample
						const float sc = in.op == 2 ? in.scale : 1.0f, of = in.op == 2 ? in.offset : 0.0f;
						auto cw = [&](uint32_t bits) {
							const float x = static_cast<float>(bits);
							return in.op == 2 ? x * sc + of : x;
						};
						if (in.op == 2) {
#pragma unroll
							for (unsigned k = 0; k < 8; k += 4) {
								const uint32_t lo = wv[k / 2], hi = wv[k / 2 + 1];
								*reinterpret_cast<float4 *>(d + k) = make_float4(static_cast<float>(lo & 0xFFFFu) * sc + of, static_cast<float>(lo >> 16) * sc + of,
								                                                 static_cast<float>(hi & 0xFFFFu) * sc + of, static_cast<float>(hi >> 16) * sc + of);
							}
						} else {
#pragma unroll
							for (unsigned k = 0; k < 8; k += 4) {
								const uint32_t lo = wv[k / 2], hi = wv[k / 2 + 1];
								*reinterpret_cast<float4 *>(d + k) = make_float4(cw(lo & 0xFFFFu), cw(lo >> 16), cw(hi & 0xFFFFu), cw(hi >> 16));
							}
						}
					} else if (esz == 2) {
#pragma unroll
						for (unsigned k = 0; k < 8; k += 4) {
							const uint32_t lo = wv[k / 2], hi = wv[k / 2 + 1];
							*reinterpret_cast<float4 *>(d + k) = make_float4(conv(lo & 0xFFFFu), conv(lo >> 16), conv(hi & 0xFFFFu), conv(hi >> 16));
						}
					} else {
#pragma unroll
						for (unsigned k = 0; k < 16; k += 4) {
							const uint32_t wd = wv[k / 4];
							*reinterpret_cast<float4 *>(d + k) =
								make_float4(conv(wd & 0xFFu), conv((wd >> 8) & 0xFFu), conv((wd >> 16) & 0xFFu), conv(wd >> 24));
						}
					}
				}
			}
		}
	}
}

// WORD staging with row pairs interleaved: dword (p, c) of the LDS window
// (pitch `pin` dwords per pair) holds biased samples (r0 + 2p, c0 + c) in its
// low half and (r0 + 2p + 1, c0 + c) in its high half, so a vertical pass
// applies two taps per v_dot2. nr is even; BYTE or WORD input. The loads
// follow stage_tile: each unit is the same 16-byte chunk of both rows of a
// pair, every load of a batch issued before any is used, none conditional,
// chunks outside the plane read at a clamped address and rebuilt sample by
// sample (rows past rows_valid as zero samples).
template <int U>
__device__ __forceinline__ void stage_tile_pairs(uint32_t *xw, unsigned pin, const char *src, long long stride, unsigned r0,
                                                 unsigned nr, unsigned rows_valid, unsigned c0, unsigned nc, unsigned src_w,
                                                 const InXform &in)
{
	const unsigned esz = in.type == 0 ? 1 : 2;
	const unsigned ce = 16 / esz;
	const unsigned cpr = nc / ce;
	const unsigned np = nr / 2;
	const unsigned total = np * cpr;
	const unsigned last_full = src_w >= ce ? (src_w - ce) / ce * ce : 0;
	const unsigned s = in.op == 1 ? in.shift : 0;
	const unsigned step_rows = kThreads / cpr, step_cols = kThreads % cpr;
	unsigned pr = threadIdx.x / cpr, col = threadIdx.x - (threadIdx.x / cpr) * cpr;
	for (unsigned base = 0; base < total; base += U * kThreads) {
		uint4 raw[U][2];
		unsigned pp[U], cc[U];
		bool full[U][2];
#pragma unroll
		for (int u = 0; u < U; ++u) {
			const bool past = pr >= np;
			pp[u] = past ? np - 1 : pr;
			cc[u] = c0 + (past ? cpr - 1 : col) * ce;
			col += step_cols;
			pr += step_rows;
			if (col >= cpr) {
				col -= cpr;
				++pr;
			}
			const unsigned cl = min(cc[u], last_full);
#pragma unroll
			for (int h = 0; h < 2; ++h) {
				const unsigned row = r0 + 2 * pp[u] + h;
				full[u][h] = row < rows_valid && cc[u] + ce <= src_w;
				raw[u][h] = *reinterpret_cast<const uint4 *>(row_ptr(src, stride, min(row, rows_valid - 1)) + static_cast<size_t>(cl) * esz);
			}
		}
#pragma unroll
		for (int u = 0; u < U; ++u) {
#pragma unroll
			for (int h = 0; h < 2; ++h) {
				const unsigned row = r0 + 2 * pp[u] + h;
				if (!full[u][h]) {
					uint32_t w[4] = { 0, 0, 0, 0 };
					if (row < rows_valid) {
						const char *srow = row_ptr(src, stride, row);
#pragma unroll
						for (unsigned k = 0; k < 16; ++k) {
							const unsigned c = cc[u] + k;
							if (k < ce && c < src_w) {
								const uint32_t v = load_as_uint(srow, c, in.type);
								if (esz == 2)
									w[(k >> 1) & 3] |= v << (16 * (k & 1));
								else
									w[k >> 2] |= v << (8 * (k & 3));
							}
						}
					}
					raw[u][h] = make_uint4(w[0], w[1], w[2], w[3]);
				}
			}
			const uint32_t ev[4] = { raw[u][0].x, raw[u][0].y, raw[u][0].z, raw[u][0].w };
			const uint32_t od[4] = { raw[u][1].x, raw[u][1].y, raw[u][1].z, raw[u][1].w };
			uint4 *d = reinterpret_cast<uint4 *>(xw + static_cast<size_t>(pp[u]) * pin + (cc[u] - c0));
			if (esz == 2) {
				// dword 2k: low halves of the two rows' dword k; 2k + 1: high halves.
				uint32_t o[8];
#pragma unroll
				for (int k = 0; k < 4; ++k) {
					o[2 * k] = shift_bias2(__builtin_amdgcn_perm(od[k], ev[k], 0x05040100u), s);
					o[2 * k + 1] = shift_bias2(__builtin_amdgcn_perm(od[k], ev[k], 0x07060302u), s);
				}
				d[0] = make_uint4(o[0], o[1], o[2], o[3]);
				d[1] = make_uint4(o[4], o[5], o[6], o[7]);
			} else {
				// dword 4k + m: byte m of the two rows' dword k, zero-extended.
				uint32_t o[16];
#pragma unroll
				for (int k = 0; k < 4; ++k)
#pragma unroll
					for (int m = 0; m < 4; ++m)
						o[4 * k + m] = shift_bias2(__builtin_amdgcn_perm(od[k], ev[k], 0x0C000C00u | (static_cast<uint32_t>(4 + m) << 16) | static_cast<uint32_t>(m)), s);
#pragma unroll
				for (int k = 0; k < 4; ++k)
					d[k] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
			}
		}
	}
}

// Horizontal taps of a thread's two output columns, in registers. WORD:
// N packed int16 pairs per column over the even-aligned window (dot2), dword
// offsets into an LDS row; pairs past the table's fwp are zero, so all N are
// applied without a branch (the staged window covers 2N samples). FLOAT: the
// first fw of N float taps, sample offsets; taps past fw leave the sum as is.
template <bool WORD, int N>
struct HRegs {
	using C = std::conditional_t<WORD, int32_t, float>;
	C c0[N], c1[N];
	int l0, l1;

	// j, j1: the two columns (equal for a lone last column); cbase: first staged sample.
	__device__ __forceinline__ void load(const PassArg &p, unsigned j, unsigned j1, unsigned cbase)
	{
		// Loads are unconditional (taps past the row re-read its last tap),
		// so they issue together with the tile's other loads.
		if constexpr (WORD) {
			l0 = static_cast<int>(p.pl[j] - cbase) / 2;
			l1 = static_cast<int>(p.pl[j1] - cbase) / 2;
#pragma unroll
			for (int k = 0; k < N; ++k) {
				const int kk = min(k, p.fwp - 1);
				const int32_t v0 = p.pc[static_cast<size_t>(j) * p.fwp + kk], v1 = p.pc[static_cast<size_t>(j1) * p.fwp + kk];
				c0[k] = k < p.fwp ? v0 : 0;
				c1[k] = k < p.fwp ? v1 : 0;
			}
		} else {
			l0 = static_cast<int>(p.left[j] - cbase);
			l1 = static_cast<int>(p.left[j1] - cbase);
#pragma unroll
			for (int k = 0; k < N; ++k) {
				const int kk = min(k, p.fw - 1);
				const float v0 = p.cf[static_cast<size_t>(j) * p.fw + kk], v1 = p.cf[static_cast<size_t>(j1) * p.fw + kk];
				c0[k] = k < p.fw ? v0 : 0.0f;
				c1[k] = k < p.fw ? v1 : 0.0f;
			}
		}
	}

	// init: the accumulators' start value (0, kBiasedInit or kWordInit).
	__device__ __forceinline__ void apply_word(const uint32_t *row, int32_t &a0, int32_t &a1, int32_t init) const
	{
		int32_t s0 = init, s1 = init;
#pragma unroll
		for (int k = 0; k < N; ++k) {
			s0 = dot2(row[l0 + k], c0[k], s0);
			s1 = dot2(row[l1 + k], c1[k], s1);
		}
		a0 = s0;
		a1 = s1;
	}

	// The first M taps, all real: the same sums without selects.
	template <int M>
	__device__ __forceinline__ void apply_float_n(const float *row, float &v0, float &v1) const
	{
		float a0 = 0.0f, a1 = 0.0f;
#pragma unroll
		for (int k = 0; k < M; ++k) {
			a0 = a0 + c0[k] * row[l0 + k];
			a1 = a1 + c1[k] * row[l1 + k];
		}
		v0 = a0;
		v1 = a1;
	}

	// Calls body(apply) once, apply(row, v0, v1) being the sum form for fw.
	// The branch on fw is taken here, outside the caller's row loop: inside
	// it, the taps of the common forms sit at fixed offsets from one address
	// per column, instead of one address per tap kept for the clamped form
	// and advanced every row.
	template <class Body>
	__device__ __forceinline__ void with_float_form(int fw, Body body) const
	{
		if (fw == N) {
			body([this](const float *row, float &v0, float &v1) { apply_float_n<N>(row, v0, v1); });
			return;
		}
		if constexpr (N > 2) {
			if (fw == N - 1) {
				body([this](const float *row, float &v0, float &v1) { apply_float_n<N - 1>(row, v0, v1); });
				return;
			}
		}
		body([this, fw](const float *row, float &v0, float &v1) { apply_float_sel(row, fw, v0, v1); });
	}

	// Any fw <= N: taps past fw leave the sums as they are.
	__device__ __forceinline__ void apply_float_sel(const float *row, int fw, float &v0, float &v1) const
	{
		float a0 = 0.0f, a1 = 0.0f;
#pragma unroll
		for (int k = 0; k < N; ++k) {
			const int kk = k < fw ? k : fw - 1; // stays inside the window; the sum ignores it
			const float x0 = row[l0 + kk], x1 = row[l1 + kk];
			const float s0 = a0 + c0[k] * x0, s1 = a1 + c1[k] * x1;
			a0 = k < fw ? s0 : a0;
			a1 = k < fw ? s1 : a1;
		}
		v0 = a0;
		v1 = a1;
	}
};

// LDS layout of the tiled kernel: per-row vertical table of the tile (start
// row, then the row's taps; copied once so the row loops read LDS, not HBM),
// the staged source window, then the intermediate.
__host__ __device__ inline unsigned tiled_vtaps(const ResizeArg &a)
{
	if (a.h_first && a.passes == 1)
		return 0;
	const PassArg &pv = a.h_first ? a.p2 : a.p1;
	return a.word ? pv.fwp : pv.fw;
}

// Start rows occupy the first round4(tile_h) words, then one row of taps per
// output row, rows back to back (the same order as the taps in HBM, so a
// tile's taps are one contiguous copy).
__host__ __device__ inline unsigned tiled_vtab_words(const ResizeArg &a)
{
	const unsigned taps = tiled_vtaps(a);
	if (!taps)
		return 0;
	return ((a.tile_h + 3) & ~3u) + a.tile_h * taps;
}

__host__ __device__ inline size_t tiled_lds_bytes(const ResizeArg &a, size_t *in_off, size_t *mid_off)
{
	const size_t elem = a.word ? 2 : 4;
	const size_t tab = (static_cast<size_t>(tiled_vtab_words(a)) * 4 + 15) & ~static_cast<size_t>(15);
	const size_t in = (elem * a.lds_in_rows * a.lds_in_cols + 15) & ~static_cast<size_t>(15);
	if (in_off)
		*in_off = tab;
	if (mid_off)
		*mid_off = tab + in;
	return tab + in + elem * a.lds_mid_rows * a.lds_mid_cols;
}

// One output tile (bx, by) of frame f; HF must equal a.h_first. Every thread
// of the block reaches the same barriers; the block must synchronise before
// reusing `smem`. Horizontal-first tiles stage fewer rows (4 chunks per
// thread in flight, fewer registers, more resident blocks); vertical-first
// tiles stage up to 8 chunks per thread at once.
template <bool WORD, int N, bool HF>
__device__ __forceinline__ void resize_tile(const ResizeArg &a, unsigned bx, unsigned by, unsigned f, unsigned char *smem)
{
	using T = lds_t<WORD>;
	size_t in_off, mid_off;
	tiled_lds_bytes(a, &in_off, &mid_off);
	uint32_t *vtab = reinterpret_cast<uint32_t *>(smem);
	T *xin = reinterpret_cast<T *>(smem + in_off);
	unsigned char *mid_base = smem + mid_off;

	const unsigned t = threadIdx.x;
	const unsigned j0 = bx * a.tile_w, i0 = by * a.tile_h;
	const unsigned nrows = min(a.tile_h, a.dst_h - i0);
	const unsigned ncols = min(a.tile_w, a.dst_w - j0);
	const char *src = frame_ptr(a.src, f);
	char *dst = frame_ptr_w(a.dst, f);
	const unsigned pin = a.lds_in_cols;

	// Vertical table rows of this tile: vtab[i] = first source row,
	// vtab[nrows + i * taps + k] = tap k (WORD: int16 pairs of the pair
	// table, whose windows start on an even row; FLOAT: float bits).
	const unsigned vtaps = tiled_vtaps(a);
	const unsigned vstride = vtaps;
	const uint32_t *vtop = vtab;
	const uint32_t *vcoef = vtab + ((a.tile_h + 3) & ~3u);
	// The tile's rows of the vertical table are contiguous in HBM (nrows
	// start rows, then nrows x vtaps taps) and keep their order in LDS: up to
	// kVtRegs x kThreads entries are loaded into registers now, beside the
	// other loads, and stored to LDS after the staging.
	constexpr int kVtRegs = 4;
	const unsigned vtot = vtaps ? nrows * (vtaps + 1) : 0;
	uint32_t vval[kVtRegs];
	unsigned vpos[kVtRegs];
	auto vt_entry = [&](unsigned e, uint32_t &val, unsigned &pos) {
		const PassArg &pv = a.h_first ? a.p2 : a.p1;
		if (e < nrows) {
			val = WORD ? pv.pl[i0 + e] : pv.left[i0 + e];
			pos = e;
		} else {
			const unsigned c = e - nrows;
			const size_t g = static_cast<size_t>(i0) * vtaps + c;
			val = WORD ? static_cast<uint32_t>(pv.pc[g]) : __float_as_uint(pv.cf[g]);
			pos = ((a.tile_h + 3) & ~3u) + c;
		}
	};
	// Register slots past the table (a uniform test) are not used.
	if (vtot) {
#pragma unroll
		for (int u = 0; u < kVtRegs; ++u)
			if (u == 0 || u * kThreads < vtot)
				vt_entry(min(t + u * kThreads, vtot - 1), vval[u], vpos[u]);
	}
	auto vt_store = [&]() {
		if (!vtot)
			return;
#pragma unroll
		for (int u = 0; u < kVtRegs; ++u)
			if (u == 0 || u * kThreads < vtot)
				vtab[vpos[u]] = vval[u];
		for (unsigned e = t + kVtRegs * kThreads; e < vtot; e += kThreads) {
			uint32_t v;
			unsigned p;
			vt_entry(e, v, p);
			vtab[p] = v;
		}
	};

	if constexpr (HF) {
		// Tile width 2 * kThreads: thread t owns output columns j0 + 2t, +1.
		const PassArg &ph = a.p1;
		const unsigned jc = 2 * t;
		const bool active = jc < ncols;
		const bool two = jc + 1 < ncols;
		const unsigned j = j0 + (active ? jc : 0);
		const unsigned j1 = two ? j + 1 : j;
		const unsigned jl = j0 + ncols - 1;
		const unsigned cb = (WORD ? ph.pl[j0] : ph.left[j0]) & ~15u;
		const unsigned ce = WORD ? ph.pl[jl] + 2 * N : ph.left[jl] + ph.fw;
		const unsigned nc = (ce - cb + 15) & ~15u;
		unsigned r0, r1;
		if (a.passes == 2) {
			const PassArg &pv = a.p2;
			const unsigned il = i0 + nrows - 1;
			r0 = WORD ? pv.pl[i0] : pv.left[i0];
			r1 = WORD ? pv.pl[il] + 2 * pv.fwp : pv.left[il] + pv.fw;
		} else {
			r0 = i0;
			r1 = i0 + nrows;
		}
		const unsigned span = r1 - r0;
		HRegs<WORD, N> h;
		h.load(ph, j, j1, cb);
		stage_tile<WORD, 4>(xin, pin, src, a.src.stride, r0, span, a.src_h, cb, nc, a.src_w, a.in);
		vt_store();
		__syncthreads();
		if constexpr (WORD) {
			const uint32_t *xw = reinterpret_cast<const uint32_t *>(xin);
			const unsigned pw = pin / 2;
			if (a.passes == 2) {
				// Row pairs (r0 even, span even) -> interleaved intermediate.
				uint32_t *mid = reinterpret_cast<uint32_t *>(mid_base);
				const unsigned pm = a.lds_mid_cols;
				for (unsigned r = 0; r < span; r += 2) {
					int32_t e0, e1, o0, o1;
					h.apply_word(xw + r * pw, e0, e1, kBiasedInit);
					h.apply_word(xw + (r + 1) * pw, o0, o1, kBiasedInit);
					const uint32_t m0 = pack16x2(pack_biased_pre(e0, a.pixel_max), pack_biased_pre(o0, a.pixel_max));
					const uint32_t m1 = pack16x2(pack_biased_pre(e1, a.pixel_max), pack_biased_pre(o1, a.pixel_max));
					*reinterpret_cast<uint2 *>(mid + (r / 2) * pm + jc) = make_uint2(m0, m1);
				}
				__syncthreads();
				// Vertical pass: four columns per thread, two groups of rows.
				const unsigned j4 = 4 * (t % (kThreads / 2));
				const unsigned g = t / (kThreads / 2);
				const unsigned n4 = j4 < ncols ? min(4u, ncols - j4) : 0;
				const bool wide = n4 == 4 && a.out.op == 0 && a.out.type == 1;
				const bool qbyte = n4 == 4 && a.out.op == 1 && a.out.type == 0;
				for (unsigned i = g; i < nrows; i += 2) {
					const unsigned row = i0 + i;
					const uint32_t *m = mid + ((vtop[i] - r0) / 2) * pm + j4;
					const uint32_t *cv = vcoef + i * vstride;
					int32_t a0 = kWordInit, a1 = kWordInit, a2 = kWordInit, a3 = kWordInit;
					if (vtaps <= static_cast<unsigned>(N)) {
						// N tap pairs unrolled, all LDS reads issued first; pairs past
						// vtaps get a zero tap over the last row (adds exactly 0).
						uint4 mm[N];
						int32_t cc[N];
#pragma unroll
						for (int k = 0; k < N; ++k) {
							const unsigned kk = min(static_cast<unsigned>(k), vtaps - 1);
							mm[k] = *reinterpret_cast<const uint4 *>(m + kk * pm);
							cc[k] = static_cast<unsigned>(k) < vtaps ? static_cast<int32_t>(cv[kk]) : 0;
						}
#pragma unroll
						for (int k = 0; k < N; ++k) {
							a0 = dot2(mm[k].x, cc[k], a0);
							a1 = dot2(mm[k].y, cc[k], a1);
							a2 = dot2(mm[k].z, cc[k], a2);
							a3 = dot2(mm[k].w, cc[k], a3);
						}
					} else {
						for (unsigned k = 0; k < vtaps; ++k) {
							const uint4 mm = *reinterpret_cast<const uint4 *>(m + k * pm);
							const int32_t c = static_cast<int32_t>(cv[k]);
							a0 = dot2(mm.x, c, a0);
							a1 = dot2(mm.y, c, a1);
							a2 = dot2(mm.z, c, a2);
							a3 = dot2(mm.w, c, a3);
						}
					}
					char *drow = row_ptr(dst, a.dst.stride, row);
					const uint16_t v0 = pack_word_pre(a0, a.pixel_max), v1 = pack_word_pre(a1, a.pixel_max);
					const uint16_t v2 = pack_word_pre(a2, a.pixel_max), v3 = pack_word_pre(a3, a.pixel_max);
					if (wide) {
						*reinterpret_cast<uint2 *>(drow + 2 * static_cast<size_t>(j0 + j4)) = make_uint2(pack16x2(v0, v1), pack16x2(v2, v3));
					} else if (qbyte) {
						// Four quantised BYTE outputs (j0 + j4 is a multiple of 4): one
						// offset load, one 4-byte store.
						const float4 d = dither_offset4(a.out, row, j0 + j4);
						const uint32_t q = quantize_d(static_cast<float>(v0), a.out, d.x) | (quantize_d(static_cast<float>(v1), a.out, d.y) << 8) |
						                   (quantize_d(static_cast<float>(v2), a.out, d.z) << 16) | (quantize_d(static_cast<float>(v3), a.out, d.w) << 24);
						*reinterpret_cast<uint32_t *>(drow + j0 + j4) = q;
					} else if (n4) {
						store_pair<true>(drow, j0 + j4, v0, v1, n4 > 1, a.out, row);
						if (n4 > 2)
							store_pair<true>(drow, j0 + j4 + 2, v2, v3, n4 > 3, a.out, row);
					}
				}
			} else {
				for (unsigned r = 0; r < span; ++r) {
					int32_t a0, a1;
					h.apply_word(xw + r * pw, a0, a1, kWordInit);
					if (active)
						store_pair<true>(row_ptr(dst, a.dst.stride, r0 + r), j, pack_word_pre(a0, a.pixel_max), pack_word_pre(a1, a.pixel_max), two, a.out, r0 + r);
				}
			}
		} else {
			if (a.passes == 2) {
				float *mid = reinterpret_cast<float *>(mid_base);
				const unsigned pm = a.lds_mid_cols;
				h.with_float_form(ph.fw, [&](auto apply) {
					for (unsigned r = 0; r < span; ++r) {
						float v0, v1;
						apply(xin + r * pin, v0, v1);
						*reinterpret_cast<float2 *>(mid + r * pm + jc) = make_float2(v0, v1);
					}
				});
				__syncthreads();
				const unsigned j4 = 4 * (t % (kThreads / 2));
				const unsigned g = t / (kThreads / 2);
				const unsigned n4 = j4 < ncols ? min(4u, ncols - j4) : 0;
				const bool wide = n4 == 4 && a.out.op == 0 && a.out.type == 3;
				for (unsigned i = g; i < nrows; i += 2) {
					const unsigned row = i0 + i;
					const float *m = mid + (vtop[i] - r0) * pm + j4;
					const uint32_t *cv = vcoef + i * vstride;
					float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
					if (vtaps == static_cast<unsigned>(N)) {
						// every tap is real: N taps unrolled, all LDS reads issued first
						float4 mm[N];
						float cc[N];
#pragma unroll
						for (int k = 0; k < N; ++k) {
							mm[k] = *reinterpret_cast<const float4 *>(m + k * pm);
							cc[k] = __uint_as_float(cv[k]);
						}
#pragma unroll
						for (int k = 0; k < N; ++k) {
							a0 = a0 + cc[k] * mm[k].x;
							a1 = a1 + cc[k] * mm[k].y;
							a2 = a2 + cc[k] * mm[k].z;
							a3 = a3 + cc[k] * mm[k].w;
						}
					} else if (vtaps < static_cast<unsigned>(N)) {
						// N taps unrolled, all LDS reads issued first; taps past vtaps
						// leave the sums as they are (select), so every kept addition
						// is the reference's, in its order.
						float4 mm[N];
						float cc[N];
#pragma unroll
						for (int k = 0; k < N; ++k) {
							const unsigned kk = min(static_cast<unsigned>(k), vtaps - 1);
							mm[k] = *reinterpret_cast<const float4 *>(m + kk * pm);
							cc[k] = __uint_as_float(cv[kk]);
						}
#pragma unroll
						for (int k = 0; k < N; ++k) {
							const bool on = static_cast<unsigned>(k) < vtaps;
							const float s0 = a0 + cc[k] * mm[k].x, s1 = a1 + cc[k] * mm[k].y;
							const float s2 = a2 + cc[k] * mm[k].z, s3 = a3 + cc[k] * mm[k].w;
							a0 = on ? s0 : a0;
							a1 = on ? s1 : a1;
							a2 = on ? s2 : a2;
							a3 = on ? s3 : a3;
						}
					} else {
						for (unsigned k = 0; k < vtaps; ++k) {
							const float4 mm = *reinterpret_cast<const float4 *>(m + k * pm);
							const float c = __uint_as_float(cv[k]);
							a0 += c * mm.x;
							a1 += c * mm.y;
							a2 += c * mm.z;
							a3 += c * mm.w;
						}
					}
					char *drow = row_ptr(dst, a.dst.stride, row);
					if (wide) {
						*reinterpret_cast<float4 *>(drow + 4 * static_cast<size_t>(j0 + j4)) = make_float4(a0, a1, a2, a3);
					} else if (n4) {
						store_pair<false>(drow, j0 + j4, a0, a1, n4 > 1, a.out, row);
						if (n4 > 2)
							store_pair<false>(drow, j0 + j4 + 2, a2, a3, n4 > 3, a.out, row);
					}
				}
			} else {
				h.with_float_form(ph.fw, [&](auto apply) {
					for (unsigned r = 0; r < span; ++r) {
						float v0, v1;
						apply(xin + r * pin, v0, v1);
						if (active)
							store_pair<false>(row_ptr(dst, a.dst.stride, r0 + r), j, v0, v1, two, a.out, r0 + r);
					}
				});
			}
		}
		return;
	}

	// Vertical pass first (or alone): stage the input rows of the tile's
	// output rows over the columns the horizontal pass (if any) reads.
	// WORD: row pairs from the even-aligned window starts (pair tables).
	const PassArg &pv = a.p1;
	const unsigned il = i0 + nrows - 1;
	const unsigned r0 = WORD ? pv.pl[i0] : pv.left[i0];
	const unsigned span = WORD ? pv.pl[il] + 2 * pv.fwp - r0 : pv.left[il] + pv.fw - r0;
	unsigned cb, nc;
	if (a.passes == 2) {
		const PassArg &ph = a.p2;
		const unsigned jl = j0 + ncols - 1;
		cb = (WORD ? ph.pl[j0] : ph.left[j0]) & ~15u;
		const unsigned ce = WORD ? ph.pl[jl] + 2 * N : ph.left[jl] + ph.fw;
		nc = (ce - cb + 15) & ~15u;
	} else {
		cb = j0;
		nc = (ncols + 15) & ~15u;
	}
	// Horizontal taps of this thread's output column pair (tile_w <= 2 x
	// kThreads: at most one pair per thread), loaded beside the staging.
	const unsigned opairs = (ncols + 1) / 2;
	const unsigned ostride = opairs >= kThreads ? kThreads : opairs;
	const unsigned ogroups = kThreads / ostride;
	const unsigned og = t / ostride;
	const unsigned oq = t - og * ostride;
	const unsigned ojc = 2 * oq;
	const bool otwo = ojc + 1 < ncols;
	const unsigned oj = j0 + ojc;
	HRegs<WORD, N> h;
	if (a.passes == 2)
		h.load(a.p2, oj, otwo ? oj + 1 : oj, cb);
	if constexpr (WORD)
		stage_tile_pairs<4>(reinterpret_cast<uint32_t *>(xin), pin, src, a.src.stride, r0, span, a.src_h, cb, nc, a.src_w, a.in);
	else
		stage_tile<WORD, 8>(xin, pin, src, a.src.stride, r0, span, a.src_h, cb, nc, a.src_w, a.in);
	vt_store();
	__syncthreads();

	const unsigned pm = a.lds_mid_cols;
	if constexpr (WORD) {
		// Vertical pass over the interleaved row pairs, two taps per v_dot2:
		// items of four columns x one output row spread over all threads.
		const uint32_t *xw = reinterpret_cast<const uint32_t *>(xin);
		const unsigned quads = nc / 4, items = quads * nrows;
		const unsigned sq = kThreads % quads, si = kThreads / quads;
		unsigned q = t % quads, i = t / quads;
		const int32_t init = a.passes == 2 ? kBiasedInit : kWordInit;
		const bool wide = a.out.op == 0 && a.out.type == 1;
		const bool qbyte = a.out.op == 1 && a.out.type == 0;
		for (unsigned it = t; it < items; it += kThreads) {
			const uint32_t *m = xw + ((vtop[i] - r0) / 2) * pin + 4 * q;
			const uint32_t *cv = vcoef + i * vstride;
			int32_t a0 = init, a1 = init, a2 = init, a3 = init;
			if (vtaps <= static_cast<unsigned>(N)) {
				// N tap pairs unrolled, all LDS reads issued first; pairs past
				// vtaps get a zero tap over the last pair (adds exactly 0).
				uint4 mm[N];
				int32_t cc[N];
#pragma unroll
				for (int k = 0; k < N; ++k) {
					const unsigned kk = min(static_cast<unsigned>(k), vtaps - 1);
					mm[k] = *reinterpret_cast<const uint4 *>(m + kk * pin);
					cc[k] = static_cast<unsigned>(k) < vtaps ? static_cast<int32_t>(cv[kk]) : 0;
				}
#pragma unroll
				for (int k = 0; k < N; ++k) {
					a0 = dot2(mm[k].x, cc[k], a0);
					a1 = dot2(mm[k].y, cc[k], a1);
					a2 = dot2(mm[k].z, cc[k], a2);
					a3 = dot2(mm[k].w, cc[k], a3);
				}
			} else {
				for (unsigned k = 0; k < vtaps; ++k) {
					const uint4 mm = *reinterpret_cast<const uint4 *>(m + k * pin);
					const int32_t c = static_cast<int32_t>(cv[k]);
					a0 = dot2(mm.x, c, a0);
					a1 = dot2(mm.y, c, a1);
					a2 = dot2(mm.z, c, a2);
					a3 = dot2(mm.w, c, a3);
				}
			}
			if (a.passes == 2) {
				*reinterpret_cast<uint2 *>(reinterpret_cast<uint32_t *>(mid_base) + i * (pm / 2) + 2 * q) =
					make_uint2(pack16x2(pack_biased_pre(a0, a.pixel_max), pack_biased_pre(a1, a.pixel_max)),
					           pack16x2(pack_biased_pre(a2, a.pixel_max), pack_biased_pre(a3, a.pixel_max)));
			} else {
				// Vertical only: cb = j0, four output columns j0 + 4q ...
				const unsigned row = i0 + i, jc = 4 * q;
				const unsigned n4 = jc < ncols ? min(4u, ncols - jc) : 0;
				char *drow = row_ptr(dst, a.dst.stride, row);
				const uint16_t v0 = pack_word_pre(a0, a.pixel_max), v1 = pack_word_pre(a1, a.pixel_max);
				const uint16_t v2 = pack_word_pre(a2, a.pixel_max), v3 = pack_word_pre(a3, a.pixel_max);
				if (n4 == 4 && wide) {
					*reinterpret_cast<uint2 *>(drow + 2 * static_cast<size_t>(j0 + jc)) = make_uint2(pack16x2(v0, v1), pack16x2(v2, v3));
				} else if (n4 == 4 && qbyte) {
					const float4 d = dither_offset4(a.out, row, j0 + jc);
					const uint32_t qv = quantize_d(static_cast<float>(v0), a.out, d.x) | (quantize_d(static_cast<float>(v1), a.out, d.y) << 8) |
					                    (quantize_d(static_cast<float>(v2), a.out, d.z) << 16) | (quantize_d(static_cast<float>(v3), a.out, d.w) << 24);
					*reinterpret_cast<uint32_t *>(drow + j0 + jc) = qv;
				} else if (n4) {
					store_pair<true>(drow, j0 + jc, v0, v1, n4 > 1, a.out, row);
					if (n4 > 2)
						store_pair<true>(drow, j0 + jc + 2, v2, v3, n4 > 3, a.out, row);
				}
			}
			q += sq;
			i += si;
			if (q >= quads) {
				q -= quads;
				++i;
			}
		}
	} else if constexpr (N <= 12) {
		// FLOAT vertical pass: items of four columns x one output row spread
		// over all threads (float4 reads of the staged rows; the taps read
		// once per four columns). Each sum is the reference's: from 0, tap
		// k added in order, no contraction.
		const float *xf = reinterpret_cast<const float *>(xin);
		const unsigned quads = nc / 4, items = quads * nrows;
		const unsigned sq = kThreads % quads, si = kThreads / quads;
		unsigned q = t % quads, i = t / quads;
		const bool wide = a.out.op == 0 && a.out.type == 3;
		for (unsigned it = t; it < items; it += kThreads) {
			const float *m = xf + (vtop[i] - r0) * pin + 4 * q;
			const uint32_t *cv = vcoef + i * vstride;
			float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
			if (vtaps == static_cast<unsigned>(N)) {
				float4 mm[N];
				float cc[N];
#pragma unroll
				for (int k = 0; k < N; ++k) {
					mm[k] = *reinterpret_cast<const float4 *>(m + k * pin);
					cc[k] = __uint_as_float(cv[k]);
				}
#pragma unroll
				for (int k = 0; k < N; ++k) {
					a0 = a0 + cc[k] * mm[k].x;
					a1 = a1 + cc[k] * mm[k].y;
					a2 = a2 + cc[k] * mm[k].z;
					a3 = a3 + cc[k] * mm[k].w;
				}
			} else if (vtaps < static_cast<unsigned>(N)) {
				// Taps past vtaps leave the sums as they are.
				float4 mm[N];
				float cc[N];
#pragma unroll
				for (int k = 0; k < N; ++k) {
					const unsigned kk = min(static_cast<unsigned>(k), vtaps - 1);
					mm[k] = *reinterpret_cast<const float4 *>(m + kk * pin);
					cc[k] = __uint_as_float(cv[kk]);
				}
#pragma unroll
				for (int k = 0; k < N; ++k) {
					const bool on = static_cast<unsigned>(k) < vtaps;
					const float s0 = a0 + cc[k] * mm[k].x, s1 = a1 + cc[k] * mm[k].y;
					const float s2 = a2 + cc[k] * mm[k].z, s3 = a3 + cc[k] * mm[k].w;
					a0 = on ? s0 : a0;
					a1 = on ? s1 : a1;
					a2 = on ? s2 : a2;
					a3 = on ? s3 : a3;
				}
			} else {
				for (unsigned k = 0; k < vtaps; ++k) {
					const float4 mm = *reinterpret_cast<const float4 *>(m + k * pin);
					const float c = __uint_as_float(cv[k]);
					a0 += c * mm.x;
					a1 += c * mm.y;
					a2 += c * mm.z;
					a3 += c * mm.w;
				}
			}
			if (a.passes == 2) {
				*reinterpret_cast<float4 *>(reinterpret_cast<float *>(mid_base) + i * pm + 4 * q) = make_float4(a0, a1, a2, a3);
			} else {
				// Vertical only: cb = j0, four output columns j0 + 4q ...
				const unsigned row = i0 + i, jc = 4 * q;
				const unsigned n4 = jc < ncols ? min(4u, ncols - jc) : 0;
				char *drow = row_ptr(dst, a.dst.stride, row);
				if (n4 == 4 && wide) {
					*reinterpret_cast<float4 *>(drow + 4 * static_cast<size_t>(j0 + jc)) = make_float4(a0, a1, a2, a3);
				} else if (n4) {
					store_pair<false>(drow, j0 + jc, a0, a1, n4 > 1, a.out, row);
					if (n4 > 2)
						store_pair<false>(drow, j0 + jc + 2, a2, a3, n4 > 3, a.out, row);
				}
			}
			q += sq;
			i += si;
			if (q >= quads) {
				q -= quads;
				++i;
			}
		}
	} else {
		// Vertical pass over column pairs (wide filters: fewer registers per
		// item): groups of threads split the rows.
		const unsigned pairs = nc / 2;
		const unsigned pstride = pairs >= kThreads ? kThreads : pairs;
		const unsigned groups = kThreads / pstride;
		const unsigned grp = t / pstride;
		if (grp < groups) {
			for (unsigned q = t - grp * pstride; q < pairs; q += kThreads) {
				for (unsigned i = grp; i < nrows; i += groups) {
					const unsigned row = i0 + i;
					const unsigned top = vtop[i] - r0;
					const uint32_t *cv = vcoef + i * vstride;
					const float *xr = xin + top * pin + 2 * q;
					float a0 = 0.0f, a1 = 0.0f;
					if (vtaps == static_cast<unsigned>(N)) {
						// every tap is real: unrolled over N taps without selects
						float2 mv[N];
						float cc[N];
#pragma unroll
						for (int k = 0; k < N; ++k) {
							mv[k] = *reinterpret_cast<const float2 *>(xr + k * pin);
							cc[k] = __uint_as_float(cv[k]);
						}
#pragma unroll
						for (int k = 0; k < N; ++k) {
							a0 = a0 + cc[k] * mv[k].x;
							a1 = a1 + cc[k] * mv[k].y;
						}
					} else if (vtaps < static_cast<unsigned>(N)) {
						// Unrolled over N taps; taps past vtaps leave the sums as they are.
						float2 mv[N];
						float cc[N];
#pragma unroll
						for (int k = 0; k < N; ++k) {
							const unsigned kk = min(static_cast<unsigned>(k), vtaps - 1);
							mv[k] = *reinterpret_cast<const float2 *>(xr + kk * pin);
							cc[k] = __uint_as_float(cv[kk]);
						}
#pragma unroll
						for (int k = 0; k < N; ++k) {
							const bool on = static_cast<unsigned>(k) < vtaps;
							const float s0 = a0 + cc[k] * mv[k].x, s1 = a1 + cc[k] * mv[k].y;
							a0 = on ? s0 : a0;
							a1 = on ? s1 : a1;
						}
					} else {
						for (unsigned k = 0; k < vtaps; ++k) {
							const float2 m = *reinterpret_cast<const float2 *>(xr + k * pin);
							const float c = __uint_as_float(cv[k]);
							a0 += c * m.x;
							a1 += c * m.y;
						}
					}
					if (a.passes == 2) {
						*reinterpret_cast<float2 *>(reinterpret_cast<float *>(mid_base) + i * pm + 2 * q) = make_float2(a0, a1);
						continue;
					}
					const unsigned jc = 2 * q;
					if (jc < ncols)
						store_pair<false>(row_ptr(dst, a.dst.stride, row), j0 + jc, a0, a1, jc + 1 < ncols, a.out, row);
				}
			}
		}
	}
	if (a.passes == 1)
		return;
	__syncthreads();

	// Horizontal pass from the intermediate rows.
	if (og >= ogroups)
		return;
	if constexpr (WORD) {
		for (unsigned i = og; i < nrows; i += ogroups) {
			const unsigned row = i0 + i;
			int32_t a0, a1;
			h.apply_word(reinterpret_cast<const uint32_t *>(mid_base) + i * (pm / 2), a0, a1, kWordInit);
			store_pair<true>(row_ptr(dst, a.dst.stride, row), oj, pack_word_pre(a0, a.pixel_max), pack_word_pre(a1, a.pixel_max), otwo, a.out, row);
		}
	} else {
		h.with_float_form(a.p2.fw, [&](auto apply) {
			for (unsigned i = og; i < nrows; i += ogroups) {
				const unsigned row = i0 + i;
				float v0, v1;
				apply(reinterpret_cast<const float *>(mid_base) + i * pm, v0, v1);
				store_pair<false>(row_ptr(dst, a.dst.stride, row), oj, v0, v1, otwo, a.out, row);
			}
		});
	}
}

template <bool WORD, int N, bool HF>
__global__ void __launch_bounds__(kThreads) k_resize_tiled(ResizeArg a)
{
	extern __shared__ __align__(16) unsigned char smem[];
	resize_tile<WORD, N, HF>(a, blockIdx.x, blockIdx.y, blockIdx.z, smem);
}

// WORD resize on the int8 matrix cores, both passes in one launch
// (resize_line_h_u16_c and resize_line_v_u16_c, resize_impl.cpp:44-59,78-95,
// with the intermediate WORD row of pack_pixel_u16 in between). The
// arithmetic of the split is described beside ImmaBlock; every product and
// sum is an exact integer, so the outputs are the reference's.
//
// A workgroup owns kImmaTileW output columns and 16 rb output rows. It
// stages its source window (16 MB rows from the first vertical window, the
// columns of its 8 horizontal windows) in LDS as samples x ^ 0x8080 (both
// bytes signed, kImmaBias), then:
// horizontal pass: wave w takes column blocks 2w, 2w + 1 for each of the MB
// 16-row blocks; A = 16 rows x [xh | xl'] (lane l: row l % 16, window
// samples 4 (l / 16 % 2) + 0..3 and + 8..11, high bytes for l < 32, low
// bytes else; the coefficient table uses the same order). The
// result (lane: rows 4 (l / 16) + r of column l % 16) is stored as bytes
// (x ^ 0x80) into two LDS planes, high and low, one column after another
// (mid_pitch bytes each), so a column's rows are contiguous. Each wave then
// copies its output rows to the staged source area and stores them as whole
// row segments.
// Vertical pass: wave w takes row blocks w and w + 4; per column block
// A = 16 columns x [yh | yl'] of the block's window rows (three dwords
// aligned with v_alignbyte), B the block's coefficients, so D holds 4
// consecutive output columns of one output row per lane. mid_pitch = 2 mod 4
// dwords: the 16 columns of a read or write fall on even banks and the
// second group of 16 lanes on odd banks. All coefficients are loaded before
// the staging, so no memory round trip sits between the passes.
typedef int32_t imma_v4i __attribute__((ext_vector_type(4)));
// Staged samples x are stored as x ^ 0x8080: the high byte is the signed high
// byte of x - 32768 and the low byte the signed low - 128, both ready as
// matrix operands.
constexpr uint32_t kImmaBias = 0x80808080u;

// The kernel asks for at least 4 waves per SIMD (at most 128 VGPRs):
// otherwise the compiler keeps the matrix results in AGPRs and copies each
// one out.
__device__ __forceinline__ int32_t imma_sum(int32_t hh, int32_t mid, int32_t ll)
{
	// ((hh << 8) + mid) << 8 + ll = 65536 hh + 256 mid + ll in wrapping 32-bit
	// arithmetic (the true sum fits), as two shift-adds: the empty asm keeps
	// the compiler from reassociating them into two shifts and a three-way
	// add. (Only compiler-generated instructions read the matrix results, so
	// the compiler inserts the wait states those reads need.)
	uint32_t t = (static_cast<uint32_t>(hh) << 8) + static_cast<uint32_t>(mid);
	asm("" : "+v"(t));
	return static_cast<int32_t>((t << 8) + static_cast<uint32_t>(ll));
}
__device__ __forceinline__ int32_t imma_pack(int32_t hh, int32_t mid, int32_t ll, int32_t pixel_max)
{
	return med3_i32(imma_sum(hh, mid, ll) >> 14, 0, pixel_max);
}
// Full 16-bit range (pixel_max 65535), accumulators started 2^29 lower (the
// 32768 output bias taken out): two results as biased int16, y - 32768 =
// sat16(s >> 14) (v_cvt_pk_i16_i32), low half first.
__device__ __forceinline__ uint32_t imma_pack2_biased(const imma_v4i &hh, const imma_v4i &md, const imma_v4i &ll, int i)
{
	return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pk_i16(imma_sum(hh[i], md[i], ll[i]) >> 14, imma_sum(hh[i + 1], md[i + 1], ll[i + 1]) >> 14));
}

// The kernel is persistent: gridDim.x workgroups (a multiple of 8) walk the
// tiles, the workgroups of XCD x (blockIdx.x % 8 == x) interleaved over one
// contiguous range of tiles, so neighbouring tiles, which share source rows
// and columns, are in flight together in one L2. While a tile is multiplied
// and stored, the next tile's source chunks and coefficient records are
// already on their way into registers, so each tile's memory round trip
// overlaps the previous tile's work.
struct ImmaPrefetch {
	uint4 raw[kImmaStageU]; // source chunks of the staged window (16 bytes each)
	uint4 hc[2];            // coefficient records of the wave's two column blocks
	int32_t hinit[2];
	int hkb[2];
	int rbase, cb;          // first staged source row and column
	unsigned bx, by, f;
};

// FULL: pixel_max is 65535, so the clamp of both passes is the saturation of
// a biased int16 pack (imma_pack2_biased).
template <int MB, bool FULL>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4, 8))) k_resize_imma(ImmaArg a)
{
	constexpr int32_t kInitShift = FULL ? (1 << 29) : 0;
	extern __shared__ __align__(16) unsigned char smem[];
	const unsigned lane = threadIdx.x & 63;
	const unsigned w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const unsigned half = (lane >> 4) & 1;
	const unsigned nx = (a.dst_w + kImmaTileW - 1) / kImmaTileW, ny = (a.dst_h + 16 * a.rb - 1) / (16 * a.rb);
	const unsigned nt = nx * ny * a.frames;
	const unsigned xcd = blockIdx.x % 8, rank = blockIdx.x / 8, g8 = gridDim.x / 8;
	const unsigned q8 = nt / 8, r8 = nt % 8;
	const unsigned tbeg = xcd < r8 ? xcd * (q8 + 1) : r8 * (q8 + 1) + (xcd - r8) * q8;
	const unsigned tend = tbeg + q8 + (xcd < r8 ? 1u : 0u);
	unsigned char *midh = smem;
	unsigned char *midl = midh + kImmaTileW * a.mid_pitch;
	uint16_t *xin = reinterpret_cast<uint16_t *>(smem + 2 * kImmaTileW * a.mid_pitch);

	// Staging geometry, the same for every tile: prefetched chunk u of this
	// thread is chunk e = u kThreads + thread of the window (16 MB rows x
	// in_cols samples); threads past the end repeat the last chunk (same
	// value, same place), chunks past kImmaStageU kThreads are loaded when
	// the tile is staged.
	const unsigned esz = a.in.type == 0 ? 1u : 2u;
	const unsigned ce = 16 / esz;
	const unsigned cpr = a.in_cols / ce;
	const unsigned total = 16 * MB * cpr;
	const unsigned last_full = (a.src_w - ce) / ce * ce; // a.src_w >= ce (host-checked)
	unsigned srow[kImmaStageU], scol[kImmaStageU];
#pragma unroll
	for (int u = 0; u < kImmaStageU; ++u) {
		const unsigned e = min(u * kThreads + threadIdx.x, total - 1);
		srow[u] = e / cpr;
		scol[u] = (e - srow[u] * cpr) * ce;
	}

	// Blocks past the plane repeat the last one; their outputs are not stored.
	auto load_h = [&](ImmaPrefetch &p) {
		const unsigned hb0 = p.bx * (kImmaTileW / 16);
#pragma unroll
		for (int q = 0; q < 2; ++q) {
			const ImmaBlock &b = a.h[min(hb0 + 2 * w + q, a.nh - 1)];
			p.hc[q] = *reinterpret_cast<const uint4 *>(b.c[lane & 31]);
			p.hinit[q] = b.init[lane & 15];
			p.hkb[q] = b.kb;
		}
	};
	// The next tile's window goes into registers right after the current
	// tile is staged; its horizontal records once the current horizontal
	// pass is done; a tile's vertical records are loaded as the tile starts.
	auto prefetch = [&](unsigned t, ImmaPrefetch &p) {
		p.bx = t % nx;
		p.by = (t / nx) % ny;
		p.f = t / (nx * ny);
		const unsigned hb0 = p.bx * (kImmaTileW / 16), vb0 = p.by * a.rb;
		p.rbase = a.v[vb0].kb;
		p.cb = a.h[hb0].kb & ~15;
		const char *src = frame_ptr(a.src, p.f);
#pragma unroll
		for (int u = 0; u < kImmaStageU; ++u) {
			// clamped into the plane; chunks that are not whole are rebuilt when stored
			const unsigned r = min(static_cast<unsigned>(p.rbase) + srow[u], a.src_h - 1);
			const unsigned c = min(static_cast<unsigned>(p.cb) + scol[u], last_full);
			p.raw[u] = *reinterpret_cast<const uint4 *>(row_ptr(src, a.src.stride, r) + static_cast<size_t>(c) * esz);
		}
	};
	// One chunk of the window as biased samples in LDS; rows and columns past
	// the plane are zero samples (they meet only zero taps), so a chunk that
	// is not whole inside the plane is rebuilt sample by sample.
	auto store_chunk = [&](const ImmaPrefetch &p, const char *src, unsigned sr, unsigned sc, uint4 v) {
		const unsigned r = static_cast<unsigned>(p.rbase) + sr, c = static_cast<unsigned>(p.cb) + sc;
		if (r >= a.src_h || c + ce > a.src_w) {
			uint32_t q[4] = { 0, 0, 0, 0 };
			if (r < a.src_h) {
				const char *row = row_ptr(src, a.src.stride, r);
#pragma unroll
				for (unsigned k = 0; k < 16; ++k)
					if (k < ce && c + k < a.src_w) {
						const uint32_t x = load_as_uint(row, c + k, a.in.type);
						if (esz == 2)
							q[(k >> 1) & 3] |= x << (16 * (k & 1));
						else
							q[k >> 2] |= x << (8 * (k & 3));
					}
			}
			v = make_uint4(q[0], q[1], q[2], q[3]);
		}
		const unsigned sh = a.in.op == 1 ? a.in.shift : 0;
		uint16_t *d = xin + static_cast<size_t>(sr) * a.in_pitch + sc;
		const uint32_t wv[4] = { v.x, v.y, v.z, v.w };
		if (esz == 2) {
			*reinterpret_cast<uint4 *>(d) = make_uint4(shift_bias2(wv[0], sh, kImmaBias), shift_bias2(wv[1], sh, kImmaBias), shift_bias2(wv[2], sh, kImmaBias),
			                                           shift_bias2(wv[3], sh, kImmaBias));
		} else {
			uint32_t o[8];
#pragma unroll
			for (int k = 0; k < 4; ++k) {
				o[2 * k] = shift_bias2((wv[k] & 0xFFu) | ((wv[k] & 0xFF00u) << 8), sh, kImmaBias);
				o[2 * k + 1] = shift_bias2(((wv[k] >> 16) & 0xFFu) | ((wv[k] >> 8) & 0xFF0000u), sh, kImmaBias);
			}
			reinterpret_cast<uint4 *>(d)[0] = make_uint4(o[0], o[1], o[2], o[3]);
			reinterpret_cast<uint4 *>(d)[1] = make_uint4(o[4], o[5], o[6], o[7]);
		}
	};
	// The prefetched chunks, then (wide windows only) the rest of the window
	// loaded here.
	auto stage = [&](const ImmaPrefetch &p) {
		const char *src = frame_ptr(a.src, p.f);
#pragma unroll
		for (int u = 0; u < kImmaStageU; ++u)
			store_chunk(p, src, srow[u], scol[u], p.raw[u]);
		for (unsigned e = kImmaStageU * kThreads + threadIdx.x; e < total; e += kThreads) {
			const unsigned sr = e / cpr, sc = (e - sr * cpr) * ce;
			const unsigned r = min(static_cast<unsigned>(p.rbase) + sr, a.src_h - 1);
			const unsigned c = min(static_cast<unsigned>(p.cb) + sc, last_full);
			store_chunk(p, src, sr, sc, *reinterpret_cast<const uint4 *>(row_ptr(src, a.src.stride, r) + static_cast<size_t>(c) * esz));
		}
	};
	// The three B operands from a record's (ch, cl): [ch | 0], [cl | ch], [0 | cl].
	const bool lo_half = lane < 32;
	auto operands = [&](const uint4 &c, long &Bh, long &Bm, long &Bl) {
		const uint64_t ch = (static_cast<uint64_t>(c.y) << 32) | c.x, cl = (static_cast<uint64_t>(c.w) << 32) | c.z;
		Bh = static_cast<long>(lo_half ? ch : 0);
		Bm = static_cast<long>(lo_half ? cl : ch);
		Bl = static_cast<long>(lo_half ? 0 : cl);
	};

	ImmaPrefetch nxt;
	unsigned t = tbeg + rank;
	if (t < tend) {
		prefetch(t, nxt);
		load_h(nxt);
	}
	for (; t < tend; t += g8) {
		const ImmaPrefetch cur = nxt;
		uint4 vc[2];
		int32_t vinit[2];
		int vkb[2];
#pragma unroll
		for (int k = 0; k < 2; ++k) {
			const ImmaBlock &b = a.v[min(cur.by * a.rb + w + 4 * k, a.nv - 1)];
			vc[k] = *reinterpret_cast<const uint4 *>(b.c[lane & 31]);
			vinit[k] = b.init[lane & 15];
			vkb[k] = b.kb;
		}
		stage(cur);
		__syncthreads();
		const bool more = t + g8 < tend;
		if (more)
			prefetch(t + g8, nxt);
		const unsigned j0 = cur.bx * kImmaTileW, i0 = cur.by * 16 * a.rb;

		// Horizontal pass.
		{
			const uint32_t sel = lane < 32 ? 0x07050301u : 0x06040200u; // high or low bytes of 4 samples
			const uint32_t *xw = reinterpret_cast<const uint32_t *>(xin);
			const unsigned pw = a.in_pitch / 2;
#pragma unroll
			for (int m = 0; m < MB; ++m) {
				const uint32_t *row = xw + (16 * m + (lane & 15)) * pw;
#pragma unroll
				for (int q = 0; q < 2; ++q) {
					// lane group g holds window samples 4 (g % 2) + 0..3 and + 8..11:
					// the two groups of a half-wave read different banks of a row
					const uint32_t *x = row + (cur.hkb[q] - cur.cb) / 2 + 2 * half;
					const uint32_t d0 = x[0], d1 = x[1], d4 = x[4], d5 = x[5];
					const uint32_t ax = __builtin_amdgcn_perm(d1, d0, sel), ay = __builtin_amdgcn_perm(d5, d4, sel);
					const long A = static_cast<long>((static_cast<uint64_t>(ay) << 32) | ax);
					long Bh, Bm, Bl;
					operands(cur.hc[q], Bh, Bm, Bl);
					const imma_v4i zero = { 0, 0, 0, 0 };
					const int32_t hi0 = cur.hinit[q] - kInitShift;
					const imma_v4i init = { hi0, hi0, hi0, hi0 };
					const imma_v4i hh = __builtin_amdgcn_mfma_i32_16x16x32_i8(A, Bh, zero, 0, 0, 0);
					const imma_v4i md = __builtin_amdgcn_mfma_i32_16x16x32_i8(A, Bm, zero, 0, 0, 0);
					const imma_v4i ll = __builtin_amdgcn_mfma_i32_16x16x32_i8(A, Bl, init, 0, 0, 0);
					uint32_t lo, hi;
					if constexpr (FULL) {
						// bytes of y - 32768: the high byte is already the signed
						// high byte, the low byte takes ^ 0x80
						const uint32_t p01 = imma_pack2_biased(hh, md, ll, 0), p23 = imma_pack2_biased(hh, md, ll, 2);
						lo = __builtin_amdgcn_perm(p23, p01, 0x06040200u) ^ 0x80808080u;
						hi = __builtin_amdgcn_perm(p23, p01, 0x07050301u);
					} else {
						const uint32_t y0 = imma_pack(hh[0], md[0], ll[0], a.pixel_max), y1 = imma_pack(hh[1], md[1], ll[1], a.pixel_max);
						const uint32_t y2 = imma_pack(hh[2], md[2], ll[2], a.pixel_max), y3 = imma_pack(hh[3], md[3], ll[3], a.pixel_max);
						const uint32_t p01 = __builtin_amdgcn_perm(y1, y0, 0x05010400u); // y0.b0 y1.b0 y0.b1 y1.b1
						const uint32_t p23 = __builtin_amdgcn_perm(y3, y2, 0x05010400u);
						lo = __builtin_amdgcn_perm(p23, p01, 0x05040100u) ^ 0x80808080u;
						hi = __builtin_amdgcn_perm(p23, p01, 0x07060302u) ^ 0x80808080u;
					}
					const unsigned off = (16 * (2 * w + q) + (lane & 15)) * a.mid_pitch + 16 * m + 4 * (lane >> 4);
					*reinterpret_cast<uint32_t *>(midh + off) = hi;
					*reinterpret_cast<uint32_t *>(midl + off) = lo;
				}
			}
		}
		__syncthreads();
		if (more)
			load_h(nxt);

		// Vertical pass. The staged source is no longer read: each wave copies
		// its output rows there before storing them.
		char *dst = frame_ptr_w(a.dst, cur.f);
		uint16_t *ot = xin + w * 16 * kImmaOutPitch;
		const unsigned char *plane = (lane < 32 ? midh : midl) + (lane & 15) * a.mid_pitch;
#pragma unroll
		for (int k = 0; k < 2; ++k) {
			const unsigned vb = w + 4 * k;
			if (vb >= a.rb)
				break;
			const unsigned o = static_cast<unsigned>(vkb[k] - cur.rbase) + 8 * half;
			const unsigned sh = o & 3;
			long Bh, Bm, Bl;
			operands(vc[k], Bh, Bm, Bl);
			const imma_v4i zero = { 0, 0, 0, 0 };
			const int32_t vi0 = vinit[k] - kInitShift;
			const imma_v4i init = { vi0, vi0, vi0, vi0 };
			// Column block c + 1 is multiplied before block c is packed into the
			// wave's LDS copy of its 16 output rows.
			imma_v4i acc[2][3];
			auto mul = [&](unsigned c, imma_v4i *r) {
				const uint32_t *p = reinterpret_cast<const uint32_t *>(plane + 16 * c * a.mid_pitch + (o & ~3u));
				const uint32_t d0 = p[0], d1 = p[1], d2 = p[2];
				const uint32_t ax = __builtin_amdgcn_alignbyte(d1, d0, sh), ay = __builtin_amdgcn_alignbyte(d2, d1, sh);
				const long A = static_cast<long>((static_cast<uint64_t>(ay) << 32) | ax);
				r[0] = __builtin_amdgcn_mfma_i32_16x16x32_i8(A, Bh, zero, 0, 0, 0);
				r[1] = __builtin_amdgcn_mfma_i32_16x16x32_i8(A, Bm, zero, 0, 0, 0);
				r[2] = __builtin_amdgcn_mfma_i32_16x16x32_i8(A, Bl, init, 0, 0, 0);
			};
			mul(0, acc[0]);
#pragma unroll
			for (unsigned c = 0; c < kImmaTileW / 16; ++c) {
				if (c + 1 < kImmaTileW / 16)
					mul(c + 1, acc[(c + 1) & 1]);
				const imma_v4i hh = acc[c & 1][0], md = acc[c & 1][1], ll = acc[c & 1][2];
				uint2 out;
				if constexpr (FULL) {
					out = make_uint2(imma_pack2_biased(hh, md, ll, 0) ^ 0x80008000u, imma_pack2_biased(hh, md, ll, 2) ^ 0x80008000u);
				} else {
					const uint32_t y0 = imma_pack(hh[0], md[0], ll[0], a.pixel_max), y1 = imma_pack(hh[1], md[1], ll[1], a.pixel_max);
					const uint32_t y2 = imma_pack(hh[2], md[2], ll[2], a.pixel_max), y3 = imma_pack(hh[3], md[3], ll[3], a.pixel_max);
					out = make_uint2(pack16x2(y0, y1), pack16x2(y2, y3));
				}
				*reinterpret_cast<uint2 *>(ot + (lane & 15) * kImmaOutPitch + 16 * c + 4 * (lane >> 4)) = out;
			}
			__builtin_amdgcn_wave_barrier();
			// Whole row segments out: 16 lanes per row, 16 bytes each.
#pragma unroll
			for (unsigned u = 0; u < 4; ++u) {
				const unsigned e = 64 * u + lane, r = e / 16, chunk = e % 16;
				const unsigned row = i0 + 16 * vb + r, col = j0 + 8 * chunk;
				const uint4 v = *reinterpret_cast<const uint4 *>(ot + r * kImmaOutPitch + 8 * chunk);
				if (row < a.dst_h) {
					uint16_t *d = reinterpret_cast<uint16_t *>(row_ptr(dst, a.dst.stride, row)) + col;
					if (col + 8 <= a.dst_w) {
						*reinterpret_cast<uint4 *>(d) = v;
					} else {
						const uint32_t vv[4] = { v.x, v.y, v.z, v.w };
#pragma unroll
						for (unsigned i = 0; i < 7; ++i)
							if (col + i < a.dst_w)
								d[i] = static_cast<uint16_t>(vv[i / 2] >> (16 * (i & 1)));
					}
				}
			}
			__builtin_amdgcn_wave_barrier();
		}
		__syncthreads();
	}
}

// Persistent batched resize: each workgroup walks the (cost-ordered) tile
// list of many WORD resize descriptors (register capacity up to
// kJobsMaxTaps pairs), one tile at a time. Descriptors and tiles are
// read-only and not aliased, so their uniform reads are scalar loads.
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4, 8))) k_resize_jobs(const ResizeArg *__restrict__ descs, const JobTile *__restrict__ tiles,
                                                          unsigned ntiles)
{
	extern __shared__ __align__(16) unsigned char smem[];
	// Workgroup b takes runs of kJobRun consecutive tiles: runs b, b + G,
	// b + 2G, ... (G = gridDim.x). A run's tiles are neighbours of one
	// descriptor (the list is sorted by cost, stably), so its fields stay in
	// the scalar cache and the shared halo rows in L2.
	for (unsigned run = blockIdx.x; run * kJobRun < ntiles; run += gridDim.x) {
		const unsigned end = min(ntiles, (run + 1) * kJobRun);
		for (unsigned it = run * kJobRun; it < end; ++it) {
			const JobTile &jt = tiles[it];
			const ResizeArg &a = descs[jt.desc];
			switch (a.fwh) {
#define ZAMD_JOB(n) case n: if (a.h_first) resize_tile<true, n, true>(a, jt.bx, jt.by, jt.f, smem); else resize_tile<true, n, false>(a, jt.bx, jt.by, jt.f, smem); break;
			ZAMD_JOB(2) ZAMD_JOB(3) ZAMD_JOB(4) ZAMD_JOB(5) ZAMD_JOB(6) ZAMD_JOB(8) ZAMD_JOB(12)
#undef ZAMD_JOB
			default: break;
			}
			__syncthreads();
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
	if (a.alpha_op) {
		// PremultiplyFilter / UnpremultiplyFilter (simple_filters.cpp:76-110), FLOAT
		// planes: x * alpha (the reference multiplies by the unclamped alpha);
		// x / alpha with alpha clamped to [0, 1] and 0 where it is 0.
		const float x = reinterpret_cast<const float *>(srow)[j];
		const float al = reinterpret_cast<const float *>(row_ptr(frame_ptr(a.alpha, f), a.alpha.stride, i))[j];
		float r;
		if (a.alpha_op == 1) {
			r = x * al;
		} else {
			const float c = fminf(fmaxf(al, 0.0f), 1.0f);
			r = c == 0.0f ? 0.0f : x / c;
		}
		reinterpret_cast<float *>(drow)[j] = r;
		return;
	}
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

// Unresize (unresize_impl.cpp:27-85): the reference's float order, one row
// (horizontal) or one column (vertical) per thread; the band coefficients
// and LU factors of an output sample are the same for every thread, so they
// are scalar loads.
__global__ void __launch_bounds__(kThreads) k_unresize_h(UnresizeArg a)
{
	const unsigned f = blockIdx.y;
	const unsigned r = blockIdx.x * kThreads + threadIdx.x;
	if (r >= a.src_h)
		return;
	const float *s = reinterpret_cast<const float *>(row_ptr(frame_ptr(a.src, f), a.src.stride, r));
	float *d = reinterpret_cast<float *>(row_ptr(frame_ptr_w(a.dst, f), a.dst.stride, r));
	float z = 0.0f;
	for (unsigned j = 0; j < a.dst_w; ++j) {
		const float *cj = a.coef + static_cast<size_t>(j) * a.row_size;
		const float *x = s + a.offset[j];
		float accum = 0.0f;
		for (unsigned k = 0; k < a.row_size; ++k)
			accum += cj[k] * x[k];
		z = (accum - a.c[j] * z) * a.l[j];
		d[j] = z;
	}
	float w = 0.0f;
	for (unsigned j = a.dst_w; j != 0; --j) {
		w = d[j - 1] - a.u[j - 1] * w;
		d[j - 1] = w;
	}
}

__global__ void __launch_bounds__(kThreads) k_unresize_v(UnresizeArg a)
{
	const unsigned f = blockIdx.y;
	const unsigned j = blockIdx.x * kThreads + threadIdx.x;
	if (j >= a.src_w)
		return;
	const char *src = frame_ptr(a.src, f);
	char *dst = frame_ptr_w(a.dst, f);
	float z = 0.0f;
	for (unsigned i = 0; i < a.dst_h; ++i) {
		const float *ci = a.coef + static_cast<size_t>(i) * a.row_size;
		const unsigned top = a.offset[i];
		float accum = 0.0f;
		for (unsigned k = 0; k < a.row_size; ++k)
			accum += ci[k] * reinterpret_cast<const float *>(row_ptr(src, a.src.stride, top + k))[j];
		z = (accum - a.c[i] * z) * a.l[i];
		reinterpret_cast<float *>(row_ptr(dst, a.dst.stride, i))[j] = z;
	}
	float w = 0.0f;
	for (unsigned i = a.dst_h; i != 0; --i) {
		float *d = reinterpret_cast<float *>(row_ptr(dst, a.dst.stride, i - 1));
		w = d[j] - a.u[i - 1] * w;
		d[j] = w;
	}
}

// Floyd-Steinberg error diffusion (dither.cpp:57-89, 286-349) as a row-skewed
// wavefront. The plane is cut into strips of 64 rows; strip s belongs to wave
// s mod kEdWaves and starts at step s D. Lane l of the strip owns row
// 64 s + l and handles column j = T - s D - 2 l - 2 at step T (every lane,
// lane 0 included, spends two steps left of column 0 taking in the row
// above's first errors), so the three errors
// it needs from the row above were produced by lane l - 1 one, two and three
// steps earlier and arrive one lane over (DPP wave shift). Between strips the
// boundary row goes through a per-wave LDS ring. D is at least
// 2 * 63 + 2 + kEdChunk (a ring entry is written a chunk before it is read)
// and at least (W + 129) / kEdWaves (a wave finishes a strip before its next
// one starts), so strips run back to back without a pipeline drain. The float
// operation order is the sequential reference's, hence bit-exact output.
// Per 16-step chunk each lane reads its 16 inputs and writes its 16 outputs
// with paired accesses, and the ring is read and written once.
constexpr int kEdWaves = 16;
constexpr int kEdChunk = 16;
static_assert(kEdMaxWidth == kEdWaves * kEdMaxRing - 129, "error-diffusion width limit follows the ring limit");
constexpr int kEdMinDelay = 2 * 63 + 2 + kEdChunk;

struct EdGeom {
	int strip;   // strip this wave works on in the chunk, -1 none
	int jb;      // this lane's column at the chunk's first step
	int row;
	bool row_ok;
};

__device__ __forceinline__ EdGeom ed_geom(int T0, int w, int lane, int D, int nstrips, int W, int H)
{
	EdGeom g{ -1, 0, 0, false };
	if (T0 < w * D)
		return g;
	const int s = w + kEdWaves * ((T0 - w * D) / (kEdWaves * D));
	if (s >= nstrips || T0 - s * D > W + 2 * 63 + 2)
		return g;
	g.strip = s;
	g.jb = T0 - s * D - 2 * lane - 2;
	g.row = 64 * s + lane;
	g.row_ok = g.row < H;
	return g;
}

// 16 consecutive inputs of one row from column jb (zero outside [0, W)).
__device__ __forceinline__ void ed_load16(const char *srow, int jb, int W, int type, bool ok, float *x)
{
	if (ok && jb >= 0 && jb + kEdChunk <= W && type == 3 && (reinterpret_cast<uintptr_t>(srow + 4 * jb) & 7) == 0) {
		const float2 *p = reinterpret_cast<const float2 *>(srow + 4 * static_cast<size_t>(jb));
#pragma unroll
		for (int k = 0; k < kEdChunk / 2; ++k) {
			const float2 v = p[k];
			x[2 * k] = v.x;
			x[2 * k + 1] = v.y;
		}
		return;
	}
	if (ok && jb >= 0 && jb + kEdChunk <= W && type == 1 && (reinterpret_cast<uintptr_t>(srow + 2 * jb) & 3) == 0) {
		const uint32_t *p = reinterpret_cast<const uint32_t *>(srow + 2 * static_cast<size_t>(jb));
#pragma unroll
		for (int k = 0; k < kEdChunk / 2; ++k) {
			const uint32_t v = p[k];
			x[2 * k] = static_cast<float>(v & 0xFFFFu);
			x[2 * k + 1] = static_cast<float>(v >> 16);
		}
		return;
	}
#pragma unroll
	for (int k = 0; k < kEdChunk; ++k) {
		const int j = jb + k;
		x[k] = (ok && j >= 0 && j < W) ? load_as_float(srow, static_cast<unsigned>(j), type) : 0.0f;
	}
}

__global__ void __launch_bounds__(kEdWaves * 64) k_error_diffusion(EdArg a)
{
	extern __shared__ __align__(16) float ed_smem[];
	const int R = static_cast<int>(a.ring);
	float *ring = ed_smem; // [kEdWaves][R]

	const unsigned f = blockIdx.x;
	const unsigned pl = blockIdx.y;
	const PlaneArg &sp = a.src[pl];
	const PlaneArg &dp = a.dst[pl];
	const int lane = threadIdx.x & 63;
	const int w = threadIdx.x >> 6;
	const int W = static_cast<int>(a.width), H = static_cast<int>(a.height);
	const int D = static_cast<int>(a.delay);
	const char *src = frame_ptr(sp, f);
	char *dst = frame_ptr_w(dp, f);
	const float maxval = static_cast<float>(1UL << a.out.bits) - 1;
	const float scale = a.scale[pl], offset = a.offset[pl];
	const int nstrips = (H + 63) / 64;
	const int steps = (nstrips - 1) * D + W + 2 * 63 + 2;
	const int nchunks = (steps + kEdChunk - 1) / kEdChunk;
	float *my_ring = ring + w * R;
	const float *up_ring = ring + ((w + kEdWaves - 1) % kEdWaves) * R;

	float e_left = 0.0f, t_m1 = 0.0f, t_0 = 0.0f, t_p1 = 0.0f, prod = 0.0f;
	int cur = -1;
	float xin[kEdChunk], xnext[kEdChunk];
	EdGeom gn = ed_geom(0, w, lane, D, nstrips, W, H);
	ed_load16(row_ptr(src, sp.stride, gn.row_ok ? gn.row : 0), gn.jb, W, a.in.type, gn.strip >= 0 && gn.row_ok, xnext);
	// The first chunk's inputs arrive before the loop. Without this wait the
	// loop's entry state holds them in flight, and the compiler waits for all
	// memory operations (the next chunk's prefetch and the stores included)
	// before the first steps of every chunk.
	__builtin_amdgcn_s_waitcnt(0);

	for (int ch = 0; ch < nchunks; ++ch) {
		const int T0 = ch * kEdChunk;
		const EdGeom g = gn;
#pragma unroll
		for (int k = 0; k < kEdChunk; ++k)
			xin[k] = xnext[k];
		// Inputs of the next chunk are fetched now, off the step-to-step path.
		gn = ed_geom(T0 + kEdChunk, w, lane, D, nstrips, W, H);
		ed_load16(row_ptr(src, sp.stride, gn.row_ok ? gn.row : 0), gn.jb, W, a.in.type, gn.strip >= 0 && gn.row_ok, xnext);
		if (g.strip >= 0) {
			if (g.strip != cur) {
				cur = g.strip;
				e_left = t_m1 = t_0 = t_p1 = prod = 0.0f;
			}
			// Row above lane 0: the previous strip's last row, columns jb0 + 1 .. jb0 + 16.
			const int jb0 = T0 - g.strip * D - 2;
			float rv = 0.0f;
			if (lane < kEdChunk && g.strip > 0) {
				const int c = jb0 + 1 + lane;
				if (c >= 0 && c < W)
					rv = up_ring[c & (R - 1)];
			}
			float eout[kEdChunk];
			uint32_t q[kEdChunk];
#pragma unroll
			for (int t = 0; t < kEdChunk; ++t) {
				const int j = g.jb + t;
				// Lane l takes lane l - 1's error, lane 0 keeps `old`, the ring value:
				// row_bcast:15 into rows 1-3 (lanes 16, 32, 48 take lanes 15, 31, 47),
				// then row_shr:1 over it (measured faster than one wave_shr:1).
				const int old = __builtin_amdgcn_readlane(__float_as_int(rv), t);
				const int cross = __builtin_amdgcn_update_dpp(old, __float_as_int(prod), 0x142, 0xE, 0xF, false);
				const float recv = __int_as_float(__builtin_amdgcn_update_dpp(cross, __float_as_int(prod), 0x111, 0xF, 0xF, false));
				t_m1 = t_0;
				t_0 = t_p1;
				t_p1 = recv;
				prod = 0.0f;
				q[t] = 0;
				eout[t] = 0.0f;
				if (g.row_ok && j >= 0 && j < W) {
					float x = xin[t] * scale + offset;
					float err = 0;
					err += e_left * (7.0f / 16.0f);
					err += t_p1 * (3.0f / 16.0f);
					err += t_0 * (5.0f / 16.0f);
					err += t_m1 * (1.0f / 16.0f);
					x += err;
					x = x < 0.0f ? 0.0f : (maxval < x ? maxval : x);
					const float qf = __builtin_rintf(x);
					q[t] = static_cast<uint32_t>(static_cast<int>(qf));
					const float e = x - qf;
					e_left = e;
					prod = e;
					eout[t] = e;
				}
			}
			// Outputs of the chunk, in column pairs (jb is even).
			if (g.row_ok) {
				char *drow = row_ptr(dst, dp.stride, g.row);
#pragma unroll
				for (int t = 0; t < kEdChunk; t += 2) {
					const int j = g.jb + t;
					const bool v0 = j >= 0 && j < W, v1 = j + 1 >= 0 && j + 1 < W;
					const uintptr_t at = reinterpret_cast<uintptr_t>(drow) + (a.out.type == 0 ? 1 : 2) * static_cast<uintptr_t>(j);
					if (v0 && v1 && (at & (a.out.type == 0 ? 1 : 3)) == 0) {
						if (a.out.type == 0)
							*reinterpret_cast<uint16_t *>(drow + j) = static_cast<uint16_t>(q[t] | (q[t + 1] << 8));
						else
							*reinterpret_cast<uint32_t *>(drow + 2 * static_cast<size_t>(j)) = q[t] | (q[t + 1] << 16);
					} else if (v0 && v1) {
						store_uint(drow, static_cast<unsigned>(j), a.out.type, q[t]);
						store_uint(drow, static_cast<unsigned>(j + 1), a.out.type, q[t + 1]);
					} else if (v0) {
						store_uint(drow, static_cast<unsigned>(j), a.out.type, q[t]);
					} else if (v1) {
						store_uint(drow, static_cast<unsigned>(j + 1), a.out.type, q[t + 1]);
					}
				}
			}
			// The strip's last row feeds the next strip through this wave's ring.
			if (lane == 63) {
#pragma unroll
				for (int t = 0; t < kEdChunk; ++t) {
					const int j = g.jb + t;
					if (j >= 0 && j < W)
						my_ring[j & (R - 1)] = eout[t];
				}
			}
		}
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
#define ZAMD_TILED(n) \
	case n: \
		if (a.h_first) hipLaunchKernelGGL((k_resize_tiled<WORD, n, true>), grid, dim3(kThreads), lds, s, a); \
		else hipLaunchKernelGGL((k_resize_tiled<WORD, n, false>), grid, dim3(kThreads), lds, s, a); \
		break;
	ZAMD_TILED(2) ZAMD_TILED(3) ZAMD_TILED(4) ZAMD_TILED(5) ZAMD_TILED(6)
	ZAMD_TILED(8) ZAMD_TILED(12) ZAMD_TILED(16) ZAMD_TILED(24) ZAMD_TILED(32)
#undef ZAMD_TILED
	default: return static_cast<int>(hipErrorInvalidValue);
	}
	return static_cast<int>(hipGetLastError());
}

} // namespace

size_t resize_lds_bytes(const ResizeArg &a)
{
	if (a.fwh)
		return tiled_lds_bytes(a, nullptr, nullptr);
	const size_t elem = a.word ? 2 : 4;
	return elem * (static_cast<size_t>(a.lds_in_rows) * a.lds_in_cols + static_cast<size_t>(a.lds_mid_rows) * a.lds_mid_cols);
}

int launch_resize(const ResizeArg &a, void *stream)
{
	return a.word ? launch_resize_t<true>(a, stream) : launch_resize_t<false>(a, stream);
}

int launch_resize_jobs(const ResizeArg *descs, const JobTile *tiles, unsigned ntiles, unsigned grid, size_t lds, void *stream)
{
	if (!ntiles)
		return 0;
	hipLaunchKernelGGL(k_resize_jobs, dim3(grid), dim3(kThreads), lds, static_cast<hipStream_t>(stream), descs, tiles, ntiles);
	return static_cast<int>(hipGetLastError());
}

int launch_resize_imma(const ImmaArg &a, void *stream)
{
	if (!a.frames || !a.dst_w || !a.dst_h)
		return 0;
	const size_t tiles = static_cast<size_t>((a.dst_w + kImmaTileW - 1) / kImmaTileW) * ((a.dst_h + 16 * a.rb - 1) / (16 * a.rb)) * a.frames;
	// Persistent grid: as many workgroups as the LDS lets every CU hold at
	// once (a multiple of 8, one share per XCD), or fewer for small batches.
	int device = 0, cus = 0;
	if (hipGetDevice(&device) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
		return static_cast<int>(hipErrorInvalidDevice);
	const size_t per_cu = std::max<size_t>(1, std::min<size_t>(8, (160 * 1024) / std::max<unsigned>(a.lds, 1)));
	size_t grid = std::min<size_t>(static_cast<size_t>(cus) * per_cu, (tiles + 7) / 8 * 8);
	grid = std::max<size_t>(8, grid / 8 * 8);
	hipStream_t s = static_cast<hipStream_t>(stream);
	switch (a.mb) {
#define ZAMD_IMMA(n) \
	case n: \
		if (a.pixel_max == 65535) \
			hipLaunchKernelGGL((k_resize_imma<n, true>), dim3(static_cast<unsigned>(grid)), dim3(kThreads), a.lds, s, a); \
		else \
			hipLaunchKernelGGL((k_resize_imma<n, false>), dim3(static_cast<unsigned>(grid)), dim3(kThreads), a.lds, s, a); \
		break;
	ZAMD_IMMA(1) ZAMD_IMMA(2) ZAMD_IMMA(3) ZAMD_IMMA(4) ZAMD_IMMA(5) ZAMD_IMMA(6)
#undef ZAMD_IMMA
	default: return static_cast<int>(hipErrorInvalidValue);
	}
	return static_cast<int>(hipGetLastError());
}

int launch_point(const PointArg &a, void *stream)
{
	dim3 grid((a.width + kThreads - 1) / kThreads, a.height, a.frames);
	hipLaunchKernelGGL(k_point, grid, dim3(kThreads), 0, static_cast<hipStream_t>(stream), a);
	return static_cast<int>(hipGetLastError());
}

// Pack / unpack (grid.y = plane, grid.z = frame): one thread per planar sample.
__global__ void __launch_bounds__(kThreads) k_pack(PackArg a)
{
	const PackPlane &p = a.p[blockIdx.y];
	const unsigned f = blockIdx.z;
	const unsigned idx = blockIdx.x * kThreads + threadIdx.x;
	if (idx >= p.width * p.height)
		return;
	const unsigned i = idx / p.width, j = idx - i * p.width;
	const size_t e = static_cast<size_t>(p.step) * j + p.pos;
	if (a.pack) {
		const char *s = row_ptr(frame_ptr(p.planar, f), p.planar.stride, i);
		char *d = row_ptr(frame_ptr_w(p.packed, f), p.packed.stride, i);
		if (p.elem == 1)
			reinterpret_cast<uint8_t *>(d)[e] = static_cast<uint8_t>(reinterpret_cast<const uint8_t *>(s)[j] << p.shift);
		else
			reinterpret_cast<uint16_t *>(d)[e] = static_cast<uint16_t>(reinterpret_cast<const uint16_t *>(s)[j] << p.shift);
	} else {
		const char *s = row_ptr(frame_ptr(p.packed, f), p.packed.stride, i);
		char *d = row_ptr(frame_ptr_w(p.planar, f), p.planar.stride, i);
		if (p.elem == 1)
			reinterpret_cast<uint8_t *>(d)[j] = static_cast<uint8_t>(reinterpret_cast<const uint8_t *>(s)[e] >> p.shift);
		else
			reinterpret_cast<uint16_t *>(d)[j] = static_cast<uint16_t>(reinterpret_cast<const uint16_t *>(s)[e] >> p.shift);
	}
}

int launch_pack(const PackArg &a, void *stream)
{
	if (!a.frames || !a.planes)
		return 0;
	unsigned most = 0;
	for (unsigned q = 0; q < a.planes; ++q)
		most = std::max(most, a.p[q].width * a.p[q].height);
	dim3 grid((most + kThreads - 1) / kThreads, a.planes, a.frames);
	hipLaunchKernelGGL(k_pack, grid, dim3(kThreads), 0, static_cast<hipStream_t>(stream), a);
	return static_cast<int>(hipGetLastError());
}

int launch_unresize(const UnresizeArg &a, void *stream)
{
	if (!a.frames)
		return 0;
	const unsigned n = a.horizontal ? a.src_h : a.src_w;
	dim3 grid((n + kThreads - 1) / kThreads, a.frames);
	if (a.horizontal)
		hipLaunchKernelGGL(k_unresize_h, grid, dim3(kThreads), 0, static_cast<hipStream_t>(stream), a);
	else
		hipLaunchKernelGGL(k_unresize_v, grid, dim3(kThreads), 0, static_cast<hipStream_t>(stream), a);
	return static_cast<int>(hipGetLastError());
}

int launch_error_diffusion(const EdArg &a, void *stream)
{
	// Strip delay: a chunk of slack past the two-column skew of 64 rows, and
	// long enough that a wave finishes a strip before its next one starts.
	const unsigned need = (a.width + 2 * 63 + 3 + kEdWaves - 1) / kEdWaves;
	const unsigned delay = std::max<unsigned>(kEdMinDelay, (need + kEdChunk - 1) / kEdChunk * kEdChunk);
	unsigned ring = 256;
	while (ring < delay)
		ring *= 2;
	if (ring > kEdMaxRing)
		return static_cast<int>(hipErrorInvalidValue); // widths above kEdMaxWidth
	EdArg b = a;
	b.delay = delay;
	b.ring = ring;
	const size_t lds = sizeof(float) * static_cast<size_t>(kEdWaves) * ring;
	if (lds > 64 * 1024) {
		// Rings above 1024 entries take more than the default 64 KiB of
		// dynamic LDS (gfx950 has 160 KiB per CU).
		static std::atomic<bool> raised{ false };
		if (!raised.load(std::memory_order_acquire)) {
			const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_error_diffusion),
			                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(sizeof(float) * kEdWaves * kEdMaxRing));
			if (e != hipSuccess)
				return static_cast<int>(e);
			raised.store(true, std::memory_order_release);
		}
	}
	hipLaunchKernelGGL(k_error_diffusion, dim3(b.frames, b.planes), dim3(kEdWaves * 64), lds, static_cast<hipStream_t>(stream), b);
	return static_cast<int>(hipGetLastError());
}

} // namespace dev
} // namespace zamd
