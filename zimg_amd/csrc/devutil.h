// Device helpers shared by the gfx950 kernel files: plane addressing,
// storage-type loads and stores, IEEE half conversions, ordered / blue-noise
// dither offsets and quantisation (the reference's depth arithmetic,
// quantize.cpp and dither.cpp, in its operation order).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bluenoise.h"
#include "kernels.h"

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

// Ordered-dither offset of an 8-bit threshold t: (t + 1) / 257 - 1/2 in float
// (dither.cpp:137-232). Constant evaluation rounds each operation to nearest
// as the device's IEEE division and subtraction do, so the tables below hold
// the reference's values bit for bit.
__host__ __device__ constexpr float dither_unit(unsigned t)
{
	return static_cast<float>(static_cast<int>(t) + 1) / static_cast<float>(255u + 2u) - 0.5f;
}

// Bayer 16x16 threshold of position (r, c) = (row & 15, column & 15) of plane
// variant p (dither.cpp:137-217): index bits interleave (row xor col, col)
// from the top; variants flip horizontally, vertically, or transpose.
__host__ __device__ constexpr unsigned bayer_threshold(int plane, unsigned r, unsigned c)
{
	switch (plane & 3) {
	case 1: c = 15 - c; break;
	case 2: r = 15 - r; break;
	case 3: { const unsigned t = r; r = c; c = t; } break;
	default: break;
	}
	const unsigned x = r ^ c;
	unsigned v = 0;
	for (unsigned b = 0; b < 4; ++b) {
		v |= ((x >> b) & 1u) << (7 - 2 * b);
		v |= ((c >> b) & 1u) << (6 - 2 * b);
	}
	return v;
}

// Offsets per Bayer variant and position (variant * 256 + r * 16 + c), and per
// 8-bit threshold, built at compile time: one load per output instead of the
// index interleave and a division.
struct alignas(16) DitherTables {
	float bayer[4 * 256];
	float unit[256];
};

constexpr DitherTables make_dither_tables()
{
	DitherTables t{};
	for (unsigned p = 0; p < 4; ++p)
		for (unsigned r = 0; r < 16; ++r)
			for (unsigned c = 0; c < 16; ++c)
				t.bayer[p * 256 + r * 16 + c] = dither_unit(bayer_threshold(static_cast<int>(p), r, c));
	for (unsigned v = 0; v < 256; ++v)
		t.unit[v] = dither_unit(v);
	return t;
}

__constant__ DitherTables c_dither = make_dither_tables();

__device__ __forceinline__ float bayer_offset(int plane, unsigned i, unsigned j)
{
	return c_dither.bayer[(static_cast<unsigned>(plane) & 3) * 256 + (i & 15) * 16 + (j & 15)];
}

__constant__ uint8_t c_blue_noise[ZAMD_BLUE_NOISE_LEN * ZAMD_BLUE_NOISE_LEN] = { ZAMD_BLUE_NOISE_BYTES };

// Blue-noise offset for output pixel (i, j) of plane variant p (dither.cpp:219-232):
// each variant starts the 64x64 map at its own (column, row) offset.
__device__ __forceinline__ float blue_noise_offset(int plane, unsigned i, unsigned j)
{
	const unsigned variant = static_cast<unsigned>(plane) & 3;
	const unsigned hoff = variant == 0 ? 0 : variant == 1 ? 32 : variant == 2 ? 16 : 48;
	const unsigned voff = variant == 0 ? 0 : variant == 1 ? 12 : variant == 2 ? 55 : 26;
	return c_dither.unit[c_blue_noise[((i + voff) % ZAMD_BLUE_NOISE_LEN) * ZAMD_BLUE_NOISE_LEN + ((hoff + j) & (ZAMD_BLUE_NOISE_LEN - 1))]];
}

__device__ __forceinline__ float dither_offset(const OutXform &o, unsigned i, unsigned j)
{
	return o.dither == 1 ? bayer_offset(o.plane, i, j) : o.dither == 2 ? blue_noise_offset(o.plane, i, j) : 0.0f;
}

// Offsets of outputs (i, j .. j + 3) for j a multiple of 4 (one 16-byte load
// for Bayer).
__device__ __forceinline__ float4 dither_offset4(const OutXform &o, unsigned i, unsigned j)
{
	if (o.dither == 1)
		return *reinterpret_cast<const float4 *>(&c_dither.bayer[(static_cast<unsigned>(o.plane) & 3) * 256 + (i & 15) * 16 + (j & 15)]);
	if (o.dither == 2)
		return make_float4(blue_noise_offset(o.plane, i, j), blue_noise_offset(o.plane, i, j + 1), blue_noise_offset(o.plane, i, j + 2),
		                   blue_noise_offset(o.plane, i, j + 3));
	return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// dither_ordered (dither.cpp:37-55) for one value already converted to float,
// with its dither offset d.
__device__ __forceinline__ unsigned quantize_d(float x, const OutXform &o, float d)
{
	x = x * o.scale + o.offset;
	x += d;
	const float maxval = static_cast<float>(1UL << o.bits) - 1;
	x = x < 0.0f ? 0.0f : (maxval < x ? maxval : x);
	return static_cast<unsigned>(static_cast<int>(__builtin_rintf(x)));
}

__device__ __forceinline__ unsigned quantize(float x, const OutXform &o, unsigned i, unsigned j)
{
	return quantize_d(x, o, dither_offset(o, i, j));
}

__device__ __forceinline__ int32_t med3_i32(int32_t x, int32_t lo, int32_t hi)
{
	int32_t r;
	asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(lo), "v"(hi));
	return r;
}

// pack_pixel_u16 (resize_impl.cpp:36-42): ((acc + 2^13) >> 14) + 32768 clamped
// to [0, pixel_max]. Adding 32768 << 14 before the shift is the same integer.
__device__ __forceinline__ uint16_t pack_word(int32_t acc, int32_t pixel_max)
{
	return static_cast<uint16_t>(med3_i32((acc + ((32768 << 14) + (1 << 13))) >> 14, 0, pixel_max));
}

// Two 16-bit values (low halves of lo, hi) in one dword.
__device__ __forceinline__ uint32_t pack16x2(uint32_t lo, uint32_t hi)
{
	return __builtin_amdgcn_perm(hi, lo, 0x05040100u);
}

} // namespace
} // namespace dev
} // namespace zamd
