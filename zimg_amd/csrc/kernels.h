// Kernel launch interfaces of the gfx950 device path (implemented in kernels.hip).
// All argument structs are plain data passed by value as kernel arguments.
#pragma once

#include <cstddef>
#include <cstdint>

#include "colorspace.h"

namespace zamd {
namespace dev {

// A batch of planes in HBM: row r of frame f starts at ptr + f * fstride + r * stride.
struct PlaneArg {
	char *ptr;
	long long stride;
	long long fstride;
};

// Depth conversion applied while loading a stage's input.
struct InXform {
	int type;        // 0 BYTE, 1 WORD, 2 HALF, 3 FLOAT (source storage)
	int op;          // 0 none, 1 left shift into WORD, 2 integer -> float (x * scale + offset),
	                 // 3 exact conversion between floating-point storage types (no scale/offset)
	unsigned shift;
	float scale, offset;
};

// Depth conversion applied while storing a stage's output.
struct OutXform {
	int type;        // storage type of the output plane (HALF stores round to nearest even)
	int op;          // 0 store as is, 1 quantise (x * scale + offset + dither, clamp, round to even)
	int dither;      // 0 none, 1 ordered (Bayer 16x16), 2 random (64x64 blue noise)
	int plane;       // dither table variant (plane index mod 4)
	unsigned bits;   // output depth
	float scale, offset;
};

// One resize pass on the device: packed coefficient rows (stride fw) and left
// offsets; for WORD also the same taps as int16 pairs over a window that
// starts on an even sample (pl = left & ~1, a zero tap in front when left is
// odd, zeros behind), fwp = fw / 2 + 1 pairs per row.
struct PassArg {
	const unsigned *left;
	const int16_t *ci;
	const float *cf;
	int fw;
	int rows;
	const unsigned *pl;
	const int32_t *pc;
	int fwp;
};

struct ResizeArg {
	PlaneArg src, dst;
	unsigned frames;
	unsigned src_w, src_h, mid_w, mid_h, dst_w, dst_h;
	int h_first;        // 1: horizontal pass first
	int passes;         // 1 or 2
	PassArg p1, p2;
	int word;           // 1: integer WORD arithmetic (Q1.14), 0: FLOAT
	int pixel_max;      // WORD clamp
	InXform in;
	OutXform out;
	unsigned tile_w, tile_h;
	unsigned lds_in_rows, lds_in_cols;   // staged input tile (H-first)
	unsigned lds_mid_rows, lds_mid_cols; // intermediate tile
	int fwh;            // 0: generic kernel; else the tiled kernel's horizontal register capacity
	                    // (WORD: tap pairs, FLOAT: taps)
};

// Horizontal register capacities of the tiled resize kernel instantiations.
constexpr int kFastTapBuckets[] = { 2, 3, 4, 5, 6, 8, 12, 16, 24, 32 };
// Output columns of a horizontal-first tile (two per thread).
constexpr unsigned kFastTileW = 512;

// WORD resize on the int8 matrix cores (horizontal pass first, both passes
// in one launch, bit-exact). A biased sample x - 32768 = 256 xh + xl' + 128
// and a Q1.14 tap c = 256 ch + cl are split into signed bytes; a 16-output
// block of a pass is three 16x16x32 products over a 16-sample window,
// A = [xh | xl'] against [ch | 0] (hh), [cl | ch] (mid) and [0 | cl] (ll),
// and the int32 sum is 65536 hh + 256 mid + ll + init, with
// init = 128 sum(c) + the rounding and unbias terms of the pack.
// One record per block of 16 outputs: c[l] holds, for lane l < 32 of the
// MFMA operand (output l % 16), the 8 signed bytes ch and then the 8 bytes cl
// of its window samples: 8 (l / 16) + 0..7 for the vertical pass,
// 4 (l / 16) + 0..3 and + 8..11 for the horizontal pass. Lanes 32-63 take
// the same samples of the xl' half, so the three operands are [ch | 0],
// [cl | ch] and [0 | cl] over lanes 0-31 | 32-63. init[n] is the
// accumulator start of output n, kb the window's first input sample.
struct alignas(16) ImmaBlock {
	uint32_t c[32][4];
	int32_t init[16];
	int32_t kb;
	int32_t pad[3];
};
// Output columns of an integer matrix-core tile (8 blocks of 16, 2 per wave).
constexpr unsigned kImmaTileW = 128;
struct ImmaArg {
	PlaneArg src, dst;
	unsigned frames, src_w, src_h, dst_w, dst_h;
	int pixel_max;
	InXform in;
	const ImmaBlock *h, *v; // blocks of the horizontal and vertical pass
	unsigned nh, nv;
	unsigned rb;            // 16-row output blocks per tile (at most 8: two per wave)
	unsigned mb;            // 16-row blocks of intermediate rows per tile (1 .. kImmaMaxMb)
	unsigned mid_pitch;     // bytes per intermediate column of one byte plane
	unsigned in_cols, in_pitch; // staged source columns and their LDS row pitch (samples)
	unsigned lds;
};
constexpr unsigned kImmaMaxMb = 6;
// 16-byte source chunks each thread prefetches for the next tile (the rest
// of a wider window is loaded when the tile is staged).
constexpr int kImmaStageU = 4;
// Row pitch (samples) of a wave's LDS copy of 16 output rows of its tile;
// the staged source area holds at least 4 such copies.
constexpr unsigned kImmaOutPitch = kImmaTileW + 8;

struct PointArg {
	PlaneArg src, dst;
	unsigned frames, width, height;
	unsigned crop_left, crop_top;
	InXform in;
	OutXform out;
	int fill;           // 1: ignore src, write fill_bits
	uint32_t fill_bits;
	int alpha_op;       // 1: FLOAT src x alpha, 2: FLOAT src / alpha (premultiply, unpremultiply)
	PlaneArg alpha;     // FLOAT alpha plane of alpha_op
};

struct CspArg {
	PlaneArg src[3], dst[3];
	unsigned frames, width, height;
	InXform in[3];
	OutXform out[3];
	CspProgram prog;
};

// Unresize of one dimension (unresize_impl.cpp:27-156), FLOAT planes:
// horizontal along rows (one thread per row), vertical down columns (one
// thread per column). Output sample j: the band coef[j][0..row_size) from
// input sample offset[j], forward substitution with c and the inverted
// diagonal l, then back substitution with u.
struct UnresizeArg {
	PlaneArg src, dst;
	unsigned frames;
	unsigned src_w, src_h, dst_w, dst_h;
	int horizontal;
	unsigned row_size;
	const float *coef;
	const unsigned *offset;
	const float *c, *l, *u;
};

// Device pack / unpack between a planar plane and its samples inside a
// packed layout (NV12 / P010 semi-planar chroma, YUY2 / UYVY 4:2:2): sample
// j of row i of the planar plane is element step * j + pos of row i of the
// packed plane, stored there shifted left by `shift` bits.
struct PackPlane {
	PlaneArg planar, packed;
	unsigned width, height;
	unsigned elem;      // bytes per sample (1 BYTE, 2 WORD)
	unsigned step, pos, shift;
};
struct PackArg {
	PackPlane p[4];
	unsigned planes, frames;
	int pack;           // 1: planar -> packed, 0: packed -> planar
};
int launch_pack(const PackArg &a, void *stream);

// Widest plane the error-diffusion wavefront takes. The LDS ring of rows
// handed between strips holds up to kEdMaxRing floats per wave (16 waves:
// 128 KiB), and a plane W wide needs a ring of at least (W + 144) / 16
// entries (rounded down), so W <= 16 kEdMaxRing - 129.
constexpr unsigned kEdMaxRing = 2048;
constexpr unsigned kEdMaxWidth = 16 * kEdMaxRing - 129;

// Error diffusion over up to 4 planes of equal size in one launch (grid.y = plane).
struct EdArg {
	PlaneArg src[4], dst[4];
	unsigned frames, width, height, planes;
	InXform in;        // load only (op must be 0)
	OutXform out;      // op 1, dither ignored
	float scale[4], offset[4];
	unsigned delay, ring; // set by the launcher: strip delay (steps) and LDS ring length
};

// Banded-contraction tables of one FLOAT resize pass for the matrix-core
// precision form: outputs in tiles of 16; tile t reads inputs kb .. kb +
// 32 nks - 1 (kb a multiple of 8); its coefficients are nonzero only below
// kend (a multiple of 8); its coefficient blocks start at block off; a block
// is one 32-wide k step: 64 lanes x 8 high 16-bit values, then 64 x 8 low
// values (lane l: output 16 t + l % 16, inputs kb + 32 s + 8 (l / 16) + i).
// Blocks of consecutive tiles are consecutive.
struct alignas(16) MfmaTile {
	int kb, kend;
	unsigned nks, off;
};
struct MfmaTab {
	const uint16_t *coef;
	const MfmaTile *tile;
};

struct MfmaArg {
	PlaneArg src, dst;
	unsigned frames;
	int horizontal;     // 1: along rows (rows x src_len -> rows x dst_len), 0: along columns
	int bf16;           // 0: FP16 hi+lo, 1: BF16 hi+lo
	unsigned src_len, dst_len;
	unsigned rows;      // horizontal pass: rows processed
	unsigned cols;      // vertical pass: columns processed
	int vec16;          // 1: source rows are 16-byte aligned (FLOAT rows may be read as float4)
	MfmaTab tab;
	InXform in;
	OutXform out;
};

// Both passes of a two-pass FLOAT resize in one launch, horizontal first.
// A workgroup owns kFusedWaves x 16 output columns and tile_h output rows;
// it streams its source window through LDS in 16-row slabs (hi/lo 16-bit
// planes, row stride slab_stride halves), keeps the intermediate columns in
// LDS (hi/lo planes, column stride mid_stride halves) next to its vertical
// coefficient blocks, and writes the output.
constexpr unsigned kFusedWaves = 4;
constexpr unsigned kFusedMaxSteps = 4;   // k steps per tile and pass
constexpr unsigned kFusedMaxWindow = 256; // source columns per workgroup
constexpr unsigned kFusedMaxRowTiles = 8; // 16-row output tiles per workgroup (tile_h <= 128)
constexpr unsigned kFusedMaxGroups = 12;  // 16-row intermediate groups per workgroup
constexpr unsigned kFusedMaxCoefBlocks = 8; // vertical coefficient blocks per workgroup
struct MfmaFusedArg {
	PlaneArg src, dst;
	unsigned frames;
	int bf16;
	unsigned src_w, src_h, dst_w, dst_h;
	unsigned tile_h;       // output rows per workgroup, a multiple of 16
	unsigned slab_stride;  // halves, >= the widest source window
	unsigned mid_stride;   // halves, >= 16 x the most intermediate row groups
	unsigned lds_bytes;
	int vec_in;            // FLOAT source without conversion, 16-byte aligned rows
	int vec_out;           // FLOAT output stored as is, 16-byte aligned rows
	int stage_out;         // with vec_in and vec_out: the output tile is written through LDS (tile_h rows of 68 floats after the slabs)
	MfmaTab th, tv;        // horizontal (src_w -> dst_w) and vertical (src_h -> dst_h) tables
	InXform in;
	OutXform out;
};

// One work item of a batched launch: output tile (bx, by) of frame f of
// resize descriptor desc.
struct JobTile {
	unsigned desc, bx, by, f;
};

// Launchers. Return hipSuccess (0) or the HIP error of the launch.
int launch_resize(const ResizeArg &a, void *stream);
int launch_point(const PointArg &a, void *stream);
int launch_colorspace(const CspArg &a, void *stream);
// Whether launch_colorspace runs the program on a compile-time chain kernel
// (straight-line code for that operation list) rather than the general one.
bool colorspace_has_chain_kernel(const CspProgram &prog);
int launch_error_diffusion(const EdArg &a, void *stream);
int launch_mfma_resize(const MfmaArg &a, void *stream);
int launch_unresize(const UnresizeArg &a, void *stream);
int launch_mfma_fused(const MfmaFusedArg &a, void *stream);
int launch_resize_imma(const ImmaArg &a, void *stream);
// Largest register capacity (tap pairs) the persistent batched kernel handles.
constexpr int kJobsMaxTaps = 12;
// Consecutive tiles of the persistent walker's list a workgroup takes at a time.
constexpr unsigned kJobRun = 8;
// Persistent launch over WORD tiled resize descriptors (device arrays):
// `grid` workgroups walk the tile list with a stride of `grid`.
int launch_resize_jobs(const ResizeArg *descs, const JobTile *tiles, unsigned ntiles, unsigned grid, size_t lds, void *stream);

// LDS bytes the resize kernel needs for a tile configuration (host-side sizing).
size_t resize_lds_bytes(const ResizeArg &a);

} // namespace dev
} // namespace zamd
