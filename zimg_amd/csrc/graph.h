// Graph planning: turns a (source format, target format, params) triple into
// the per-plane chain of stages, as the reference's GraphBuilder does
// (src/zimg/graph/graphbuilder.cpp:595-1261), and keeps that chain as a flat
// list of nodes for the device lowering.
#pragma once

#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "colorspace.h"
#include "common.h"
#include "filter.h"

namespace zamd {

enum class ColorFamily : int { GREY = 0, RGB = 1, YUV = 2 };
enum class FieldParity : int { PROGRESSIVE = 0, TOP = 1, BOTTOM = 2 };
enum class ChromaLocW : int { LEFT, CENTER };
enum class ChromaLocH : int { CENTER, TOP, BOTTOM };
enum class AlphaType : int { NONE = 0, STRAIGHT = 1, PREMULTIPLIED = 2 };

// GraphBuilder::state (graphbuilder.h:92-131).
struct ImageState {
	unsigned width = 0, height = 0;
	PixelType type = PixelType::BYTE;
	unsigned subsample_w = 0, subsample_h = 0;
	ColorFamily color = ColorFamily::GREY;
	Colorspace colorspace;
	unsigned depth = 0;
	bool fullrange = false;
	FieldParity parity = FieldParity::PROGRESSIVE;
	ChromaLocW chroma_w = ChromaLocW::LEFT;
	ChromaLocH chroma_h = ChromaLocH::CENTER;
	double active_left = 0, active_top = 0, active_width = 0, active_height = 0;
	AlphaType alpha = AlphaType::NONE;
};

// GraphBuilder::params (graphbuilder.h:133-172).
struct BuildParams {
	ResampleFilter filter{ ZIMG_RESIZE_BICUBIC, 0.0, 0.5 };
	ResampleFilter filter_uv{ ZIMG_RESIZE_BILINEAR, 0.0, 0.0 };
	Dither dither = Dither::NONE;
	double peak_luminance = NAN;
	bool approximate_gamma = false;
	bool scene_referred = false;
	bool chromatic_adaptation = false;
	bool unresize = false; // resample_filter ZIMG_RESIZE_UNRESIZE (zimg.cpp:21,322,450)
};

// PREMUL / UNPREMUL: FLOAT colour plane (dep 0) times / divided by the FLOAT
// alpha plane (dep 1), PremultiplyFilter / UnpremultiplyFilter (simple_filters.cpp:76-110).
// UNRESIZE_H / UNRESIZE_V: FLOAT unresize of one dimension (UnresizeImplH / V,
// unresize/unresize_impl.cpp:27-156).
enum class NodeKind : int { SOURCE, DEPTH, RESIZE_H, RESIZE_V, COLORSPACE, COPY_RECT, FILL, PREMUL, UNPREMUL, UNRESIZE_H, UNRESIZE_V };

// A node produces one plane (COLORSPACE produces three; consumers name the output index).
struct Node {
	NodeKind kind = NodeKind::SOURCE;
	int dep[3] = { -1, -1, -1 };
	int dep_out[3] = { 0, 0, 0 };
	unsigned width = 0, height = 0; // output plane size
	PixelFormat out;
	int plane = 0;                  // plane index the node was created for (dither table variant)

	// DEPTH
	PixelFormat in;
	enum DepthOp : int { SHIFT, TO_FLOAT, DITHER } depth_op = SHIFT;
	Dither dither = Dither::NONE;
	float scale = 1.0f, offset = 0.0f;
	unsigned shift = 0;

	// RESIZE_H / RESIZE_V
	std::shared_ptr<FilterTable> table;
	// UNRESIZE_H / UNRESIZE_V
	std::shared_ptr<UnresizeTable> utable;
	unsigned depth_bits = 16;       // WORD pixel maximum = 2^depth_bits - 1

	// COLORSPACE
	CspProgram csp;

	// COPY_RECT
	unsigned crop_left = 0, crop_top = 0;

	// FILL
	uint32_t fill_bits = 0;
};

struct PlaneRef {
	int node = -1;
	int out = 0;
};

struct Plan {
	std::vector<Node> nodes;
	PlaneRef sink[4];  // by buffer plane index (Y/R, U/G, V/B, A); node -1 = plane absent
	unsigned src_w[4] = {}, src_h[4] = {}, dst_w[4] = {}, dst_h[4] = {}; // 0 = plane absent
	PixelType src_type = PixelType::BYTE, dst_type = PixelType::BYTE;
	std::vector<std::string> trace; // observer-style description of every lowering decision

	std::string describe() const;
};

void validate_state(const ImageState &s);

// GraphBuilder::set_source + connect (graphbuilder.cpp:1083-1113,1214-1243).
Plan build_plan(const ImageState &src, const ImageState &dst, const BuildParams &params);

} // namespace zamd
