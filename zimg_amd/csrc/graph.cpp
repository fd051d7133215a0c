// Graph planning. The decisions (which planes are resized in which pixel
// format, with which shift and sub-window, where depth conversions and the
// colourspace stage go) follow the reference's GraphBuilder step for step
// (graphbuilder.cpp:64-188 siting, :190-235 validation, :595-1072 lowering),
// because they fix the arithmetic and therefore the output bits. The result is
// a flat node list that the device layer lowers and fuses into kernels.
#include "graph.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <limits>

namespace zamd {

namespace {

constexpr int PY = 0, PU = 1, PV = 2, PA = 3;
using Mask = std::array<bool, 4>;
constexpr Mask LUMA{ true, false, false, false };
constexpr Mask CHROMA{ false, true, true, false };
constexpr Mask ALPHA{ false, false, false, true };
constexpr size_t DIM_MAX = static_cast<size_t>(1U) << (std::numeric_limits<size_t>::digits / 2 - 2);

Mask operator|(const Mask &a, const Mask &b) { return { a[0] || b[0], a[1] || b[1], a[2] || b[2], a[3] || b[3] }; }

// Offset of a chroma sample from the centred position, in chroma pixels (graphbuilder.cpp:64-107).
double chroma_offset_w(ChromaLocW loc, double subsampling) { return loc == ChromaLocW::LEFT ? -0.5 + 0.5 * subsampling : 0.0; }

double chroma_offset_h(ChromaLocH loc, double subsampling)
{
	if (loc == ChromaLocH::TOP) return -0.5 + 0.5 * subsampling;
	if (loc == ChromaLocH::BOTTOM) return 0.5 - 0.5 * subsampling;
	return 0.0;
}

// Field offsets relative to a progressive frame of field size (graphbuilder.cpp:109-188).
double field_adjust(FieldParity parity, double offset)
{
	if (parity == FieldParity::TOP) return -0.25 + offset / 2;
	if (parity == FieldParity::BOTTOM) return 0.25 + offset / 2;
	return offset;
}

double luma_field_offset(FieldParity parity)
{
	return parity == FieldParity::TOP ? -0.25 : parity == FieldParity::BOTTOM ? 0.25 : 0.0;
}

struct PlaneState {
	unsigned width = 0, height = 0;
	PixelFormat format;
	double left = 0, top = 0, aw = 0, ah = 0; // active region

	friend bool operator==(const PlaneState &a, const PlaneState &b)
	{
		return a.width == b.width && a.height == b.height && a.format == b.format && a.left == b.left && a.top == b.top &&
		       a.aw == b.aw && a.ah == b.ah;
	}
	friend bool operator!=(const PlaneState &a, const PlaneState &b) { return !(a == b); }
};

struct State {
	PlaneState p[4];
	ColorFamily color = ColorFamily::GREY;
	Colorspace csp;
	AlphaType alpha = AlphaType::NONE;

	bool has_chroma() const { return color != ColorFamily::GREY; }
	bool has_alpha() const { return alpha != AlphaType::NONE; }

	void chroma_from_luma(unsigned sw, unsigned sh)
	{
		const double scw = 1.0 / (1UL << sw), sch = 1.0 / (1UL << sh);
		p[PU] = p[PY];
		p[PU].width >>= sw;
		p[PU].height >>= sh;
		p[PU].format.chroma = color == ColorFamily::YUV;
		p[PU].left *= scw;
		p[PU].top *= sch;
		p[PU].aw *= scw;
		p[PU].ah *= sch;
		p[PV] = p[PU];
	}

	void chroma_from_luma_444()
	{
		p[PU] = p[PY];
		p[PU].format.chroma = color == ColorFamily::YUV;
		p[PV] = p[PU];
	}

	void alpha_from_luma()
	{
		p[PA] = p[PY];
		p[PA].format.fullrange = pixel_is_integer(p[PA].format.type);
	}

	State() = default;

	explicit State(const ImageState &s) : color(s.color), csp(s.colorspace), alpha(s.alpha)
	{
		p[PY].width = s.width;
		p[PY].height = s.height;
		p[PY].format.type = s.type;
		p[PY].format.depth = s.depth;
		p[PY].format.fullrange = pixel_is_integer(s.type) && s.fullrange;
		p[PY].format.chroma = false;
		p[PY].format.ycgco = csp.matrix == Matrix::YCGCO;
		p[PY].left = s.active_left;
		p[PY].top = s.active_top;
		p[PY].aw = s.active_width;
		p[PY].ah = s.active_height;
		if (color != ColorFamily::GREY)
			chroma_from_luma(s.subsample_w, s.subsample_h);
		if (alpha != AlphaType::NONE)
			alpha_from_luma();

		// Pixel siting (graphbuilder.cpp:348-366).
		p[PY].top -= luma_field_offset(s.parity);
		if (color != ColorFamily::GREY) {
			const double scw = static_cast<double>(p[PU].width) / p[PY].width;
			const double sch = static_cast<double>(p[PU].height) / p[PY].height;
			const double ow = chroma_offset_w(s.chroma_w, scw);
			const double oh = field_adjust(s.parity, chroma_offset_h(s.chroma_h, sch));
			p[PU].left -= ow;
			p[PU].top -= oh;
			p[PV].left -= ow;
			p[PV].top -= oh;
		}
		if (alpha != AlphaType::NONE)
			p[PA].top -= luma_field_offset(s.parity);
	}

	friend bool operator==(const State &a, const State &b)
	{
		return a.color == b.color && a.csp == b.csp && a.alpha == b.alpha && a.p[PY] == b.p[PY] &&
		       (!a.has_chroma() || a.p[PU] == b.p[PU]) && (!a.has_chroma() || a.p[PV] == b.p[PV]) &&
		       (!a.has_alpha() || a.p[PA] == b.p[PA]);
	}
	friend bool operator!=(const State &a, const State &b) { return !(a == b); }
};

std::string fmt_pixel(const PixelFormat &f)
{
	char buf[48];
	std::snprintf(buf, sizeof(buf), "%d/%u %c:%c%s", static_cast<int>(f.type), f.depth, f.fullrange ? 'f' : 'l', f.chroma ? 'c' : 'l',
	              f.ycgco ? " ycgco" : "");
	return buf;
}

[[noreturn]] void internal(const char *what) { fail(ZIMG_ERROR_UNKNOWN, std::string("invalid graph state: ") + what); }

class Planner {
	Plan &m_plan;
	const BuildParams &m_params;
	PlaneRef m_ids[4];
	State m_state;

	int add_node(Node n)
	{
		m_plan.nodes.push_back(std::move(n));
		return static_cast<int>(m_plan.nodes.size()) - 1;
	}

	template <class F>
	void each(const Mask &mask, F f)
	{
		for (int q = 0; q < 4; ++q)
			if (mask[q])
				f(q);
	}

	void trace(const std::string &s) { m_plan.trace.push_back(s); }

	State float_444(const State &s, bool include_alpha)
	{
		State r = s;
		r.p[PY].format = PixelType::FLOAT;
		if (r.has_chroma()) r.chroma_from_luma_444();
		if (r.has_alpha() && include_alpha) r.alpha_from_luma();
		return r;
	}

	bool needs_colorspace(const State &target)
	{
		if (m_state.csp == target.csp)
			return false;
		if (m_state.csp.transfer == target.csp.transfer && m_state.csp.primaries == target.csp.primaries) {
			if (m_state.color == ColorFamily::GREY) return false;
			if (m_state.color != ColorFamily::RGB && target.color == ColorFamily::GREY) return false;
		}
		return true;
	}

	bool needs_interpolation_plane(const State &t, int q)
	{
		const PlaneState &s = m_state.p[q], &d = t.p[q];
		auto phase = [](double x) { return std::modf(x, &x); };
		if (s.aw != d.aw || s.ah != d.ah) return true;
		if (phase(s.left) != phase(d.left) || phase(s.top) != phase(d.top)) return true;
		return false;
	}

	bool needs_interpolation(const State &t)
	{
		if (needs_interpolation_plane(t, PY)) return true;
		if (m_state.has_chroma() && t.has_chroma() && (needs_interpolation_plane(t, PU) || needs_interpolation_plane(t, PV))) return true;
		if (m_state.has_alpha() && t.has_alpha() && needs_interpolation_plane(t, PA)) return true;
		return false;
	}

	bool can_copy_tile(const State &t, int q)
	{
		auto is_int = [](double x) { return std::trunc(x) == x; };
		const PlaneState &s = m_state.p[q], &d = t.p[q];
		if (s.aw == s.width || s.ah == s.height) return false;
		if (d.aw != d.width || d.ah != d.height) return false;
		if (s.aw != d.width || s.ah != d.height) return false;
		const double l = s.left - d.left, tp = s.top - d.top;
		if (!is_int(l) || !is_int(tp)) return false;
		if (l < 0 || tp < 0) return false;
		if (l + s.aw > s.width || tp + s.ah > s.height) return false;
		return true;
	}

	bool needs_resize_plane(const State &t, int q)
	{
		PlaneState pl = t.p[q];
		pl.format = m_state.p[q].format;
		return m_state.p[q] != pl;
	}

	bool needs_premul(const State &t)
	{
		if (m_state.alpha != AlphaType::STRAIGHT) return false;
		return t.alpha != AlphaType::STRAIGHT || needs_colorspace(t) || needs_interpolation(t);
	}

	void yuv_to_grey()
	{
		trace("yuv_to_grey");
		m_ids[PU] = m_ids[PV] = PlaneRef{};
		m_state.color = ColorFamily::GREY;
	}

	void grey_to_rgb(Matrix m)
	{
		trace("grey_to_rgb");
		m_ids[PU] = m_ids[PY];
		m_ids[PV] = m_ids[PY];
		m_state.color = ColorFamily::RGB;
		m_state.csp.matrix = m;
		m_state.chroma_from_luma_444();
	}

	void grey_to_yuv(const State &t)
	{
		trace("grey_to_yuv");
		PixelFormat f = m_state.p[PY].format;
		f.chroma = true;
		Node n;
		n.kind = NodeKind::FILL;
		n.width = t.p[PU].width;
		n.height = t.p[PU].height;
		n.out = f;
		n.plane = PU;
		switch (f.type) {
		case PixelType::BYTE:
		case PixelType::WORD: n.fill_bits = 1U << (f.depth - 1); break;
		case PixelType::HALF: n.fill_bits = 0; break;
		case PixelType::FLOAT: n.fill_bits = 0; break;
		}
		int id = add_node(n);
		m_ids[PU] = PlaneRef{ id, 0 };
		m_ids[PV] = PlaneRef{ id, 0 };
		m_state.color = ColorFamily::YUV;
		m_state.csp.matrix = t.csp.matrix;
		m_state.p[PU] = t.p[PU];
		m_state.p[PU].format = f;
		m_state.p[PV] = t.p[PV];
		m_state.p[PV].format = f;
	}

	// premultiply / unpremultiply (graphbuilder.cpp:661-692): one node per
	// colour plane, reading the colour plane and the alpha plane (FLOAT 4:4:4).
	void alpha_op(NodeKind kind, AlphaType result)
	{
		if (m_state.p[PY].format.type != PixelType::FLOAT || m_state.p[PA].format.type != PixelType::FLOAT ||
		    (m_state.has_chroma() && (m_state.p[PU].width != m_state.p[PY].width || m_state.p[PU].height != m_state.p[PY].height)))
			internal("alpha operation outside FLOAT 4:4:4");
		trace(kind == NodeKind::PREMUL ? "premultiply" : "unpremultiply");
		for (int q = 0; q < (m_state.has_chroma() ? 3 : 1); ++q) {
			Node n;
			n.kind = kind;
			n.width = m_state.p[q].width;
			n.height = m_state.p[q].height;
			n.out = m_state.p[q].format;
			n.plane = q;
			n.dep[0] = m_ids[q].node;
			n.dep_out[0] = m_ids[q].out;
			n.dep[1] = m_ids[PA].node;
			n.dep_out[1] = m_ids[PA].out;
			m_ids[q] = PlaneRef{ add_node(n), 0 };
		}
		m_state.alpha = result;
	}

	void add_opaque_alpha(AlphaType type)
	{
		trace("add_opaque");
		PixelFormat f = m_state.p[PY].format;
		Node n;
		n.kind = NodeKind::FILL;
		n.width = m_state.p[PY].width;
		n.height = m_state.p[PY].height;
		n.out = f;
		n.plane = PA;
		switch (f.type) {
		case PixelType::BYTE: n.fill_bits = 0xFFU >> (8 - f.depth); break;
		case PixelType::WORD: n.fill_bits = 0xFFFFU >> (16 - f.depth); break;
		case PixelType::HALF: n.fill_bits = 0x3C00; break;
		case PixelType::FLOAT: n.fill_bits = 0x3F800000U; break;
		}
		m_ids[PA] = PlaneRef{ add_node(n), 0 };
		m_state.alpha = type;
		m_state.alpha_from_luma();
	}

	// graphbuilder.cpp:721-751
	PixelFormat choose_resize_format(const State &t, int q)
	{
		if (m_params.unresize)
			return PixelType::FLOAT;
		auto supported = [](PixelType ty) { return ty == PixelType::WORD || ty == PixelType::FLOAT; };
		const double src_pels = static_cast<double>(m_state.p[q].width) * m_state.p[q].height;
		const double dst_pels = static_cast<double>(t.p[q].width) * t.p[q].height;
		const PixelFormat sf = m_state.p[q].format, df = t.p[q].format;

		if (supported(sf.type) && supported(df.type)) return src_pels < dst_pels ? df : sf;
		if (supported(sf.type)) return sf;
		if (supported(df.type)) return df;
		if (sf.type == PixelType::BYTE && !sf.fullrange) return PixelFormat{ PixelType::WORD, 16, false, sf.chroma, sf.ycgco };
		return PixelType::FLOAT;
	}

	void add_resize_pass(const Mask &mask, bool horizontal, unsigned src_w, unsigned src_h, unsigned dst_dim, double shift,
	                     double subwidth, const ResampleFilter &f, const PixelFormat &fmt)
	{
		const unsigned src_dim = horizontal ? src_w : src_h;
		auto table = std::make_shared<FilterTable>(compute_filter(f, src_dim, dst_dim, shift, subwidth));
		each(mask, [&](int q) {
			Node n;
			n.kind = horizontal ? NodeKind::RESIZE_H : NodeKind::RESIZE_V;
			n.dep[0] = m_ids[q].node;
			n.dep_out[0] = m_ids[q].out;
			n.width = horizontal ? dst_dim : src_w;
			n.height = horizontal ? src_h : dst_dim;
			n.out = fmt;
			n.plane = q;
			n.table = table;
			n.depth_bits = fmt.depth;
			m_ids[q] = PlaneRef{ add_node(n), 0 };
		});
	}

	void add_unresize_pass(const Mask &mask, bool horizontal, unsigned up_w, unsigned up_h, unsigned orig_dim, double shift)
	{
		auto table = std::make_shared<UnresizeTable>(compute_unresize(horizontal ? up_w : up_h, orig_dim, shift));
		each(mask, [&](int q) {
			Node n;
			n.kind = horizontal ? NodeKind::UNRESIZE_H : NodeKind::UNRESIZE_V;
			n.dep[0] = m_ids[q].node;
			n.dep_out[0] = m_ids[q].out;
			n.width = horizontal ? orig_dim : up_w;
			n.height = horizontal ? up_h : orig_dim;
			n.out = m_state.p[q].format;
			n.plane = q;
			n.utable = table;
			m_ids[q] = PlaneRef{ add_node(n), 0 };
		});
	}

	// graphbuilder.cpp:753-824 and ResizeConversion::create (resize.cpp:38-102)
	void resize_plane(const State &t, const Mask &mask, int q)
	{
		if (!needs_resize_plane(t, q))
			return;
		if (m_params.unresize) {
			unresize_plane(t, mask, q);
			return;
		}
		const PlaneState &sp = m_state.p[q], &dp = t.p[q];
		const double scale_w = static_cast<double>(dp.aw) / sp.aw;
		const double scale_h = static_cast<double>(dp.ah) / sp.ah;
		const double shift_w = sp.left - dp.left / scale_w;
		const double shift_h = sp.top - dp.top / scale_h;
		const double subwidth = sp.aw * (dp.width / dp.aw);
		const double subheight = sp.ah * (dp.height / dp.ah);
		const ResampleFilter &f = (q == PU || q == PV) ? m_params.filter_uv : m_params.filter;
		const PixelFormat fmt = sp.format;

		char buf[160];
		std::snprintf(buf, sizeof(buf), "resize[%d]: [%u, %u] => [%u, %u] (%f, %f, %f, %f)", q, sp.width, sp.height, dp.width, dp.height,
		              shift_w, shift_h, subwidth, subheight);
		trace(buf);

		if (pixel_is_integer(fmt.type) && fmt.type != PixelType::WORD)
			internal("resize format");
		if (fmt.type == PixelType::HALF)
			fail(ZIMG_ERROR_UNSUPPORTED_OPERATION, "HALF resize not supported");

		const unsigned sw = sp.width, sh = sp.height, dw = dp.width, dh = dp.height;
		const bool skip_h = sw == dw && shift_w == 0 && subwidth == sw;
		const bool skip_v = sh == dh && shift_h == 0 && subheight == sh;

		if (!(skip_h && skip_v)) {
			if (skip_h) {
				add_resize_pass(mask, false, sw, sh, dh, shift_h, subheight, f, fmt);
			} else if (skip_v) {
				add_resize_pass(mask, true, sw, sh, dw, shift_w, subwidth, f, fmt);
			} else {
				const double xs = static_cast<double>(dw) / subwidth, ys = static_cast<double>(dh) / subheight;
				const double h_first_cost = std::max(xs, 1.0) * 2.0 + xs * std::max(ys, 1.0);
				const double v_first_cost = std::max(ys, 1.0) + ys * std::max(xs, 1.0) * 2.0;
				if (h_first_cost < v_first_cost) {
					add_resize_pass(mask, true, sw, sh, dw, shift_w, subwidth, f, fmt);
					add_resize_pass(mask, false, dw, sh, dh, shift_h, subheight, f, fmt);
				} else {
					add_resize_pass(mask, false, sw, sh, dh, shift_h, subheight, f, fmt);
					add_resize_pass(mask, true, sw, dh, dw, shift_w, subwidth, f, fmt);
				}
			}
		}
		each(mask, [&](int r) {
			PixelFormat keep = m_state.p[r].format;
			m_state.p[r] = t.p[r];
			m_state.p[r].format = keep;
		});
	}

	// graphbuilder.cpp:761-796 with UnresizeConversion::create (unresize/unresize.cpp:36-88).
	void unresize_plane(const State &t, const Mask &mask, int q)
	{
		const PlaneState &sp = m_state.p[q], &dp = t.p[q];
		if (sp.width != sp.aw || sp.height != sp.ah || dp.width != dp.aw || dp.height != dp.ah)
			fail(ZIMG_ERROR_RESAMPLING_NOT_AVAILABLE, "unresize not supported for for given subregion");
		const double scale_w = static_cast<double>(dp.aw) / sp.aw;
		const double scale_h = static_cast<double>(dp.ah) / sp.ah;
		const double shift_w = sp.left - dp.left / scale_w;
		const double shift_h = sp.top - dp.top / scale_h;
		if (sp.format.type != PixelType::FLOAT)
			internal("unresize format");

		char buf[160];
		std::snprintf(buf, sizeof(buf), "unresize[%d]: [%u, %u] => [%u, %u] (%f, %f)", q, sp.width, sp.height, dp.width, dp.height,
		              shift_w, shift_h);
		trace(buf);

		const unsigned uw = sp.width, uh = sp.height, ow = dp.width, oh = dp.height;
		const bool skip_h = uw == ow && shift_w == 0;
		const bool skip_v = uh == oh && shift_h == 0;
		if (skip_h && !skip_v) {
			add_unresize_pass(mask, false, uw, uh, oh, shift_h);
		} else if (skip_v && !skip_h) {
			add_unresize_pass(mask, true, uw, uh, ow, shift_w);
		} else if (!skip_h && !skip_v) {
			const double xs = static_cast<double>(ow) / uw, ys = static_cast<double>(oh) / uh;
			const double h_first_cost = std::max(xs, 1.0) * 2.0 + xs * std::max(ys, 1.0);
			const double v_first_cost = std::max(ys, 1.0) + ys * std::max(xs, 1.0) * 2.0;
			if (h_first_cost < v_first_cost) {
				add_unresize_pass(mask, true, uw, uh, ow, shift_w);
				add_unresize_pass(mask, false, ow, uh, oh, shift_h);
			} else {
				add_unresize_pass(mask, false, uw, uh, oh, shift_h);
				add_unresize_pass(mask, true, uw, oh, ow, shift_w);
			}
		}
		each(mask, [&](int r) {
			PixelFormat keep = m_state.p[r].format;
			m_state.p[r] = t.p[r];
			m_state.p[r].format = keep;
		});
	}

	void convert_colorspace(const Colorspace &csp)
	{
		if (m_state.color == ColorFamily::GREY) internal("colorspace on grey");
		if (m_state.csp == csp) return;

		CspParams cp;
		cp.peak_luminance = std::isnan(m_params.peak_luminance) ? 100.0 : m_params.peak_luminance;
		cp.approximate_gamma = m_params.approximate_gamma;
		cp.scene_referred = m_params.scene_referred;
		cp.chromatic_adaptation = m_params.chromatic_adaptation;

		char buf[160];
		std::snprintf(buf, sizeof(buf), "colorspace: [%d, %d, %d] => [%d, %d, %d] (%f)", static_cast<int>(m_state.csp.matrix),
		              static_cast<int>(m_state.csp.transfer), static_cast<int>(m_state.csp.primaries), static_cast<int>(csp.matrix),
		              static_cast<int>(csp.transfer), static_cast<int>(csp.primaries), cp.peak_luminance);
		trace(buf);

		CspProgram prog = build_colorspace_program(m_state.csp, csp, cp);
		if (prog.n) {
			Node n;
			n.kind = NodeKind::COLORSPACE;
			for (int i = 0; i < 3; ++i) {
				n.dep[i] = m_ids[i].node;
				n.dep_out[i] = m_ids[i].out;
			}
			n.width = m_state.p[PY].width;
			n.height = m_state.p[PY].height;
			n.out = PixelType::FLOAT;
			n.csp = prog;
			int id = add_node(n);
			for (int i = 0; i < 3; ++i)
				m_ids[i] = PlaneRef{ id, i };
		}
		if (csp.matrix == Matrix::RGB) {
			m_state.color = ColorFamily::RGB;
			m_state.p[PU].format.chroma = false;
			m_state.p[PV].format.chroma = false;
		} else {
			m_state.color = ColorFamily::YUV;
			m_state.p[PU].format.chroma = true;
			m_state.p[PV].format.chroma = true;
		}
		m_state.csp = csp;
	}

	// DepthConversion::create (depth.cpp:50-67) with the scale/offset of quantize.h:17-44.
	void convert_pixel_format(const PixelFormat &fmt, const Mask &mask, int q)
	{
		const PixelFormat in = m_state.p[q].format;
		if (in == fmt)
			return;
		trace("depth[" + std::to_string(q) + "]: [" + fmt_pixel(in) + "] => [" + fmt_pixel(fmt) + "]");

		if ((!in.fullrange && in.depth < 8 && pixel_is_integer(in.type)) || (!fmt.fullrange && fmt.depth < 8 && pixel_is_integer(fmt.type)))
			fail(ZIMG_ERROR_BIT_DEPTH_OVERFLOW, "bit depth must be at least 8 for limited range");

		auto int_offset = [](const PixelFormat &f) -> double {
			return pixel_is_float(f.type) ? 0 : f.chroma ? (1L << (f.depth - 1)) : !f.fullrange ? (16L << (f.depth - 8)) : 0;
		};
		auto int_range = [](const PixelFormat &f) -> double {
			return pixel_is_float(f.type) ? 1 : f.fullrange ? ((1L << f.depth) - 1) : (f.chroma && !f.ycgco) ? (224L << (f.depth - 8)) : (219L << (f.depth - 8));
		};
		const double ri = int_range(in), oi = int_offset(in), ro = int_range(fmt), oo = int_offset(fmt);

		Node base;
		base.kind = NodeKind::DEPTH;
		base.in = in;
		base.out = fmt;
		base.scale = static_cast<float>(ro / ri);
		base.offset = static_cast<float>(-oi * ro / ri + oo);

		const bool lossless = pixel_is_integer(in.type) && pixel_is_integer(fmt.type) && !in.fullrange && !fmt.fullrange &&
		                      in.chroma == fmt.chroma && fmt.depth >= in.depth;
		if (lossless) {
			base.depth_op = Node::SHIFT;
			base.shift = fmt.depth - in.depth;
		} else if (pixel_is_float(fmt.type)) {
			base.depth_op = Node::TO_FLOAT;
		} else {
			base.depth_op = Node::DITHER;
			base.dither = m_params.dither;
		}
		each(mask, [&](int r) {
			Node n = base;
			n.dep[0] = m_ids[r].node;
			n.dep_out[0] = m_ids[r].out;
			n.width = m_state.p[r].width;
			n.height = m_state.p[r].height;
			n.plane = r;
			m_ids[r] = PlaneRef{ add_node(n), 0 };
		});
		each(mask, [&](int r) { m_state.p[r].format = fmt; });
	}

	enum class Mode { LUMA, CHROMA, ALPHA };

	// graphbuilder.cpp:892-951
	void connect_plane(const State &target, Mode mode, bool reinterpret_range)
	{
		Mask mask{};
		State tmp = target;
		bool reinterpreted = false;
		int q;
		if (mode == Mode::LUMA) {
			mask = m_state.color == ColorFamily::RGB ? (LUMA | CHROMA) : LUMA;
			q = PY;
		} else if (mode == Mode::CHROMA) {
			mask = CHROMA;
			q = PU;
		} else {
			mask = ALPHA;
			q = PA;
		}

		if (reinterpret_range) {
			const PixelFormat sf = m_state.p[q].format, df = tmp.p[q].format;
			if (sf.type == PixelType::BYTE && sf.fullrange && df.fullrange && sf.depth == df.depth) {
				each(mask, [&](int r) { m_state.p[r].format.fullrange = false; });
				each(mask, [&](int r) { tmp.p[r].format.fullrange = false; });
				reinterpreted = true;
			}
		}

		if (can_copy_tile(tmp, q)) {
			const PixelFormat f = m_state.p[q].format;
			const unsigned l = static_cast<unsigned>(m_state.p[q].left - tmp.p[q].left);
			const unsigned t = static_cast<unsigned>(m_state.p[q].top - tmp.p[q].top);
			char buf[128];
			std::snprintf(buf, sizeof(buf), "subrectangle[%d]: [%u, %u, %u, %u]", q, l, t, tmp.p[q].width, tmp.p[q].height);
			trace(buf);
			each(mask, [&](int r) {
				Node n;
				n.kind = NodeKind::COPY_RECT;
				n.dep[0] = m_ids[r].node;
				n.dep_out[0] = m_ids[r].out;
				n.width = tmp.p[q].width;
				n.height = tmp.p[q].height;
				n.out = f;
				n.plane = r;
				n.crop_left = l;
				n.crop_top = t;
				m_ids[r] = PlaneRef{ add_node(n), 0 };
			});
			each(mask, [&](int r) {
				m_state.p[r] = tmp.p[r];
				m_state.p[r].format = f;
			});
			if (needs_resize_plane(tmp, q)) internal("copy tile left a resize");
		}

		if (needs_resize_plane(tmp, q)) {
			PixelFormat f = choose_resize_format(tmp, q);
			convert_pixel_format(f, mask, q);
			resize_plane(tmp, mask, q);
		}
		if (m_state.p[q].format != tmp.p[q].format)
			convert_pixel_format(tmp.p[q].format, mask, q);

		if (reinterpreted)
			each(mask, [&](int r) { m_state.p[r].format.fullrange = true; });

		if (m_state.p[q] != target.p[q]) internal("plane not connected");
	}

	void connect_color_planar(const State &t, bool reinterpret_range)
	{
		connect_plane(t, Mode::LUMA, reinterpret_range);
		if (m_state.color == ColorFamily::YUV)
			connect_plane(t, Mode::CHROMA, reinterpret_range);
	}

	// graphbuilder.cpp:967-1019
	void connect_color_channels(const State &target)
	{
		if (needs_colorspace(target)) {
			State tmp = float_444(m_state, false);
			const State &w = m_state.p[PY].width < target.p[PY].width ? m_state : target;
			const State &h = m_state.p[PY].height < target.p[PY].height ? m_state : target;
			tmp.p[PY].width = w.p[PY].width;
			tmp.p[PY].height = h.p[PY].height;
			tmp.p[PY].left = w.p[PY].left;
			tmp.p[PY].top = h.p[PY].top;
			tmp.p[PY].aw = w.p[PY].aw;
			tmp.p[PY].ah = h.p[PY].ah;
			if (tmp.has_chroma())
				tmp.chroma_from_luma_444();

			connect_color_planar(tmp, false);
			if (!m_state.has_chroma())
				grey_to_rgb(target.color == ColorFamily::RGB ? target.csp.matrix : Matrix::RGB);
			convert_colorspace(target.csp);
		}
		if (m_state.has_chroma() && !target.has_chroma())
			yuv_to_grey();
		if (!m_state.has_chroma() && !target.has_chroma() && m_state.csp.transfer == target.csp.transfer &&
		    m_state.csp.primaries == target.csp.primaries)
			m_state.csp.matrix = target.csp.matrix;

		connect_color_planar(target, true);

		if (m_state.color == ColorFamily::GREY && target.color == ColorFamily::RGB)
			grey_to_rgb(target.csp.matrix);
		if (m_state.color == ColorFamily::GREY && target.color == ColorFamily::YUV)
			grey_to_yuv(target);

		if (m_state.color != target.color || m_state.csp != target.csp || m_state.p[PY] != target.p[PY] ||
		    (m_state.has_chroma() && (m_state.p[PU] != target.p[PU] || m_state.p[PV] != target.p[PV])))
			internal("color channels not connected");
	}

public:
	Planner(Plan &plan, const BuildParams &params) : m_plan(plan), m_params(params) {}

	void run(const ImageState &src, const ImageState &dst)
	{
		m_state = State(src);
		State target(dst);

		auto add_source = [&](int q, unsigned w, unsigned h) {
			Node n;
			n.kind = NodeKind::SOURCE;
			n.width = w;
			n.height = h;
			n.out = m_state.p[q].format;
			n.plane = q;
			m_ids[q] = PlaneRef{ add_node(n), 0 };
		};
		add_source(PY, m_state.p[PY].width, m_state.p[PY].height);
		if (m_state.has_chroma()) {
			add_source(PU, m_state.p[PU].width, m_state.p[PU].height);
			add_source(PV, m_state.p[PV].width, m_state.p[PV].height);
		}
		if (m_state.has_alpha())
			add_source(PA, m_state.p[PA].width, m_state.p[PA].height);

		// connect_internal (graphbuilder.cpp:1021-1072).
		if (needs_premul(target)) {
			const auto orig_plane = m_state.p[PA];
			const PlaneRef orig_ref = m_ids[PA];
			const State tmp = float_444(m_state, true);
			connect_color_channels(tmp);
			connect_plane(tmp, Mode::ALPHA, false);
			alpha_op(NodeKind::PREMUL, AlphaType::PREMULTIPLIED);
			if (target.has_alpha() && target.p[PA] == orig_plane) {
				m_ids[PA] = orig_ref;
				m_state.p[PA] = orig_plane;
			}
		}
		if (m_state.has_alpha() && !target.has_alpha()) {
			trace("discard_alpha");
			m_ids[PA] = PlaneRef{};
			m_state.alpha = AlphaType::NONE;
		}
		if (m_state.alpha == AlphaType::PREMULTIPLIED && target.alpha == AlphaType::STRAIGHT) {
			const auto orig_plane = m_state.p[PA];
			const PlaneRef orig_ref = m_ids[PA];
			const State tmp = float_444(target, true);
			connect_color_channels(tmp);
			connect_plane(tmp, Mode::ALPHA, false);
			alpha_op(NodeKind::UNPREMUL, AlphaType::STRAIGHT);
			if (target.has_alpha() && target.p[PA] == orig_plane) {
				m_ids[PA] = orig_ref;
				m_state.p[PA] = orig_plane;
			}
		}

		connect_color_channels(target);
		if (m_state.has_alpha()) {
			if (m_state.alpha != target.alpha) internal("alpha type");
			connect_plane(target, Mode::ALPHA, true);
		}
		if (!m_state.has_alpha() && target.has_alpha())
			add_opaque_alpha(target.alpha);
		if (m_state != target)
			fail(ZIMG_ERROR_UNKNOWN, "failed to connect graph");

		// Sinks by buffer plane index: Y/R, U/G, V/B, A (plane 3 even for grey + alpha, zimg.h:390-418).
		m_plan.sink[PY] = m_ids[PY];
		if (m_state.has_chroma()) {
			m_plan.sink[PU] = m_ids[PU];
			m_plan.sink[PV] = m_ids[PV];
		}
		if (m_state.has_alpha())
			m_plan.sink[PA] = m_ids[PA];
	}
};

} // namespace

void validate_state(const ImageState &s)
{
	// graphbuilder.cpp:190-235
	if (!s.width || !s.height)
		fail(ZIMG_ERROR_INVALID_IMAGE_SIZE, "image dimensions must be non-zero");
	if (s.width > DIM_MAX || s.height > DIM_MAX)
		fail(ZIMG_ERROR_INVALID_IMAGE_SIZE, "image dimensions exceed implementation limit");
	if (static_cast<size_t>(s.width) * pixel_size(s.type) > static_cast<size_t>(std::numeric_limits<ptrdiff_t>::max()))
		fail(ZIMG_ERROR_INVALID_IMAGE_SIZE, "image width exceeds memory addressing limit");

	if (s.color == ColorFamily::GREY) {
		if (s.subsample_w || s.subsample_h)
			fail(ZIMG_ERROR_GREYSCALE_SUBSAMPLING, "cannot subsample greyscale image");
		if (s.colorspace.matrix == Matrix::RGB)
			fail(ZIMG_ERROR_COLOR_FAMILY_MISMATCH, "GREY color family cannot have RGB matrix coefficients");
	}
	if (s.color == ColorFamily::RGB) {
		if (s.subsample_w || s.subsample_h)
			fail(ZIMG_ERROR_UNSUPPORTED_SUBSAMPLING, "subsampled RGB image not supported");
		if (s.colorspace.matrix != Matrix::UNSPECIFIED && s.colorspace.matrix != Matrix::RGB)
			fail(ZIMG_ERROR_COLOR_FAMILY_MISMATCH, "RGB color family cannot have YUV matrix coefficients");
	}
	if (s.color == ColorFamily::YUV && s.colorspace.matrix == Matrix::RGB)
		fail(ZIMG_ERROR_COLOR_FAMILY_MISMATCH, "YUV color family cannot have RGB matrix coefficients");

	if (s.subsample_h > 1 && s.parity != FieldParity::PROGRESSIVE)
		fail(ZIMG_ERROR_UNSUPPORTED_SUBSAMPLING, "interlaced vertical subsampling greater than 2x is not supported");
	if (s.subsample_w > 2 || s.subsample_h > 2)
		fail(ZIMG_ERROR_UNSUPPORTED_SUBSAMPLING, "subsampling greater than 4x is not supported");
	if (s.width % (1 << s.subsample_w) || s.height % (1 << s.subsample_h))
		fail(ZIMG_ERROR_IMAGE_NOT_DIVISIBLE, "image dimensions must be divisible by subsampling factor");

	if (s.depth > pixel_depth(s.type))
		fail(ZIMG_ERROR_BIT_DEPTH_OVERFLOW, "bit depth exceeds limits of type");
	if (!s.fullrange && s.depth < 8)
		fail(ZIMG_ERROR_BIT_DEPTH_OVERFLOW, "bit depth must be at least 8 for limited range");

	if (!std::isfinite(s.active_left) || !std::isfinite(s.active_top) || !std::isfinite(s.active_width) || !std::isfinite(s.active_height))
		fail(ZIMG_ERROR_INVALID_IMAGE_SIZE, "active window must be finite");
	if (s.active_width <= 0 || s.active_height <= 0)
		fail(ZIMG_ERROR_INVALID_IMAGE_SIZE, "active window must be positive");
}

Plan build_plan(const ImageState &src, const ImageState &dst, const BuildParams &params)
{
	validate_state(src);
	validate_state(dst);
	if (dst.active_left != 0 || dst.active_top != 0 || dst.active_width != dst.width || dst.active_height != dst.height)
		fail(ZIMG_ERROR_RESAMPLING_NOT_AVAILABLE, "active subregion not supported on target image");

	Plan plan;
	Planner(plan, params).run(src, dst);

	plan.src_type = src.type;
	plan.dst_type = dst.type;
	auto dims = [](const ImageState &s, unsigned *w, unsigned *h) {
		w[PY] = s.width;
		h[PY] = s.height;
		if (s.color != ColorFamily::GREY) {
			w[PU] = w[PV] = s.width >> s.subsample_w;
			h[PU] = h[PV] = s.height >> s.subsample_h;
		}
		if (s.alpha != AlphaType::NONE) {
			w[PA] = s.width;
			h[PA] = s.height;
		}
	};
	dims(src, plan.src_w, plan.src_h);
	dims(dst, plan.dst_w, plan.dst_h);
	return plan;
}

std::string Plan::describe() const
{
	std::string s;
	for (auto &t : trace) {
		s += t;
		s += '\n';
	}
	return s;
}

} // namespace zamd
