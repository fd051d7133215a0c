// Device lowering and execution. The planned node list is turned into one
// launch per node, then adjacent launches are fused when the intermediate
// plane has a single reader, so a typical graph runs as a few kernels whose
// HBM traffic is close to "read the input once, write the output once":
//   depth-in -> resize H -> resize V -> dither   => one k_resize launch
//   to-float x3 -> colourspace -> dither x3      => one k_colorspace launch
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <array>
#include <cstring>

#include "device.h"
#include "zmath.h"

namespace zamd {

namespace {

constexpr size_t kPitchAlign = 256;
constexpr size_t kLdsBudget = 64 * 1024;

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

void check_hip(hipError_t e, const char *what)
{
	if (e != hipSuccess)
		fail(ZIMG_ERROR_UNKNOWN, std::string(what) + ": " + hipGetErrorString(e));
}

int type_code(PixelType t) { return static_cast<int>(t); }

} // namespace

DeviceProgram::DeviceProgram(const Plan &plan) : m_plan(plan)
{
	lower();
	fuse();
	for (auto &s : m_steps)
		if (!s.dead && s.kind == Step::RESIZE) {
			configure_resize(s);
			s.imma = imma_plan(s);
		}
	m_steps.erase(std::remove_if(m_steps.begin(), m_steps.end(), [](const Step &s) { return s.dead; }), m_steps.end());

	// Drop scratch slots that fusion made unreachable and renumber the rest.
	std::vector<int> remap(m_slots.size(), -1);
	std::vector<Slot> live;
	for (auto &s : m_steps) {
		for (Buf *b : { &s.in[0], &s.in[1], &s.in[2], &s.out[0], &s.out[1], &s.out[2] }) {
			if (b->kind != Buf::TMP)
				continue;
			if (remap[b->idx] < 0) {
				remap[b->idx] = static_cast<int>(live.size());
				live.push_back(m_slots[b->idx]);
			}
			b->idx = remap[b->idx];
		}
	}
	m_slots = std::move(live);
}

DeviceProgram::~DeviceProgram()
{
	for (auto &kv : m_tables) {
		if (kv.second && kv.second->mem)
			(void)hipFree(kv.second->mem);
		if (kv.second && kv.second->lut_mem)
			(void)hipFree(kv.second->lut_mem);
		if (kv.second && kv.second->unresize_mem)
			(void)hipFree(kv.second->unresize_mem);
	}
	for (auto &kv : m_mfma_tables)
		if (kv.second && kv.second->mem)
			(void)hipFree(kv.second->mem);
	for (auto &kv : m_imma_tables)
		if (kv.second && kv.second->mem)
			(void)hipFree(kv.second->mem);
}

void DeviceProgram::lower()
{
	const auto &nodes = m_plan.nodes;
	auto key = [](int node, int out) { return node * 4 + out; };
	std::map<int, int> readers;
	std::map<int, std::vector<int>> sink_planes;

	for (const Node &n : nodes) {
		const int ndeps = n.kind == NodeKind::COLORSPACE ? 3
		                  : (n.kind == NodeKind::PREMUL || n.kind == NodeKind::UNPREMUL) ? 2
		                  : (n.kind == NodeKind::SOURCE || n.kind == NodeKind::FILL ? 0 : 1);
		for (int d = 0; d < ndeps; ++d)
			readers[key(n.dep[d], n.dep_out[d])]++;
	}
	for (int q = 0; q < 4; ++q)
		if (m_plan.sink[q].node >= 0)
			sink_planes[key(m_plan.sink[q].node, m_plan.sink[q].out)].push_back(q);

	std::map<int, Buf> where;
	auto new_slot = [&](unsigned w, unsigned h, unsigned bpp) {
		m_slots.push_back(Slot{ w, h, bpp, align_up(static_cast<size_t>(w) * bpp, kPitchAlign) });
		return Buf{ Buf::TMP, static_cast<int>(m_slots.size()) - 1 };
	};
	auto place = [&](int id, int out, const PixelFormat &fmt) {
		const Node &n = nodes[id];
		const int k = key(id, out);
		auto sp = sink_planes.find(k);
		Buf b;
		if (n.kind == NodeKind::SOURCE)
			b = Buf{ Buf::SRC, n.plane };
		else if (sp != sink_planes.end() && sp->second.size() == 1 && readers[k] == 0)
			b = Buf{ Buf::DST, sp->second[0] };
		else
			b = new_slot(n.width, n.height, pixel_size(fmt.type));
		where[k] = b;
	};

	for (int id = 0; id < static_cast<int>(nodes.size()); ++id) {
		const Node &n = nodes[id];
		if (n.kind == NodeKind::COLORSPACE) {
			for (int o = 0; o < 3; ++o)
				place(id, o, PixelFormat{ PixelType::FLOAT });
		} else {
			place(id, 0, n.out);
		}
		if (n.kind == NodeKind::SOURCE)
			continue;

		Step s;
		s.width = n.width;
		s.height = n.height;
		s.out[0] = where[key(id, 0)];
		char label[96];

		if (n.kind == NodeKind::COLORSPACE) {
			s.kind = Step::CSP;
			s.prog = n.csp;
			for (int k = 0; k < 3; ++k) {
				s.in[k] = where[key(n.dep[k], n.dep_out[k])];
				s.ix[k] = dev::InXform{ type_code(PixelType::FLOAT), 0, 0, 1.0f, 0.0f };
				s.out[k] = where[key(id, k)];
				s.ox[k] = dev::OutXform{ type_code(PixelType::FLOAT), 0, 0, 0, 32, 1.0f, 0.0f };
			}
			s.label = "colorspace[" + n.csp.describe() + "]";
			m_steps.push_back(s);
			continue;
		}

		const Node *dep = n.dep[0] >= 0 ? &nodes[n.dep[0]] : nullptr;
		const PixelType in_type = dep ? (dep->kind == NodeKind::COLORSPACE ? PixelType::FLOAT : dep->out.type) : PixelType::BYTE;
		if (dep)
			s.in[0] = where[key(n.dep[0], n.dep_out[0])];
		s.ix[0] = dev::InXform{ type_code(in_type), 0, 0, 1.0f, 0.0f };
		s.ox[0] = dev::OutXform{ type_code(n.out.type), 0, 0, n.plane, n.out.depth, 1.0f, 0.0f };

		switch (n.kind) {
		case NodeKind::DEPTH:
			if (n.depth_op == Node::SHIFT) {
				s.kind = Step::POINT;
				s.ix[0].op = 1;
				s.ix[0].shift = n.shift;
				std::snprintf(label, sizeof(label), "shift<<%u[%d]", n.shift, n.plane);
			} else if (n.depth_op == Node::TO_FLOAT) {
				// select_depth_convert_func (depth_convert.cpp:80-100): integers take
				// x * scale + offset; HALF <-> FLOAT is an exact conversion.
				s.kind = Step::POINT;
				if (pixel_is_float(in_type)) {
					if (in_type == n.out.type)
						fail(ZIMG_ERROR_UNKNOWN, "no conversion between pixel types");
					s.ix[0].op = 3;
				} else {
					s.ix[0].op = 2;
					s.ix[0].scale = n.scale;
					s.ix[0].offset = n.offset;
				}
				std::snprintf(label, sizeof(label), "%s[%d]", n.out.type == PixelType::HALF ? "to_half" : "to_float", n.plane);
			} else if (n.dither == Dither::ERROR_DIFFUSION) {
				if (n.width > dev::kEdMaxWidth)
					fail(ZIMG_ERROR_UNSUPPORTED_OPERATION, "error diffusion supports planes up to " + std::to_string(dev::kEdMaxWidth) +
					                                           " samples wide");
				s.kind = Step::ED;
				s.ed_scale[0] = n.scale;
				s.ed_offset[0] = n.offset;
				s.ox[0].op = 1;
				std::snprintf(label, sizeof(label), "error_diffusion[%d]", n.plane);
			} else {
				s.kind = Step::POINT;
				s.ox[0].op = 1;
				s.ox[0].dither = n.dither == Dither::ORDERED ? 1 : n.dither == Dither::RANDOM ? 2 : 0;
				s.ox[0].scale = n.scale;
				s.ox[0].offset = n.offset;
				std::snprintf(label, sizeof(label), "%s[%d]",
				              n.dither == Dither::ORDERED ? "dither_bayer" : n.dither == Dither::RANDOM ? "dither_blue" : "quantize", n.plane);
			}
			s.label = label;
			break;
		case NodeKind::RESIZE_H:
		case NodeKind::RESIZE_V:
			s.kind = Step::RESIZE;
			s.passes = 1;
			s.h_first = n.kind == NodeKind::RESIZE_H;
			s.tab[0] = id;
			s.word = n.out.type == PixelType::WORD;
			s.pixel_max = static_cast<int>((1UL << n.depth_bits) - 1);
			s.in_w = dep->width;
			s.in_h = dep->height;
			s.mid_w = s.in_w;
			s.mid_h = s.in_h;
			std::snprintf(label, sizeof(label), "resize_%c(fw=%u)[%d]", s.h_first ? 'h' : 'v', n.table->filter_width, n.plane);
			s.label = label;
			break;
		case NodeKind::UNRESIZE_H:
		case NodeKind::UNRESIZE_V:
			if (in_type != PixelType::FLOAT)
				fail(ZIMG_ERROR_UNKNOWN, "unresize of a non-FLOAT plane");
			s.kind = Step::UNRESIZE;
			s.h_first = n.kind == NodeKind::UNRESIZE_H;
			s.tab[0] = id;
			s.in_w = dep->width;
			s.in_h = dep->height;
			std::snprintf(label, sizeof(label), "unresize_%c[%d]", s.h_first ? 'h' : 'v', n.plane);
			s.label = label;
			break;
		case NodeKind::COPY_RECT:
			s.kind = Step::POINT;
			s.crop_left = n.crop_left;
			s.crop_top = n.crop_top;
			s.label = "copy_rect";
			break;
		case NodeKind::FILL:
			s.kind = Step::POINT;
			s.fill = 1;
			s.fill_bits = n.fill_bits;
			s.label = "fill";
			break;
		case NodeKind::PREMUL:
		case NodeKind::UNPREMUL:
			if (in_type != PixelType::FLOAT || n.out.type != PixelType::FLOAT)
				fail(ZIMG_ERROR_UNKNOWN, "alpha operation on a non-FLOAT plane");
			s.kind = Step::POINT;
			s.alpha_op = n.kind == NodeKind::PREMUL ? 1 : 2;
			s.in[1] = where[key(n.dep[1], n.dep_out[1])];
			s.ix[1] = dev::InXform{ type_code(PixelType::FLOAT), 0, 0, 1.0f, 0.0f };
			std::snprintf(label, sizeof(label), "%s[%d]", s.alpha_op == 1 ? "premultiply" : "unpremultiply", n.plane);
			s.label = label;
			break;
		default:
			fail(ZIMG_ERROR_UNKNOWN, "unexpected node");
		}
		m_steps.push_back(s);
	}

	// Sinks whose value lives elsewhere (a source plane, a shared node) get a copy.
	for (int q = 0; q < 4; ++q) {
		const PlaneRef &r = m_plan.sink[q];
		if (r.node < 0)
			continue;
		Buf b = where[key(r.node, r.out)];
		if (b.kind == Buf::DST && b.idx == q)
			continue;
		const Node &n = nodes[r.node];
		const PixelType t = n.kind == NodeKind::COLORSPACE ? PixelType::FLOAT : n.out.type;
		Step s;
		s.kind = Step::POINT;
		s.in[0] = b;
		s.out[0] = Buf{ Buf::DST, q };
		s.width = n.width;
		s.height = n.height;
		s.ix[0] = dev::InXform{ type_code(t), 0, 0, 1.0f, 0.0f };
		s.ox[0] = dev::OutXform{ type_code(t), 0, 0, q, 0, 1.0f, 0.0f };
		s.label = "copy[" + std::to_string(q) + "]";
		m_steps.push_back(s);
	}
}

void DeviceProgram::fuse()
{
	auto same = [](const Buf &a, const Buf &b) { return a.kind == b.kind && a.idx == b.idx; };
	auto readers_of = [&](const Buf &b) {
		int n = 0;
		for (auto &s : m_steps) {
			if (s.dead) continue;
			const int nin = s.kind == Step::CSP ? 3 : s.alpha_op ? 2 : 1;
			for (int k = 0; k < nin; ++k)
				if (same(s.in[k], b)) ++n;
		}
		return n;
	};
	auto plain_in = [](const dev::InXform &x) { return x.op == 0; };
	auto plain_out = [](const dev::OutXform &x) { return x.op == 0; };
	constexpr int kFloat = 3, kHalf = 2;
	auto is_point_depth_in = [&](const Step &s) {
		return s.kind == Step::POINT && !s.alpha_op && !s.fill && !s.crop_left && !s.crop_top && s.ix[0].op != 0 && s.ox[0].op == 0 &&
		       (s.ix[0].op == 1 || s.ox[0].type == kFloat);
	};
	auto is_point_to_half = [&](const Step &s) {
		return s.kind == Step::POINT && !s.alpha_op && !s.fill && !s.crop_left && !s.crop_top && s.ix[0].op == 3 && s.ix[0].type == kFloat &&
		       s.ox[0].op == 0 && s.ox[0].type == kHalf;
	};
	auto is_point_quant = [](const Step &s) {
		return s.kind == Step::POINT && !s.alpha_op && !s.fill && !s.crop_left && !s.crop_top && s.ix[0].op == 0 && s.ox[0].op == 1;
	};

	bool changed = true;
	while (changed) {
		changed = false;
		for (size_t ia = 0; ia < m_steps.size() && !changed; ++ia) {
			Step &A = m_steps[ia];
			if (A.dead)
				continue;
			const int nout = A.kind == Step::CSP ? 3 : 1;
			for (int ko = 0; ko < nout && !changed; ++ko) {
				const Buf mid = A.out[ko];
				if (mid.kind != Buf::TMP || readers_of(mid) != 1)
					continue;
				for (size_t ib = ia + 1; ib < m_steps.size() && !changed; ++ib) {
					Step &B = m_steps[ib];
					if (B.dead)
						continue;
					const int nin = B.kind == Step::CSP ? 3 : B.alpha_op ? 2 : 1;
					int kin = -1;
					for (int k = 0; k < nin; ++k)
						if (same(B.in[k], mid)) kin = k;
					if (kin < 0)
						continue;

					// depth-in (shift / to-float) folded into a resize or colourspace input
					if (is_point_depth_in(A) && (B.kind == Step::RESIZE || B.kind == Step::CSP) && plain_in(B.ix[kin])) {
						const bool to_float = A.ix[0].op == 2 || A.ix[0].op == 3;
						const bool ok = B.kind == Step::CSP ? to_float : (B.word ? A.ix[0].op == 1 : to_float);
						if (ok) {
							B.in[kin] = A.in[0];
							B.ix[kin] = A.ix[0];
							B.label = A.label + "+" + B.label;
							A.dead = true;
							changed = true;
						}
					}
					// two single resize passes in opposite directions -> one fused launch
					else if (A.kind == Step::RESIZE && B.kind == Step::RESIZE && A.passes == 1 && B.passes == 1 && A.h_first != B.h_first &&
					         A.word == B.word && plain_out(A.ox[0]) && plain_in(B.ix[0]) && m_plan.nodes[A.tab[0]].table->sorted() &&
					         m_plan.nodes[B.tab[0]].table->sorted()) {
						B.passes = 2;
						B.h_first = A.h_first;
						B.tab[1] = B.tab[0];
						B.tab[0] = A.tab[0];
						B.in[0] = A.in[0];
						B.ix[0] = A.ix[0];
						B.in_w = A.in_w;
						B.in_h = A.in_h;
						B.mid_w = A.width;
						B.mid_h = A.height;
						B.label = A.label + "+" + B.label;
						A.dead = true;
						changed = true;
					}
					// quantisation / dither / FLOAT -> HALF folded into the producer's store
					else if ((A.kind == Step::RESIZE || A.kind == Step::CSP) && plain_out(A.ox[ko]) &&
					         (is_point_quant(B) || (is_point_to_half(B) && A.ox[ko].type == kFloat))) {
						A.out[ko] = B.out[0];
						A.ox[ko] = B.ox[0];
						A.label += "+" + B.label;
						B.dead = true;
						changed = true;
					}
				}
			}
		}
	}

	// Error diffusion of equally sized planes runs as one launch (one workgroup
	// per plane and frame), provided nothing in between touches their buffers.
	for (size_t ia = 0; ia < m_steps.size(); ++ia) {
		Step &A = m_steps[ia];
		if (A.dead || A.kind != Step::ED)
			continue;
		for (size_t ib = ia + 1; ib < m_steps.size() && A.ed_planes < 3; ++ib) {
			Step &B = m_steps[ib];
			if (B.dead || B.kind != Step::ED || B.width != A.width || B.height != A.height || B.ix[0].type != A.ix[0].type ||
			    B.ox[0].type != A.ox[0].type || B.ox[0].bits != A.ox[0].bits)
				continue;
			bool clash = false;
			for (size_t k = ia + 1; k < ib && !clash; ++k) {
				const Step &M = m_steps[k];
				if (M.dead)
					continue;
				for (int o = 0; o < 3; ++o)
					if (same(M.out[o], B.in[0]) || same(M.out[o], B.out[0]) || same(M.in[o], B.out[0]))
						clash = true;
			}
			if (clash)
				continue;
			const int p = A.ed_planes++;
			A.in[p] = B.in[0];
			A.out[p] = B.out[0];
			A.ed_scale[p] = B.ed_scale[0];
			A.ed_offset[p] = B.ed_offset[0];
			A.label += "+" + B.label;
			B.dead = true;
		}
	}
}

namespace {

// Even-aligned start of a WORD pass row and its pair count (the tiled kernel's table form).
unsigned pair_left(const FilterTable &t, unsigned o) { return t.left[o] & ~1u; }
unsigned pair_count(const FilterTable &t) { return t.filter_width / 2 + 1; }

} // namespace

// Every LDS and table index the tiled kernel forms for this configuration,
// recomputed on the host with the kernel's formulas: the staged windows fit
// the LDS arrays and every tap window lies inside its staged window.
void DeviceProgram::check_tiled(const Step &s, const TiledConfig &c, const FilterTable *th_tab, const FilterTable *tv_tab) const
{
	auto bad = [](const char *what) { fail(ZIMG_ERROR_UNKNOWN, std::string("tiled resize configuration: ") + what); };
	const unsigned tw = c.tile_w, th = c.tile_h;
	if (s.h_first && tw != dev::kFastTileW)
		bad("tile width");
	if (!s.h_first && !th_tab && tw % 16)
		bad("vertical-only tile width");
	const bool w = s.word;
	for (unsigned j0 = 0; j0 < s.width; j0 += tw) {
		const unsigned nc_out = std::min(tw, s.width - j0), jl = j0 + nc_out - 1;
		unsigned cb = j0, nc = (nc_out + 15) & ~15u;
		if (th_tab) {
			cb = (w ? pair_left(*th_tab, j0) : th_tab->left[j0]) & ~15u;
			nc = ((w ? pair_left(*th_tab, jl) + 2 * static_cast<unsigned>(c.fwh) : th_tab->left[jl] + th_tab->filter_width) - cb + 15) & ~15u;
			for (unsigned j = j0; j <= jl; ++j) {
				const unsigned lo = w ? pair_left(*th_tab, j) : th_tab->left[j];
				const unsigned width = w ? 2 * static_cast<unsigned>(c.fwh) : th_tab->filter_width;
				if (lo < cb || lo - cb + width > nc || (w && (lo - cb) % 2))
					bad("horizontal window");
			}
			if (static_cast<int>(w ? pair_count(*th_tab) : th_tab->filter_width) > c.fwh)
				bad("horizontal register capacity");
		}
		if (nc > c.in_cols || (th_tab && !s.h_first && nc > c.mid_cols) || cb % 16 || nc % 16)
			bad("staged columns");
	}
	for (unsigned i0 = 0; i0 < s.height; i0 += th) {
		const unsigned nr = std::min(th, s.height - i0), il = i0 + nr - 1;
		unsigned r0 = i0, span = nr;
		if (tv_tab) {
			const bool pairs = w;
			r0 = pairs ? pair_left(*tv_tab, i0) : tv_tab->left[i0];
			span = (pairs ? pair_left(*tv_tab, il) + 2 * pair_count(*tv_tab) : tv_tab->left[il] + tv_tab->filter_width) - r0;
			for (unsigned i = i0; i <= il; ++i) {
				const unsigned top = pairs ? pair_left(*tv_tab, i) : tv_tab->left[i];
				if (top < r0 || top - r0 + (pairs ? 2 * pair_count(*tv_tab) : tv_tab->filter_width) > span || (pairs && (top - r0) % 2))
					bad("vertical window");
			}
			if (pairs && (span % 2 || r0 % 2))
				bad("row pairs");
		}
		if (span > c.in_rows || (s.h_first && tv_tab && span > c.mid_rows) || (!s.h_first && th_tab && nr > c.mid_rows))
			bad("staged rows");
	}
	if (s.h_first && tv_tab && c.mid_cols < dev::kFastTileW)
		bad("intermediate width");
}

void DeviceProgram::configure_resize(Step &s)
{
	const FilterTable &t1 = *m_plan.nodes[s.tab[0]].table;
	const FilterTable *t2 = s.passes == 2 ? m_plan.nodes[s.tab[1]].table.get() : nullptr;
	const size_t elem = s.word ? 2 : 4;
	const bool sorted1 = t1.sorted();

	// Tiled kernel: sorted tables, horizontal taps within a register bucket.
	const FilterTable *th_tab = s.h_first ? &t1 : t2;
	const FilterTable *tv_tab = s.h_first ? t2 : &t1;
	const bool sorted_all = sorted1 && (!t2 || t2->sorted());
	int fwh = 0;
	if (sorted_all) {
		// The horizontal taps set the register bucket; a vertical-only step takes
		// it from its vertical taps (its tap loop is unrolled over the bucket,
		// wider filters take the kernel's tap loop in the smallest bucket).
		const unsigned need = th_tab ? (s.word ? pair_count(*th_tab) : th_tab->filter_width)
		                             : tv_tab ? (s.word ? pair_count(*tv_tab) : tv_tab->filter_width) : 1;
		for (int b : dev::kFastTapBuckets) {
			if (need <= static_cast<unsigned>(b)) {
				fwh = b;
				break;
			}
		}
		if (!fwh && !th_tab)
			fwh = dev::kFastTapBuckets[0];
	}
	if (fwh) {
		// Horizontal window [cb, cb + nc) of output columns [j0, j0 + nc_out) exactly as the kernel computes it.
		auto h_cols = [&](const FilterTable &t, unsigned j0, unsigned ncols) {
			const unsigned jl = j0 + ncols - 1;
			const unsigned cb = (s.word ? pair_left(t, j0) : t.left[j0]) & ~15u;
			const unsigned ce = s.word ? pair_left(t, jl) + 2 * static_cast<unsigned>(fwh) : t.left[jl] + t.filter_width;
			return (ce - cb + 15) & ~15u;
		};
		// Per-row vertical taps held in LDS for a tile row (start row + taps).
		const unsigned vtaps = (s.h_first && !t2) ? 0 : (s.word ? pair_count(*tv_tab) : tv_tab->filter_width);
		auto ceil_div = [](size_t a, size_t b) { return (a + b - 1) / b; };
		std::vector<unsigned> widths = s.h_first ? std::vector<unsigned>{ dev::kFastTileW } : std::vector<unsigned>{ 512, 256, 128, 64, 32, 16 };
		for (unsigned tw : widths) {
			for (unsigned th : { 64u, 32u, 16u, 8u, 4u, 2u, 1u }) {
				unsigned in_rows = 0, in_cols = 0, mid_rows = 0, mid_cols = 0;
				for (unsigned j0 = 0; j0 < s.width; j0 += tw) {
					const unsigned nc = std::min(tw, s.width - j0);
					in_cols = std::max(in_cols, th_tab ? h_cols(*th_tab, j0, nc) : (nc + 15) & ~15u);
				}
				for (unsigned i0 = 0; i0 < s.height; i0 += th) {
					const unsigned nr = std::min(th, s.height - i0);
					const unsigned il = i0 + nr - 1;
					unsigned span;
					if (!tv_tab)
						span = nr;
					else if (s.word)
						span = pair_left(*tv_tab, il) + 2 * pair_count(*tv_tab) - pair_left(*tv_tab, i0);
					else
						span = tv_tab->left[il] + tv_tab->filter_width - tv_tab->left[i0];
					in_rows = std::max(in_rows, span);
				}
				if (t2) {
					mid_rows = s.h_first ? in_rows : th;
					mid_cols = s.h_first ? tw : in_cols;
				}
				const size_t vtab = vtaps ? align_up((align_up(th, 4) + static_cast<size_t>(th) * vtaps) * 4, 16) : 0;
				const size_t bytes = vtab + align_up(elem * in_rows * in_cols, 16) + elem * mid_rows * mid_cols;
				if (bytes > kLdsBudget)
					continue;
				// Cost model: per-block work in thread iterations (staging, first and
				// second pass, weighted by taps), scaled by how many blocks share a CU.
				const size_t blocks = ceil_div(s.width, tw) * ceil_div(s.height, th);
				const double stage = static_cast<double>(ceil_div(static_cast<size_t>(in_rows) * in_cols * elem / 16, 256)) * 6.0 + 40.0;
				const double fwv = tv_tab ? (s.word ? pair_count(*tv_tab) : tv_tab->filter_width) : 0;
				const double fwhh = th_tab ? fwh : 0;
				double work;
				if (s.h_first) {
					work = in_rows * (fwhh + 4) + (t2 ? th * (fwv + 4) : 0);
				} else {
					const size_t pairs = in_cols / 2;
					const size_t groups = pairs >= 256 ? 1 : 256 / pairs;
					work = static_cast<double>(ceil_div(pairs, 256) * ceil_div(th, groups)) * (fwv + 4);
					if (t2) {
						const size_t opairs = (tw + 1) / 2;
						const size_t og = opairs >= 256 ? 1 : 256 / opairs;
						work += static_cast<double>(ceil_div(opairs, 256) * ceil_div(th, og)) * (fwhh + 4);
					}
				}
				const size_t per_cu = std::min<size_t>(4, (160 * 1024) / std::max<size_t>(bytes, 1));
				const double cost = static_cast<double>(blocks) * (stage + work) / static_cast<double>(std::max<size_t>(per_cu, 1));
				TiledConfig c;
				c.fwh = fwh;
				c.tile_w = tw;
				c.tile_h = th;
				c.in_rows = in_rows;
				c.in_cols = in_cols;
				c.mid_rows = mid_rows;
				c.mid_cols = mid_cols;
				c.cost = cost;
				s.fast_options.push_back(c);
			}
		}
		std::stable_sort(s.fast_options.begin(), s.fast_options.end(), [](const TiledConfig &x, const TiledConfig &y) { return x.cost < y.cost; });
		if (s.fast_options.size() > 6)
			s.fast_options.resize(6);
		for (const TiledConfig &c : s.fast_options)
			check_tiled(s, c, th_tab, tv_tab);
		if (!s.fast_options.empty())
			s.fast = s.fast_options.front();
	}

	static const unsigned kWidths[] = { 256, 128, 64, 32, 16, 8, 4, 2, 1 };
	static const unsigned kHeights[] = { 32, 16, 8, 4, 2, 1 };

	if (s.h_first && !sorted1) {
		// Non-monotonic left offsets (possible with extreme shifts): single pass,
		// taps read straight from HBM without a staged tile.
		if (s.passes != 1)
			fail(ZIMG_ERROR_UNKNOWN, "unsorted filter in a fused resize");
		if (t1.filter_width > 64)
			fail(ZIMG_ERROR_RESAMPLING_NOT_AVAILABLE, "unsorted filter wider than 64 taps");
		s.tile_w = 256;
		s.tile_h = 32;
		s.lds_in_rows = s.lds_in_cols = s.lds_mid_rows = s.lds_mid_cols = 0;
		return;
	}

	for (unsigned tw : kWidths) {
		for (unsigned th : kHeights) {
			unsigned in_rows = 0, in_cols = 0, mid_rows = 0, mid_cols = 0;
			if (s.h_first) {
				if (!sorted1)
					break;
				for (unsigned j0 = 0; j0 < s.width; j0 += tw) {
					const unsigned nc = std::min(tw, s.width - j0);
					in_cols = std::max(in_cols, t1.left[j0 + nc - 1] + t1.filter_width - t1.left[j0]);
				}
				for (unsigned i0 = 0; i0 < s.height; i0 += th) {
					const unsigned nr = std::min(th, s.height - i0);
					in_rows = std::max(in_rows, t2 ? t2->left[i0 + nr - 1] + t2->filter_width - t2->left[i0] : nr);
				}
				if (t2) {
					mid_rows = in_rows;
					mid_cols = tw;
				}
			} else if (t2) {
				mid_rows = th;
				for (unsigned j0 = 0; j0 < s.width; j0 += tw) {
					const unsigned nc = std::min(tw, s.width - j0);
					mid_cols = std::max(mid_cols, t2->left[j0 + nc - 1] + t2->filter_width - t2->left[j0]);
				}
			}
			const size_t bytes = elem * (static_cast<size_t>(in_rows) * in_cols + static_cast<size_t>(mid_rows) * mid_cols);
			if (bytes <= kLdsBudget) {
				s.tile_w = tw;
				s.tile_h = th;
				s.lds_in_rows = in_rows;
				s.lds_in_cols = in_cols;
				s.lds_mid_rows = mid_rows;
				s.lds_mid_cols = mid_cols;
				return;
			}
		}
	}
	fail(ZIMG_ERROR_RESAMPLING_NOT_AVAILABLE, "resize tile does not fit in LDS");
}

size_t DeviceProgram::tmp_size(unsigned frames) const
{
	size_t total = 0;
	for (auto &sl : m_slots)
		total += align_up(sl.pitch * sl.height * frames, kPitchAlign);
	for (auto &s : m_steps) {
		if (!s.dead && mfma_step(s))
			total += mfma_mid_bytes(s, frames);
		total += surrogate_bytes(s, frames);
	}
	return total;
}

void DeviceProgram::set_transfer_surrogates(const std::vector<const learned::Surrogate *> &s)
{
	std::set<int> funcs;
	for (const learned::Surrogate *m : s) {
		if (!m)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null transfer surrogate");
		if (!funcs.insert(m->func()).second)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "two surrogates for one transfer function");
	}
	m_surrogates = s;
}

std::vector<DeviceProgram::CspSegment> DeviceProgram::csp_segments(const Step &s, const CspProgram &prog) const
{
	std::vector<CspSegment> segs;
	if (m_surrogates.empty() || s.dead || s.kind != Step::CSP)
		return segs;
	auto find = [&](int func) -> const learned::Surrogate * {
		for (const learned::Surrogate *m : m_surrogates)
			if (m->func() == func)
				return m;
		return nullptr;
	};
	bool any = false;
	CspSegment cur;
	auto flush = [&]() {
		if (cur.prog.n) {
			segs.push_back(cur);
			cur = CspSegment{};
		}
	};
	for (int k = 0; k < prog.n; ++k) {
		const CspOp &op = prog.op[k];
		const learned::Surrogate *m = op.kind == CSP_GAMMA ? find(op.func) : op.kind == CSP_B67 ? find(zm::F_B67_OETF) : nullptr;
		if (!m) {
			cur.prog.op[cur.prog.n++] = op;
			continue;
		}
		any = true;
		if (op.kind == CSP_B67) {
			// inverse OOTF scaling stays a colourspace operation; the OETF
			// applied to its result (unit scales) is the learned part
			CspOp scale = op;
			scale.kind = CSP_B67_SCALE;
			cur.prog.op[cur.prog.n++] = scale;
			flush();
			segs.push_back(CspSegment{ m, CspProgram{}, 1.0f, 1.0f });
		} else {
			flush();
			segs.push_back(CspSegment{ m, CspProgram{}, op.pre, op.post });
		}
	}
	flush();
	if (!any)
		segs.clear();
	return segs;
}

size_t DeviceProgram::surrogate_bytes(const Step &s, unsigned frames) const
{
	if (csp_segments(s, s.prog).empty())
		return 0;
	const size_t pitch = align_up(static_cast<size_t>(s.width) * 4, kPitchAlign);
	return 3 * align_up(pitch * s.height * frames, kPitchAlign);
}

void DeviceProgram::run_learned_csp(const Step &s, const std::vector<CspSegment> &segs, const dev::PlaneArg *src,
                                    const dev::PlaneArg *dst, char *mid, unsigned frames, void *stream) const
{
	// Intermediates: three FLOAT planes; colourspace and surrogate launches
	// are elementwise, so they run in place there.
	const size_t pitch = align_up(static_cast<size_t>(s.width) * 4, kPitchAlign);
	const size_t plane = align_up(pitch * s.height * frames, kPitchAlign);
	dev::PlaneArg t[3];
	for (int p = 0; p < 3; ++p)
		t[p] = dev::PlaneArg{ mid + p * plane, static_cast<long long>(pitch), static_cast<long long>(pitch * s.height) };
	const dev::InXform plain_in{ 3, 0, 0, 1.0f, 0.0f };
	const dev::OutXform plain_out{ 3, 0, 0, 0, 32, 1.0f, 0.0f };
	bool in_plain = true, out_plain = true;
	for (int p = 0; p < 3; ++p) {
		in_plain = in_plain && s.ix[p].type == 3 && s.ix[p].op == 0;
		out_plain = out_plain && s.ox[p].type == 3 && s.ox[p].op == 0;
	}
	std::vector<CspSegment> list = segs;
	if (list.front().sur && !in_plain)
		list.insert(list.begin(), CspSegment{});
	if (list.back().sur && !out_plain)
		list.push_back(CspSegment{});
	for (size_t i = 0; i < list.size(); ++i) {
		const bool first = i == 0, last = i + 1 == list.size();
		const dev::PlaneArg *in = first ? src : t;
		const dev::PlaneArg *out = last ? dst : t;
		const CspSegment &g = list[i];
		if (g.sur) {
			learned::PlaneView v[3];
			for (int p = 0; p < 3; ++p)
				v[p] = learned::PlaneView{ in[p].ptr, out[p].ptr, in[p].stride, in[p].fstride, out[p].stride, out[p].fstride };
			learned::apply_planes(g.sur->device_model(g.pre, g.post), v, 3, s.width, s.height, frames, stream);
			continue;
		}
		dev::CspArg a{};
		for (int p = 0; p < 3; ++p) {
			a.src[p] = in[p];
			a.dst[p] = out[p];
			a.in[p] = first ? s.ix[p] : plain_in;
			a.out[p] = last ? s.ox[p] : plain_out;
		}
		a.frames = frames;
		a.width = s.width;
		a.height = s.height;
		a.prog = g.prog;
		if (const int rc = dev::launch_colorspace(a, stream))
			fail(ZIMG_ERROR_UNKNOWN, std::string("device launch failed: ") + hipGetErrorString(static_cast<hipError_t>(rc)));
	}
}

void DeviceProgram::set_resize_precision(int mode)
{
	if (mode < 0 || mode > 2)
		fail(ZIMG_ERROR_ENUM_OUT_OF_RANGE, "resize precision mode");
	m_precision = mode;
}

namespace {

// Input window of each 16-output tile of a pass on matrix cores: inputs
// [kb, kend) hold every nonzero coefficient (both multiples of 8), read in
// nks steps of 32 from kb.
struct Band {
	std::vector<int> kb, kend;
	std::vector<unsigned> nks;
};

Band band_of(const FilterTable &t)
{
	Band b;
	const unsigned tiles = (t.filter_rows + 15) / 16;
	for (unsigned tt = 0; tt < tiles; ++tt) {
		unsigned lo = UINT_MAX, hi = 0;
		for (unsigned o = 16 * tt; o < std::min(16 * tt + 16, t.filter_rows); ++o) {
			lo = std::min(lo, t.left[o]);
			hi = std::max(hi, t.left[o] + t.filter_width);
		}
		const unsigned kb = lo & ~7u;
		b.kb.push_back(static_cast<int>(kb));
		b.kend.push_back(static_cast<int>((hi + 7) & ~7u));
		b.nks.push_back((hi - kb + 31) / 32);
	}
	return b;
}

// Widest input window over runs of `group` consecutive tiles.
unsigned group_window(const Band &b, unsigned group)
{
	unsigned widest = 0;
	for (size_t t0 = 0; t0 < b.kb.size(); t0 += group) {
		int lo = INT_MAX, hi = 0;
		for (size_t t = t0; t < std::min(b.kb.size(), t0 + group); ++t) {
			lo = std::min(lo, b.kb[t]);
			hi = std::max(hi, b.kend[t]);
		}
		widest = std::max(widest, static_cast<unsigned>(hi - lo));
	}
	return widest;
}

// Smallest LDS row stride (16-bit units) of at least n that is an odd multiple
// of 8: rows then start in distinct 16-byte bank groups.
unsigned odd_stride(unsigned n)
{
	unsigned v = (n + 7) & ~7u;
	return (v % 16 == 0) ? v + 8 : v;
}

} // namespace

namespace {

// Window starts of the 16-output blocks of a WORD pass for the integer
// matrix-core form (even starts when `even`: the horizontal pass), or an
// empty vector when a block
// does not fit one 16-sample window, the starts are not monotonic, a tap
// does not split into two signed bytes (c >= 32640) or a row's sum could
// leave the int32 range (sum |c| * 32768 + 2^29 + 2^13 >= 2^31).
std::vector<int> imma_windows(const FilterTable &t, bool even)
{
	std::vector<int> kb;
	const unsigned fw = t.filter_width;
	for (unsigned o = 0; o < t.filter_rows; ++o) {
		long mag = 0;
		for (unsigned k = 0; k < fw; ++k) {
			const int c = t.data_i16[static_cast<size_t>(o) * fw + k];
			if (c >= 32640)
				return {};
			mag += c < 0 ? -c : c;
		}
		if (mag > 49000)
			return {};
	}
	for (unsigned o0 = 0; o0 < t.filter_rows; o0 += 16) {
		const int k = static_cast<int>(even ? t.left[o0] & ~1u : t.left[o0]);
		for (unsigned o = o0; o < std::min(t.filter_rows, o0 + 16); ++o)
			if (static_cast<int>(t.left[o]) < k || static_cast<int>(t.left[o] + fw) > k + 16)
				return {};
		if (!kb.empty() && k < kb.back())
			return {};
		kb.push_back(k);
	}
	return kb;
}

// horizontal: lane group g holds window samples 4 g + 0..3 and + 8..11 (the
// kernel's LDS read order); vertical: 8 g + 0..7.
void imma_blocks(const FilterTable &t, const std::vector<int> &kb, bool horizontal, std::vector<dev::ImmaBlock> &out)
{
	const unsigned fw = t.filter_width;
	for (size_t b = 0; b < kb.size(); ++b) {
		dev::ImmaBlock blk{};
		blk.kb = kb[b];
		for (unsigned n = 0; n < 16; ++n) {
			const unsigned o = 16 * static_cast<unsigned>(b) + n;
			int32_t sum = 0;
			for (unsigned k = 0; o < t.filter_rows && k < fw; ++k)
				sum += t.data_i16[static_cast<size_t>(o) * fw + k];
			blk.init[n] = (1 << 29) + (1 << 13) + 128 * sum;
		}
		for (unsigned lane = 0; lane < 32; ++lane) {
			const unsigned o = 16 * static_cast<unsigned>(b) + lane % 16, g = lane / 16;
			uint8_t bytes[16];
			for (unsigned i = 0; i < 8; ++i) {
				const unsigned pos = horizontal ? 4 * g + (i & 3) + 8 * (i >> 2) : 8 * g + i;
				const long smp = kb[b] + static_cast<long>(pos);
				int c = 0;
				if (o < t.filter_rows && smp >= t.left[o] && smp < static_cast<long>(t.left[o] + fw))
					c = t.data_i16[static_cast<size_t>(o) * fw + (smp - t.left[o])];
				const int cl = static_cast<int8_t>(c & 0xFF);
				bytes[i] = static_cast<uint8_t>((c - cl) / 256);
				bytes[8 + i] = static_cast<uint8_t>(cl);
			}
			std::memcpy(blk.c[lane], bytes, 16);
		}
		out.push_back(blk);
	}
}

} // namespace

ImmaPlan DeviceProgram::imma_plan(const Step &s) const
{
	ImmaPlan p;
	if (!s.word || s.passes != 2 || !s.h_first)
		return p;
	if (s.ox[0].op != 0 || s.ox[0].type != 1 || s.ix[0].type > 1 || s.ix[0].op > 1)
		return p;
	const FilterTable &th = *m_plan.nodes[s.tab[0]].table;
	const FilterTable &tv = *m_plan.nodes[s.tab[1]].table;
	if (th.filter_rows != s.width || tv.filter_rows != s.height)
		return p;
	const std::vector<int> kh = imma_windows(th, true), kv = imma_windows(tv, false);
	if (kh.empty() || kv.empty())
		return p;
	const unsigned nh = static_cast<unsigned>(kh.size()), nv = static_cast<unsigned>(kv.size());
	const unsigned cpt = dev::kImmaTileW / 16;
	unsigned in_cols = 0;
	for (unsigned b0 = 0; b0 < nh; b0 += cpt) {
		const int cb = kh[b0] & ~15;
		for (unsigned b = b0; b < std::min(nh, b0 + cpt); ++b)
			in_cols = std::max(in_cols, static_cast<unsigned>((kh[b] + 16 - cb + 15) & ~15));
	}
	const unsigned rb = 8;
	unsigned mb = 0;
	for (unsigned b0 = 0; b0 < nv; b0 += rb)
		for (unsigned b = b0; b < std::min(nv, b0 + rb); ++b)
			mb = std::max(mb, static_cast<unsigned>(kv[b] + 16 - kv[b0] + 15) / 16);
	if (mb > dev::kImmaMaxMb)
		return p;
	p.rb = rb;
	p.mb = mb;
	p.mid_pitch = 16 * mb + 8; // 2 mod 4 dwords (bank spread of the kernel's reads and writes)
	p.in_cols = in_cols;
	p.in_pitch = odd_stride(in_cols);
	p.lds = 2 * dev::kImmaTileW * p.mid_pitch + std::max(2 * 16 * mb * p.in_pitch, 2 * 4 * 16 * dev::kImmaOutPitch);
	// A source row holds at least one whole 16-byte chunk (the kernel's
	// clamped loads stay inside the row).
	const unsigned esz = s.ix[0].type == 0 ? 1 : 2;
	p.ok = p.lds <= 64 * 1024 && s.in_w * esz >= 16;
	if (p.ok) {
		// Every LDS index the kernel forms, re-derived with its formulas: the
		// horizontal windows (2 x 4 dwords from (kb - cb) / 2 + 2 (g % 2)) lie in the
		// staged columns, the vertical windows (three dwords from byte
		// (kb - rbase + 8 (g % 2)) & ~3) in the computed intermediate rows and
		// in the column's mid_pitch bytes. Blocks past the plane repeat the last.
		for (unsigned t0 = 0; t0 < nh; t0 += cpt) {
			const int cb = kh[t0] & ~15;
			for (unsigned k = 0; k < cpt; ++k) {
				const int b = kh[std::min(t0 + k, nh - 1)];
				if (b < cb || (b - cb) % 2 || static_cast<unsigned>(b - cb) / 2 + 2 + 5 >= p.in_pitch / 2 ||
				    static_cast<unsigned>(b - cb) + 16 > in_cols)
					fail(ZIMG_ERROR_UNKNOWN, "integer matrix-core plan: horizontal window outside the staged columns");
			}
		}
		for (unsigned t0 = 0; t0 < nv; t0 += rb) {
			for (unsigned k = 0; k < rb; ++k) {
				const int d = kv[std::min(t0 + k, nv - 1)] - kv[t0];
				if (d < 0 || static_cast<unsigned>(d) + 16 > 16 * mb || ((static_cast<unsigned>(d) + 8) & ~3u) + 12 > p.mid_pitch)
					fail(ZIMG_ERROR_UNKNOWN, "integer matrix-core plan: vertical window outside the intermediate rows");
			}
		}
	}
	return p;
}

const DeviceProgram::ImmaTables &DeviceProgram::imma_tables_for_current_device() const
{
	int device = 0;
	check_hip(hipGetDevice(&device), "hipGetDevice");
	std::lock_guard<std::mutex> lock(m_mutex);
	auto it = m_imma_tables.find(device);
	if (it != m_imma_tables.end())
		return *it->second;
	std::vector<int> ids;
	std::vector<dev::ImmaBlock> blocks;
	std::vector<size_t> first;
	for (auto &s : m_steps) {
		if (s.dead || s.kind != Step::RESIZE || !s.imma.ok)
			continue;
		for (int q = 0; q < 2; ++q) {
			const int key = 2 * s.tab[q] + q;
			if (std::find(ids.begin(), ids.end(), key) != ids.end())
				continue;
			const FilterTable &t = *m_plan.nodes[s.tab[q]].table;
			ids.push_back(key);
			first.push_back(blocks.size());
			imma_blocks(t, imma_windows(t, q == 0), q == 0, blocks);
		}
	}
	auto tab = std::make_shared<ImmaTables>();
	if (!blocks.empty()) {
		const size_t bytes = blocks.size() * sizeof(dev::ImmaBlock);
		check_hip(hipMalloc(&tab->mem, bytes), "hipMalloc(integer matrix-core tables)");
		check_hip(hipMemcpy(tab->mem, blocks.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy(integer matrix-core tables)");
		for (size_t n = 0; n < ids.size(); ++n)
			tab->pass[ids[n]] = static_cast<const dev::ImmaBlock *>(tab->mem) + first[n];
	}
	m_imma_tables[device] = tab;
	return *tab;
}

DeviceProgram::FusedPlan DeviceProgram::mfma_fused_plan(const Step &s) const
{
	FusedPlan p;
	if (s.passes != 2)
		return p;
	const FilterTable &th = *m_plan.nodes[s.tab[s.h_first ? 0 : 1]].table;
	const FilterTable &tv = *m_plan.nodes[s.tab[s.h_first ? 1 : 0]].table;
	if (th.filter_rows != s.width || tv.filter_rows != s.height)
		return p;
	const Band bh = band_of(th), bv = band_of(tv);
	const auto max_nks = [](const Band &b) { return *std::max_element(b.nks.begin(), b.nks.end()); };
	if (max_nks(bh) > dev::kFusedMaxSteps || max_nks(bv) > dev::kFusedMaxSteps)
		return p;
	const unsigned wc = group_window(bh, dev::kFusedWaves);
	if (wc > dev::kFusedMaxWindow)
		return p;
	// Tallest tile whose LDS leaves room for two workgroups per CU.
	for (unsigned tile_h : { 128u, 64u, 32u, 16u }) {
		const unsigned group = tile_h / 16;
		const unsigned wr = (group_window(bv, group) + 15) & ~15u;
		unsigned blocks = 0; // vertical coefficient blocks of one workgroup
		for (size_t t0 = 0; t0 < bv.kb.size(); t0 += group) {
			unsigned n = 0;
			for (size_t t = t0; t < std::min(bv.kb.size(), t0 + group); ++t)
				n += bv.nks[t];
			blocks = std::max(blocks, n);
		}
		if (wr / 16 > dev::kFusedMaxGroups || blocks > dev::kFusedMaxCoefBlocks)
			continue;
		const unsigned sc = odd_stride(wc), sr = odd_stride(wr);
		const unsigned lds = 2 * (4 * 16 * sc + 2 * 16 * dev::kFusedWaves * sr) + 2048 * blocks;
		if (lds <= 80 * 1024) {
			p.ok = true;
			p.tile_h = tile_h;
			p.slab_stride = sc;
			p.mid_stride = sr;
			p.coef_blocks = blocks;
			p.lds = lds;
			// Output tile through LDS: tile_h rows of 16 kFusedWaves + 4 floats after the slabs.
			const unsigned staged = std::max(lds, 2 * 4 * 16 * sc + tile_h * (16 * dev::kFusedWaves + 4) * 4);
			p.lds_staged = staged <= 80 * 1024 ? staged : 0;
			return p;
		}
	}
	return p;
}

// FP32 intermediate of a two-pass FLOAT resize run on matrix cores as two launches.
size_t DeviceProgram::mfma_mid_bytes(const Step &s, unsigned frames) const
{
	if (s.passes != 2 || mfma_fused_plan(s).ok)
		return 0;
	return align_up(align_up(static_cast<size_t>(s.mid_w) * 4, kPitchAlign) * s.mid_h * frames, kPitchAlign);
}

namespace {

uint16_t to_half_bits(float x) { return __builtin_bit_cast(uint16_t, static_cast<_Float16>(x)); }
float half_bits_to_float(uint16_t h) { return static_cast<float>(__builtin_bit_cast(_Float16, h)); }

uint16_t to_bf16_bits(float x)
{
	uint32_t u;
	std::memcpy(&u, &x, 4);
	u += 0x7FFFu + ((u >> 16) & 1u); // round to nearest even
	return static_cast<uint16_t>(u >> 16);
}
float bf16_bits_to_float(uint16_t h)
{
	const uint32_t u = static_cast<uint32_t>(h) << 16;
	float x;
	std::memcpy(&x, &u, 4);
	return x;
}

} // namespace

const DeviceProgram::MfmaTables &DeviceProgram::mfma_tables_for_current_device() const
{
	int device = 0;
	check_hip(hipGetDevice(&device), "hipGetDevice");
	const int key = device * 4 + m_precision;
	std::lock_guard<std::mutex> lock(m_mutex);
	auto it = m_mfma_tables.find(key);
	if (it != m_mfma_tables.end())
		return *it->second;

	const bool bf16 = m_precision == 2;
	struct Host {
		int id;
		std::vector<unsigned> off;
		Band band;
		std::vector<uint16_t> coef;
	};
	std::vector<Host> hosts;
	for (auto &s : m_steps) {
		if (s.dead || !mfma_step(s))
			continue;
		for (int p = 0; p < s.passes; ++p) {
			const int id = s.tab[p];
			if (std::any_of(hosts.begin(), hosts.end(), [&](const Host &h) { return h.id == id; }))
				continue;
			const FilterTable &t = *m_plan.nodes[id].table;
			Host h;
			h.id = id;
			h.band = band_of(t);
			const unsigned tiles = static_cast<unsigned>(h.band.kb.size());
			unsigned blocks = 0;
			for (unsigned tt = 0; tt < tiles; ++tt) {
				const int kb = h.band.kb[tt];
				const unsigned nks = h.band.nks[tt];
				h.off.push_back(blocks);
				for (unsigned st = 0; st < nks; ++st) {
					uint16_t hiv[64 * 8], lov[64 * 8];
					for (unsigned lane = 0; lane < 64; ++lane) {
						for (unsigned i = 0; i < 8; ++i) {
							const unsigned o = 16 * tt + lane % 16;
							const long k = kb + 32 * st + 8 * (lane / 16) + i;
							double c = 0.0;
							if (o < t.filter_rows && k >= t.left[o] && k < static_cast<long>(t.left[o] + t.filter_width))
								c = t.data64[static_cast<size_t>(o) * t.filter_width + (k - t.left[o])];
							const float cf = static_cast<float>(c);
							const uint16_t hb = bf16 ? to_bf16_bits(cf) : to_half_bits(cf);
							const double rem = c - static_cast<double>(bf16 ? bf16_bits_to_float(hb) : half_bits_to_float(hb));
							const float rf = static_cast<float>(rem);
							hiv[lane * 8 + i] = hb;
							lov[lane * 8 + i] = bf16 ? to_bf16_bits(rf) : to_half_bits(rf);
						}
					}
					h.coef.insert(h.coef.end(), hiv, hiv + 512);
					h.coef.insert(h.coef.end(), lov, lov + 512);
				}
				blocks += nks;
			}
			hosts.push_back(std::move(h));
		}
	}

	auto tab = std::make_shared<MfmaTables>();
	size_t bytes = 0;
	std::vector<size_t> offs;
	for (auto &h : hosts) {
		offs.push_back(bytes);
		bytes += align_up(h.coef.size() * 2, 256) + align_up(h.off.size() * sizeof(dev::MfmaTile), 256);
	}
	if (bytes) {
		std::vector<char> host(bytes, 0);
		for (size_t n = 0; n < hosts.size(); ++n) {
			const Host &h = hosts[n];
			size_t o = offs[n];
			std::memcpy(&host[o], h.coef.data(), h.coef.size() * 2);
			o += align_up(h.coef.size() * 2, 256);
			for (size_t t = 0; t < h.off.size(); ++t) {
				const dev::MfmaTile rec{ h.band.kb[t], h.band.kend[t], h.band.nks[t], h.off[t] };
				std::memcpy(&host[o + t * sizeof(rec)], &rec, sizeof(rec));
			}
		}
		check_hip(hipMalloc(&tab->mem, bytes), "hipMalloc(matrix-core tables)");
		check_hip(hipMemcpy(tab->mem, host.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy(matrix-core tables)");
		char *base = static_cast<char *>(tab->mem);
		for (size_t n = 0; n < hosts.size(); ++n) {
			const Host &h = hosts[n];
			const size_t o = offs[n] + align_up(h.coef.size() * 2, 256);
			tab->pass[h.id] = dev::MfmaTab{ reinterpret_cast<const uint16_t *>(base + offs[n]), reinterpret_cast<const dev::MfmaTile *>(base + o) };
		}
	}
	m_mfma_tables[key] = tab;
	return *tab;
}

void DeviceProgram::run_mfma_resize(const Step &s, const dev::PlaneArg &src, const dev::PlaneArg &dst, char *mid, unsigned frames,
                                    void *stream) const
{
	const MfmaTables &mt = mfma_tables_for_current_device();
	const FusedPlan fp = mfma_fused_plan(s);
	if (fp.ok) {
		dev::MfmaFusedArg a{};
		a.src = src;
		a.dst = dst;
		a.frames = frames;
		a.bf16 = m_precision == 2 ? 1 : 0;
		a.src_w = s.in_w;
		a.src_h = s.in_h;
		a.dst_w = s.width;
		a.dst_h = s.height;
		a.tile_h = fp.tile_h;
		a.slab_stride = fp.slab_stride;
		a.mid_stride = fp.mid_stride;
		a.lds_bytes = fp.lds;
		a.vec_in = (s.ix[0].type == 3 && s.ix[0].op == 0 && aligned16(src)) ? 1 : 0;
		a.vec_out = (s.ox[0].type == 3 && s.ox[0].op == 0 && aligned16(dst)) ? 1 : 0;
		a.stage_out = (a.vec_in && a.vec_out && fp.lds_staged) ? 1 : 0;
		if (a.stage_out)
			a.lds_bytes = fp.lds_staged;
		a.th = mt.pass.at(s.tab[s.h_first ? 0 : 1]);
		a.tv = mt.pass.at(s.tab[s.h_first ? 1 : 0]);
		a.in = s.ix[0];
		a.out = s.ox[0];
		check_hip(static_cast<hipError_t>(dev::launch_mfma_fused(a, stream)), "matrix-core fused resize launch");
		return;
	}
	const dev::InXform plain_in{ 3, 0, 0, 1.0f, 0.0f };
	const dev::OutXform plain_out{ 3, 0, 0, 0, 32, 1.0f, 0.0f };
	const size_t mid_pitch = align_up(static_cast<size_t>(s.mid_w) * 4, kPitchAlign);
	const dev::PlaneArg midp{ mid, static_cast<long long>(mid_pitch), static_cast<long long>(mid_pitch * s.mid_h) };
	unsigned w = s.in_w, h = s.in_h;
	for (int p = 0; p < s.passes; ++p) {
		const bool horizontal = (p == 0) == s.h_first;
		const FilterTable &t = *m_plan.nodes[s.tab[p]].table;
		dev::MfmaArg a{};
		a.src = p == 0 ? src : midp;
		a.dst = p == s.passes - 1 ? dst : midp;
		a.frames = frames;
		a.horizontal = horizontal ? 1 : 0;
		a.bf16 = m_precision == 2 ? 1 : 0;
		a.src_len = horizontal ? w : h;
		a.dst_len = t.filter_rows;
		a.rows = h;
		a.cols = w;
		a.tab = mt.pass.at(s.tab[p]);
		a.in = p == 0 ? s.ix[0] : plain_in;
		a.out = p == s.passes - 1 ? s.ox[0] : plain_out;
		a.vec16 = (reinterpret_cast<uintptr_t>(a.src.ptr) % 16 == 0 && a.src.stride % 16 == 0 && a.src.fstride % 16 == 0) ? 1 : 0;
		check_hip(static_cast<hipError_t>(dev::launch_mfma_resize(a, stream)), "matrix-core resize launch");
		if (horizontal)
			w = t.filter_rows;
		else
			h = t.filter_rows;
	}
}

const DeviceProgram::Tables &DeviceProgram::tables_for_current_device() const
{
	int device = 0;
	check_hip(hipGetDevice(&device), "hipGetDevice");
	std::lock_guard<std::mutex> lock(m_mutex);
	auto it = m_tables.find(device);
	if (it != m_tables.end())
		return *it->second;

	// One device allocation holds every pass table of the program.
	std::vector<int> ids;
	for (auto &s : m_steps)
		if (s.kind == Step::RESIZE)
			for (int p = 0; p < s.passes; ++p)
				if (std::find(ids.begin(), ids.end(), s.tab[p]) == ids.end())
					ids.push_back(s.tab[p]);

	size_t bytes = 0;
	std::vector<size_t> offs;
	for (int id : ids) {
		const FilterTable &t = *m_plan.nodes[id].table;
		offs.push_back(bytes);
		bytes += align_up(t.left.size() * 4, 256) + align_up(t.data_i16.size() * 2, 256) + align_up(t.data.size() * 4, 256) +
		         align_up(t.left.size() * 4, 256) + align_up(static_cast<size_t>(t.filter_rows) * pair_count(t) * 4, 256);
	}
	auto tab = std::make_shared<Tables>();
	if (bytes) {
		std::vector<char> host(bytes, 0);
		for (size_t n = 0; n < ids.size(); ++n) {
			const FilterTable &t = *m_plan.nodes[ids[n]].table;
			size_t o = offs[n];
			std::memcpy(&host[o], t.left.data(), t.left.size() * 4);
			const size_t o_i16 = o + align_up(t.left.size() * 4, 256);
			std::memcpy(&host[o_i16], t.data_i16.data(), t.data_i16.size() * 2);
			const size_t o_f32 = o_i16 + align_up(t.data_i16.size() * 2, 256);
			std::memcpy(&host[o_f32], t.data.data(), t.data.size() * 4);
			// Pair form: window from left & ~1, a zero tap first when left is odd.
			const size_t o_pl = o_f32 + align_up(t.data.size() * 4, 256);
			const size_t o_pc = o_pl + align_up(t.left.size() * 4, 256);
			const unsigned fwp = pair_count(t);
			for (unsigned r = 0; r < t.filter_rows; ++r) {
				const unsigned pl = pair_left(t, r);
				const unsigned d = t.left[r] - pl;
				std::memcpy(&host[o_pl + 4 * r], &pl, 4);
				auto tap = [&](int k) -> uint16_t {
					return (k >= 0 && k < static_cast<int>(t.filter_width)) ? static_cast<uint16_t>(t.data_i16[static_cast<size_t>(r) * t.filter_width + k]) : 0;
				};
				for (unsigned p = 0; p < fwp; ++p) {
					const uint32_t w = tap(static_cast<int>(2 * p) - static_cast<int>(d)) | (static_cast<uint32_t>(tap(static_cast<int>(2 * p + 1) - static_cast<int>(d))) << 16);
					std::memcpy(&host[o_pc + 4 * (static_cast<size_t>(r) * fwp + p)], &w, 4);
				}
			}
		}
		check_hip(hipMalloc(&tab->mem, bytes), "hipMalloc(tables)");
		check_hip(hipMemcpy(tab->mem, host.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy(tables)");
		char *base = static_cast<char *>(tab->mem);
		for (size_t n = 0; n < ids.size(); ++n) {
			const FilterTable &t = *m_plan.nodes[ids[n]].table;
			const size_t o = offs[n];
			const size_t o_i16 = o + align_up(t.left.size() * 4, 256);
			const size_t o_f32 = o_i16 + align_up(t.data_i16.size() * 2, 256);
			const size_t o_pl = o_f32 + align_up(t.data.size() * 4, 256);
			const size_t o_pc = o_pl + align_up(t.left.size() * 4, 256);
			tab->pass[ids[n]] = dev::PassArg{ reinterpret_cast<const unsigned *>(base + o), reinterpret_cast<const int16_t *>(base + o_i16),
			                                  reinterpret_cast<const float *>(base + o_f32), static_cast<int>(t.filter_width),
			                                  static_cast<int>(t.filter_rows), reinterpret_cast<const unsigned *>(base + o_pl),
			                                  reinterpret_cast<const int32_t *>(base + o_pc), static_cast<int>(pair_count(t)) };
		}
	}
	// Gamma lookup tables of the approximate-gamma operations, one allocation.
	std::vector<float> luts;
	std::vector<std::pair<int, size_t>> lut_at;
	for (size_t si = 0; si < m_steps.size(); ++si) {
		const Step &st = m_steps[si];
		if (st.dead || st.kind != Step::CSP)
			continue;
		for (int k = 0; k < st.prog.n; ++k) {
			std::vector<float> l = build_gamma_lut(st.prog.op[k]);
			if (l.empty())
				continue;
			lut_at.emplace_back(static_cast<int>(si) * 8 + k, luts.size());
			luts.insert(luts.end(), l.begin(), l.end());
			luts.resize(align_up(luts.size(), 64), 0.0f);
		}
	}
	if (!luts.empty()) {
		check_hip(hipMalloc(&tab->lut_mem, luts.size() * 4), "hipMalloc(gamma tables)");
		check_hip(hipMemcpy(tab->lut_mem, luts.data(), luts.size() * 4, hipMemcpyHostToDevice), "hipMemcpy(gamma tables)");
		for (auto &e : lut_at)
			tab->lut[e.first] = static_cast<const float *>(tab->lut_mem) + e.second;
	}
	// Unresize tables: band, offsets and LU factors of each pass, one allocation.
	std::vector<char> ubuf;
	std::vector<std::pair<int, std::array<size_t, 5>>> uat;
	for (const Step &st : m_steps) {
		if (st.dead || st.kind != Step::UNRESIZE)
			continue;
		const UnresizeTable &t = *m_plan.nodes[st.tab[0]].utable;
		std::array<size_t, 5> o{};
		auto put = [&](const void *p, size_t n) {
			const size_t at = align_up(ubuf.size(), 256);
			ubuf.resize(at + n);
			std::memcpy(ubuf.data() + at, p, n);
			return at;
		};
		o[0] = put(t.coef.data(), t.coef.size() * 4);
		o[1] = put(t.offset.data(), t.offset.size() * 4);
		o[2] = put(t.lu_c.data(), t.lu_c.size() * 4);
		o[3] = put(t.lu_l.data(), t.lu_l.size() * 4);
		o[4] = put(t.lu_u.data(), t.lu_u.size() * 4);
		uat.emplace_back(st.tab[0], o);
	}
	if (!ubuf.empty()) {
		check_hip(hipMalloc(&tab->unresize_mem, ubuf.size()), "hipMalloc(unresize tables)");
		check_hip(hipMemcpy(tab->unresize_mem, ubuf.data(), ubuf.size(), hipMemcpyHostToDevice), "hipMemcpy(unresize tables)");
		const char *base = static_cast<const char *>(tab->unresize_mem);
		for (auto &e : uat) {
			const UnresizeTable &t = *m_plan.nodes[e.first].utable;
			dev::UnresizeArg a{};
			a.row_size = t.row_size;
			a.coef = reinterpret_cast<const float *>(base + e.second[0]);
			a.offset = reinterpret_cast<const unsigned *>(base + e.second[1]);
			a.c = reinterpret_cast<const float *>(base + e.second[2]);
			a.l = reinterpret_cast<const float *>(base + e.second[3]);
			a.u = reinterpret_cast<const float *>(base + e.second[4]);
			tab->unresize[e.first] = a;
		}
	}
	m_tables[device] = tab;
	return *tab;
}

void DeviceProgram::prepare_current_device() const
{
	int device = 0;
	check_hip(hipGetDevice(&device), "hipGetDevice");
	tables_for_current_device();
	for (const Step &s : m_steps) {
		if (s.dead || s.kind != Step::RESIZE)
			continue;
		if (mfma_step(s))
			mfma_tables_for_current_device();
		else if (s.imma.ok)
			imma_tables_for_current_device();
	}
	std::lock_guard<std::mutex> lock(m_mutex);
	m_prepared.insert(device * 4 + m_precision);
}

bool DeviceProgram::resident_on_current_device() const
{
	int device = 0;
	if (hipGetDevice(&device) != hipSuccess)
		return false;
	std::lock_guard<std::mutex> lock(m_mutex);
	return m_prepared.count(device * 4 + m_precision) != 0;
}

bool DeviceProgram::aligned16(const dev::PlaneArg &p)
{
	return (reinterpret_cast<uintptr_t>(p.ptr) % 16) == 0 && (p.stride % 16) == 0 && (p.fstride % 16) == 0;
}

dev::ResizeArg DeviceProgram::resize_arg(const Step &s, const dev::PlaneArg &src, const dev::PlaneArg &dst, unsigned frames,
                                         const Tables &tabs, const TiledConfig *cfg) const
{
	dev::ResizeArg a{};
	a.src = src;
	a.dst = dst;
	a.frames = frames;
	a.src_w = s.in_w;
	a.src_h = s.in_h;
	a.mid_w = s.mid_w;
	a.mid_h = s.mid_h;
	a.dst_w = s.width;
	a.dst_h = s.height;
	a.h_first = s.h_first ? 1 : 0;
	a.passes = s.passes;
	a.p1 = tabs.pass.at(s.tab[0]);
	if (s.passes == 2)
		a.p2 = tabs.pass.at(s.tab[1]);
	a.word = s.word ? 1 : 0;
	a.pixel_max = s.pixel_max;
	a.in = s.ix[0];
	a.out = s.ox[0];
	if (!cfg && s.fast.fwh && aligned16(a.src) && aligned16(a.dst))
		cfg = &s.fast;
	if (cfg) {
		a.fwh = cfg->fwh;
		a.tile_w = cfg->tile_w;
		a.tile_h = cfg->tile_h;
		a.lds_in_rows = cfg->in_rows;
		a.lds_in_cols = cfg->in_cols;
		a.lds_mid_rows = cfg->mid_rows;
		a.lds_mid_cols = cfg->mid_cols;
	} else {
		a.fwh = 0;
		a.tile_w = s.tile_w;
		a.tile_h = s.tile_h;
		a.lds_in_rows = s.lds_in_rows;
		a.lds_in_cols = s.lds_in_cols;
		a.lds_mid_rows = s.lds_mid_rows;
		a.lds_mid_cols = s.lds_mid_cols;
	}
	return a;
}

unsigned DeviceProgram::autotune(const DeviceBuffers &src, const DeviceBuffers &dst, void *tmp, unsigned frames, void *stream)
{
	if (!frames)
		return 0;
	hipStream_t hs = static_cast<hipStream_t>(stream);
	// One event before each step and one after the last (run()'s profiling
	// form): a candidate is timed on its own step, the minimum of 5 runs.
	std::vector<hipEvent_t> ev(m_steps.size() + 1, nullptr);
	auto destroy = [&]() {
		for (hipEvent_t e : ev)
			if (e)
				(void)hipEventDestroy(e);
	};
	unsigned changed = 0;
	try {
		for (hipEvent_t &e : ev)
			check_hip(hipEventCreate(&e), "hipEventCreate");
		auto time_step = [&](size_t si) -> float {
			run(src, dst, tmp, frames, stream); // warm-up
			float best = 0;
			for (int r = 0; r < 5; ++r) {
				run(src, dst, tmp, frames, stream, reinterpret_cast<void *const *>(ev.data()));
				check_hip(hipEventSynchronize(ev[si + 1]), "hipEventSynchronize");
				float ms = 0;
				check_hip(hipEventElapsedTime(&ms, ev[si], ev[si + 1]), "hipEventElapsedTime");
				best = r == 0 ? ms : std::min(best, ms);
			}
			check_hip(hipStreamSynchronize(hs), "hipStreamSynchronize");
			return best;
		};
		auto same = [](const TiledConfig &x, const TiledConfig &y) {
			return x.fwh == y.fwh && x.tile_w == y.tile_w && x.tile_h == y.tile_h && x.in_rows == y.in_rows && x.in_cols == y.in_cols &&
			       x.mid_rows == y.mid_rows && x.mid_cols == y.mid_cols;
		};
		for (size_t si = 0; si < m_steps.size(); ++si) {
			Step &s = m_steps[si];
			// Steps on the matrix cores (precision form, integer WORD) do not use the tiled shape.
			if (s.dead || s.kind != Step::RESIZE || !s.fast.fwh || s.fast_options.size() < 2 || s.imma.ok || mfma_step(s))
				continue;
			const TiledConfig orig = s.fast;
			TiledConfig best = orig;
			float best_ms = time_step(si);
			for (const TiledConfig &c : s.fast_options) {
				if (same(c, orig))
					continue;
				s.fast = c;
				const float ms = time_step(si);
				if (ms < best_ms * 0.98f) {
					best = c;
					best_ms = ms;
				}
			}
			s.fast = best;
			if (!same(best, orig))
				++changed;
		}
	} catch (...) {
		destroy();
		throw;
	}
	destroy();
	return changed;
}

std::vector<double> DeviceProgram::step_bytes() const
{
	auto elem = [](int type) { return type == 0 ? 1.0 : type == 3 ? 4.0 : 2.0; };
	std::vector<double> out;
	for (const Step &s : m_steps) {
		double b = 0;
		const double px = static_cast<double>(s.width) * s.height;
		switch (s.kind) {
		case Step::RESIZE:
			b = static_cast<double>(s.in_w) * s.in_h * elem(s.ix[0].type) + px * elem(s.ox[0].type);
			break;
		case Step::POINT:
			b = (s.fill ? 0.0 : px * elem(s.ix[0].type)) + px * elem(s.ox[0].type);
			break;
		case Step::CSP:
			for (int k = 0; k < 3; ++k)
				b += px * (elem(s.ix[k].type) + elem(s.ox[k].type));
			break;
		case Step::ED:
			b = s.ed_planes * px * (elem(s.ix[0].type) + elem(s.ox[0].type));
			break;
		case Step::UNRESIZE:
			b = static_cast<double>(s.in_w) * s.in_h * 4 + px * 4;
			break;
		}
		out.push_back(b);
	}
	return out;
}

void DeviceProgram::run(const DeviceBuffers &src, const DeviceBuffers &dst, void *tmp, unsigned frames, void *stream,
                        void *const *events) const
{
	if (!frames)
		return;
	if (frames > 65535)
		fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "too many frames in one batch");
	const Tables &tabs = tables_for_current_device();

	std::vector<size_t> slot_off(m_slots.size());
	size_t off = 0;
	for (size_t k = 0; k < m_slots.size(); ++k) {
		slot_off[k] = off;
		off += align_up(m_slots[k].pitch * m_slots[k].height * frames, kPitchAlign);
	}
	std::vector<size_t> mfma_off(m_steps.size(), 0);
	for (size_t k = 0; k < m_steps.size(); ++k) {
		if (!m_steps[k].dead && mfma_step(m_steps[k])) {
			mfma_off[k] = off;
			off += mfma_mid_bytes(m_steps[k], frames);
		}
	}
	std::vector<size_t> sur_off(m_steps.size(), 0);
	for (size_t k = 0; k < m_steps.size(); ++k) {
		if (const size_t b = surrogate_bytes(m_steps[k], frames)) {
			sur_off[k] = off;
			off += b;
		}
	}
	if (off && !tmp)
		fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "device tmp buffer required");

	auto resolve = [&](const Buf &b) {
		dev::PlaneArg p{};
		switch (b.kind) {
		case Buf::SRC:
			p = dev::PlaneArg{ src.ptr[b.idx], src.stride[b.idx], src.fstride[b.idx] };
			break;
		case Buf::DST:
			p = dev::PlaneArg{ dst.ptr[b.idx], dst.stride[b.idx], dst.fstride[b.idx] };
			break;
		case Buf::TMP: {
			const Slot &sl = m_slots[b.idx];
			p = dev::PlaneArg{ static_cast<char *>(tmp) + slot_off[b.idx], static_cast<long long>(sl.pitch),
			                   static_cast<long long>(sl.pitch * sl.height) };
			break;
		}
		default:
			break;
		}
		if (b.kind != Buf::NONE && !p.ptr)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "missing plane buffer");
		return p;
	};

	for (size_t si = 0; si < m_steps.size(); ++si) {
		const Step &s = m_steps[si];
		int rc = 0;
		if (events)
			check_hip(hipEventRecord(static_cast<hipEvent_t>(events[si]), static_cast<hipStream_t>(stream)), "hipEventRecord");
		if (!s.dead && mfma_step(s)) {
			run_mfma_resize(s, resolve(s.in[0]), resolve(s.out[0]), static_cast<char *>(tmp) + mfma_off[si], frames, stream);
			continue;
		}
		switch (s.kind) {
		case Step::RESIZE: {
			const dev::PlaneArg in = resolve(s.in[0]), out = resolve(s.out[0]);
			if (s.imma.ok && aligned16(in) && aligned16(out)) {
				const ImmaTables &it = imma_tables_for_current_device();
				dev::ImmaArg a{};
				a.src = in;
				a.dst = out;
				a.frames = frames;
				a.src_w = s.in_w;
				a.src_h = s.in_h;
				a.dst_w = s.width;
				a.dst_h = s.height;
				a.pixel_max = s.pixel_max;
				a.in = s.ix[0];
				a.h = it.pass.at(2 * s.tab[0]);
				a.v = it.pass.at(2 * s.tab[1] + 1);
				a.nh = (s.width + 15) / 16;
				a.nv = (s.height + 15) / 16;
				a.rb = s.imma.rb;
				a.mb = s.imma.mb;
				a.mid_pitch = s.imma.mid_pitch;
				a.in_cols = s.imma.in_cols;
				a.in_pitch = s.imma.in_pitch;
				a.lds = s.imma.lds;
				rc = dev::launch_resize_imma(a, stream);
				break;
			}
			const dev::ResizeArg a = resize_arg(s, in, out, frames, tabs);
			rc = dev::launch_resize(a, stream);
			break;
		}
		case Step::POINT: {
			dev::PointArg a{};
			if (!s.fill)
				a.src = resolve(s.in[0]);
			a.dst = resolve(s.out[0]);
			a.frames = frames;
			a.width = s.width;
			a.height = s.height;
			a.crop_left = s.crop_left;
			a.crop_top = s.crop_top;
			a.in = s.ix[0];
			a.out = s.ox[0];
			a.fill = s.fill;
			a.fill_bits = s.fill_bits;
			a.alpha_op = s.alpha_op;
			if (s.alpha_op)
				a.alpha = resolve(s.in[1]);
			rc = dev::launch_point(a, stream);
			break;
		}
		case Step::CSP: {
			dev::CspArg a{};
			for (int k = 0; k < 3; ++k) {
				a.src[k] = resolve(s.in[k]);
				a.dst[k] = resolve(s.out[k]);
				a.in[k] = s.ix[k];
				a.out[k] = s.ox[k];
			}
			a.frames = frames;
			a.width = s.width;
			a.height = s.height;
			a.prog = s.prog;
			for (int k = 0; k < a.prog.n; ++k)
				if (a.prog.op[k].kind == CSP_LUT_LINEAR || a.prog.op[k].kind == CSP_LUT_GAMMA)
					a.prog.op[k].lut = tabs.lut.at(static_cast<int>(si) * 8 + k);
			const std::vector<CspSegment> segs = csp_segments(s, a.prog);
			if (!segs.empty()) {
				run_learned_csp(s, segs, a.src, a.dst, static_cast<char *>(tmp) + sur_off[si], frames, stream);
				break;
			}
			rc = dev::launch_colorspace(a, stream);
			break;
		}
		case Step::UNRESIZE: {
			dev::UnresizeArg a = tabs.unresize.at(s.tab[0]);
			a.src = resolve(s.in[0]);
			a.dst = resolve(s.out[0]);
			a.frames = frames;
			a.src_w = s.in_w;
			a.src_h = s.in_h;
			a.dst_w = s.width;
			a.dst_h = s.height;
			a.horizontal = s.h_first ? 1 : 0;
			rc = dev::launch_unresize(a, stream);
			break;
		}
		case Step::ED: {
			dev::EdArg a{};
			for (int p = 0; p < s.ed_planes; ++p) {
				a.src[p] = resolve(s.in[p]);
				a.dst[p] = resolve(s.out[p]);
				a.scale[p] = s.ed_scale[p];
				a.offset[p] = s.ed_offset[p];
			}
			a.planes = static_cast<unsigned>(s.ed_planes);
			a.frames = frames;
			a.width = s.width;
			a.height = s.height;
			a.in = s.ix[0];
			a.out = s.ox[0];
			rc = dev::launch_error_diffusion(a, stream);
			break;
		}
		}
		if (rc != 0)
			fail(ZIMG_ERROR_UNKNOWN, std::string("kernel launch failed: ") + s.label + ": " + hipGetErrorString(static_cast<hipError_t>(rc)));
	}
	if (events)
		check_hip(hipEventRecord(static_cast<hipEvent_t>(events[m_steps.size()]), static_cast<hipStream_t>(stream)), "hipEventRecord");
}

std::string DeviceProgram::describe() const
{
	std::string r;
	for (auto &s : m_steps) {
		char buf[256];
		static const char *kinds[] = { "resize", "point", "colorspace", "error_diffusion", "unresize" };
		if (s.kind == Step::RESIZE)
			std::snprintf(buf, sizeof(buf), "%s %ux%u->%ux%u passes=%d %s tile=%ux%u lds=%zu tiled=%d(%ux%u lds=%zu): %s\n", kinds[s.kind],
			              s.in_w, s.in_h, s.width, s.height, s.passes, s.h_first ? "h_first" : "v_first", s.tile_w, s.tile_h,
			              (s.word ? 2 : 4) * (static_cast<size_t>(s.lds_in_rows) * s.lds_in_cols + static_cast<size_t>(s.lds_mid_rows) * s.lds_mid_cols),
			              s.fast.fwh, s.fast.tile_w, s.fast.tile_h,
			              (s.word ? 2 : 4) * (static_cast<size_t>(s.fast.in_rows) * s.fast.in_cols + static_cast<size_t>(s.fast.mid_rows) * s.fast.mid_cols),
			              s.label.c_str());
		else
			std::snprintf(buf, sizeof(buf), "%s %ux%u: %s\n", kinds[s.kind], s.width, s.height, s.label.c_str());
		r += buf;
		if (s.kind == Step::CSP) {
			const std::vector<CspSegment> segs = csp_segments(s, s.prog);
			if (!segs.empty()) {
				std::string t = "  learned transfers, launches:";
				for (const CspSegment &g : segs)
					t += g.sur ? " surrogate(f" + std::to_string(g.sur->func()) + ")" : " colorspace[" + g.prog.describe() + "]";
				r += t + "\n";
			} else if (dev::colorspace_has_chain_kernel(s.prog)) {
				r += "  chain kernel: operations fixed at compile time\n";
			}
		}
		if (s.kind == Step::RESIZE && s.imma.ok) {
			std::snprintf(buf, sizeof(buf), "  integer matrix cores: one launch, %ux%u tiles, lds=%u\n", dev::kImmaTileW, 16 * s.imma.rb,
			              s.imma.lds);
			r += buf;
		}
		if (!s.dead && mfma_step(s)) {
			const FusedPlan fp = mfma_fused_plan(s);
			if (fp.ok)
				std::snprintf(buf, sizeof(buf), "  matrix cores %s: one launch, %ux%u tiles, lds=%u\n", m_precision == 2 ? "bf16" : "fp16",
				              16 * dev::kFusedWaves, fp.tile_h, fp.lds);
			else
				std::snprintf(buf, sizeof(buf), "  matrix cores %s: %d launch(es) through an FP32 intermediate\n", m_precision == 2 ? "bf16" : "fp16",
				              s.passes);
			r += buf;
		}
	}
	return r;
}

} // namespace zamd
