// The C API (include/zimg.h). Enumeration translation, versioned structure
// import and error reporting follow the reference's api/zimg.cpp (:33-687) so
// that callers observe the same codes, defaults and version semantics; graph
// construction and execution are this library's (planner + device program).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "device.h"
#include "jobs.h"
#include "learned.h"
#include "graph.h"
#include "zimg.h"

namespace {

// Staging of one host-buffer call (zimg_filter_graph_process): pinned host
// rows of every plane, their device copies plus the program's device scratch,
// and the stream the call runs on. Created on first need, then reused.
struct HostStaging {
	int device = -1;
	char *host = nullptr;
	size_t host_size = 0;
	void *dev = nullptr;
	size_t dev_size = 0;
	hipStream_t stream = nullptr;

	HostStaging(const HostStaging &) = delete;
	HostStaging &operator=(const HostStaging &) = delete;
	HostStaging() = default;
	~HostStaging()
	{
		int cur = 0;
		if (hipGetDevice(&cur) != hipSuccess)
			return;
		(void)hipSetDevice(device);
		if (stream)
			(void)hipStreamDestroy(stream);
		if (dev)
			(void)hipFree(dev);
		if (host)
			(void)hipHostFree(host);
		(void)hipSetDevice(cur);
	}
};

} // namespace

// Device scratch of the allocation-free entry point
// (zimg_filter_graph_process_device_batch_noalloc): the intermediate planes a
// graph keeps in its device tmp, for callers that cannot allocate. Every GPU
// the library runs on holds one copy until the process ends; in host code
// the array's name is not a device address, so launches use the address
// hipGetSymbolAddress returns for the current device.
constexpr size_t kNoAllocScratchBytes = size_t{ 128 } << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

std::mutex g_scratch_mutex;
std::vector<void *> g_scratch_ptr; // by HIP device

// Address of this device's copy of `scratch`, looked up once per device with
// that device current.
void *scratch_for_current_device()
{
	int device = 0;
	if (hipGetDevice(&device) != hipSuccess)
		zamd::fail(ZIMG_ERROR_UNKNOWN, "no HIP device");
	std::lock_guard<std::mutex> lock(g_scratch_mutex);
	if (static_cast<size_t>(device) >= g_scratch_ptr.size())
		g_scratch_ptr.resize(static_cast<size_t>(device) + 1, nullptr);
	if (!g_scratch_ptr[device]) {
		void *p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p)
			zamd::fail(ZIMG_ERROR_UNKNOWN, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
		g_scratch_ptr[device] = p;
	}
	return g_scratch_ptr[device];
}

} // namespace

struct zimg_filter_graph {
	std::unique_ptr<zamd::DeviceProgram> prog;
	unsigned src_subsample_h = 0, dst_subsample_h = 0;
	unsigned src_bpp = 1, dst_bpp = 1;

	// Staging sets of zimg_filter_graph_process. A call takes a free set of its
	// device (or makes one), so concurrent calls run on their own streams and
	// a steady sequence of calls allocates nothing; the lock covers only the
	// hand-over.
	mutable std::mutex pool_mutex;
	mutable std::vector<std::unique_ptr<HostStaging>> pool;
	mutable unsigned pool_sets = 0;
};

namespace {

using namespace zamd;

constexpr unsigned API_2_0 = ZIMG_MAKE_API_VERSION(2, 0);
constexpr unsigned API_2_1 = ZIMG_MAKE_API_VERSION(2, 1);
constexpr unsigned API_2_2 = ZIMG_MAKE_API_VERSION(2, 2);
constexpr unsigned API_2_4 = ZIMG_MAKE_API_VERSION(2, 4);
constexpr unsigned API_2_5 = ZIMG_MAKE_API_VERSION(2, 5);
constexpr unsigned VERSION[3] = { 3, 0, 6 };

thread_local zimg_error_code_e g_last_error = ZIMG_ERROR_SUCCESS;
thread_local std::string g_last_message;

zimg_error_code_e record(zimg_error_code_e code, const char *msg)
{
	g_last_error = code;
	try {
		g_last_message = msg ? msg : "";
	} catch (...) {
		g_last_message.clear();
	}
	return code;
}

zimg_error_code_e handle_current_exception()
{
	try {
		throw;
	} catch (const Error &e) {
		return record(e.code, e.what());
	} catch (const std::bad_alloc &) {
		return record(ZIMG_ERROR_OUT_OF_MEMORY, "out of memory");
	} catch (const std::exception &e) {
		return record(ZIMG_ERROR_UNKNOWN, e.what());
	} catch (...) {
		return record(ZIMG_ERROR_UNKNOWN, "unknown error");
	}
}

[[noreturn]] void enum_error(const char *msg) { fail(ZIMG_ERROR_ENUM_OUT_OF_RANGE, msg); }

PixelType tr_pixel(zimg_pixel_type_e t)
{
	switch (t) {
	case ZIMG_PIXEL_BYTE: return PixelType::BYTE;
	case ZIMG_PIXEL_WORD: return PixelType::WORD;
	case ZIMG_PIXEL_HALF: return PixelType::HALF;
	case ZIMG_PIXEL_FLOAT: return PixelType::FLOAT;
	default: enum_error("unrecognized pixel type");
	}
}

bool tr_range(zimg_pixel_range_e r)
{
	switch (r) {
	case ZIMG_RANGE_LIMITED: return false;
	case ZIMG_RANGE_FULL: return true;
	default: enum_error("unrecognized pixel range");
	}
}

ColorFamily tr_family(zimg_color_family_e f)
{
	switch (f) {
	case ZIMG_COLOR_GREY: return ColorFamily::GREY;
	case ZIMG_COLOR_RGB: return ColorFamily::RGB;
	case ZIMG_COLOR_YUV: return ColorFamily::YUV;
	default: enum_error("unrecognized color family");
	}
}

AlphaType tr_alpha(zimg_alpha_type_e a)
{
	switch (a) {
	case ZIMG_ALPHA_NONE: return AlphaType::NONE;
	case ZIMG_ALPHA_STRAIGHT: return AlphaType::STRAIGHT;
	case ZIMG_ALPHA_PREMULTIPLIED: return AlphaType::PREMULTIPLIED;
	default: enum_error("unrecognized alpha type");
	}
}

FieldParity tr_parity(zimg_field_parity_e p)
{
	switch (p) {
	case ZIMG_FIELD_PROGRESSIVE: return FieldParity::PROGRESSIVE;
	case ZIMG_FIELD_TOP: return FieldParity::TOP;
	case ZIMG_FIELD_BOTTOM: return FieldParity::BOTTOM;
	default: enum_error("unrecognized field parity");
	}
}

void tr_chroma(zimg_chroma_location_e c, ChromaLocW &w, ChromaLocH &h)
{
	switch (c) {
	case ZIMG_CHROMA_LEFT: w = ChromaLocW::LEFT; h = ChromaLocH::CENTER; return;
	case ZIMG_CHROMA_CENTER: w = ChromaLocW::CENTER; h = ChromaLocH::CENTER; return;
	case ZIMG_CHROMA_TOP_LEFT: w = ChromaLocW::LEFT; h = ChromaLocH::TOP; return;
	case ZIMG_CHROMA_TOP: w = ChromaLocW::CENTER; h = ChromaLocH::TOP; return;
	case ZIMG_CHROMA_BOTTOM_LEFT: w = ChromaLocW::LEFT; h = ChromaLocH::BOTTOM; return;
	case ZIMG_CHROMA_BOTTOM: w = ChromaLocW::CENTER; h = ChromaLocH::BOTTOM; return;
	default: enum_error("unrecognized chroma location");
	}
}

// ITU enumerations: out of [0, 255] is a range error, an unknown code inside it
// means "no conversion" (zimg.cpp:114-124).
template <class E>
void itu_check(E v, const char *msg)
{
	if (static_cast<int>(v) < 0 || static_cast<int>(v) > 255)
		enum_error(msg);
}

Matrix tr_matrix(zimg_matrix_coefficients_e m)
{
	itu_check(m, "unrecognized matrix coefficients");
	switch (m) {
	case ZIMG_MATRIX_RGB: return Matrix::RGB;
	case ZIMG_MATRIX_BT709: return Matrix::REC_709;
	case ZIMG_MATRIX_UNSPECIFIED: return Matrix::UNSPECIFIED;
	case ZIMG_MATRIX_FCC: return Matrix::FCC;
	case ZIMG_MATRIX_BT470_BG:
	case ZIMG_MATRIX_ST170_M: return Matrix::REC_601;
	case ZIMG_MATRIX_ST240_M: return Matrix::SMPTE_240M;
	case ZIMG_MATRIX_YCGCO: return Matrix::YCGCO;
	case ZIMG_MATRIX_BT2020_NCL: return Matrix::REC_2020_NCL;
	case ZIMG_MATRIX_BT2020_CL: return Matrix::REC_2020_CL;
	case ZIMG_MATRIX_CHROMATICITY_DERIVED_NCL: return Matrix::CHROMATICITY_DERIVED_NCL;
	case ZIMG_MATRIX_CHROMATICITY_DERIVED_CL: return Matrix::CHROMATICITY_DERIVED_CL;
	case ZIMG_MATRIX_BT2100_ICTCP: return Matrix::REC_2100_ICTCP;
	default: fail(ZIMG_ERROR_NO_COLORSPACE_CONVERSION, "unrecognized matrix coefficients");
	}
}

Transfer tr_transfer(zimg_transfer_characteristics_e t)
{
	itu_check(t, "unrecognized transfer characteristics");
	switch (t) {
	case ZIMG_TRANSFER_BT709:
	case ZIMG_TRANSFER_BT601:
	case ZIMG_TRANSFER_BT2020_10:
	case ZIMG_TRANSFER_BT2020_12: return Transfer::REC_709;
	case ZIMG_TRANSFER_UNSPECIFIED: return Transfer::UNSPECIFIED;
	case ZIMG_TRANSFER_ST240_M: return Transfer::SMPTE_240M;
	case ZIMG_TRANSFER_BT470_M: return Transfer::REC_470_M;
	case ZIMG_TRANSFER_BT470_BG: return Transfer::REC_470_BG;
	case ZIMG_TRANSFER_IEC_61966_2_4: return Transfer::XVYCC;
	case ZIMG_TRANSFER_BT1361: return Transfer::REC_1361;
	case ZIMG_TRANSFER_IEC_61966_2_1: return Transfer::SRGB;
	case ZIMG_TRANSFER_LINEAR: return Transfer::LINEAR;
	case ZIMG_TRANSFER_LOG_100: return Transfer::LOG_100;
	case ZIMG_TRANSFER_LOG_316: return Transfer::LOG_316;
	case ZIMG_TRANSFER_ST2084: return Transfer::ST_2084;
	case ZIMG_TRANSFER_ST428: return Transfer::ST_428;
	case ZIMG_TRANSFER_ARIB_B67: return Transfer::ARIB_B67;
	default: fail(ZIMG_ERROR_NO_COLORSPACE_CONVERSION, "unrecognized transfer characteristics");
	}
}

Primaries tr_primaries(zimg_color_primaries_e p)
{
	itu_check(p, "unrecognized color primaries");
	switch (p) {
	case ZIMG_PRIMARIES_BT470_M: return Primaries::REC_470_M;
	case ZIMG_PRIMARIES_BT470_BG: return Primaries::REC_470_BG;
	case ZIMG_PRIMARIES_BT709: return Primaries::REC_709;
	case ZIMG_PRIMARIES_UNSPECIFIED: return Primaries::UNSPECIFIED;
	case ZIMG_PRIMARIES_ST170_M:
	case ZIMG_PRIMARIES_ST240_M: return Primaries::SMPTE_C;
	case ZIMG_PRIMARIES_FILM: return Primaries::FILM;
	case ZIMG_PRIMARIES_BT2020: return Primaries::REC_2020;
	case ZIMG_PRIMARIES_ST428: return Primaries::XYZ;
	case ZIMG_PRIMARIES_ST431_2: return Primaries::DCI_P3;
	case ZIMG_PRIMARIES_ST432_1: return Primaries::DCI_P3_D65;
	case ZIMG_PRIMARIES_EBU3213_E: return Primaries::EBU_3213_E;
	default: fail(ZIMG_ERROR_NO_COLORSPACE_CONVERSION, "unrecognized color primaries");
	}
}

Dither tr_dither(zimg_dither_type_e d)
{
	switch (d) {
	case ZIMG_DITHER_NONE: return Dither::NONE;
	case ZIMG_DITHER_ORDERED: return Dither::ORDERED;
	case ZIMG_DITHER_RANDOM: return Dither::RANDOM;
	case ZIMG_DITHER_ERROR_DIFFUSION: return Dither::ERROR_DIFFUSION;
	default: enum_error("unrecognized dither type");
	}
}

void tr_cpu(zimg_cpu_type_e c)
{
	const int v = static_cast<int>(c);
	if (!(v >= 0 && v <= 2) && !(v >= 1000 && v <= 1016))
		enum_error("unrecognized cpu type");
}

ResampleFilter tr_filter(zimg_resample_filter_e f, double a, double b)
{
	// zimg.cpp:320-351
	ResampleFilter r;
	r.type = f;
	switch (f) {
	case ZIMG_RESIZE_POINT:
	case ZIMG_RESIZE_BILINEAR:
	case ZIMG_RESIZE_SPLINE16:
	case ZIMG_RESIZE_SPLINE36:
	case ZIMG_RESIZE_SPLINE64:
		return r;
	case ZIMG_RESIZE_BICUBIC:
		r.param_a = std::isnan(a) ? 0.0 : a;
		r.param_b = std::isnan(b) ? 0.5 : b;
		return r;
	case ZIMG_RESIZE_LANCZOS: {
		const double taps = std::isnan(a) ? 3.0 : std::clamp(a, 1.0, static_cast<double>(UINT_MAX));
		const unsigned t = static_cast<unsigned>(taps);
		if (t >= 16)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "lanczos tap count too high");
		r.param_a = t;
		return r;
	}
	default:
		// ZIMG_RESIZE_UNRESIZE (zimg.cpp:21,320-323): no filter object; the
		// builder unresizes instead (graphbuilder.cpp:761-796).
		if (static_cast<int>(f) == -1)
			return r;
		enum_error("unrecognized resampling filter");
	}
}

ImageState import_state_common(const zimg_image_format &f)
{
	ImageState s;
	if (f.version >= API_2_0) {
		s.width = f.width;
		s.height = f.height;
		s.type = tr_pixel(f.pixel_type);
		s.subsample_w = f.subsample_w;
		s.subsample_h = f.subsample_h;
		s.color = tr_family(f.color_family);
		s.depth = f.depth ? f.depth : pixel_depth(s.type);
		s.fullrange = tr_range(f.pixel_range);
		s.parity = tr_parity(f.field_parity);
		tr_chroma(f.chroma_location, s.chroma_w, s.chroma_h);
	}
	if (f.version >= API_2_1) {
		s.active_left = std::isnan(f.active_region.left) ? 0 : f.active_region.left;
		s.active_top = std::isnan(f.active_region.top) ? 0 : f.active_region.top;
		s.active_width = std::isnan(f.active_region.width) ? f.width : f.active_region.width;
		s.active_height = std::isnan(f.active_region.height) ? f.height : f.active_region.height;
	} else {
		s.active_left = 0;
		s.active_top = 0;
		s.active_width = f.width;
		s.active_height = f.height;
	}
	if (f.version >= API_2_4)
		s.alpha = tr_alpha(f.alpha);
	return s;
}

void import_states(const zimg_image_format &src, const zimg_image_format &dst, ImageState &s, ImageState &d)
{
	if (src.version < API_2_0 || dst.version < API_2_0)
		fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "API version invalid");
	s = import_state_common(src);
	d = import_state_common(dst);
	// Unenumerated colourspaces are accepted when they form the no-op case (zimg.cpp:414-431).
	if (src.color_family == dst.color_family && src.matrix_coefficients == dst.matrix_coefficients &&
	    src.transfer_characteristics == dst.transfer_characteristics && src.color_primaries == dst.color_primaries) {
		s.colorspace = Colorspace{};
		d.colorspace = Colorspace{};
	} else {
		s.colorspace = { tr_matrix(src.matrix_coefficients), tr_transfer(src.transfer_characteristics), tr_primaries(src.color_primaries) };
		d.colorspace = { tr_matrix(dst.matrix_coefficients), tr_transfer(dst.transfer_characteristics), tr_primaries(dst.color_primaries) };
	}
}

BuildParams import_params(const zimg_graph_builder_params &p)
{
	if (p.version < API_2_0)
		fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "API version invalid");
	BuildParams r;
	if (p.version >= API_2_0) {
		r.filter = tr_filter(p.resample_filter, p.filter_param_a, p.filter_param_b);
		r.filter_uv = tr_filter(p.resample_filter_uv, p.filter_param_a_uv, p.filter_param_b_uv);
		r.dither = tr_dither(p.dither_type);
		r.unresize = static_cast<int>(p.resample_filter) == -1;
		if (!r.unresize && static_cast<int>(p.resample_filter_uv) == -1)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "unresize is selected by resample_filter");
		tr_cpu(p.cpu_type);
	}
	if (p.version >= API_2_2) {
		r.peak_luminance = p.nominal_peak_luminance;
		r.approximate_gamma = !!p.allow_approximate_gamma;
	}
	if (p.version >= API_2_5) {
		r.scene_referred = !!p.scene_referred;
		r.chromatic_adaptation = !!p.chromatic_adaptation;
	}
	return r;
}

const zamd::DeviceProgram &program(const zimg_filter_graph *g)
{
	if (!g || !g->prog)
		fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null graph");
	return *g->prog;
}

template <class Buf>
DeviceBuffers import_device_buffer(const Buf *b, const ptrdiff_t *fstride, const unsigned *plane_w, bool is_src)
{
	if (!b)
		fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null buffer");
	if (b->version < API_2_0)
		fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "API version invalid");
	const unsigned nplanes = b->version >= API_2_4 ? 4 : 3;
	DeviceBuffers d;
	for (unsigned p = 0; p < nplanes; ++p) {
		if (!plane_w[p])
			continue;
		if (b->plane[p].mask != ZIMG_BUFFER_MAX)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "device buffers must be whole planes (mask = ZIMG_BUFFER_MAX)");
		d.ptr[p] = static_cast<char *>(const_cast<void *>(static_cast<const void *>(b->plane[p].data)));
		d.stride[p] = b->plane[p].stride;
		d.fstride[p] = fstride ? fstride[p] : 0;
		if (!d.ptr[p])
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, is_src ? "missing source plane" : "missing destination plane");
	}
	return d;
}

// Where each planar sample of a graph-side plane lives in a packed layout
// (zimg_device_packed_buffer): packed plane index, element step and position.
struct PackMap {
	int packed = -1; // -1: the plane is absent
	unsigned step = 1, pos = 0;
};

void pack_maps(int packing, const Plan &plan, bool src, PixelType type, PackMap map[4])
{
	const unsigned *w = src ? plan.src_w : plan.dst_w;
	const unsigned *h = src ? plan.src_h : plan.dst_h;
	for (int q = 0; q < 4; ++q)
		map[q] = PackMap{};
	if (packing == ZIMG_PACKING_PLANAR) {
		for (int q = 0; q < 4; ++q)
			if (w[q])
				map[q] = PackMap{ q, 1, 0 };
		return;
	}
	if (!w[1] || !w[2] || w[3])
		fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "packed layouts carry YUV without alpha");
	if (packing == ZIMG_PACKING_NV12) {
		if (w[1] * 2 != w[0] || h[1] * 2 != h[0] || (type != PixelType::BYTE && type != PixelType::WORD))
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "NV12 / P010 layout needs 4:2:0 BYTE or WORD");
		map[0] = PackMap{ 0, 1, 0 };
		map[1] = PackMap{ 1, 2, 0 };
		map[2] = PackMap{ 1, 2, 1 };
	} else if (packing == ZIMG_PACKING_YUY2 || packing == ZIMG_PACKING_UYVY) {
		if (w[1] * 2 != w[0] || h[1] != h[0] || type != PixelType::BYTE)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "YUY2 / UYVY layout needs 4:2:2 BYTE");
		const bool yuy2 = packing == ZIMG_PACKING_YUY2;
		map[0] = PackMap{ 0, 2, yuy2 ? 0u : 1u };
		map[1] = PackMap{ 0, 4, yuy2 ? 1u : 0u };
		map[2] = PackMap{ 0, 4, yuy2 ? 3u : 2u };
	} else {
		fail(ZIMG_ERROR_ENUM_OUT_OF_RANGE, "unrecognized packing");
	}
}

// Bytes of planar staging one side of a packed call needs per frame (planes
// used in place need none), and each staged plane's offset and pitch.
size_t packed_staging(const PackMap map[4], const unsigned *w, const unsigned *h, unsigned bpp, unsigned shift, bool stage[4],
                      size_t off[4], size_t pitch[4], unsigned frames, size_t base)
{
	size_t o = base;
	for (int q = 0; q < 4; ++q) {
		stage[q] = map[q].packed >= 0 && (map[q].step != 1 || shift != 0);
		pitch[q] = (static_cast<size_t>(w[q]) * bpp + 255) / 256 * 256;
		off[q] = o;
		if (stage[q])
			o += pitch[q] * h[q] * frames;
	}
	return o - base;
}

} // namespace

extern "C" {

void zimg_get_version_info(unsigned *major, unsigned *minor, unsigned *micro)
{
	*major = VERSION[0];
	*minor = VERSION[1];
	*micro = VERSION[2];
}

unsigned zimg_get_api_version(unsigned *major, unsigned *minor)
{
	if (major) *major = ZIMG_API_VERSION_MAJOR;
	if (minor) *minor = ZIMG_API_VERSION_MINOR;
	return ZIMG_API_VERSION;
}

zimg_error_code_e zimg_get_last_error(char *err_msg, size_t n)
{
	if (err_msg && n) {
		std::strncpy(err_msg, g_last_message.c_str(), n);
		err_msg[n - 1] = '\0';
	}
	return g_last_error;
}

void zimg_clear_last_error(void)
{
	g_last_error = ZIMG_ERROR_SUCCESS;
	g_last_message.clear();
	g_last_message.shrink_to_fit();
}

unsigned zimg_select_buffer_mask(unsigned count)
{
	if (count <= 1)
		return 0;
	return ZIMG_BUFFER_MAX >> __builtin_clz(count - 1);
}

void zimg_filter_graph_free(zimg_filter_graph *ptr) { delete ptr; }

zimg_error_code_e zimg_filter_graph_get_tmp_size(const zimg_filter_graph *ptr, size_t *out)
{
	try {
		program(ptr);
		// The host path stages planes through the graph's own pinned buffers.
		*out = 0;
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

zimg_error_code_e zimg_filter_graph_get_input_buffering(const zimg_filter_graph *ptr, unsigned *out)
{
	try {
		program(ptr);
		*out = 1U << ptr->src_subsample_h; // rows are copied out as soon as each unpack returns
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

zimg_error_code_e zimg_filter_graph_get_output_buffering(const zimg_filter_graph *ptr, unsigned *out)
{
	try {
		program(ptr);
		*out = 1U << ptr->dst_subsample_h;
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

zimg_error_code_e zimg_filter_graph_process(const zimg_filter_graph *ptr, const zimg_image_buffer_const *src, const zimg_image_buffer *dst, void *tmp,
                                            zimg_filter_graph_callback unpack_cb, void *unpack_user,
                                            zimg_filter_graph_callback pack_cb, void *pack_user)
{
	(void)tmp; // zimg_filter_graph_get_tmp_size is 0: staging belongs to the graph
	try {
		const DeviceProgram &prog = program(ptr);
		const Plan &plan = prog.plan();
		if (!src || !dst)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null buffer");
		if (src->version < API_2_0 || dst->version < API_2_0)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "API version invalid");

		size_t src_off[4] = {}, dst_off[4] = {};
		size_t o = 0;
		for (int q = 0; q < 4; ++q) {
			src_off[q] = o;
			o += static_cast<size_t>(plan.src_w[q]) * plan.src_h[q] * ptr->src_bpp;
		}
		for (int q = 0; q < 4; ++q) {
			dst_off[q] = o;
			o += static_cast<size_t>(plan.dst_w[q]) * plan.dst_h[q] * ptr->dst_bpp;
		}
		const size_t src_bytes = dst_off[0];
		const size_t dst_bytes = o - src_bytes;
		const size_t dtmp_off = (src_bytes + dst_bytes + 255) / 256 * 256;
		const size_t dev_need = dtmp_off + prog.tmp_size(1);

		int device = 0;
		if (hipGetDevice(&device) != hipSuccess)
			fail(ZIMG_ERROR_UNKNOWN, "no HIP device");

		// Take a staging set of this device, or make one (first use, or more
		// concurrent callers than before).
		std::unique_ptr<HostStaging> st;
		{
			std::lock_guard<std::mutex> lock(ptr->pool_mutex);
			for (auto it = ptr->pool.begin(); it != ptr->pool.end(); ++it) {
				if ((*it)->device == device && (*it)->host_size >= o && (*it)->dev_size >= dev_need) {
					st = std::move(*it);
					ptr->pool.erase(it);
					break;
				}
			}
		}
		if (!st) {
			st = std::make_unique<HostStaging>();
			st->device = device;
			if (hipHostMalloc(reinterpret_cast<void **>(&st->host), std::max<size_t>(o, 1), hipHostMallocDefault) != hipSuccess)
				fail(ZIMG_ERROR_OUT_OF_MEMORY, "pinned staging allocation failed");
			st->host_size = o;
			if (hipMalloc(&st->dev, std::max<size_t>(dev_need, 1)) != hipSuccess)
				fail(ZIMG_ERROR_OUT_OF_MEMORY, "device staging allocation failed");
			st->dev_size = dev_need;
			if (hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking) != hipSuccess)
				fail(ZIMG_ERROR_UNKNOWN, "hipStreamCreate failed");
			std::lock_guard<std::mutex> lock(ptr->pool_mutex);
			++ptr->pool_sets;
		}
		// The set goes back to the pool however the call ends.
		struct Return {
			const zimg_filter_graph *g;
			std::unique_ptr<HostStaging> &st;
			~Return()
			{
				if (!st)
					return;
				std::lock_guard<std::mutex> lock(g->pool_mutex);
				g->pool.push_back(std::move(st));
			}
		} give_back{ ptr, st };
		char *host = st->host;
		char *dmem = static_cast<char *>(st->dev);

		// Unpack: rows in groups of the vertical subsampling unit (zimg.h:455-482).
		const unsigned ssh = ptr->src_subsample_h;
		for (unsigned i = 0; i < plan.src_h[0]; i += 1U << ssh) {
			if (unpack_cb && unpack_cb(unpack_user, i, 0, plan.src_w[0]))
				fail(ZIMG_ERROR_USER_CALLBACK_FAILED, "user callback failed");
			for (int q = 0; q < 4; ++q) {
				if (!plan.src_w[q])
					continue;
				const unsigned sh = (q == 1 || q == 2) ? ssh : 0;
				const size_t rowbytes = static_cast<size_t>(plan.src_w[q]) * ptr->src_bpp;
				for (unsigned r = i >> sh; r < ((i + (1U << ssh)) >> sh) && r < plan.src_h[q]; ++r) {
					const char *row = static_cast<const char *>(src->plane[q].data) + static_cast<ptrdiff_t>(r & src->plane[q].mask) * src->plane[q].stride;
					std::memcpy(host + src_off[q] + r * rowbytes, row, rowbytes);
				}
			}
		}

		if (hipMemcpyAsync(dmem, host, src_bytes, hipMemcpyHostToDevice, st->stream) != hipSuccess)
			fail(ZIMG_ERROR_UNKNOWN, "upload failed");
		DeviceBuffers ds, dd;
		for (int q = 0; q < 4; ++q) {
			if (plan.src_w[q]) {
				ds.ptr[q] = dmem + src_off[q];
				ds.stride[q] = static_cast<long long>(plan.src_w[q]) * ptr->src_bpp;
			}
			if (plan.dst_w[q]) {
				dd.ptr[q] = dmem + dst_off[q];
				dd.stride[q] = static_cast<long long>(plan.dst_w[q]) * ptr->dst_bpp;
			}
		}
		prog.run(ds, dd, dmem + dtmp_off, 1, st->stream);
		if (hipMemcpyAsync(host + src_bytes, dmem + src_bytes, dst_bytes, hipMemcpyDeviceToHost, st->stream) != hipSuccess)
			fail(ZIMG_ERROR_UNKNOWN, "download failed");
		if (hipStreamSynchronize(st->stream) != hipSuccess)
			fail(ZIMG_ERROR_UNKNOWN, "device execution failed");

		const unsigned dsh = ptr->dst_subsample_h;
		for (unsigned i = 0; i < plan.dst_h[0]; i += 1U << dsh) {
			for (int q = 0; q < 4; ++q) {
				if (!plan.dst_w[q])
					continue;
				const unsigned sh = (q == 1 || q == 2) ? dsh : 0;
				const size_t rowbytes = static_cast<size_t>(plan.dst_w[q]) * ptr->dst_bpp;
				for (unsigned r = i >> sh; r < ((i + (1U << dsh)) >> sh) && r < plan.dst_h[q]; ++r) {
					char *row = static_cast<char *>(dst->plane[q].data) + static_cast<ptrdiff_t>(r & dst->plane[q].mask) * dst->plane[q].stride;
					std::memcpy(row, host + dst_off[q] + r * rowbytes, rowbytes);
				}
			}
			if (pack_cb && pack_cb(pack_user, i, 0, plan.dst_w[0]))
				fail(ZIMG_ERROR_USER_CALLBACK_FAILED, "user callback failed");
		}
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

void zimg_image_format_default(zimg_image_format *ptr, unsigned version)
{
	ptr->version = version;
	if (version >= API_2_0) {
		ptr->width = 0;
		ptr->height = 0;
		ptr->pixel_type = static_cast<zimg_pixel_type_e>(-1);
		ptr->subsample_w = 0;
		ptr->subsample_h = 0;
		ptr->color_family = ZIMG_COLOR_GREY;
		ptr->matrix_coefficients = ZIMG_MATRIX_UNSPECIFIED;
		ptr->transfer_characteristics = ZIMG_TRANSFER_UNSPECIFIED;
		ptr->color_primaries = ZIMG_PRIMARIES_UNSPECIFIED;
		ptr->depth = 0;
		ptr->pixel_range = ZIMG_RANGE_LIMITED;
		ptr->field_parity = ZIMG_FIELD_PROGRESSIVE;
		ptr->chroma_location = ZIMG_CHROMA_LEFT;
	}
	if (version >= API_2_1) {
		ptr->active_region.left = NAN;
		ptr->active_region.top = NAN;
		ptr->active_region.width = NAN;
		ptr->active_region.height = NAN;
	}
	if (version >= API_2_4)
		ptr->alpha = ZIMG_ALPHA_NONE;
}

void zimg_graph_builder_params_default(zimg_graph_builder_params *ptr, unsigned version)
{
	ptr->version = version;
	if (version >= API_2_0) {
		ptr->resample_filter = ZIMG_RESIZE_BICUBIC;
		ptr->filter_param_a = NAN;
		ptr->filter_param_b = NAN;
		ptr->resample_filter_uv = ZIMG_RESIZE_BILINEAR;
		ptr->filter_param_a_uv = NAN;
		ptr->filter_param_b_uv = NAN;
		ptr->dither_type = ZIMG_DITHER_NONE;
		ptr->cpu_type = ZIMG_CPU_AUTO;
	}
	if (version >= API_2_2) {
		ptr->nominal_peak_luminance = NAN;
		ptr->allow_approximate_gamma = 0;
	}
	if (version >= API_2_5) {
		ptr->scene_referred = 0;
		ptr->chromatic_adaptation = 0;
	}
}

zimg_filter_graph *zimg_filter_graph_build(const zimg_image_format *src_format, const zimg_image_format *dst_format, const zimg_graph_builder_params *params)
{
	try {
		if (!src_format || !dst_format)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null format");
		ImageState s, d;
		import_states(*src_format, *dst_format, s, d);
		BuildParams bp;
		if (params)
			bp = import_params(*params);
		Plan plan = build_plan(s, d, bp);
		auto g = std::make_unique<zimg_filter_graph>();
		g->prog = std::make_unique<DeviceProgram>(plan);
		// Coefficient tables go to the current device now, so that processing
		// on it allocates nothing (no device: they upload on first use).
		int ndev = 0;
		if (hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0)
			g->prog->prepare_current_device();
		else
			(void)hipGetLastError();
		g->src_subsample_h = s.color == ColorFamily::GREY ? 0 : s.subsample_h;
		g->dst_subsample_h = d.color == ColorFamily::GREY ? 0 : d.subsample_h;
		g->src_bpp = pixel_size(s.type);
		g->dst_bpp = pixel_size(d.type);
		return g.release();
	} catch (...) {
		handle_current_exception();
		return nullptr;
	}
}

zimg_error_code_e zimg_filter_graph_get_device_tmp_size(const zimg_filter_graph *ptr, unsigned count, size_t *out)
{
	try {
		*out = program(ptr).tmp_size(count ? count : 1);
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

zimg_error_code_e zimg_filter_graph_process_device_batch(const zimg_filter_graph *ptr, unsigned count,
                                                         const zimg_image_buffer_const *src, const ptrdiff_t src_frame_stride[4],
                                                         const zimg_image_buffer *dst, const ptrdiff_t dst_frame_stride[4],
                                                         void *tmp, void *stream)
{
	try {
		const DeviceProgram &prog = program(ptr);
		const Plan &plan = prog.plan();
		DeviceBuffers ds = import_device_buffer(src, src_frame_stride, plan.src_w, true);
		DeviceBuffers dd = import_device_buffer(dst, dst_frame_stride, plan.dst_w, false);
		prog.run(ds, dd, tmp, count, stream);
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

zimg_error_code_e zimg_filter_graph_autotune_device(zimg_filter_graph *ptr, unsigned count, const zimg_image_buffer_const *src,
                                                   const ptrdiff_t src_frame_stride[4], const zimg_image_buffer *dst,
                                                   const ptrdiff_t dst_frame_stride[4], void *tmp, void *stream, unsigned *changed)
{
	try {
		if (!ptr || !ptr->prog)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null graph");
		const Plan &plan = ptr->prog->plan();
		DeviceBuffers ds = import_device_buffer(src, src_frame_stride, plan.src_w, true);
		DeviceBuffers dd = import_device_buffer(dst, dst_frame_stride, plan.dst_w, false);
		const unsigned n = ptr->prog->autotune(ds, dd, tmp, count, stream);
		if (changed)
			*changed = n;
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

zimg_error_code_e zimg_filter_graph_process_device_batch_noalloc(const zimg_filter_graph *ptr, unsigned count,
                                                                 const zimg_image_buffer_const *src, const ptrdiff_t src_frame_stride[4],
                                                                 const zimg_image_buffer *dst, const ptrdiff_t dst_frame_stride[4],
                                                                 void *stream)
{
	try {
		const DeviceProgram &prog = program(ptr);
		const Plan &plan = prog.plan();
		const size_t need = prog.tmp_size(count ? count : 1);
		if (need > kNoAllocScratchBytes)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "the graph needs " + std::to_string(need) + " bytes of device scratch for " + std::to_string(count) +
			                                      " frame(s); the allocation-free entry point holds " + std::to_string(kNoAllocScratchBytes));
		if (!prog.resident_on_current_device())
			fail(ZIMG_ERROR_UNSUPPORTED_OPERATION,
			     "coefficient tables are not resident on the current device: build the graph (or set its precision) with this device current");
		DeviceBuffers ds = import_device_buffer(src, src_frame_stride, plan.src_w, true);
		DeviceBuffers dd = import_device_buffer(dst, dst_frame_stride, plan.dst_w, false);
		// Every intermediate plane is written in full before it is read, so
		// nothing in the scratch needs clearing between calls.
		prog.run(ds, dd, need ? scratch_for_current_device() : nullptr, count, stream);
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

zimg_error_code_e zimg_filter_graph_get_device_packed_tmp_size(const zimg_filter_graph *ptr, unsigned count, int src_packing,
                                                               int dst_packing, size_t *out)
{
	try {
		const DeviceProgram &prog = program(ptr);
		const Plan &plan = prog.plan();
		const unsigned n = count ? count : 1;
		PackMap ms[4], md[4];
		pack_maps(src_packing, plan, true, plan.src_type, ms);
		pack_maps(dst_packing, plan, false, plan.dst_type, md);
		bool st[4];
		size_t off[4], pitch[4];
		// shift 1 stands for "any nonzero": the luma plane is then staged too
		size_t total = packed_staging(ms, plan.src_w, plan.src_h, ptr->src_bpp, src_packing == ZIMG_PACKING_PLANAR ? 0 : 1, st, off, pitch, n, 0);
		total += packed_staging(md, plan.dst_w, plan.dst_h, ptr->dst_bpp, dst_packing == ZIMG_PACKING_PLANAR ? 0 : 1, st, off, pitch, n, 0);
		*out = (total + 255) / 256 * 256 + prog.tmp_size(n);
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

zimg_error_code_e zimg_filter_graph_process_device_packed(const zimg_filter_graph *ptr, unsigned count, const zimg_device_packed_buffer *src,
                                                          const zimg_device_packed_buffer *dst, void *tmp, void *stream)
{
	try {
		const DeviceProgram &prog = program(ptr);
		const Plan &plan = prog.plan();
		if (!src || !dst)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null buffer");
		if (src->version < API_2_0 || dst->version < API_2_0)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "API version invalid");
		const unsigned n = count ? count : 1;
		if (src->shift >= 8 * ptr->src_bpp || dst->shift >= 8 * ptr->dst_bpp ||
		    (ptr->src_bpp != 2 && src->shift) || (ptr->dst_bpp != 2 && dst->shift))
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "sample shift only for WORD samples, below 16");
		PackMap ms[4], md[4];
		pack_maps(src->packing, plan, true, plan.src_type, ms);
		pack_maps(dst->packing, plan, false, plan.dst_type, md);
		bool sst[4], dst_st[4];
		size_t soff[4], spitch[4], doff[4], dpitch[4];
		const unsigned sshift = src->packing == ZIMG_PACKING_PLANAR ? 0 : src->shift;
		const unsigned dshift = dst->packing == ZIMG_PACKING_PLANAR ? 0 : dst->shift;
		size_t used = packed_staging(ms, plan.src_w, plan.src_h, ptr->src_bpp, sshift, sst, soff, spitch, n, 0);
		used += packed_staging(md, plan.dst_w, plan.dst_h, ptr->dst_bpp, dshift, dst_st, doff, dpitch, n, used);
		const size_t prog_off = (used + 255) / 256 * 256;
		if ((used || prog.tmp_size(n)) && !tmp)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "device tmp buffer required");
		char *t = static_cast<char *>(tmp);

		// graph-side planar views and the pack launches around the program
		DeviceBuffers ds, dd;
		dev::PackArg unpack{}, pack{};
		unpack.frames = pack.frames = n;
		unpack.pack = 0;
		pack.pack = 1;
		auto side = [&](const zimg_device_packed_buffer *b, const PackMap *m, const bool *st, const size_t *off, const size_t *pitch,
		                const unsigned *w, const unsigned *h, unsigned bpp, unsigned shift, DeviceBuffers &out, dev::PackArg &pa) {
			for (int q = 0; q < 4; ++q) {
				if (m[q].packed < 0)
					continue;
				const int pp = m[q].packed;
				if (!b->data[pp])
					fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "missing packed plane");
				const dev::PlaneArg packed{ static_cast<char *>(b->data[pp]), b->stride[pp], b->frame_stride[pp] };
				if (!st[q]) {
					out.ptr[q] = packed.ptr;
					out.stride[q] = packed.stride;
					out.fstride[q] = packed.fstride;
					continue;
				}
				const dev::PlaneArg planar{ t + off[q], static_cast<long long>(pitch[q]), static_cast<long long>(pitch[q] * h[q]) };
				out.ptr[q] = planar.ptr;
				out.stride[q] = planar.stride;
				out.fstride[q] = planar.fstride;
				pa.p[pa.planes++] = dev::PackPlane{ planar, packed, w[q], h[q], bpp, m[q].step, m[q].pos, shift };
			}
		};
		side(src, ms, sst, soff, spitch, plan.src_w, plan.src_h, ptr->src_bpp, sshift, ds, unpack);
		side(dst, md, dst_st, doff, dpitch, plan.dst_w, plan.dst_h, ptr->dst_bpp, dshift, dd, pack);
		int rc = dev::launch_pack(unpack, stream);
		if (rc)
			fail(ZIMG_ERROR_UNKNOWN, std::string("unpack launch failed: ") + hipGetErrorString(static_cast<hipError_t>(rc)));
		prog.run(ds, dd, t ? t + prog_off : nullptr, n, stream);
		rc = dev::launch_pack(pack, stream);
		if (rc)
			fail(ZIMG_ERROR_UNKNOWN, std::string("pack launch failed: ") + hipGetErrorString(static_cast<hipError_t>(rc)));
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

zimg_error_code_e zimg_filter_graph_process_device(const zimg_filter_graph *ptr, const zimg_image_buffer_const *src,
                                                   const zimg_image_buffer *dst, void *tmp, void *stream)
{
	return zimg_filter_graph_process_device_batch(ptr, 1, src, nullptr, dst, nullptr, tmp, stream);
}

zimg_error_code_e zimg_filter_graph_process_device_profiled(const zimg_filter_graph *ptr, unsigned count,
                                                            const zimg_image_buffer_const *src, const ptrdiff_t src_frame_stride[4],
                                                            const zimg_image_buffer *dst, const ptrdiff_t dst_frame_stride[4],
                                                            void *tmp, void *stream, unsigned max_stages, float *stage_ms,
                                                            double *stage_bytes, unsigned *stages)
{
	std::vector<hipEvent_t> ev;
	try {
		const DeviceProgram &prog = program(ptr);
		const Plan &plan = prog.plan();
		DeviceBuffers ds = import_device_buffer(src, src_frame_stride, plan.src_w, true);
		DeviceBuffers dd = import_device_buffer(dst, dst_frame_stride, plan.dst_w, false);
		const size_t n = prog.steps().size();
		ev.resize(n + 1, nullptr);
		for (auto &e : ev)
			if (hipEventCreate(&e) != hipSuccess)
				fail(ZIMG_ERROR_UNKNOWN, "hipEventCreate");
		std::vector<void *> evp(ev.begin(), ev.end());
		prog.run(ds, dd, tmp, count, stream, evp.data());
		if (hipEventSynchronize(ev[n]) != hipSuccess)
			fail(ZIMG_ERROR_UNKNOWN, "hipEventSynchronize");
		const std::vector<double> bytes = prog.step_bytes();
		for (size_t k = 0; k < n && k < max_stages; ++k) {
			float ms = 0;
			(void)hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
			if (stage_ms) stage_ms[k] = ms;
			if (stage_bytes) stage_bytes[k] = bytes[k];
		}
		if (stages) *stages = static_cast<unsigned>(n);
		for (auto &e : ev)
			(void)hipEventDestroy(e);
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		for (auto &e : ev)
			if (e)
				(void)hipEventDestroy(e);
		return handle_current_exception();
	}
}

zimg_error_code_e zimg_filter_graph_set_resize_precision(zimg_filter_graph *ptr, int mode)
{
	try {
		if (!ptr || !ptr->prog)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null graph");
		ptr->prog->set_resize_precision(mode);
		int ndev = 0;
		if (hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0)
			ptr->prog->prepare_current_device();
		else
			(void)hipGetLastError();
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

zimg_error_code_e zimg_filter_graph_set_transfer_surrogates(zimg_filter_graph *ptr, const zimg_transfer_surrogate *const *s, unsigned count)
{
	try {
		if (!ptr || !ptr->prog)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null graph");
		if (count && !s)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null surrogate list");
		std::vector<const zamd::learned::Surrogate *> v;
		for (unsigned k = 0; k < count; ++k)
			v.push_back(reinterpret_cast<const zamd::learned::Surrogate *>(s[k]));
		ptr->prog->set_transfer_surrogates(v);
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

zimg_error_code_e zimg_device_jobs_get_tmp_size(const zimg_filter_graph *const *graphs, unsigned count, size_t *out)
{
	try {
		if (!out || (count && !graphs))
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null argument");
		std::vector<const DeviceProgram *> progs;
		for (unsigned k = 0; k < count; ++k)
			progs.push_back(&program(graphs[k]));
		*out = zamd::DeviceJobs::tmp_size(progs);
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

zimg_device_jobs *zimg_device_jobs_create(const zimg_filter_graph *const *graphs, unsigned count, const zimg_image_buffer_const *src,
                                          const zimg_image_buffer *dst, void *tmp, unsigned flags, void *stream)
{
	try {
		if (count && (!graphs || !src || !dst))
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null argument");
		std::vector<zamd::DeviceJobs::Job> jobs;
		for (unsigned k = 0; k < count; ++k) {
			const DeviceProgram &prog = program(graphs[k]);
			const Plan &plan = prog.plan();
			jobs.push_back(zamd::DeviceJobs::Job{ &prog, import_device_buffer(&src[k], nullptr, plan.src_w, true),
			                                      import_device_buffer(&dst[k], nullptr, plan.dst_w, false) });
		}
		return reinterpret_cast<zimg_device_jobs *>(new zamd::DeviceJobs(std::move(jobs), tmp, flags, stream));
	} catch (...) {
		handle_current_exception();
		return nullptr;
	}
}

zimg_error_code_e zimg_device_jobs_run(zimg_device_jobs *jobs, void *stream)
{
	try {
		if (!jobs)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null jobs");
		reinterpret_cast<zamd::DeviceJobs *>(jobs)->run(stream);
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

size_t zimg_device_jobs_describe(const zimg_device_jobs *jobs, char *buf, size_t n)
{
	if (!jobs)
		return 0;
	const std::string s = reinterpret_cast<const zamd::DeviceJobs *>(jobs)->describe();
	if (buf && n) {
		std::strncpy(buf, s.c_str(), n);
		buf[n - 1] = '\0';
	}
	return s.size();
}

void zimg_device_jobs_free(zimg_device_jobs *jobs)
{
	delete reinterpret_cast<zamd::DeviceJobs *>(jobs);
}

zimg_transfer_surrogate *zimg_transfer_surrogate_train(int transfer, int to_linear, double tolerance, int relative, unsigned steps,
                                                       unsigned seed, void *stream)
{
	try {
		zamd::learned::Config cfg;
		cfg.transfer = transfer;
		cfg.to_linear = to_linear ? 1 : 0;
		cfg.tolerance = tolerance;
		cfg.relative = relative ? 1 : 0;
		cfg.steps = steps;
		cfg.seed = seed;
		return reinterpret_cast<zimg_transfer_surrogate *>(new zamd::learned::Surrogate(cfg, stream));
	} catch (...) {
		handle_current_exception();
		return nullptr;
	}
}

zimg_error_code_e zimg_transfer_surrogate_apply(const zimg_transfer_surrogate *s, float *values, size_t n, float pre, float post,
                                                void *stream)
{
	try {
		if (!s || (n && !values))
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null argument");
		reinterpret_cast<const zamd::learned::Surrogate *>(s)->apply(values, n, pre, post, stream);
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

zimg_error_code_e zimg_transfer_exact_apply(int transfer, int to_linear, float *values, size_t n, float pre, float post, void *stream)
{
	try {
		if (n && !values)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null argument");
		zamd::learned::apply_exact(zamd::learned::function_for(transfer, to_linear), values, n, pre, post, stream);
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

zimg_error_code_e zimg_transfer_surrogate_info(const zimg_transfer_surrogate *s, double *coverage, double *max_error, double *train_loss)
{
	try {
		if (!s)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null surrogate");
		const auto *m = reinterpret_cast<const zamd::learned::Surrogate *>(s);
		if (coverage) *coverage = m->coverage();
		if (max_error) *max_error = m->max_validated_error();
		if (train_loss) *train_loss = m->train_loss();
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

unsigned zimg_transfer_surrogate_param_count(void)
{
	return static_cast<unsigned>(zamd::learned::param_count());
}

zimg_error_code_e zimg_transfer_surrogate_get_params(const zimg_transfer_surrogate *s, float *params)
{
	try {
		if (!s || !params)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null argument");
		reinterpret_cast<const zamd::learned::Surrogate *>(s)->get_params(params);
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

zimg_error_code_e zimg_transfer_surrogate_gradients(const float *params, const float *u, const float *t, unsigned n, float *grad)
{
	try {
		if (!params || !u || !t || !grad)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null argument");
		zamd::learned::gradients(params, u, t, n, grad, nullptr);
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

zimg_error_code_e zimg_transfer_surrogate_eval(const float *params, const float *u, float *v, unsigned n)
{
	try {
		if (!params || !u || !v)
			fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "null argument");
		zamd::learned::eval(params, u, v, n, nullptr);
		return ZIMG_ERROR_SUCCESS;
	} catch (...) {
		return handle_current_exception();
	}
}

void zimg_transfer_surrogate_free(zimg_transfer_surrogate *s)
{
	delete reinterpret_cast<zamd::learned::Surrogate *>(s);
}

size_t zimg_filter_graph_describe(const zimg_filter_graph *ptr, char *buf, size_t n)
{
	try {
		const DeviceProgram &prog = program(ptr);
		std::string s = "plan:\n" + prog.plan().describe() + "device:\n" + prog.describe();
		unsigned sets = 0;
		{
			std::lock_guard<std::mutex> lock(ptr->pool_mutex);
			sets = ptr->pool_sets;
		}
		s += "host staging sets: " + std::to_string(sets) + "\n";
		if (buf && n) {
			std::strncpy(buf, s.c_str(), n);
			buf[n - 1] = '\0';
		}
		return s.size();
	} catch (...) {
		handle_current_exception();
		return 0;
	}
}

} // extern "C"
