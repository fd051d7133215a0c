// Test-only bridge over the reference's own leaf sources, compiled from where
// they lie under the reference tree by oracle/Makefile into oracle/_ref/.
// It exposes extern "C" entry points so the tests can compare the oracle's
// restatement with the reference bit for bit. Nothing here is shipped.
//
// Sources compiled (all as-is, no stand-in headers): resize/filter.cpp,
// common/matrix.cpp, common/libm_wrapper.cpp, common/cpuinfo.cpp,
// common/x86/cpuinfo_x86.cpp, colorspace/{gamma, colorspace_param,matrix3,
// operation,operation_impl,colorspace_graph}.cpp, colorspace/x86/{operation_impl_x86,
// operation_impl_avx2,operation_impl_avx512,gamma_constants_avx512}.cpp,
// depth/{quantize,blue}.cpp, unresize/bilinear.cpp and the vendored musl libm
// of test/extra/musl-libm.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "common/cpuinfo.h"
#include "common/x86/cpuinfo_x86.h"
#include "common/libm_wrapper.h"
#include "common/pixel.h"
#include "colorspace/colorspace.h"
#include "colorspace/colorspace_graph.h"
#include "colorspace/colorspace_param.h"
#include "colorspace/gamma.h"
#include "colorspace/operation.h"
#include "depth/blue.h"
#include "depth/quantize.h"
#include "resize/filter.h"
#include "unresize/bilinear.h"
#include "musl-libm/mymath.h"

namespace {

std::unique_ptr<zimg::resize::Filter> make_filter(int type, double a, double b)
{
	using namespace zimg::resize;
	switch (type) {
	case 0: return std::make_unique<PointFilter>();
	case 1: return std::make_unique<BilinearFilter>();
	case 2: return std::make_unique<BicubicFilter>(a, b);
	case 3: return std::make_unique<Spline16Filter>();
	case 4: return std::make_unique<Spline36Filter>();
	case 5: return std::make_unique<LanczosFilter>(static_cast<unsigned>(a));
	case 6: return std::make_unique<Spline64Filter>();
	default: return nullptr;
	}
}

} // namespace

extern "C" {

// Mirrors test/main.cpp:10-18 (musl) or restores the library defaults (libm).
void ref_set_libm(int musl)
{
	if (musl) {
		zimg_x_expf = myexpf;
		zimg_x_logf = mylogf;
		zimg_x_log10f = mylog10f;
		zimg_x_powf = mypowf;
		zimg_x_sin = mysin;
		zimg_x_cos = mycos;
	} else {
		zimg_x_expf = expf;
		zimg_x_logf = logf;
		zimg_x_log10f = log10f;
		zimg_x_powf = powf;
		zimg_x_sin = sin;
		zimg_x_cos = cos;
	}
}

void ref_powf_n(const float *x, const float *y, float *out, size_t n)
{
	for (size_t i = 0; i < n; ++i) out[i] = zimg_x_powf(x[i], y[i]);
}

void ref_expf_n(const float *x, float *out, size_t n)
{
	for (size_t i = 0; i < n; ++i) out[i] = zimg_x_expf(x[i]);
}

void ref_logf_n(const float *x, float *out, size_t n)
{
	for (size_t i = 0; i < n; ++i) out[i] = zimg_x_logf(x[i]);
}

double ref_filter_eval(int type, double a, double b, double x)
{
	auto f = make_filter(type, a, b);
	return (*f)(x);
}

int ref_compute_filter(int type, double a, double b, unsigned src_dim, unsigned dst_dim, double shift, double width,
                       unsigned *fw, float *data, int16_t *data_i16, unsigned *left)
{
	try {
		auto f = make_filter(type, a, b);
		zimg::resize::FilterContext ctx = zimg::resize::compute_filter(*f, src_dim, dst_dim, shift, width);
		*fw = ctx.filter_width;
		if (data) {
			for (unsigned i = 0; i < ctx.filter_rows; ++i) {
				for (unsigned j = 0; j < ctx.filter_width; ++j) {
					data[static_cast<size_t>(i) * ctx.filter_width + j] = ctx.data[static_cast<size_t>(i) * ctx.stride + j];
					data_i16[static_cast<size_t>(i) * ctx.filter_width + j] = ctx.data_i16[static_cast<size_t>(i) * ctx.stride_i16 + j];
				}
				left[i] = ctx.left[i];
			}
		}
		return 0;
	} catch (...) {
		return 1;
	}
}

// Evaluates select_transfer_function(...).to_linear / .to_gamma (no scale).
float ref_transfer(int transfer, int to_linear, int scene_referred, double peak, float x)
{
	auto tf = zimg::colorspace::select_transfer_function(static_cast<zimg::colorspace::TransferCharacteristics>(transfer), peak, scene_referred != 0);
	return to_linear ? tf.to_linear(x) : tf.to_gamma(x);
}

void ref_transfer_scales(int transfer, double peak, int scene_referred, float *to_linear_scale, float *to_gamma_scale)
{
	auto tf = zimg::colorspace::select_transfer_function(static_cast<zimg::colorspace::TransferCharacteristics>(transfer), peak, scene_referred != 0);
	*to_linear_scale = tf.to_linear_scale;
	*to_gamma_scale = tf.to_gamma_scale;
}

// Runs the operation chain the reference builds for (in -> out) over three
// float planes in place, mirroring ColorspaceConversion::create and
// ColorspaceConversionImpl::process (colorspace.cpp:45-122) with CPUClass::NONE.
// Returns the number of operations, 0 for identity, -1 if no path.
static int colorspace_apply(int m_in, int t_in, int p_in, int m_out, int t_out, int p_out, double peak, int scene_referred,
                            int approximate_gamma, int chromatic_adaptation, zimg::CPUClass cpu, float *p0, float *p1, float *p2,
                            size_t n);

int ref_colorspace_apply(int m_in, int t_in, int p_in, int m_out, int t_out, int p_out,
                         double peak, int scene_referred, int approximate_gamma, int chromatic_adaptation,
                         float *p0, float *p1, float *p2, size_t n)
{
	return colorspace_apply(m_in, t_in, p_in, m_out, t_out, p_out, peak, scene_referred, approximate_gamma, chromatic_adaptation,
	                        zimg::CPUClass::NONE, p0, p1, p2, n);
}

// The same with the reference's CPU dispatch (cpu AUTO): on an x86 host with
// AVX2, FMA and F16C its AVX2 operations run. Returns -2 when the host lacks them.
int ref_colorspace_apply_auto(int m_in, int t_in, int p_in, int m_out, int t_out, int p_out,
                              double peak, int scene_referred, int approximate_gamma, int chromatic_adaptation,
                              float *p0, float *p1, float *p2, size_t n)
{
	const zimg::X86Capabilities caps = zimg::query_x86_capabilities();
	if (!caps.avx2 || !caps.fma || !caps.f16c)
		return -2;
	return colorspace_apply(m_in, t_in, p_in, m_out, t_out, p_out, peak, scene_referred, approximate_gamma, chromatic_adaptation,
	                        zimg::CPUClass::AUTO, p0, p1, p2, n);
}

static int colorspace_apply(int m_in, int t_in, int p_in, int m_out, int t_out, int p_out, double peak, int scene_referred,
                            int approximate_gamma, int chromatic_adaptation, zimg::CPUClass cpu, float *p0, float *p1, float *p2,
                            size_t n)
{
	using namespace zimg::colorspace;
	try {
		ColorspaceDefinition in{ static_cast<MatrixCoefficients>(m_in), static_cast<TransferCharacteristics>(t_in), static_cast<ColorPrimaries>(p_in) };
		ColorspaceDefinition out{ static_cast<MatrixCoefficients>(m_out), static_cast<TransferCharacteristics>(t_out), static_cast<ColorPrimaries>(p_out) };
		if (!scene_referred) {
			if (in.transfer == TransferCharacteristics::SMPTE_240M) in.transfer = TransferCharacteristics::REC_709;
			if (out.transfer == TransferCharacteristics::SMPTE_240M) out.transfer = TransferCharacteristics::REC_709;
		}
		if (in == out)
			return 0;
		OperationParams params;
		params.set_peak_luminance(peak).set_approximate_gamma(approximate_gamma != 0)
		      .set_scene_referred(scene_referred != 0).set_chromatic_adaptation(chromatic_adaptation != 0);
		auto path = get_operation_path(in, out);
		std::vector<std::unique_ptr<Operation>> ops;
		for (auto &f : path)
			ops.push_back(f(params, cpu));
		float *planes[3] = { p0, p1, p2 };
		for (auto &op : ops)
			op->process(planes, planes, 0, static_cast<unsigned>(n));
		return static_cast<int>(ops.size());
	} catch (...) {
		return -1;
	}
}

// create_bilinear_context (unresize/bilinear.cpp:112-172) of unresizing
// up_dim samples to orig_dim: with coef == NULL only *row_size is written.
int ref_unresize_context(unsigned up_dim, unsigned orig_dim, double shift, unsigned *row_size, float *coef, unsigned *offset,
                         float *lu_c, float *lu_l, float *lu_u)
{
	try {
		zimg::unresize::BilinearContext ctx = zimg::unresize::create_bilinear_context(orig_dim, up_dim, shift);
		*row_size = ctx.matrix_row_size;
		if (coef) {
			for (unsigned i = 0; i < ctx.output_width; ++i) {
				for (unsigned k = 0; k < ctx.matrix_row_size; ++k)
					coef[static_cast<size_t>(i) * ctx.matrix_row_size + k] = ctx.matrix_coefficients[static_cast<size_t>(i) * ctx.matrix_row_stride + k];
				offset[i] = ctx.matrix_row_offsets[i];
				lu_c[i] = ctx.lu_c[i];
				lu_l[i] = ctx.lu_l[i];
				lu_u[i] = ctx.lu_u[i];
			}
		}
		return 0;
	} catch (...) {
		return -1;
	}
}

void ref_ncl_rgb_to_yuv(int matrix, double out[9])
{
	auto m = zimg::colorspace::ncl_rgb_to_yuv_matrix(static_cast<zimg::colorspace::MatrixCoefficients>(matrix));
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			out[i * 3 + j] = m[i][j];
}

float ref_half_to_float(uint16_t h) { return zimg::depth::half_to_float(h); }
uint16_t ref_float_to_half(float f) { return zimg::depth::float_to_half(f); }

void ref_blue_noise(uint8_t out[4096])
{
	std::memcpy(out, &zimg::depth::blue_noise_table[0][0], 4096);
}

void ref_scale_offset(int type_in, unsigned depth_in, int full_in, int chroma_in, int ycgco_in,
                      int type_out, unsigned depth_out, int full_out, int chroma_out, int ycgco_out,
                      float *scale, float *offset)
{
	zimg::PixelFormat in{ static_cast<zimg::PixelType>(type_in), depth_in, full_in != 0, chroma_in != 0, ycgco_in != 0 };
	zimg::PixelFormat out{ static_cast<zimg::PixelType>(type_out), depth_out, full_out != 0, chroma_out != 0, ycgco_out != 0 };
	auto so = zimg::depth::get_scale_offset(in, out);
	*scale = so.first;
	*offset = so.second;
}

} // extern "C"
