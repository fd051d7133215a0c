// Colourspace conversion planning (host): shortest operation path and the
// per-operation constants handed to the device kernels.
// Reference: src/zimg/colorspace/{colorspace,colorspace_graph,operation,
// operation_impl,colorspace_param,matrix3}.cpp.
#pragma once

#include <string>
#include <vector>

#include "common.h"

namespace zamd {

// CSP_MATRIX_FMA, CSP_LUT_LINEAR and CSP_LUT_GAMMA are the operations the
// reference runs with allow_approximate_gamma on an x86 CPU with AVX2, FMA
// and F16C (operation_impl_avx2.cpp:43-68, 122-240, 258-382): a matrix as a
// product and two fused multiply-adds; inverse gamma from a 2^16 + 1 entry
// table over [-0.5, 1.5]; gamma from a 2^16 entry table indexed by the
// binary16 bits of the value.
//
// CSP_B67_SCALE is the first half of CSP_B67 (the inverse OOTF scaling by
// Ys^((1 - gamma) / gamma), without the OETF): a program splits CSP_B67 into
// it and a learned OETF when a transfer surrogate takes over the OETF.
enum CspOpKind : int { CSP_MATRIX = 0, CSP_GAMMA = 1, CSP_B67 = 2, CSP_B67_INV = 3, CSP_CL_TO_RGB = 4, CSP_CL_TO_YUV = 5,
                       CSP_MATRIX_FMA = 6, CSP_LUT_LINEAR = 7, CSP_LUT_GAMMA = 8, CSP_B67_SCALE = 9 };

// Entries of the lookup tables of CSP_LUT_LINEAR and CSP_LUT_GAMMA.
constexpr unsigned kLutLinearSize = 65537, kLutGammaSize = 65536;

// One in-place operation on three float planes (plain old data, copied to
// kernel arguments as is).
struct CspOp {
	int kind;
	int func;     // transfer function (zm::TransferFunc); CL: the linearising / gamma function
	float m[9];   // CSP_MATRIX, row major
	float pre;    // input scale (gamma, B67, CL->YUV)
	float post;   // output scale (gamma, inverse B67, CL->RGB)
	float kr, kg, kb;
	float nb, pb, nr, pr;
	const float *lut; // CSP_LUT_*: the table on the device running the kernel (set at launch)
};

// The table of a CSP_LUT_* operation, entry by entry as the reference builds it.
std::vector<float> build_gamma_lut(const CspOp &op);

struct CspProgram {
	int n = 0;
	CspOp op[6];
	std::string describe() const;
};

struct CspParams {
	double peak_luminance = 100.0;
	bool approximate_gamma = false;
	bool scene_referred = false;
	bool chromatic_adaptation = false;
};

// ColorspaceConversion::create (colorspace.cpp:95-122): empty program for identity.
// Throws Error(ZIMG_ERROR_NO_COLORSPACE_CONVERSION) when no path exists.
CspProgram build_colorspace_program(Colorspace in, Colorspace out, const CspParams &params);

// FP64 helpers exposed for the learned-transfer and precision tooling.
void ncl_rgb_to_yuv_matrix(Matrix m, double out[3][3]);

} // namespace zamd
