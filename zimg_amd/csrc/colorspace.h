// Colourspace conversion planning (host): shortest operation path and the
// per-operation constants handed to the device kernels.
// Reference: src/zimg/colorspace/{colorspace,colorspace_graph,operation,
// operation_impl,colorspace_param,matrix3}.cpp.
#pragma once

#include <string>

#include "common.h"

namespace zamd {

enum CspOpKind : int { CSP_MATRIX = 0, CSP_GAMMA = 1, CSP_B67 = 2, CSP_B67_INV = 3, CSP_CL_TO_RGB = 4, CSP_CL_TO_YUV = 5 };

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
};

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
