// Resampling filters and coefficient tables (host, FP64).
// Reference: src/zimg/resize/filter.{h,cpp}.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

namespace zamd {

// A resampling kernel evaluated on the host in double precision.
struct ResampleFilter {
	int type = 2;        // zimg_resample_filter_e
	double param_a = 0;  // bicubic b / lanczos taps
	double param_b = 0.5;// bicubic c

	unsigned support() const;
	double operator()(double x) const;
	const char *name() const;
};

// Packed coefficient table of one resize pass (FilterContext, filter.h:133-165).
// Rows are dense with stride == filter_width.
struct FilterTable {
	unsigned filter_width = 0;
	unsigned filter_rows = 0;
	unsigned input_width = 0;
	std::vector<float> data;       // [rows * filter_width]
	std::vector<int16_t> data_i16; // Q1.14, [rows * filter_width]
	std::vector<double> data64;    // the unrounded FP64 matrix entries, same layout
	std::vector<unsigned> left;    // [rows]

	bool sorted() const;
};

// compute_filter (filter.cpp:249-306) followed by matrix_to_filter (:43-127).
FilterTable compute_filter(const ResampleFilter &f, unsigned src_dim, unsigned dst_dim, double shift, double width);

// Unresize of one dimension (BilinearContext, unresize/bilinear.h:12-72): the
// least-squares inverse of a bilinear upscale from orig_dim to up_dim samples.
// Output sample j is the solution of (A'A) x = A'y with A the bilinear
// matrix: a band of A' (row_size taps from offset[j]) then a tridiagonal LU
// solve, forward with c and the inverted diagonal l, backward with u.
struct UnresizeTable {
	unsigned up_dim = 0, orig_dim = 0;
	unsigned row_size = 0;
	std::vector<float> coef;       // [orig_dim * row_size]
	std::vector<unsigned> offset;  // [orig_dim]
	std::vector<float> lu_c, lu_l, lu_u; // [orig_dim]
};

// create_bilinear_context (unresize/bilinear.cpp:112-172). shift is the
// conversion's shift of the output grid (graphbuilder.cpp:773-774).
UnresizeTable compute_unresize(unsigned up_dim, unsigned orig_dim, double shift);

} // namespace zamd
