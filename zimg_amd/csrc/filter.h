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

} // namespace zamd
