// Host-side resampling kernels and coefficient generation (FP64).
// Reference: src/zimg/resize/filter.cpp. The steps and rounding rules are kept
// identical because the device kernels must reproduce the reference's integer
// output exactly, which fixes every coefficient bit.
#include "filter.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <limits>

#include "common.h"

namespace zamd {

namespace {

constexpr double kPi = 3.14159265358979323846;

double sinc(double x) { return x == 0.0 ? 1.0 : std::sin(x * kPi) / (x * kPi); } // filter.cpp:21-25

double horner3(double x, double c0, double c1, double c2, double c3) { return c0 + x * (c1 + x * (c2 + x * c3)); }

// Pixel-grid rounding that keeps round(x - 1) == round(x) - 1 (filter.cpp:32-40).
double grid_round(double x) { return x < 0 ? std::floor(x + 0.5) : std::floor(x + 0.49999999999999994); }

// One output row of the sparse resampling matrix. Matches the storage rule of
// the reference's RowMatrix (common/matrix.cpp:14-21,104-129): a column joins
// the stored span only when an assignment changes its value.
class SparseRow {
	std::vector<double> &m_dense;
	size_t m_lo = 0, m_hi = 0;
public:
	explicit SparseRow(std::vector<double> &dense) : m_dense(dense) {}

	double get(size_t j) const { return (j >= m_lo && j < m_hi) ? m_dense[j] : 0.0; }

	void set(size_t j, double v)
	{
		if (get(j) == v)
			return;
		if (m_lo == m_hi) {
			m_lo = j;
			m_hi = j + 1;
		} else {
			m_lo = std::min(m_lo, j);
			m_hi = std::max(m_hi, j + 1);
		}
		m_dense[j] = v;
	}

	size_t lo() const { return m_lo; }
	size_t hi() const { return m_hi; }

	void extract(std::vector<double> &out) const { out.assign(m_dense.begin() + m_lo, m_dense.begin() + m_hi); }
	void clear() { std::fill(m_dense.begin() + m_lo, m_dense.begin() + m_hi, 0.0); m_lo = m_hi = 0; }
};

} // namespace

unsigned ResampleFilter::support() const
{
	switch (type) {
	case ZIMG_RESIZE_POINT: return 0;
	case ZIMG_RESIZE_BILINEAR: return 1;
	case ZIMG_RESIZE_BICUBIC: return 2;
	case ZIMG_RESIZE_SPLINE16: return 2;
	case ZIMG_RESIZE_SPLINE36: return 3;
	case ZIMG_RESIZE_SPLINE64: return 4;
	case ZIMG_RESIZE_LANCZOS: return static_cast<unsigned>(param_a);
	default: return 0;
	}
}

const char *ResampleFilter::name() const
{
	switch (type) {
	case ZIMG_RESIZE_POINT: return "point";
	case ZIMG_RESIZE_BILINEAR: return "bilinear";
	case ZIMG_RESIZE_BICUBIC: return "bicubic";
	case ZIMG_RESIZE_SPLINE16: return "spline16";
	case ZIMG_RESIZE_SPLINE36: return "spline36";
	case ZIMG_RESIZE_SPLINE64: return "spline64";
	case ZIMG_RESIZE_LANCZOS: return "lanczos";
	default: return "?";
	}
}

double ResampleFilter::operator()(double x) const
{
	// Kernel definitions of filter.cpp:134-246.
	switch (type) {
	case ZIMG_RESIZE_POINT:
		return 1.0;
	case ZIMG_RESIZE_BILINEAR:
		return std::max(1.0 - std::abs(x), 0.0);
	case ZIMG_RESIZE_BICUBIC: {
		const double b = param_a, c = param_b;
		const double p0 = (6.0 - 2.0 * b) / 6.0;
		const double p2 = (-18.0 + 12.0 * b + 6.0 * c) / 6.0;
		const double p3 = (12.0 - 9.0 * b - 6.0 * c) / 6.0;
		const double q0 = (8.0 * b + 24.0 * c) / 6.0;
		const double q1 = (-12.0 * b - 48.0 * c) / 6.0;
		const double q2 = (6.0 * b + 30.0 * c) / 6.0;
		const double q3 = (-b - 6.0 * c) / 6.0;
		x = std::abs(x);
		if (x < 1.0) return horner3(x, p0, 0.0, p2, p3);
		if (x < 2.0) return horner3(x, q0, q1, q2, q3);
		return 0.0;
	}
	case ZIMG_RESIZE_SPLINE16:
		x = std::abs(x);
		if (x < 1.0) return horner3(x, 1.0, -1.0 / 5.0, -9.0 / 5.0, 1.0);
		if (x < 2.0) return horner3(x - 1.0, 0.0, -7.0 / 15.0, 4.0 / 5.0, -1.0 / 3.0);
		return 0.0;
	case ZIMG_RESIZE_SPLINE36:
		x = std::abs(x);
		if (x < 1.0) return horner3(x, 1.0, -3.0 / 209.0, -453.0 / 209.0, 13.0 / 11.0);
		if (x < 2.0) return horner3(x - 1.0, 0.0, -156.0 / 209.0, 270.0 / 209.0, -6.0 / 11.0);
		if (x < 3.0) return horner3(x - 2.0, 0.0, 26.0 / 209.0, -45.0 / 209.0, 1.0 / 11.0);
		return 0.0;
	case ZIMG_RESIZE_SPLINE64:
		x = std::abs(x);
		if (x < 1.0) return horner3(x, 1.0, -3.0 / 2911.0, -6387.0 / 2911.0, 49.0 / 41.0);
		if (x < 2.0) return horner3(x - 1.0, 0.0, -2328.0 / 2911.0, 4032.0 / 2911.0, -24.0 / 41.0);
		if (x < 3.0) return horner3(x - 2.0, 0.0, 582.0 / 2911.0, -1008.0 / 2911.0, 6.0 / 41.0);
		if (x < 4.0) return horner3(x - 3.0, 0.0, -97.0 / 2911.0, 168.0 / 2911.0, -1.0 / 41.0);
		return 0.0;
	case ZIMG_RESIZE_LANCZOS: {
		const unsigned taps = static_cast<unsigned>(param_a);
		x = std::abs(x);
		return x < taps ? sinc(x) * sinc(x / taps) : 0.0;
	}
	default:
		return 0.0;
	}
}

bool FilterTable::sorted() const { return std::is_sorted(left.begin(), left.end()); }

FilterTable compute_filter(const ResampleFilter &f, unsigned src_dim, unsigned dst_dim, double shift, double width)
{
	const double scale = static_cast<double>(dst_dim) / width;
	const double step = std::min(scale, 1.0);
	const double support = static_cast<double>(f.support()) / step;
	const unsigned taps = std::max(static_cast<unsigned>(std::ceil(support)) * 2U, 1U);

	if (support > static_cast<unsigned>(UINT_MAX / 2))
		fail(ZIMG_ERROR_RESAMPLING_NOT_AVAILABLE, "filter width too great");

	// Pass 1: the FP64 resampling matrix, one sparse row per output sample.
	std::vector<double> dense(src_dim, 0.0);
	std::vector<size_t> row_lo(dst_dim), row_hi(dst_dim);
	std::vector<std::vector<double>> rows(dst_dim);
	const double clamp_hi = std::nextafter(static_cast<double>(src_dim), -INFINITY);

	for (unsigned i = 0; i < dst_dim; ++i) {
		SparseRow row(dense);
		const double pos = (i + 0.5) / scale + shift;
		const double begin = grid_round(pos - taps / 2.0) + 0.5;

		double total = 0.0;
		for (unsigned j = 0; j < taps; ++j)
			total += f((begin + j - pos) * step);

		size_t leftmost = SIZE_MAX;
		for (unsigned j = 0; j < taps; ++j) {
			const double xpos = begin + j;
			// Mirror at the image edges, then clamp what is still outside.
			double src_pos = xpos < 0.0 ? -xpos : xpos >= src_dim ? 2.0 * src_dim - xpos : xpos;
			src_pos = std::clamp(src_pos, 0.0, clamp_hi);

			const size_t idx = static_cast<size_t>(std::floor(src_pos));
			row.set(idx, row.get(idx) + f((xpos - pos) * step) / total);
			leftmost = std::min(leftmost, idx);
		}
		// The leftmost tap always belongs to the stored span (filter.cpp:295-299).
		if (row.get(leftmost) == 0.0) {
			row.set(leftmost, DBL_EPSILON);
			row.set(leftmost, 0.0);
		}

		row_lo[i] = row.lo();
		row_hi[i] = row.hi();
		row.extract(rows[i]);
		row.clear();
	}

	// Pass 2: pack to a fixed width, rounding to f32 and Q1.14 with error feedback.
	size_t fw = 0;
	for (unsigned i = 0; i < dst_dim; ++i)
		fw = std::max(fw, row_hi[i] - row_lo[i]);
	if (!fw)
		fail(ZIMG_ERROR_UNKNOWN, "empty resampling matrix");

	FilterTable t;
	t.filter_width = static_cast<unsigned>(fw);
	t.filter_rows = dst_dim;
	t.input_width = src_dim;
	t.data.resize(fw * dst_dim);
	t.data_i16.resize(fw * dst_dim);
	t.data64.resize(fw * dst_dim);
	t.left.resize(dst_dim);

	for (unsigned i = 0; i < dst_dim; ++i) {
		const size_t left = std::min(row_lo[i], static_cast<size_t>(src_dim) - fw);
		double err_f32 = 0.0, err_i16 = 0.0;
		int16_t sum_i16 = 0, peak_i16 = 0;
		size_t peak_idx = 0;
		float *drow = &t.data[i * fw];
		int16_t *irow = &t.data_i16[i * fw];

		for (size_t j = 0; j < fw; ++j) {
			const size_t col = left + j;
			const double c = (col >= row_lo[i] && col < row_hi[i]) ? rows[i][col - row_lo[i]] : 0.0;
			const double want_f32 = c - err_f32;
			const double want_i16 = c * (1 << 14) - err_i16;
			const float q_f32 = static_cast<float>(want_f32);
			const int16_t q_i16 = static_cast<int16_t>(std::lrint(want_i16));
			err_f32 = static_cast<double>(q_f32) - want_f32;
			err_i16 = static_cast<double>(q_i16) - want_i16;

			if (std::abs(q_i16) > std::abs(peak_i16)) {
				peak_i16 = q_i16;
				peak_idx = j;
			}
			sum_i16 = static_cast<int16_t>(sum_i16 + q_i16);
			drow[j] = q_f32;
			irow[j] = q_i16;
			t.data64[i * fw + j] = c;
		}
		// Integer taps sum to exactly 1.0 in Q1.14: the residue goes to the largest tap.
		irow[peak_idx] = static_cast<int16_t>(irow[peak_idx] + ((1 << 14) - sum_i16));
		t.left[i] = static_cast<unsigned>(left);
	}
	return t;
}

namespace {

// One row of a sparse matrix as the reference's RowMatrix keeps it: the
// stored range [left, left + v.size()) and its values (zeros outside).
struct BandRow {
	size_t left = 0;
	std::vector<double> v;
	double at(size_t j) const { return j < left || j >= left + v.size() ? 0.0 : v[j - left]; }
	size_t right() const { return left + v.size(); }
};

} // namespace

UnresizeTable compute_unresize(unsigned up_dim, unsigned orig_dim, double shift)
{
	// create_bilinear_context(in = orig_dim, out = up_dim, shift) (bilinear.cpp:112-172)
	const unsigned in = orig_dim, out = up_dim;
	if (in > out)
		fail(ZIMG_ERROR_RESAMPLING_NOT_AVAILABLE, "unresize can not upscale");
	const double sh = -shift * in / out;

	// bilinear_weights (bilinear.cpp:64-108): out rows of one or two entries.
	std::vector<BandRow> m(out);
	const double leftmost = 0.5 + sh, rightmost = in - 0.5 + sh;
	const unsigned leftmost_idx = static_cast<unsigned>(std::clamp(std::floor(leftmost), 0.0, in - 1.0));
	const unsigned rightmost_idx = static_cast<unsigned>(std::clamp(std::floor(rightmost), 0.0, in - 1.0));
	for (unsigned i = 0; i < out; ++i) {
		const double position = (i + 0.5) * in / out;
		if (position <= leftmost) {
			m[i].left = leftmost_idx;
			m[i].v = { 1.0 };
		} else if (position >= rightmost) {
			m[i].left = rightmost_idx;
			m[i].v = { 1.0 };
		} else {
			const unsigned left_idx = static_cast<unsigned>(std::floor(position - leftmost));
			const double distance = position - left_idx - leftmost;
			m[i].left = left_idx;
			m[i].v = { 1.0 - distance, distance };
		}
	}

	// Transpose, its rows trimmed to their first and last nonzero entries
	// (RowMatrix operator~ and compress, common/matrix.cpp:154-200).
	std::vector<BandRow> t(in);
	for (unsigned c = 0; c < in; ++c) {
		size_t first = out, last = 0;
		for (unsigned i = 0; i < out; ++i) {
			if (m[i].at(c) != 0) {
				first = std::min<size_t>(first, i);
				last = i;
			}
		}
		if (first <= last && first < out) {
			t[c].left = first;
			for (size_t i = first; i <= last; ++i)
				t[c].v.push_back(m[i].at(c));
		}
	}

	// The tridiagonal of A'A (operator*, matrix.cpp:202-219: each entry summed
	// over the lhs row's stored range in order), then its LU factors
	// (tridiagonal_decompose, bilinear.cpp:31-59).
	auto prod = [&](size_t i, size_t j) {
		double acc = 0;
		for (size_t k = t[i].left; k < t[i].right(); ++k)
			acc += t[i].at(k) * m[k].at(j);
		return acc;
	};
	const double eps = std::numeric_limits<double>::epsilon();
	std::vector<double> lc(in, 0.0), ll(in, 0.0), lu(in, 0.0);
	if (in == 1) {
		ll[0] = prod(0, 0);
	} else {
		lc[0] = 0;
		ll[0] = prod(0, 0);
		lu[0] = prod(0, 1) / (ll[0] + eps);
		for (size_t i = 1; i + 1 < in; ++i) {
			lc[i] = prod(i, i - 1);
			ll[i] = prod(i, i) - lc[i] * lu[i - 1];
			lu[i] = prod(i, i + 1) / (ll[i] + eps);
		}
		lc[in - 1] = prod(in - 1, in - 2);
		ll[in - 1] = prod(in - 1, in - 1) - lc[in - 1] * lu[in - 2];
		lu[in - 1] = 0;
	}

	UnresizeTable r;
	r.up_dim = out;
	r.orig_dim = in;
	size_t rowsize = 0;
	for (unsigned i = 0; i < in; ++i)
		rowsize = std::max(rowsize, t[i].right() - t[i].left);
	if (!rowsize)
		fail(ZIMG_ERROR_UNKNOWN, "unresize: empty matrix");
	r.row_size = static_cast<unsigned>(rowsize);
	r.coef.assign(static_cast<size_t>(in) * rowsize, 0.0f);
	r.offset.resize(in);
	r.lu_c.resize(in);
	r.lu_l.resize(in);
	r.lu_u.resize(in);
	for (unsigned i = 0; i < in; ++i) {
		const size_t left = std::min(t[i].left, static_cast<size_t>(out) - rowsize);
		for (size_t j = 0; j < t[i].right() - left; ++j)
			r.coef[i * rowsize + j] = static_cast<float>(t[i].at(left + j));
		r.offset[i] = static_cast<unsigned>(left);
		r.lu_c[i] = static_cast<float>(lc[i]);
		r.lu_l[i] = static_cast<float>(1.0 / (ll[i] + eps));
		r.lu_u[i] = static_cast<float>(lu[i]);
	}
	return r;
}

} // namespace zamd
