// Colourspace planning: FP64 matrix derivation, transfer selection and the
// breadth-first search over colourspace states. Kept step-for-step equal to
// the reference (colorspace_param.cpp, matrix3.cpp, gamma.cpp:341-409,
// operation.cpp, operation_impl.cpp:290-383, colorspace_graph.cpp:115-243),
// because the float constants cast from these FP64 results, and the order of
// the chosen operations, decide the output bits.
#include "colorspace.h"

#include <array>
#include <cmath>
#include <cstdio>
#include <deque>
#include <vector>

#include "zmath.h"

namespace zamd {

namespace {

using Vec3 = std::array<double, 3>;
using Mat3 = std::array<Vec3, 3>;

Mat3 identity3() { return { Vec3{ 1, 0, 0 }, Vec3{ 0, 1, 0 }, Vec3{ 0, 0, 1 } }; }

double det2(double a00, double a01, double a10, double a11) { return a00 * a11 - a01 * a10; }

Mat3 mul(const Mat3 &a, const Mat3 &b)
{
	Mat3 r{};
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j) {
			double acc = 0;
			for (int k = 0; k < 3; ++k)
				acc += a[i][k] * b[k][j];
			r[i][j] = acc;
		}
	return r;
}

Vec3 mul(const Mat3 &a, const Vec3 &v)
{
	Vec3 r{};
	for (int i = 0; i < 3; ++i) {
		double acc = 0;
		for (int k = 0; k < 3; ++k)
			acc += a[i][k] * v[k];
		r[i] = acc;
	}
	return r;
}

Mat3 inverse(const Mat3 &m)
{
	// Cofactor inverse (matrix3.cpp:68-95).
	double det = 0;
	det += m[0][0] * det2(m[1][1], m[1][2], m[2][1], m[2][2]);
	det -= m[0][1] * det2(m[1][0], m[1][2], m[2][0], m[2][2]);
	det += m[0][2] * det2(m[1][0], m[1][1], m[2][0], m[2][1]);
	Mat3 r{};
	r[0][0] = det2(m[1][1], m[1][2], m[2][1], m[2][2]) / det;
	r[0][1] = det2(m[0][2], m[0][1], m[2][2], m[2][1]) / det;
	r[0][2] = det2(m[0][1], m[0][2], m[1][1], m[1][2]) / det;
	r[1][0] = det2(m[1][2], m[1][0], m[2][2], m[2][0]) / det;
	r[1][1] = det2(m[0][0], m[0][2], m[2][0], m[2][2]) / det;
	r[1][2] = det2(m[0][2], m[0][0], m[1][2], m[1][0]) / det;
	r[2][0] = det2(m[1][0], m[1][1], m[2][0], m[2][1]) / det;
	r[2][1] = det2(m[0][1], m[0][0], m[2][1], m[2][0]) / det;
	r[2][2] = det2(m[0][0], m[0][1], m[1][0], m[1][1]) / det;
	return r;
}

Mat3 transpose(const Mat3 &m)
{
	Mat3 r{};
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			r[i][j] = m[j][i];
	return r;
}

Vec3 cross(const Vec3 &a, const Vec3 &b)
{
	return { a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0] };
}

double dot(const Vec3 &a, const Vec3 &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

Vec3 xy_to_xyz(double x, double y) { return { x / y, 1.0, (1.0 - x - y) / y }; }

// Standard constants (colorspace_param.h:14-39): ITU-R BT.601/709/2020, SMPTE 240M, FCC.
void kr_kb(Matrix m, double &kr, double &kb)
{
	switch (m) {
	case Matrix::RGB: kr = 0; kb = 0; return;
	case Matrix::FCC: kr = 0.3; kb = 0.11; return;
	case Matrix::SMPTE_240M: kr = 0.212; kb = 0.087; return;
	case Matrix::REC_601: kr = 0.299; kb = 0.114; return;
	case Matrix::REC_709: kr = 0.2126; kb = 0.0722; return;
	case Matrix::REC_2020_NCL:
	case Matrix::REC_2020_CL: kr = 0.2627; kb = 0.0593; return;
	default: fail(ZIMG_ERROR_UNKNOWN, "unrecognized matrix coefficients");
	}
}

Vec3 white_point(Primaries p)
{
	switch (p) {
	case Primaries::REC_470_M:
	case Primaries::FILM: return xy_to_xyz(0.31, 0.316);              // illuminant C
	case Primaries::XYZ: return xy_to_xyz(1.0 / 3.0, 1.0 / 3.0);      // illuminant E
	case Primaries::DCI_P3: return xy_to_xyz(0.314, 0.351);           // DCI white
	default: return xy_to_xyz(0.3127, 0.3290);                        // D65
	}
}

std::array<std::array<double, 2>, 3> primaries_xy(Primaries p)
{
	switch (p) {
	case Primaries::REC_470_M: return { { { 0.670, 0.330 }, { 0.210, 0.710 }, { 0.140, 0.080 } } };
	case Primaries::REC_470_BG: return { { { 0.640, 0.330 }, { 0.290, 0.600 }, { 0.150, 0.060 } } };
	case Primaries::SMPTE_C: return { { { 0.630, 0.340 }, { 0.310, 0.595 }, { 0.155, 0.070 } } };
	case Primaries::REC_709: return { { { 0.640, 0.330 }, { 0.300, 0.600 }, { 0.150, 0.060 } } };
	case Primaries::FILM: return { { { 0.681, 0.319 }, { 0.243, 0.692 }, { 0.145, 0.049 } } };
	case Primaries::REC_2020: return { { { 0.708, 0.292 }, { 0.170, 0.797 }, { 0.131, 0.046 } } };
	case Primaries::DCI_P3:
	case Primaries::DCI_P3_D65: return { { { 0.680, 0.320 }, { 0.265, 0.690 }, { 0.150, 0.060 } } };
	case Primaries::EBU_3213_E: return { { { 0.630, 0.340 }, { 0.295, 0.605 }, { 0.155, 0.077 } } };
	default: fail(ZIMG_ERROR_UNKNOWN, "unrecognized primaries");
	}
}

Mat3 rgb_to_yuv_kr_kb(double kr, double kb)
{
	const double kg = 1.0 - kr - kb;
	const double us = 1.0 / (2.0 - 2.0 * kb);
	const double vs = 1.0 / (2.0 - 2.0 * kr);
	Mat3 r{};
	r[0] = { kr, kg, kb };
	r[1] = { -kr * us, -kg * us, (1.0 - kb) * us };
	r[2] = { (1.0 - kr) * vs, -kg * vs, -kb * vs };
	return r;
}

Mat3 ncl_rgb_to_yuv(Matrix m)
{
	if (m == Matrix::YCGCO)
		return { Vec3{ 0.25, 0.5, 0.25 }, Vec3{ -0.25, 0.5, -0.25 }, Vec3{ 0.5, 0, -0.5 } };
	if (m == Matrix::REC_2100_LMS)
		return { Vec3{ 1688.0 / 4096.0, 2146.0 / 4096.0, 262.0 / 4096.0 },
		         Vec3{ 683.0 / 4096.0, 2951.0 / 4096.0, 462.0 / 4096.0 },
		         Vec3{ 99.0 / 4096.0, 309.0 / 4096.0, 3688.0 / 4096.0 } };
	double kr, kb;
	kr_kb(m, kr, kb);
	return rgb_to_yuv_kr_kb(kr, kb);
}

void kr_kb_from_primaries(Primaries p, double &kr, double &kb)
{
	// ITU-T H.265 (E-22)..(E-27).
	auto xy = primaries_xy(p);
	Vec3 r = xy_to_xyz(xy[0][0], xy[0][1]), g = xy_to_xyz(xy[1][0], xy[1][1]), b = xy_to_xyz(xy[2][0], xy[2][1]);
	Vec3 w = white_point(p);
	Vec3 xr{ r[0], g[0], b[0] }, yr{ r[1], g[1], b[1] }, zr{ r[2], g[2], b[2] };
	kr = dot(w, cross(g, b)) / dot(xr, cross(yr, zr));
	kb = dot(w, cross(r, g)) / dot(xr, cross(yr, zr));
}

Mat3 ncl_rgb_to_yuv_from_primaries(Primaries p)
{
	if (p == Primaries::REC_709) return ncl_rgb_to_yuv(Matrix::REC_709);
	if (p == Primaries::REC_2020) return ncl_rgb_to_yuv(Matrix::REC_2020_NCL);
	double kr, kb;
	kr_kb_from_primaries(p, kr, kb);
	return rgb_to_yuv_kr_kb(kr, kb);
}

Mat3 ncl_yuv_to_rgb_from_primaries(Primaries p)
{
	if (p == Primaries::REC_709) return inverse(ncl_rgb_to_yuv(Matrix::REC_709));
	if (p == Primaries::REC_2020) return inverse(ncl_rgb_to_yuv(Matrix::REC_2020_NCL));
	double kr, kb;
	kr_kb_from_primaries(p, kr, kb);
	return inverse(rgb_to_yuv_kr_kb(kr, kb));
}

Mat3 lms_to_ictcp(Transfer t)
{
	if (t == Transfer::ARIB_B67)
		return { Vec3{ 0.5, 0.5, 0.0 }, Vec3{ 3625.0 / 4096.0, -7465.0 / 4096.0, 3840.0 / 4096.0 },
		         Vec3{ 9500.0 / 4096.0, -9212.0 / 4096.0, -288.0 / 4096.0 } };
	return { Vec3{ 0.5, 0.5, 0.0 }, Vec3{ 6610.0 / 4096.0, -13613.0 / 4096.0, 7003.0 / 4096.0 },
	         Vec3{ 17933.0 / 4096.0, -17390.0 / 4096.0, -543.0 / 4096.0 } };
}

Mat3 rgb_to_xyz(Primaries p)
{
	// Bruce Lindbloom, RGB/XYZ matrices from chromaticities.
	if (p == Primaries::XYZ) return identity3();
	auto xy = primaries_xy(p);
	Mat3 x = transpose(Mat3{ xy_to_xyz(xy[0][0], xy[0][1]), xy_to_xyz(xy[1][0], xy[1][1]), xy_to_xyz(xy[2][0], xy[2][1]) });
	Vec3 s = mul(inverse(x), white_point(p));
	Mat3 m{};
	for (int i = 0; i < 3; ++i)
		for (int k = 0; k < 3; ++k)
			m[i][k] = x[i][k] * s[k];
	return m;
}

Mat3 xyz_to_rgb(Primaries p) { return p == Primaries::XYZ ? identity3() : inverse(rgb_to_xyz(p)); }

Mat3 bradford_adaptation(Primaries in, Primaries out)
{
	const Mat3 bradford = { Vec3{ 0.8951, 0.2664, -0.1614 }, Vec3{ -0.7502, 1.7135, 0.0367 }, Vec3{ 0.0389, -0.0685, 1.0296 } };
	Vec3 wi = white_point(in), wo = white_point(out);
	if (wi == wo) return identity3();
	Vec3 ri = mul(bradford, wi), ro = mul(bradford, wo);
	Mat3 d{};
	d[0][0] = ro[0] / ri[0];
	d[1][1] = ro[1] / ri[1];
	d[2][2] = ro[2] / ri[2];
	return mul(mul(inverse(bradford), d), bradford);
}

struct TransferPair {
	int to_linear, to_gamma;
	float to_linear_scale = 1.0f, to_gamma_scale = 1.0f;
};

TransferPair select_transfer(Transfer t, double peak, bool scene)
{
	using namespace zm;
	TransferPair f{};
	switch (t) {
	case Transfer::LOG_100: f.to_linear = F_LOG100_INV_OETF; f.to_gamma = F_LOG100_OETF; break;
	case Transfer::LOG_316: f.to_linear = F_LOG316_INV_OETF; f.to_gamma = F_LOG316_OETF; break;
	case Transfer::REC_709:
		f.to_linear = scene ? F_709_INV_OETF : F_1886_EOTF;
		f.to_gamma = scene ? F_709_OETF : F_1886_INV_EOTF;
		break;
	case Transfer::REC_470_M: f.to_linear = F_470M_OETF; f.to_gamma = F_470M_INV_OETF; break;
	case Transfer::REC_470_BG: f.to_linear = F_470BG_OETF; f.to_gamma = F_470BG_INV_OETF; break;
	case Transfer::SMPTE_240M:
		f.to_linear = scene ? F_240M_INV_OETF : F_1886_EOTF;
		f.to_gamma = scene ? F_240M_OETF : F_1886_INV_EOTF;
		break;
	case Transfer::REC_1361:
		f.to_linear = scene ? F_1361_INV_OETF : F_1361_EOTF;
		f.to_gamma = scene ? F_1361_OETF : F_1361_INV_EOTF;
		break;
	case Transfer::XVYCC:
		f.to_linear = scene ? F_XVYCC_INV_OETF : F_XVYCC_EOTF;
		f.to_gamma = scene ? F_XVYCC_OETF : F_XVYCC_INV_EOTF;
		break;
	case Transfer::SRGB: f.to_linear = F_SRGB_EOTF; f.to_gamma = F_SRGB_INV_EOTF; break;
	case Transfer::ST_2084:
		f.to_linear = scene ? F_2084_INV_OETF : F_2084_EOTF;
		f.to_gamma = scene ? F_2084_OETF : F_2084_INV_EOTF;
		f.to_linear_scale = static_cast<float>(10000.0f / peak);
		f.to_gamma_scale = static_cast<float>(peak / 10000.0f);
		break;
	case Transfer::ST_428: f.to_linear = F_428_EOTF; f.to_gamma = F_428_INV_EOTF; break;
	case Transfer::ARIB_B67:
		f.to_linear = scene ? F_B67_INV_OETF : F_B67_EOTF;
		f.to_gamma = scene ? F_B67_OETF : F_B67_INV_EOTF;
		f.to_linear_scale = scene ? 12.0f : static_cast<float>(1000.0 / peak);
		f.to_gamma_scale = scene ? 1.0f / 12.0f : static_cast<float>(peak / 1000.0);
		break;
	default:
		fail(ZIMG_ERROR_UNKNOWN, "invalid transfer characteristics");
	}
	return f;
}

enum Edge : int { E_RGB_TO_YUV, E_YUV_TO_RGB, E_LIN_TO_GAMMA, E_GAMMA_TO_LIN, E_CL_RGB_TO_YUV, E_CL_YUV_TO_RGB,
                  E_GAMUT, E_LMS_TO_ICTCP, E_ICTCP_TO_LMS };

constexpr int NM = static_cast<int>(Matrix::COUNT), NT = static_cast<int>(Transfer::COUNT), NP = static_cast<int>(Primaries::COUNT);

int encode(const Colorspace &c) { return (static_cast<int>(c.matrix) * NT + static_cast<int>(c.transfer)) * NP + static_cast<int>(c.primaries); }

Colorspace decode(int i)
{
	return { static_cast<Matrix>(i / (NT * NP)), static_cast<Transfer>((i / NP) % NT), static_cast<Primaries>(i % NP) };
}

bool valid(const Colorspace &c)
{
	// colorspace_graph.cpp:67-100
	const bool cl_ok = c.matrix == Matrix::REC_2020_CL && c.transfer == Transfer::REC_709;
	const bool ictcp_ok = c.matrix == Matrix::REC_2100_ICTCP && (c.transfer == Transfer::ST_2084 || c.transfer == Transfer::ARIB_B67) &&
	                      c.primaries == Primaries::REC_2020;
	const bool lms_ok = c.matrix == Matrix::REC_2100_LMS &&
	                    (c.transfer == Transfer::LINEAR || c.transfer == Transfer::ST_2084 || c.transfer == Transfer::ARIB_B67) &&
	                    c.primaries == Primaries::REC_2020;
	return !(c.matrix == Matrix::REC_2020_CL && !cl_ok) &&
	       !(c.matrix == Matrix::CHROMATICITY_DERIVED_NCL && c.primaries == Primaries::UNSPECIFIED) &&
	       !(c.matrix == Matrix::CHROMATICITY_DERIVED_CL && c.primaries == Primaries::UNSPECIFIED) &&
	       !(c.matrix == Matrix::REC_2100_ICTCP && !ictcp_ok) &&
	       !(c.matrix == Matrix::REC_2100_LMS && !lms_ok);
}

// Outgoing edges of a state, in the reference's enumeration order
// (colorspace_graph.cpp:115-193); BFS tie-breaking depends on this order.
void neighbours(const Colorspace &c, std::vector<std::pair<Colorspace, Edge>> &out)
{
	out.clear();
	auto add = [&](const Colorspace &to, Edge e) { out.emplace_back(to, e); };

	if (c.matrix == Matrix::RGB) {
		for (int m = 0; m < NM; ++m) {
			Matrix mm = static_cast<Matrix>(m);
			if (mm == Matrix::UNSPECIFIED || mm == Matrix::RGB || mm == Matrix::REC_2020_CL || mm == Matrix::CHROMATICITY_DERIVED_NCL ||
			    mm == Matrix::CHROMATICITY_DERIVED_CL || mm == Matrix::REC_2100_LMS || mm == Matrix::REC_2100_ICTCP)
				continue;
			add(c.to(mm), E_RGB_TO_YUV);
		}
		if (c.primaries != Primaries::UNSPECIFIED)
			add(c.to(Matrix::CHROMATICITY_DERIVED_NCL), E_RGB_TO_YUV);

		if (c.transfer == Transfer::LINEAR) {
			for (int t = 0; t < NT; ++t) {
				Transfer tt = static_cast<Transfer>(t);
				if (tt == c.transfer || tt == Transfer::UNSPECIFIED)
					continue;
				add(c.to(tt), E_LIN_TO_GAMMA);
				if (c.primaries != Primaries::UNSPECIFIED)
					add(c.to(tt).to(Matrix::CHROMATICITY_DERIVED_CL), E_CL_RGB_TO_YUV);
			}
			if (c.primaries != Primaries::UNSPECIFIED) {
				for (int p = 0; p < NP; ++p) {
					Primaries pp = static_cast<Primaries>(p);
					if (pp != c.primaries && pp != Primaries::UNSPECIFIED)
						add(c.to(pp), E_GAMUT);
				}
			}
			add(c.to(Matrix::REC_2020_CL).to(Transfer::REC_709), E_CL_RGB_TO_YUV);
			if (c.primaries == Primaries::REC_2020)
				add(c.to(Matrix::REC_2100_LMS), E_RGB_TO_YUV);
		} else if (c.transfer != Transfer::UNSPECIFIED) {
			add(c.to(Transfer::LINEAR), E_GAMMA_TO_LIN);
		}
	} else if (c.matrix == Matrix::REC_2020_CL || c.matrix == Matrix::CHROMATICITY_DERIVED_CL) {
		add(c.to(Matrix::RGB).to(Transfer::LINEAR), E_CL_YUV_TO_RGB);
	} else if (c.matrix == Matrix::REC_2100_LMS) {
		if (c.transfer == Transfer::ST_2084 || c.transfer == Transfer::ARIB_B67) {
			add(c.to(Matrix::REC_2100_ICTCP), E_LMS_TO_ICTCP);
			add(c.to(Transfer::LINEAR), E_GAMMA_TO_LIN);
		}
		if (c.transfer == Transfer::LINEAR) {
			add(c.to(Matrix::RGB), E_YUV_TO_RGB);
			add(c.to(Transfer::ST_2084), E_LIN_TO_GAMMA);
			add(c.to(Transfer::ARIB_B67), E_LIN_TO_GAMMA);
		}
	} else if (c.matrix == Matrix::REC_2100_ICTCP) {
		if (c.transfer == Transfer::ST_2084 || c.transfer == Transfer::ARIB_B67)
			add(c.to(Matrix::REC_2100_LMS), E_ICTCP_TO_LMS);
	} else if (c.matrix != Matrix::UNSPECIFIED) {
		add(c.to(Matrix::RGB), E_YUV_TO_RGB);
	}
}

void set_matrix(CspOp &op, const Mat3 &m)
{
	op = CspOp{};
	op.kind = CSP_MATRIX;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			op.m[i * 3 + j] = static_cast<float>(m[i][j]);
}

CspOp make_op(Edge e, const Colorspace &in, const Colorspace &out, const CspParams &p)
{
	CspOp op{};
	const bool display_b67_in = in.primaries != Primaries::UNSPECIFIED && !p.approximate_gamma && !p.scene_referred;
	const bool display_b67_out = out.primaries != Primaries::UNSPECIFIED && !p.approximate_gamma && !p.scene_referred;

	switch (e) {
	case E_YUV_TO_RGB:
		set_matrix(op, in.matrix == Matrix::CHROMATICITY_DERIVED_NCL ? ncl_yuv_to_rgb_from_primaries(in.primaries) : inverse(ncl_rgb_to_yuv(in.matrix)));
		break;
	case E_RGB_TO_YUV:
		set_matrix(op, out.matrix == Matrix::CHROMATICITY_DERIVED_NCL ? ncl_rgb_to_yuv_from_primaries(out.primaries) : ncl_rgb_to_yuv(out.matrix));
		break;
	case E_ICTCP_TO_LMS:
		set_matrix(op, inverse(lms_to_ictcp(in.transfer)));
		break;
	case E_LMS_TO_ICTCP:
		set_matrix(op, lms_to_ictcp(in.transfer));
		break;
	case E_GAMUT:
		set_matrix(op, p.chromatic_adaptation ? mul(mul(xyz_to_rgb(out.primaries), bradford_adaptation(in.primaries, out.primaries)), rgb_to_xyz(in.primaries))
		                                      : mul(xyz_to_rgb(out.primaries), rgb_to_xyz(in.primaries)));
		break;
	case E_GAMMA_TO_LIN:
		if (in.transfer == Transfer::ARIB_B67 && display_b67_in) {
			// Display-referred HLG with luminance OOTF (operation_impl.cpp:117-158).
			Mat3 m = ncl_rgb_to_yuv_from_primaries(in.primaries);
			op.kind = CSP_B67_INV;
			op.kr = static_cast<float>(m[0][0]);
			op.kg = static_cast<float>(m[0][1]);
			op.kb = static_cast<float>(m[0][2]);
			op.post = select_transfer(Transfer::ARIB_B67, p.peak_luminance, false).to_linear_scale;
		} else {
			TransferPair f = select_transfer(in.transfer, p.peak_luminance, p.scene_referred);
			op.kind = CSP_GAMMA;
			op.func = f.to_linear;
			op.pre = 1.0f;
			op.post = f.to_linear_scale;
		}
		break;
	case E_LIN_TO_GAMMA:
		if (out.transfer == Transfer::ARIB_B67 && display_b67_out) {
			Mat3 m = ncl_rgb_to_yuv_from_primaries(out.primaries);
			op.kind = CSP_B67;
			op.kr = static_cast<float>(m[0][0]);
			op.kg = static_cast<float>(m[0][1]);
			op.kb = static_cast<float>(m[0][2]);
			op.pre = select_transfer(Transfer::ARIB_B67, p.peak_luminance, false).to_gamma_scale;
		} else {
			TransferPair f = select_transfer(out.transfer, p.peak_luminance, p.scene_referred);
			op.kind = CSP_GAMMA;
			op.func = f.to_gamma;
			op.pre = f.to_gamma_scale;
			op.post = 1.0f;
		}
		break;
	case E_CL_YUV_TO_RGB:
	case E_CL_RGB_TO_YUV: {
		// Constant-luminance BT.2020 is always scene-referred BT.709 transfer:
		// the factories assert it (operation_impl.cpp:364,377).
		const Colorspace &c = e == E_CL_YUV_TO_RGB ? in : out;
		TransferPair f = select_transfer(Transfer::REC_709, p.peak_luminance, true);
		Mat3 m = c.matrix == Matrix::CHROMATICITY_DERIVED_CL ? ncl_rgb_to_yuv_from_primaries(c.primaries) : ncl_rgb_to_yuv(c.matrix);
		op.kind = e == E_CL_YUV_TO_RGB ? CSP_CL_TO_RGB : CSP_CL_TO_YUV;
		op.func = e == E_CL_YUV_TO_RGB ? f.to_linear : f.to_gamma;
		op.kr = static_cast<float>(m[0][0]);
		op.kg = static_cast<float>(m[0][1]);
		op.kb = static_cast<float>(m[0][2]);
		op.nb = zm::transfer(f.to_gamma, 1.0f - static_cast<float>(m[0][2]));
		op.pb = 1.0f - zm::transfer(f.to_gamma, static_cast<float>(m[0][2]));
		op.nr = zm::transfer(f.to_gamma, 1.0f - static_cast<float>(m[0][0]));
		op.pr = 1.0f - zm::transfer(f.to_gamma, static_cast<float>(m[0][0]));
		op.pre = f.to_gamma_scale;
		op.post = f.to_linear_scale;
		break;
	}
	}
	if (p.approximate_gamma) {
		// The reference's AVX2 operations (operation_impl_x86.cpp:12-77): the
		// constant-luminance and HLG operations have none and stay exact.
		if (op.kind == CSP_MATRIX)
			op.kind = CSP_MATRIX_FMA;
		else if (op.kind == CSP_GAMMA)
			op.kind = e == E_GAMMA_TO_LIN ? CSP_LUT_LINEAR : CSP_LUT_GAMMA;
	}
	return op;
}

const char *op_name(int k)
{
	switch (k) {
	case CSP_MATRIX: return "matrix";
	case CSP_MATRIX_FMA: return "matrix_fma";
	case CSP_LUT_LINEAR: return "lut_linear";
	case CSP_LUT_GAMMA: return "lut_gamma";
	case CSP_GAMMA: return "gamma";
	case CSP_B67: return "hlg_oetf_ootf";
	case CSP_B67_SCALE: return "hlg_ootf_scale";
	case CSP_B67_INV: return "hlg_eotf";
	case CSP_CL_TO_RGB: return "cl_to_rgb";
	case CSP_CL_TO_YUV: return "cl_to_yuv";
	default: return "?";
	}
}

} // namespace

std::vector<float> build_gamma_lut(const CspOp &op)
{
	std::vector<float> lut;
	if (op.kind == CSP_LUT_LINEAR) {
		// operation_impl_avx2.cpp:258-274: entry i = f(i / 2^16 * 2 - 0.5) * post
		lut.resize(kLutLinearSize);
		for (unsigned i = 0; i < kLutLinearSize; ++i) {
			const float x = static_cast<float>(i) / 65536.0f * 2.0f - 0.5f;
			lut[i] = zm::transfer(op.func, x) * op.post;
		}
	} else if (op.kind == CSP_LUT_GAMMA) {
		// :304-320: entry h = f(half(h) * pre)
		lut.resize(kLutGammaSize);
		for (unsigned h = 0; h < kLutGammaSize; ++h) {
			const float x = zm::half_to_float(static_cast<uint16_t>(h));
			lut[h] = zm::transfer(op.func, x * op.pre);
		}
	}
	return lut;
}

std::string CspProgram::describe() const
{
	std::string s;
	for (int i = 0; i < n; ++i) {
		char buf[64];
		std::snprintf(buf, sizeof(buf), "%s%s(f%d)", i ? "," : "", op_name(op[i].kind), op[i].func);
		s += buf;
	}
	return s;
}

void ncl_rgb_to_yuv_matrix(Matrix m, double out[3][3])
{
	Mat3 r = ncl_rgb_to_yuv(m);
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			out[i][j] = r[i][j];
}

CspProgram build_colorspace_program(Colorspace in, Colorspace out, const CspParams &params)
{
	if (!params.scene_referred) {
		// SMPTE 240M transfer is treated as BT.709 unless scene-referred (colorspace.cpp:103-108).
		if (in.transfer == Transfer::SMPTE_240M) in.transfer = Transfer::REC_709;
		if (out.transfer == Transfer::SMPTE_240M) out.transfer = Transfer::REC_709;
	}
	CspProgram prog;
	if (in == out)
		return prog;
	if (!valid(in) || !valid(out))
		fail(ZIMG_ERROR_NO_COLORSPACE_CONVERSION, "invalid colorspace definition");

	const int nstates = NM * NT * NP;
	std::vector<int> parent(nstates, -1), via(nstates, -1);
	std::vector<char> seen(nstates, 0);
	std::deque<int> queue;
	std::vector<std::pair<Colorspace, Edge>> edges;
	const int src = encode(in), dst = encode(out);
	int v = -1;

	seen[src] = 1;
	queue.push_back(src);
	while (!queue.empty()) {
		v = queue.front();
		queue.pop_front();
		if (v == dst)
			break;
		neighbours(decode(v), edges);
		for (auto &e : edges) {
			int w = encode(e.first);
			if (seen[w])
				continue;
			seen[w] = 1;
			queue.push_back(w);
			parent[w] = v;
			via[w] = e.second;
		}
	}
	if (v != dst)
		fail(ZIMG_ERROR_NO_COLORSPACE_CONVERSION, "no path between colorspaces");

	std::vector<int> chain;
	for (int x = dst; x != src; x = parent[x])
		chain.push_back(x);
	if (chain.size() > 6)
		fail(ZIMG_ERROR_UNKNOWN, "too many operations");

	for (size_t i = 0; i < chain.size(); ++i) {
		int to = chain[chain.size() - 1 - i];
		prog.op[i] = make_op(static_cast<Edge>(via[to]), decode(parent[to]), decode(to), params);
	}
	prog.n = static_cast<int>(chain.size());
	return prog;
}

} // namespace zamd
