// Shared host-side types of the MI355X zimg library: pixel formats, the
// internal colourspace enumerations and the exception hierarchy that the C API
// maps onto zimg_error_code_e.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "zimg.h"

namespace zamd {

// Pixel types in zimg numbering (common/pixel.h:22-27).
enum class PixelType : int { BYTE = 0, WORD = 1, HALF = 2, FLOAT = 3 };

inline unsigned pixel_size(PixelType t) { return t == PixelType::BYTE ? 1 : t == PixelType::FLOAT ? 4 : 2; }
inline unsigned pixel_depth(PixelType t) { return t == PixelType::BYTE ? 8 : t == PixelType::FLOAT ? 32 : 16; }
inline bool pixel_is_integer(PixelType t) { return t == PixelType::BYTE || t == PixelType::WORD; }
inline bool pixel_is_float(PixelType t) { return !pixel_is_integer(t); }

// Pixel format and its equality rule (common/pixel.h:118-199): floating-point
// formats compare type and chroma only.
struct PixelFormat {
	PixelType type = PixelType::BYTE;
	unsigned depth = 0;
	bool fullrange = false;
	bool chroma = false;
	bool ycgco = false;

	PixelFormat() = default;
	PixelFormat(PixelType t) : type(t), depth(pixel_depth(t)) {}
	PixelFormat(PixelType t, unsigned d, bool full = false, bool chr = false, bool yc = false)
		: type(t), depth(d), fullrange(full), chroma(chr), ycgco(yc) {}

	friend bool operator==(const PixelFormat &a, const PixelFormat &b)
	{
		return pixel_is_float(a.type) ? a.type == b.type && a.chroma == b.chroma
		                              : a.type == b.type && a.depth == b.depth && a.fullrange == b.fullrange && a.chroma == b.chroma;
	}
	friend bool operator!=(const PixelFormat &a, const PixelFormat &b) { return !(a == b); }
};

// Internal colourspace enumerations (colorspace/colorspace.h:18-63).
enum class Matrix : int { UNSPECIFIED, RGB, REC_601, REC_709, FCC, SMPTE_240M, YCGCO, REC_2020_NCL, REC_2020_CL,
                          CHROMATICITY_DERIVED_NCL, CHROMATICITY_DERIVED_CL, REC_2100_LMS, REC_2100_ICTCP, COUNT };
enum class Transfer : int { UNSPECIFIED, LINEAR, LOG_100, LOG_316, REC_709, REC_470_M, REC_470_BG, SMPTE_240M, XVYCC,
                            REC_1361, SRGB, ST_2084, ST_428, ARIB_B67, COUNT };
enum class Primaries : int { UNSPECIFIED, REC_470_M, REC_470_BG, SMPTE_C, REC_709, FILM, REC_2020, XYZ, DCI_P3,
                             DCI_P3_D65, EBU_3213_E, COUNT };

struct Colorspace {
	Matrix matrix = Matrix::UNSPECIFIED;
	Transfer transfer = Transfer::UNSPECIFIED;
	Primaries primaries = Primaries::UNSPECIFIED;

	Colorspace to(Matrix m) const { Colorspace c = *this; c.matrix = m; return c; }
	Colorspace to(Transfer t) const { Colorspace c = *this; c.transfer = t; return c; }
	Colorspace to(Primaries p) const { Colorspace c = *this; c.primaries = p; return c; }
	friend bool operator==(const Colorspace &a, const Colorspace &b)
	{
		return a.matrix == b.matrix && a.transfer == b.transfer && a.primaries == b.primaries;
	}
	friend bool operator!=(const Colorspace &a, const Colorspace &b) { return !(a == b); }
};

enum class Dither : int { NONE = 0, ORDERED = 1, RANDOM = 2, ERROR_DIFFUSION = 3 };

// Exceptions carrying the C API error code (zimg.cpp:61-102 maps the
// reference's exception types to the same codes).
struct Error : std::runtime_error {
	zimg_error_code_e code;
	Error(zimg_error_code_e c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

[[noreturn]] inline void fail(zimg_error_code_e code, const std::string &msg) { throw Error(code, msg); }

} // namespace zamd
