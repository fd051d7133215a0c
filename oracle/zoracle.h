/*
 * zoracle — CPU restatement of zimg's per-plane arithmetic, used ONLY as the
 * test oracle (tests/, __graft_entry__.smoke(), bench.py's cpu_baseline leg).
 * Nothing in the shipped library links or calls this code.
 *
 * Every function cites the reference file:line it restates. Reference paths
 * are relative to the zimg source tree (src/zimg/...).
 *
 * Build: plain C99, -O2 -ffp-contract=off -fno-fast-math (the reference's unit
 * tests force the same flags, configure.ac:37-39), so float expressions round
 * exactly as the reference's portable C path does.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - filter kernels: reference KATs (test/resize/filter_test.cpp:8-106);
 *   - compute_filter / matrix_to_filter tables: bit-compared with the compiled
 *     reference resize/filter.cpp (oracle/_ref) over many (src, dst, shift);
 *   - powf / expf / logf: bit-compared with the reference's vendored musl
 *     (test/extra/musl-libm) compiled into oracle/_ref;
 *   - transfer functions, 3x3 matrices, colourspace operation chains:
 *     bit-compared with the compiled reference colorspace sources;
 *   - resize, depth, ordered dither, error diffusion kernels: closed forms and
 *     brute-force checks (the reference's resize_impl.cpp / dither.cpp need
 *     the absent graphengine headers and cannot be compiled here).
 *   - log10f: fdlibm method, bit-compared with the same vendored musl.
 */
#ifndef ZORACLE_H_
#define ZORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- libm restatement (ARM optimized-routines powf/expf/logf and fdlibm
 *      log10f, the methods of the reference's vendored musl 1.2.0) ---- */
float orc_powf(float x, float y);
float orc_expf(float x);
float orc_logf(float x);
float orc_log10f(float x);

/* ---- resampling filters (resize/filter.cpp:134-246) ---- */
enum {
	ORC_FILTER_POINT = 0,
	ORC_FILTER_BILINEAR = 1,
	ORC_FILTER_BICUBIC = 2,
	ORC_FILTER_SPLINE16 = 3,
	ORC_FILTER_SPLINE36 = 4,
	ORC_FILTER_LANCZOS = 5,
	ORC_FILTER_SPLINE64 = 6,
};

/* param_a / param_b: bicubic b, c; lanczos taps in param_a. */
double orc_filter_eval(int type, double param_a, double param_b, double x);
unsigned orc_filter_support(int type, double param_a);

/* compute_filter (resize/filter.cpp:249-306) + matrix_to_filter (:43-127).
 * First call with data == NULL to get filter_width; then call again with
 * data[rows*fw], data_i16[rows*fw], left[rows]. Rows are packed (stride = fw).
 * Returns 0 on success, nonzero on error. */
int orc_compute_filter(int type, double param_a, double param_b,
                       unsigned src_dim, unsigned dst_dim, double shift, double width,
                       unsigned *filter_width,
                       float *data, int16_t *data_i16, unsigned *left);

/* As orc_compute_filter, also returning the unrounded FP64 entries of the
 * resampling matrix in the packed layout (data64, may be NULL). */
int orc_compute_filter64(int type, double param_a, double param_b,
                         unsigned src_dim, unsigned dst_dim, double shift, double width,
                         unsigned *filter_width,
                         float *data, int16_t *data_i16, double *data64, unsigned *left);

/* ---- resize kernels (resize/resize_impl.cpp:31-114), whole planes ---- */
void orc_resize_h_u16(const int16_t *coeffs, const unsigned *left, unsigned fw, unsigned dst_w,
                      const uint16_t *src, ptrdiff_t src_stride, uint16_t *dst, ptrdiff_t dst_stride,
                      unsigned height, unsigned depth);
void orc_resize_v_u16(const int16_t *coeffs, const unsigned *left, unsigned fw, unsigned dst_h,
                      const uint16_t *src, ptrdiff_t src_stride, uint16_t *dst, ptrdiff_t dst_stride,
                      unsigned width, unsigned depth);
void orc_resize_h_f32(const float *coeffs, const unsigned *left, unsigned fw, unsigned dst_w,
                      const float *src, ptrdiff_t src_stride, float *dst, ptrdiff_t dst_stride,
                      unsigned height);
void orc_resize_v_f32(const float *coeffs, const unsigned *left, unsigned fw, unsigned dst_h,
                      const float *src, ptrdiff_t src_stride, float *dst, ptrdiff_t dst_stride,
                      unsigned width);

/* ---- depth (depth/quantize.h, depth_convert.cpp, dither.cpp, quantize.cpp) ----
 * Pixel types follow zimg: 0 BYTE, 1 WORD, 2 HALF, 3 FLOAT. */
void orc_scale_offset(int type_in, unsigned depth_in, int full_in, int chroma_in, int ycgco_in,
                      int type_out, unsigned depth_out, int full_out, int chroma_out, int ycgco_out,
                      float *scale, float *offset);
float orc_half_to_float(uint16_t h);
uint16_t orc_float_to_half(float f);

/* Strides are in bytes. */
void orc_convert_to_float(int type_in, const void *src, ptrdiff_t src_stride,
                          float *dst, ptrdiff_t dst_stride, unsigned width, unsigned height,
                          float scale, float offset);
/* Conversion into HALF storage (depth_convert.cpp:32-63): integers take
 * float_to_half(x * scale + offset), FLOAT input float_to_half(x). */
void orc_convert_to_half(int type_in, const void *src, ptrdiff_t src_stride,
                         uint16_t *dst, ptrdiff_t dst_stride, unsigned width, unsigned height,
                         float scale, float offset);
void orc_left_shift(int type_in, const void *src, ptrdiff_t src_stride,
                    int type_out, void *dst, ptrdiff_t dst_stride,
                    unsigned width, unsigned height, unsigned shift);

/* dither_type: 0 NONE, 1 ORDERED (Bayer), 2 RANDOM (blue noise; requires
 * blue_noise[64*64] bytes), plane selects the Bayer/blue-noise variant. */
void orc_dither_ordered(int dither_type, const uint8_t *blue_noise, unsigned plane,
                        int type_in, const void *src, ptrdiff_t src_stride,
                        int type_out, void *dst, ptrdiff_t dst_stride,
                        unsigned width, unsigned height, float scale, float offset, unsigned bits);
void orc_dither_ed(int type_in, const void *src, ptrdiff_t src_stride,
                   int type_out, void *dst, ptrdiff_t dst_stride,
                   unsigned width, unsigned height, float scale, float offset, unsigned bits);
/* Bayer 16x16 table (dither.cpp:140-157) regenerated by bit interleaving. */
void orc_bayer_table(uint8_t out[256]);

/* ---- colourspace (colorspace sources) ----
 * Enumerations are the reference's internal ones (colorspace.h:18-63). */
int orc_transfer_eval(int transfer, int to_linear, int scene_referred, float x, float *y);
void orc_transfer_scales(int transfer, double peak_luminance, int scene_referred,
                         float *to_linear_scale, float *to_gamma_scale);

/* Build the operation chain between two colourspaces (colorspace_graph.cpp:198-243).
 * Returns the number of operations (<= 6), or -1 if no path. The opaque
 * descriptor is filled for orc_colorspace_apply. A short text trace of the ops
 * is written to trace (may be NULL). */
typedef struct orc_csp_ops {
	int n;
	int kind[6];
	float m[6][3][3];
	int func[6];
	float pre[6], post[6];
	float kr[6], kg[6], kb[6];
	float nb[6], pb[6], nr[6], pr[6];
	int aux_func[6];
} orc_csp_ops;

int orc_colorspace_build(int m_in, int t_in, int p_in, int m_out, int t_out, int p_out,
                         double peak_luminance, int scene_referred, int chromatic_adaptation,
                         orc_csp_ops *ops, char *trace, size_t trace_len);
/* The chain the reference runs with allow_approximate_gamma on an x86 CPU
 * with AVX2, FMA and F16C (cpu AUTO): FMA matrices, lookup-table gamma,
 * scene-referred HLG. Same arguments and result as orc_colorspace_build. */
int orc_colorspace_build_approx(int m_in, int t_in, int p_in, int m_out, int t_out, int p_out,
                                double peak_luminance, int scene_referred, int chromatic_adaptation,
                                orc_csp_ops *ops, char *trace, size_t trace_len);
void orc_colorspace_apply(const orc_csp_ops *ops, float *p0, float *p1, float *p2, size_t n);
/* ---- unresize (unresize/bilinear.cpp, unresize_impl.cpp) ----
 * Context of unresizing up_dim samples to orig_dim (call first with coef ==
 * NULL for row_size): A' band rows of row_size taps from offset[i] (stride
 * row_size), and the LU factors of A'A (lu_l inverted). */
int orc_unresize_context(unsigned up_dim, unsigned orig_dim, double shift, unsigned *row_size,
                         float *coef, unsigned *offset, float *lu_c, float *lu_l, float *lu_u);
void orc_unresize_h_f32(const float *coef, const unsigned *offset, unsigned row_size, const float *lu_c, const float *lu_l,
                        const float *lu_u, unsigned out_w, const float *src, ptrdiff_t src_stride, float *dst,
                        ptrdiff_t dst_stride, unsigned height);
void orc_unresize_v_f32(const float *coef, const unsigned *offset, unsigned row_size, const float *lu_c, const float *lu_l,
                        const float *lu_u, unsigned out_h, const float *src, ptrdiff_t src_stride, float *dst,
                        ptrdiff_t dst_stride, unsigned width);

/* Double-precision matrix helpers exposed for tests. */
void orc_ncl_rgb_to_yuv(int matrix, double out[9]);
void orc_gamut_matrix(int p_in, int p_out, int adapt, double out[9]);

#ifdef __cplusplus
}
#endif

#endif
