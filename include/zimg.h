/*
 * zimg.h — C ABI of the MI355X (gfx950) per-plane image processing library.
 *
 * Drop-in for the z.lib C API 2.4 (reference: src/zimg/api/zimg.h). Every
 * type, enumerator value, struct layout and function signature below matches
 * the reference header so that existing callers (VapourSynth, FFmpeg zscale,
 * the reference's doc/example programs) link against this library unchanged.
 * Each declaration cites the reference declaration it replaces.
 *
 * Additions, marked "device extension", take device (HBM) pointers and a HIP
 * stream (passed as void * so this header stays free of HIP types) and run the
 * graph asynchronously on the GPU.
 */
#ifndef ZIMG_H_
#define ZIMG_H_

#include <stddef.h>

#if defined(_WIN32) || defined(__CYGWIN__)
  #define ZIMG_VISIBILITY
#elif defined(__GNUC__)
  #define ZIMG_VISIBILITY __attribute__((visibility("default")))
#else
  #define ZIMG_VISIBILITY
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define ZIMG_MAKE_API_VERSION(x, y) (((x) << 8) | (y))
#define ZIMG_API_VERSION_MAJOR 2
#define ZIMG_API_VERSION_MINOR 4
#define ZIMG_API_VERSION ZIMG_MAKE_API_VERSION(ZIMG_API_VERSION_MAJOR, ZIMG_API_VERSION_MINOR)

/* zimg.h:60 */
ZIMG_VISIBILITY
void zimg_get_version_info(unsigned *major, unsigned *minor, unsigned *micro);

/* zimg.h:71 */
ZIMG_VISIBILITY
unsigned zimg_get_api_version(unsigned *major, unsigned *minor);

/* zimg.h:79-121 */
typedef enum zimg_error_code_e {
	ZIMG_ERROR_UNKNOWN = -1,
	ZIMG_ERROR_SUCCESS = 0,

	ZIMG_ERROR_OUT_OF_MEMORY        = 1,
	ZIMG_ERROR_USER_CALLBACK_FAILED = 2,

	ZIMG_ERROR_LOGIC                 = (1 << 10),
	ZIMG_ERROR_GREYSCALE_SUBSAMPLING = ZIMG_ERROR_LOGIC + 1,
	ZIMG_ERROR_COLOR_FAMILY_MISMATCH = ZIMG_ERROR_LOGIC + 2,
	ZIMG_ERROR_IMAGE_NOT_DIVISIBLE   = ZIMG_ERROR_LOGIC + 3,
	ZIMG_ERROR_BIT_DEPTH_OVERFLOW    = ZIMG_ERROR_LOGIC + 4,

	ZIMG_ERROR_ILLEGAL_ARGUMENT   = (2 << 10),
	ZIMG_ERROR_ENUM_OUT_OF_RANGE  = ZIMG_ERROR_ILLEGAL_ARGUMENT + 1,
	ZIMG_ERROR_INVALID_IMAGE_SIZE = ZIMG_ERROR_ILLEGAL_ARGUMENT + 2,

	ZIMG_ERROR_UNSUPPORTED_OPERATION      = (3 << 10),
	ZIMG_ERROR_UNSUPPORTED_SUBSAMPLING    = ZIMG_ERROR_UNSUPPORTED_OPERATION + 1,
	ZIMG_ERROR_NO_COLORSPACE_CONVERSION   = ZIMG_ERROR_UNSUPPORTED_OPERATION + 2,
	ZIMG_ERROR_NO_FIELD_PARITY_CONVERSION = ZIMG_ERROR_UNSUPPORTED_OPERATION + 3,
	ZIMG_ERROR_RESAMPLING_NOT_AVAILABLE   = ZIMG_ERROR_UNSUPPORTED_OPERATION + 4
} zimg_error_code_e;

/* zimg.h:136 */
ZIMG_VISIBILITY
zimg_error_code_e zimg_get_last_error(char *err_msg, size_t n);

/* zimg.h:142 */
ZIMG_VISIBILITY
void zimg_clear_last_error(void);

/* zimg.h:145-171. The x86 values are accepted and treated like AUTO: the
 * device path does not depend on the host CPU class. */
typedef enum zimg_cpu_type_e {
	ZIMG_CPU_NONE     = 0,
	ZIMG_CPU_AUTO     = 1,
	ZIMG_CPU_AUTO_64B = 2
#if defined(__i386) || defined(_M_IX86) || defined(_M_X64) || defined(__x86_64__)
	,ZIMG_CPU_X86_MMX       = 1000,
	ZIMG_CPU_X86_SSE        = 1001,
	ZIMG_CPU_X86_SSE2       = 1002,
	ZIMG_CPU_X86_SSE3       = 1003,
	ZIMG_CPU_X86_SSSE3      = 1004,
	ZIMG_CPU_X86_SSE41      = 1005,
	ZIMG_CPU_X86_SSE42      = 1006,
	ZIMG_CPU_X86_AVX        = 1007,
	ZIMG_CPU_X86_F16C       = 1008,
	ZIMG_CPU_X86_AVX2       = 1009,
	ZIMG_CPU_X86_AVX512F    = 1010,
	ZIMG_CPU_X86_AVX512_SKX = 1011,
	ZIMG_CPU_X86_AVX512_CLX = 1012,
	ZIMG_CPU_X86_AVX512_PMC = 1013,
	ZIMG_CPU_X86_AVX512_SNC = 1014,
	ZIMG_CPU_X86_AVX512_WLC = 1015,
	ZIMG_CPU_X86_AVX512_GLC = 1016
#endif
} zimg_cpu_type_e;

/* zimg.h:173-178 */
typedef enum zimg_pixel_type_e {
	ZIMG_PIXEL_BYTE  = 0,
	ZIMG_PIXEL_WORD  = 1,
	ZIMG_PIXEL_HALF  = 2,
	ZIMG_PIXEL_FLOAT = 3
} zimg_pixel_type_e;

/* zimg.h:180-184 */
typedef enum zimg_pixel_range_e {
	ZIMG_RANGE_INTERNAL = -1,
	ZIMG_RANGE_LIMITED  = 0,
	ZIMG_RANGE_FULL     = 1
} zimg_pixel_range_e;

/* zimg.h:186-190 */
typedef enum zimg_color_family_e {
	ZIMG_COLOR_GREY = 0,
	ZIMG_COLOR_RGB  = 1,
	ZIMG_COLOR_YUV  = 2
} zimg_color_family_e;

/* zimg.h:192-196 */
typedef enum zimg_alpha_type_e {
	ZIMG_ALPHA_NONE          = 0,
	ZIMG_ALPHA_STRAIGHT      = 1,
	ZIMG_ALPHA_PREMULTIPLIED = 2
} zimg_alpha_type_e;

/* zimg.h:198-202 */
typedef enum zimg_field_parity_e {
	ZIMG_FIELD_PROGRESSIVE = 0,
	ZIMG_FIELD_TOP         = 1,
	ZIMG_FIELD_BOTTOM      = 2
} zimg_field_parity_e;

/* zimg.h:204-212 */
typedef enum zimg_chroma_location_e {
	ZIMG_CHROMA_INTERNAL    = -1,
	ZIMG_CHROMA_LEFT        = 0,
	ZIMG_CHROMA_CENTER      = 1,
	ZIMG_CHROMA_TOP_LEFT    = 2,
	ZIMG_CHROMA_TOP         = 3,
	ZIMG_CHROMA_BOTTOM_LEFT = 4,
	ZIMG_CHROMA_BOTTOM      = 5
} zimg_chroma_location_e;

/* zimg.h:214-245 */
typedef enum zimg_matrix_coefficients_e {
	ZIMG_MATRIX_INTERNAL                 = -1,
	ZIMG_MATRIX_RGB                      = 0,
	ZIMG_MATRIX_BT709                    = 1,
	ZIMG_MATRIX_UNSPECIFIED              = 2,
	ZIMG_MATRIX_FCC                      = 4,
	ZIMG_MATRIX_BT470_BG                 = 5,
	ZIMG_MATRIX_ST170_M                  = 6,
	ZIMG_MATRIX_ST240_M                  = 7,
	ZIMG_MATRIX_YCGCO                    = 8,
	ZIMG_MATRIX_BT2020_NCL               = 9,
	ZIMG_MATRIX_BT2020_CL                = 10,
	ZIMG_MATRIX_ST2085_YDZDX             = 11,
	ZIMG_MATRIX_CHROMATICITY_DERIVED_NCL = 12,
	ZIMG_MATRIX_CHROMATICITY_DERIVED_CL  = 13,
	ZIMG_MATRIX_BT2100_ICTCP             = 14,
	ZIMG_MATRIX_ST2128_IPT_C2            = 15,
	ZIMG_MATRIX_YCGCO_RE                 = 16,
	ZIMG_MATRIX_YCGCO_RO                 = 17

#define ZIMG_MATRIX_709      ZIMG_MATRIX_BT709
#define ZIMG_MATRIX_470BG    ZIMG_MATRIX_BT470_BG
#define ZIMG_MATRIX_170M     ZIMG_MATRIX_ST170_M
#define ZIMG_MATRIX_240M     ZIMG_MATRIX_ST240_M
#define ZIMG_MATRIX_2020_NCL ZIMG_MATRIX_BT2020_NCL
#define ZIMG_MATRIX_2020_CL  ZIMG_MATRIX_BT2020_CL
#define ZIMG_MATRIX_ICTCP    ZIMG_MATRIX_BT2100_ICTCP
} zimg_matrix_coefficients_e;

/* zimg.h:247-276 */
typedef enum zimg_transfer_characteristics_e {
	ZIMG_TRANSFER_INTERNAL      = -1,
	ZIMG_TRANSFER_BT709         = 1,
	ZIMG_TRANSFER_UNSPECIFIED   = 2,
	ZIMG_TRANSFER_BT470_M       = 4,
	ZIMG_TRANSFER_BT470_BG      = 5,
	ZIMG_TRANSFER_BT601         = 6,
	ZIMG_TRANSFER_ST240_M       = 7,
	ZIMG_TRANSFER_LINEAR        = 8,
	ZIMG_TRANSFER_LOG_100       = 9,
	ZIMG_TRANSFER_LOG_316       = 10,
	ZIMG_TRANSFER_IEC_61966_2_4 = 11,
	ZIMG_TRANSFER_BT1361        = 12,
	ZIMG_TRANSFER_IEC_61966_2_1 = 13,
	ZIMG_TRANSFER_BT2020_10     = 14,
	ZIMG_TRANSFER_BT2020_12     = 15,
	ZIMG_TRANSFER_ST2084        = 16,
	ZIMG_TRANSFER_ST428         = 17,
	ZIMG_TRANSFER_ARIB_B67      = 18

#define ZIMG_TRANSFER_709     ZIMG_TRANSFER_BT709
#define ZIMG_TRANSFER_470_M   ZIMG_TRANSFER_BT470_M
#define ZIMG_TRANSFER_470_BG  ZIMG_TRANSFER_BT470_BG
#define ZIMG_TRANSFER_601     ZIMG_TRANSFER_BT601
#define ZIMG_TRANSFER_240M    ZIMG_TRANSFER_ST240_M
#define ZIMG_TRANSFER_2020_10 ZIMG_TRANSFER_BT2020_10
#define ZIMG_TRANSFER_2020_12 ZIMG_TRANSFER_BT2020_12
} zimg_transfer_characteristics_e;

/* zimg.h:278-302 */
typedef enum zimg_color_primaries_e {
	ZIMG_PRIMARIES_INTERNAL    = -1,
	ZIMG_PRIMARIES_BT709       = 1,
	ZIMG_PRIMARIES_UNSPECIFIED = 2,
	ZIMG_PRIMARIES_BT470_M     = 4,
	ZIMG_PRIMARIES_BT470_BG    = 5,
	ZIMG_PRIMARIES_ST170_M     = 6,
	ZIMG_PRIMARIES_ST240_M     = 7,
	ZIMG_PRIMARIES_FILM        = 8,
	ZIMG_PRIMARIES_BT2020      = 9,
	ZIMG_PRIMARIES_ST428       = 10,
	ZIMG_PRIMARIES_ST431_2     = 11,
	ZIMG_PRIMARIES_ST432_1     = 12,
	ZIMG_PRIMARIES_EBU3213_E   = 22

#define ZIMG_PRIMARIES_709    ZIMG_PRIMARIES_BT709
#define ZIMG_PRIMARIES_470_M  ZIMG_PRIMARIES_BT470_M
#define ZIMG_PRIMARIES_470_BG ZIMG_PRIMARIES_BT470_BG
#define ZIMG_PRIMARIES_170M   ZIMG_PRIMARIES_ST170_M
#define ZIMG_PRIMARIES_240M   ZIMG_PRIMARIES_ST240_M
#define ZIMG_PRIMARIES_2020   ZIMG_PRIMARIES_BT2020
} zimg_color_primaries_e;

/* zimg.h:304-309 */
typedef enum zimg_dither_type_e {
	ZIMG_DITHER_NONE            = 0,
	ZIMG_DITHER_ORDERED         = 1,
	ZIMG_DITHER_RANDOM          = 2,
	ZIMG_DITHER_ERROR_DIFFUSION = 3
} zimg_dither_type_e;

/* zimg.h:371-379 */
typedef enum zimg_resample_filter_e {
	ZIMG_RESIZE_POINT    = 0,
	ZIMG_RESIZE_BILINEAR = 1,
	ZIMG_RESIZE_BICUBIC  = 2,
	ZIMG_RESIZE_SPLINE16 = 3,
	ZIMG_RESIZE_SPLINE36 = 4,
	ZIMG_RESIZE_SPLINE64 = 6,
	ZIMG_RESIZE_LANCZOS  = 5
} zimg_resample_filter_e;

#define ZIMG_BUFFER_MAX (~0U)

/* zimg.h:409-430. Row i of plane p starts at data + (i & mask) * stride. */
typedef struct zimg_image_buffer_const {
	unsigned version;

	struct {
		const void *data;
		ptrdiff_t stride;
		unsigned mask;
	} plane[4];
} zimg_image_buffer_const;

typedef struct zimg_image_buffer {
	unsigned version;

	struct {
		void *data;
		ptrdiff_t stride;
		unsigned mask;
	} plane[4];
} zimg_image_buffer;

/* zimg.h:438 */
ZIMG_VISIBILITY
unsigned zimg_select_buffer_mask(unsigned count);

/* zimg.h:451 */
typedef struct zimg_filter_graph zimg_filter_graph;

/* zimg.h:484 */
typedef int (*zimg_filter_graph_callback)(void *user, unsigned i, unsigned left, unsigned right);

/* zimg.h:492 */
ZIMG_VISIBILITY
void zimg_filter_graph_free(zimg_filter_graph *ptr);

/* zimg.h:507. Size of the HOST temporary buffer for zimg_filter_graph_process. */
ZIMG_VISIBILITY
zimg_error_code_e zimg_filter_graph_get_tmp_size(const zimg_filter_graph *ptr, size_t *out);

/* zimg.h:521 */
ZIMG_VISIBILITY
zimg_error_code_e zimg_filter_graph_get_input_buffering(const zimg_filter_graph *ptr, unsigned *out);

/* zimg.h:533 */
ZIMG_VISIBILITY
zimg_error_code_e zimg_filter_graph_get_output_buffering(const zimg_filter_graph *ptr, unsigned *out);

/* zimg.h:550-552. Host buffers (ring buffers and callbacks allowed); the
 * frame is staged to HBM, processed by the device kernels on the library's
 * stream, and staged back before the call returns. */
ZIMG_VISIBILITY
zimg_error_code_e zimg_filter_graph_process(const zimg_filter_graph *ptr, const zimg_image_buffer_const *src, const zimg_image_buffer *dst, void *tmp,
                                            zimg_filter_graph_callback unpack_cb, void *unpack_user,
                                            zimg_filter_graph_callback pack_cb, void *pack_user);

/* zimg.h:558-588 */
typedef struct zimg_image_format {
	unsigned version;

	unsigned width;
	unsigned height;
	zimg_pixel_type_e pixel_type;

	unsigned subsample_w;
	unsigned subsample_h;

	zimg_color_family_e color_family;
	zimg_matrix_coefficients_e matrix_coefficients;
	zimg_transfer_characteristics_e transfer_characteristics;
	zimg_color_primaries_e color_primaries;

	unsigned depth;
	zimg_pixel_range_e pixel_range;

	zimg_field_parity_e field_parity;
	zimg_chroma_location_e chroma_location;

	struct {
		double left;
		double top;
		double width;
		double height;
	} active_region;

	zimg_alpha_type_e alpha;
} zimg_image_format;

/* zimg.h:593-647 */
typedef struct zimg_graph_builder_params {
	unsigned version;

	zimg_resample_filter_e resample_filter;
	double filter_param_a;
	double filter_param_b;

	zimg_resample_filter_e resample_filter_uv;
	double filter_param_a_uv;
	double filter_param_b_uv;

	zimg_dither_type_e dither_type;
	zimg_cpu_type_e cpu_type;

	double nominal_peak_luminance;
	char allow_approximate_gamma;

	char scene_referred;
	char chromatic_adaptation;
} zimg_graph_builder_params;

/* zimg.h:655 */
ZIMG_VISIBILITY
void zimg_image_format_default(zimg_image_format *ptr, unsigned version);

/* zimg.h:663 */
ZIMG_VISIBILITY
void zimg_graph_builder_params_default(zimg_graph_builder_params *ptr, unsigned version);

/* zimg.h:679 */
ZIMG_VISIBILITY
zimg_filter_graph *zimg_filter_graph_build(const zimg_image_format *src_format, const zimg_image_format *dst_format, const zimg_graph_builder_params *params);


/* ---------------------------------------------------------------------------
 * Device extensions (no counterpart in the reference; they replace the
 * reference's graphengine row scheduler, filtergraph.cpp:96-122, with one
 * asynchronous launch sequence on a HIP stream).
 *
 * Device buffers: plane[p].data are HBM pointers, plane[p].mask must be
 * ZIMG_BUFFER_MAX, strides are positive multiples of 4 bytes. `stream` is a
 * hipStream_t (NULL = the default stream). Calls return once the work is
 * enqueued; results are valid when the stream has drained. The caller owns
 * tmp (device memory of zimg_filter_graph_get_device_tmp_size bytes); calls
 * that share one tmp buffer must be ordered on the same stream.
 * ------------------------------------------------------------------------ */

/* device extension: bytes of device scratch needed for `count` frames. */
ZIMG_VISIBILITY
zimg_error_code_e zimg_filter_graph_get_device_tmp_size(const zimg_filter_graph *ptr, unsigned count, size_t *out);

/* device extension: one frame, device buffers. */
ZIMG_VISIBILITY
zimg_error_code_e zimg_filter_graph_process_device(const zimg_filter_graph *ptr, const zimg_image_buffer_const *src,
                                                   const zimg_image_buffer *dst, void *tmp, void *stream);

/* device extension: `count` frames of one graph in one launch sequence.
 * Frame f of plane p is at plane[p].data + f * frame_stride[p] bytes. */
ZIMG_VISIBILITY
zimg_error_code_e zimg_filter_graph_process_device_batch(const zimg_filter_graph *ptr, unsigned count,
                                                         const zimg_image_buffer_const *src, const ptrdiff_t src_frame_stride[4],
                                                         const zimg_image_buffer *dst, const ptrdiff_t dst_frame_stride[4],
                                                         void *tmp, void *stream);

/* device extension: arithmetic of FLOAT resize stages on the device.
 * ZIMG_RESIZE_PRECISION_EXACT: the reference's FP32 evaluation order, bit for
 * bit (default). ZIMG_RESIZE_PRECISION_MFMA_FP16 / _BF16: each pass as a
 * banded contraction on matrix cores with FP32 accumulation, operands split
 * into a 16-bit high part and a 16-bit remainder (FP16 requires |samples| <
 * 65504). Integer (WORD) resizes and dithered outputs always use the exact
 * path. Call before zimg_filter_graph_get_device_tmp_size and processing; not
 * safe against concurrent processing of the same graph. */
typedef enum zimg_resize_precision_e {
	ZIMG_RESIZE_PRECISION_EXACT     = 0,
	ZIMG_RESIZE_PRECISION_MFMA_FP16 = 1,
	ZIMG_RESIZE_PRECISION_MFMA_BF16 = 2
} zimg_resize_precision_e;

ZIMG_VISIBILITY
zimg_error_code_e zimg_filter_graph_set_resize_precision(zimg_filter_graph *ptr, int mode);

/* device extension, no counterpart in the reference (as a tile-shape
 * autotuner for callers that process many batches of one format pair): time
 * the candidate tile shapes of the graph's tiled resize steps on these
 * buffers with `count` frames (arguments as zimg_filter_graph_process_device_batch;
 * dst is overwritten) and keep the fastest for later calls. Outputs are the
 * same with any shape. *changed (may be NULL) receives the number of steps
 * whose shape changed. Not safe against concurrent processing of the same
 * graph. */
ZIMG_VISIBILITY
zimg_error_code_e zimg_filter_graph_autotune_device(zimg_filter_graph *ptr, unsigned count, const zimg_image_buffer_const *src,
                                                   const ptrdiff_t src_frame_stride[4], const zimg_image_buffer *dst,
                                                   const ptrdiff_t dst_frame_stride[4], void *tmp, void *stream, unsigned *changed);

/* device extension, no counterpart in the reference: the allocation-free form
 * of zimg_filter_graph_process_device_batch, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). The graph's intermediate planes live in a library
 * scratch array of 128 MiB per GPU (held until the process ends) instead of
 * a caller tmp; a call whose zimg_filter_graph_get_device_tmp_size(count)
 * exceeds it fails with ZIMG_ERROR_ILLEGAL_ARGUMENT. Nothing is allocated:
 * the graph's coefficient tables must already be on the current device
 * (zimg_filter_graph_build and _set_resize_precision upload them to the
 * device current at the time), else ZIMG_ERROR_UNSUPPORTED_OPERATION. All
 * calls on one GPU share its scratch: calls that overlap in time on one GPU
 * (other streams, other threads) race and must be serialised by the caller. */
ZIMG_VISIBILITY
zimg_error_code_e zimg_filter_graph_process_device_batch_noalloc(const zimg_filter_graph *ptr, unsigned count,
                                                                 const zimg_image_buffer_const *src, const ptrdiff_t src_frame_stride[4],
                                                                 const zimg_image_buffer *dst, const ptrdiff_t dst_frame_stride[4],
                                                                 void *stream);

/* device extension: packed layouts in HBM, the device counterpart of the
 * reference's pack / unpack callbacks (zimg.h:455-484, the YUY2 and planar
 * layouts of src/testapp/frame.cpp:30-100). A buffer is planar
 * (ZIMG_PACKING_PLANAR: data[p] as zimg_image_buffer planes), semi-planar
 * 4:2:0 (ZIMG_PACKING_NV12: data[0] luma, data[1] U and V samples
 * alternating; BYTE is NV12, WORD with shift 6 is P010, shift 0 is P016) or
 * packed 4:2:2 BYTE in data[0] (ZIMG_PACKING_YUY2: Y0 U Y1 V per pixel pair;
 * ZIMG_PACKING_UYVY: U Y0 V Y1). WORD samples sit in each 16-bit word
 * shifted left by `shift` bits (unpacking shifts right, dropping the low
 * bits). The graph's format must match the layout (YUV, 4:2:0 for NV12, 4:2:2
 * BYTE for YUY2 / UYVY, no alpha). Unpacking and packing run as launches
 * around the graph's own, through tmp (zimg_filter_graph_get_device_packed_tmp_size
 * bytes; a luma plane stored contiguously with shift 0 is used in place). */
#define ZIMG_PACKING_PLANAR 0
#define ZIMG_PACKING_NV12 1
#define ZIMG_PACKING_YUY2 2
#define ZIMG_PACKING_UYVY 3

typedef struct zimg_device_packed_buffer {
	unsigned version;         /* ZIMG_API_VERSION */
	int packing;              /* ZIMG_PACKING_* */
	unsigned shift;           /* WORD samples: bits below the value in each 16-bit word */
	void *data[4];            /* device pointers */
	ptrdiff_t stride[4];      /* bytes per row */
	ptrdiff_t frame_stride[4];/* bytes between frames of a batch */
} zimg_device_packed_buffer;

ZIMG_VISIBILITY
zimg_error_code_e zimg_filter_graph_get_device_packed_tmp_size(const zimg_filter_graph *ptr, unsigned count, int src_packing,
                                                               int dst_packing, size_t *out);

ZIMG_VISIBILITY
zimg_error_code_e zimg_filter_graph_process_device_packed(const zimg_filter_graph *ptr, unsigned count, const zimg_device_packed_buffer *src,
                                                          const zimg_device_packed_buffer *dst, void *tmp, void *stream);

/* device extension: zimg_filter_graph_process_device_batch with HIP events
 * recorded around each device stage on `stream`; waits for the stream, then
 * writes up to max_stages stage durations (ms) and algorithmic bytes per
 * frame (each stage's input planes read once, output planes written once),
 * and the number of stages to *stages. */
ZIMG_VISIBILITY
zimg_error_code_e zimg_filter_graph_process_device_profiled(const zimg_filter_graph *ptr, unsigned count,
                                                            const zimg_image_buffer_const *src, const ptrdiff_t src_frame_stride[4],
                                                            const zimg_image_buffer *dst, const ptrdiff_t dst_frame_stride[4],
                                                            void *tmp, void *stream, unsigned max_stages, float *stage_ms,
                                                            double *stage_bytes, unsigned *stages);

/* device extension: batched jobs. `count` independent graphs, job k running
 * graphs[k] on src[k] -> dst[k] (device buffers as above, one frame each),
 * are planned once into one persistent launch over the resize tiles of all
 * jobs whose stages are all tiled integer resizes (the others run their own
 * launch sequences first), captured into a HIP graph at creation and
 * replayed by zimg_device_jobs_run. Buffers are fixed at creation, and the
 * graphs must outlive the jobs object (the captured launches read their
 * coefficient tables). tmp holds
 * zimg_device_jobs_get_tmp_size bytes (may be NULL when that is 0).
 * ZIMG_DEVICE_JOBS_AUTOTUNE times the cost model's candidate tile shapes for
 * each distinct stage shape on the given planes at creation, in the
 * persistent kernel with its LDS limit (results are cached per process);
 * without it the cost model's first choice is taken.
 * ZIMG_DEVICE_JOBS_NO_GRAPH skips the capture (each run
 * enqueues the launches). `stream` at creation is used for autotuning. */
typedef struct zimg_device_jobs zimg_device_jobs;

#define ZIMG_DEVICE_JOBS_AUTOTUNE 1u
#define ZIMG_DEVICE_JOBS_NO_GRAPH 2u

ZIMG_VISIBILITY
zimg_error_code_e zimg_device_jobs_get_tmp_size(const zimg_filter_graph *const *graphs, unsigned count, size_t *out);

ZIMG_VISIBILITY
zimg_device_jobs *zimg_device_jobs_create(const zimg_filter_graph *const *graphs, unsigned count, const zimg_image_buffer_const *src,
                                          const zimg_image_buffer *dst, void *tmp, unsigned flags, void *stream);

ZIMG_VISIBILITY
zimg_error_code_e zimg_device_jobs_run(zimg_device_jobs *jobs, void *stream);

ZIMG_VISIBILITY
size_t zimg_device_jobs_describe(const zimg_device_jobs *jobs, char *buf, size_t n);

ZIMG_VISIBILITY
void zimg_device_jobs_free(zimg_device_jobs *jobs);

/* device extension: learned transfer-function surrogate. A 1-32-32-1 SiLU
 * MLP for one transfer function (transfer: zimg_transfer_characteristics_e;
 * ST2084, ARIB_B67 and the BT.1886 family; to_linear selects the direction),
 * trained on the current device by the library's own forward and backward
 * kernels on samples of the exact function, then validated per input bin
 * against the FP64 function: bins whose error exceeds `tolerance` (absolute,
 * or relative when `relative` is nonzero) are evaluated exactly. */
typedef struct zimg_transfer_surrogate zimg_transfer_surrogate;

ZIMG_VISIBILITY
zimg_transfer_surrogate *zimg_transfer_surrogate_train(int transfer, int to_linear, double tolerance, int relative, unsigned steps,
                                                       unsigned seed, void *stream);

/* values[k] <- post * f(values[k] * pre) for n device floats. */
ZIMG_VISIBILITY
zimg_error_code_e zimg_transfer_surrogate_apply(const zimg_transfer_surrogate *s, float *values, size_t n, float pre, float post,
                                                void *stream);

/* The same map with the exact transfer function only. */
ZIMG_VISIBILITY
zimg_error_code_e zimg_transfer_exact_apply(int transfer, int to_linear, float *values, size_t n, float pre, float post, void *stream);

/* Fraction of validated bins, largest validated error, final training loss. */
ZIMG_VISIBILITY
zimg_error_code_e zimg_transfer_surrogate_info(const zimg_transfer_surrogate *s, double *coverage, double *max_error, double *train_loss);

/* Network parameters (zimg_transfer_surrogate_param_count() floats: w1[32],
 * b1[32], W2[32][32] row-major by output, b2[32], w3[32], b3). */
ZIMG_VISIBILITY
unsigned zimg_transfer_surrogate_param_count(void);

ZIMG_VISIBILITY
zimg_error_code_e zimg_transfer_surrogate_get_params(const zimg_transfer_surrogate *s, float *params);

/* The training kernels' gradient of mean((v - t)^2) over n host samples
 * (n a multiple of 64); grad receives param_count() + 1 floats, the last one
 * the loss. The inference layers' outputs v at inputs u. For verification. */
ZIMG_VISIBILITY
zimg_error_code_e zimg_transfer_surrogate_gradients(const float *params, const float *u, const float *t, unsigned n, float *grad);

ZIMG_VISIBILITY
zimg_error_code_e zimg_transfer_surrogate_eval(const float *params, const float *u, float *v, unsigned n);

ZIMG_VISIBILITY
void zimg_transfer_surrogate_free(zimg_transfer_surrogate *s);

/* device extension: route a graph's transfer functions through learned
 * surrogates (count of them, at most one per transfer function and
 * direction; count 0 restores the exact transfers). Each colourspace
 * conversion's gamma operation whose function has a surrogate, and the HLG
 * OETF that follows the inverse OOTF (with to_linear 0 for ARIB_B67), run
 * through the surrogate kernel with its exact fallback, between colourspace
 * launches, in three FLOAT planes of the device tmp buffer (size tmp after
 * this call). The reference has no counterpart: its GammaOperationC
 * (colorspace/operation_impl.cpp:53-77) evaluates every transfer exactly.
 * The surrogates must outlive every call that uses them and live on the
 * device that processes. */
ZIMG_VISIBILITY
zimg_error_code_e zimg_filter_graph_set_transfer_surrogates(zimg_filter_graph *ptr, const zimg_transfer_surrogate *const *s, unsigned count);

/* device extension: a short text description of the lowered device stages
 * (the analogue of the reference's FilterObserver trace, graphbuilder.h:45-64).
 * Writes at most n bytes including the terminator; returns the full length. */
ZIMG_VISIBILITY
size_t zimg_filter_graph_describe(const zimg_filter_graph *ptr, char *buf, size_t n);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* ZIMG_H_ */
