"""Python binding of the MI355X zimg C ABI (include/zimg.h) over ctypes.

It mirrors the reference's C++ wrapper (src/zimg/api/zimg++.hpp:143-252):
``ImageFormat`` / ``GraphBuilderParams`` default themselves through the C
library, ``FilterGraph`` owns a graph handle. The device entry points take raw
device pointers, so any allocator works (the tests and bench use torch tensors).

The library is the built ``zimg_amd/libzimg_amd.so``; importing this module
raises if it is missing — there is no fallback implementation.
"""
import ctypes
import math
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libzimg_amd.so")

API_VERSION = (2 << 8) | 4
BUFFER_MAX = 0xFFFFFFFF

# enumerations (zimg.h)
PIXEL_BYTE, PIXEL_WORD, PIXEL_HALF, PIXEL_FLOAT = 0, 1, 2, 3
RANGE_LIMITED, RANGE_FULL = 0, 1
COLOR_GREY, COLOR_RGB, COLOR_YUV = 0, 1, 2
ALPHA_NONE, ALPHA_STRAIGHT, ALPHA_PREMULTIPLIED = 0, 1, 2
FIELD_PROGRESSIVE, FIELD_TOP, FIELD_BOTTOM = 0, 1, 2
CHROMA_LEFT, CHROMA_CENTER, CHROMA_TOP_LEFT, CHROMA_TOP, CHROMA_BOTTOM_LEFT, CHROMA_BOTTOM = range(6)
MATRIX_RGB, MATRIX_BT709, MATRIX_UNSPECIFIED = 0, 1, 2
MATRIX_FCC, MATRIX_BT470_BG, MATRIX_ST170_M, MATRIX_ST240_M, MATRIX_YCGCO = 4, 5, 6, 7, 8
MATRIX_BT2020_NCL, MATRIX_BT2020_CL = 9, 10
MATRIX_CHROMATICITY_DERIVED_NCL, MATRIX_CHROMATICITY_DERIVED_CL, MATRIX_BT2100_ICTCP = 12, 13, 14
TRANSFER_BT709, TRANSFER_UNSPECIFIED, TRANSFER_BT470_M, TRANSFER_BT470_BG = 1, 2, 4, 5
TRANSFER_BT601, TRANSFER_ST240_M, TRANSFER_LINEAR, TRANSFER_LOG_100, TRANSFER_LOG_316 = 6, 7, 8, 9, 10
TRANSFER_IEC_61966_2_4, TRANSFER_BT1361, TRANSFER_IEC_61966_2_1 = 11, 12, 13
TRANSFER_BT2020_10, TRANSFER_BT2020_12, TRANSFER_ST2084, TRANSFER_ST428, TRANSFER_ARIB_B67 = 14, 15, 16, 17, 18
PRIMARIES_BT709, PRIMARIES_UNSPECIFIED, PRIMARIES_BT470_M, PRIMARIES_BT470_BG = 1, 2, 4, 5
PRIMARIES_ST170_M, PRIMARIES_ST240_M, PRIMARIES_FILM, PRIMARIES_BT2020 = 6, 7, 8, 9
PRIMARIES_ST428, PRIMARIES_ST431_2, PRIMARIES_ST432_1, PRIMARIES_EBU3213_E = 10, 11, 12, 22
DITHER_NONE, DITHER_ORDERED, DITHER_RANDOM, DITHER_ERROR_DIFFUSION = 0, 1, 2, 3
RESIZE_POINT, RESIZE_BILINEAR, RESIZE_BICUBIC, RESIZE_SPLINE16, RESIZE_SPLINE36, RESIZE_LANCZOS, RESIZE_SPLINE64 = 0, 1, 2, 3, 4, 5, 6
CPU_NONE, CPU_AUTO, CPU_AUTO_64B = 0, 1, 2

ERROR_SUCCESS = 0
ERROR_UNKNOWN = -1
ERROR_OUT_OF_MEMORY = 1
ERROR_USER_CALLBACK_FAILED = 2
ERROR_LOGIC = 1 << 10
ERROR_GREYSCALE_SUBSAMPLING = ERROR_LOGIC + 1
ERROR_COLOR_FAMILY_MISMATCH = ERROR_LOGIC + 2
ERROR_IMAGE_NOT_DIVISIBLE = ERROR_LOGIC + 3
ERROR_BIT_DEPTH_OVERFLOW = ERROR_LOGIC + 4
ERROR_ILLEGAL_ARGUMENT = 2 << 10
ERROR_ENUM_OUT_OF_RANGE = ERROR_ILLEGAL_ARGUMENT + 1
ERROR_INVALID_IMAGE_SIZE = ERROR_ILLEGAL_ARGUMENT + 2
ERROR_UNSUPPORTED_OPERATION = 3 << 10
ERROR_UNSUPPORTED_SUBSAMPLING = ERROR_UNSUPPORTED_OPERATION + 1
ERROR_NO_COLORSPACE_CONVERSION = ERROR_UNSUPPORTED_OPERATION + 2
ERROR_NO_FIELD_PARITY_CONVERSION = ERROR_UNSUPPORTED_OPERATION + 3
ERROR_RESAMPLING_NOT_AVAILABLE = ERROR_UNSUPPORTED_OPERATION + 4

# every symbol include/zimg.h declares
EXPORTS = [
    "zimg_get_version_info", "zimg_get_api_version", "zimg_get_last_error", "zimg_clear_last_error",
    "zimg_select_buffer_mask", "zimg_filter_graph_free", "zimg_filter_graph_get_tmp_size",
    "zimg_filter_graph_get_input_buffering", "zimg_filter_graph_get_output_buffering", "zimg_filter_graph_process",
    "zimg_image_format_default", "zimg_graph_builder_params_default", "zimg_filter_graph_build",
    "zimg_filter_graph_get_device_tmp_size", "zimg_filter_graph_process_device",
    "zimg_filter_graph_process_device_batch", "zimg_filter_graph_process_device_batch_noalloc", "zimg_filter_graph_describe",
    "zimg_filter_graph_autotune_device",
    "zimg_filter_graph_get_device_packed_tmp_size", "zimg_filter_graph_process_device_packed",
    "zimg_filter_graph_set_resize_precision", "zimg_filter_graph_set_transfer_surrogates",
    "zimg_filter_graph_process_device_profiled",
    "zimg_device_jobs_get_tmp_size", "zimg_device_jobs_create", "zimg_device_jobs_run", "zimg_device_jobs_describe",
    "zimg_device_jobs_free",
    "zimg_transfer_surrogate_train", "zimg_transfer_surrogate_apply", "zimg_transfer_exact_apply", "zimg_transfer_surrogate_info",
    "zimg_transfer_surrogate_param_count", "zimg_transfer_surrogate_get_params", "zimg_transfer_surrogate_gradients",
    "zimg_transfer_surrogate_eval", "zimg_transfer_surrogate_free",
]

DEVICE_JOBS_AUTOTUNE, DEVICE_JOBS_NO_GRAPH = 1, 2

RESIZE_PRECISION_EXACT, RESIZE_PRECISION_MFMA_FP16, RESIZE_PRECISION_MFMA_BF16 = 0, 1, 2


class _Region(ctypes.Structure):
    _fields_ = [("left", ctypes.c_double), ("top", ctypes.c_double), ("width", ctypes.c_double), ("height", ctypes.c_double)]


class ImageFormatStruct(ctypes.Structure):
    """zimg_image_format (zimg.h:558-588)."""
    _fields_ = [
        ("version", ctypes.c_uint), ("width", ctypes.c_uint), ("height", ctypes.c_uint), ("pixel_type", ctypes.c_int),
        ("subsample_w", ctypes.c_uint), ("subsample_h", ctypes.c_uint), ("color_family", ctypes.c_int),
        ("matrix_coefficients", ctypes.c_int), ("transfer_characteristics", ctypes.c_int), ("color_primaries", ctypes.c_int),
        ("depth", ctypes.c_uint), ("pixel_range", ctypes.c_int), ("field_parity", ctypes.c_int),
        ("chroma_location", ctypes.c_int), ("active_region", _Region), ("alpha", ctypes.c_int),
    ]


class GraphBuilderParamsStruct(ctypes.Structure):
    """zimg_graph_builder_params (zimg.h:593-647)."""
    _fields_ = [
        ("version", ctypes.c_uint), ("resample_filter", ctypes.c_int), ("filter_param_a", ctypes.c_double),
        ("filter_param_b", ctypes.c_double), ("resample_filter_uv", ctypes.c_int), ("filter_param_a_uv", ctypes.c_double),
        ("filter_param_b_uv", ctypes.c_double), ("dither_type", ctypes.c_int), ("cpu_type", ctypes.c_int),
        ("nominal_peak_luminance", ctypes.c_double), ("allow_approximate_gamma", ctypes.c_char),
        ("scene_referred", ctypes.c_char), ("chromatic_adaptation", ctypes.c_char),
    ]


class _Plane(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("stride", ctypes.c_ssize_t), ("mask", ctypes.c_uint)]


class ImageBufferStruct(ctypes.Structure):
    """zimg_image_buffer / zimg_image_buffer_const (zimg.h:409-430)."""
    _fields_ = [("version", ctypes.c_uint), ("plane", _Plane * 4)]


PACKING_PLANAR, PACKING_NV12, PACKING_YUY2, PACKING_UYVY = 0, 1, 2, 3


class PackedBufferStruct(ctypes.Structure):
    """zimg_device_packed_buffer (include/zimg.h)."""
    _fields_ = [("version", ctypes.c_uint), ("packing", ctypes.c_int), ("shift", ctypes.c_uint),
                ("data", ctypes.c_void_p * 4), ("stride", ctypes.c_ssize_t * 4), ("frame_stride", ctypes.c_ssize_t * 4)]


def packed_buffer(packing, planes, shift=0, version=None):
    """planes: up to 4 (device ptr, row stride, frame stride) or None."""
    b = PackedBufferStruct()
    b.version = API_VERSION if version is None else version
    b.packing = packing
    b.shift = shift
    for i, pl in enumerate(planes):
        if pl is None:
            continue
        b.data[i] = pl[0]
        b.stride[i] = pl[1]
        b.frame_stride[i] = pl[2] if len(pl) > 2 else 0
    return b


CALLBACK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint)

_lib = None


def lib():
    """Load the built library; raise if it is absent (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run `make -C zimg_amd` (or __graft_entry__.build())")
        # One HIP runtime per process: when torch (which bundles its own
        # libamdhip64.so.7) provides the device buffers, load it first so the
        # library binds to the same runtime instance by soname.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        L.zimg_get_version_info.argtypes = [ctypes.POINTER(ctypes.c_uint)] * 3
        L.zimg_get_api_version.restype = ctypes.c_uint
        L.zimg_get_api_version.argtypes = [ctypes.POINTER(ctypes.c_uint)] * 2
        L.zimg_get_last_error.restype = ctypes.c_int
        L.zimg_get_last_error.argtypes = [ctypes.c_char_p, sz]
        L.zimg_select_buffer_mask.restype = ctypes.c_uint
        L.zimg_select_buffer_mask.argtypes = [ctypes.c_uint]
        L.zimg_filter_graph_free.argtypes = [vp]
        L.zimg_filter_graph_get_tmp_size.restype = ctypes.c_int
        L.zimg_filter_graph_get_tmp_size.argtypes = [vp, ctypes.POINTER(sz)]
        L.zimg_filter_graph_get_input_buffering.restype = ctypes.c_int
        L.zimg_filter_graph_get_input_buffering.argtypes = [vp, ctypes.POINTER(ctypes.c_uint)]
        L.zimg_filter_graph_get_output_buffering.restype = ctypes.c_int
        L.zimg_filter_graph_get_output_buffering.argtypes = [vp, ctypes.POINTER(ctypes.c_uint)]
        L.zimg_filter_graph_process.restype = ctypes.c_int
        L.zimg_filter_graph_process.argtypes = [vp, ctypes.POINTER(ImageBufferStruct), ctypes.POINTER(ImageBufferStruct), vp,
                                                CALLBACK, vp, CALLBACK, vp]
        L.zimg_image_format_default.argtypes = [ctypes.POINTER(ImageFormatStruct), ctypes.c_uint]
        L.zimg_graph_builder_params_default.argtypes = [ctypes.POINTER(GraphBuilderParamsStruct), ctypes.c_uint]
        L.zimg_filter_graph_build.restype = vp
        L.zimg_filter_graph_build.argtypes = [ctypes.POINTER(ImageFormatStruct), ctypes.POINTER(ImageFormatStruct),
                                              ctypes.POINTER(GraphBuilderParamsStruct)]
        L.zimg_filter_graph_get_device_tmp_size.restype = ctypes.c_int
        L.zimg_filter_graph_get_device_tmp_size.argtypes = [vp, ctypes.c_uint, ctypes.POINTER(sz)]
        L.zimg_filter_graph_process_device.restype = ctypes.c_int
        L.zimg_filter_graph_process_device.argtypes = [vp, ctypes.POINTER(ImageBufferStruct), ctypes.POINTER(ImageBufferStruct), vp, vp]
        L.zimg_filter_graph_process_device_batch.restype = ctypes.c_int
        L.zimg_filter_graph_process_device_batch.argtypes = [vp, ctypes.c_uint, ctypes.POINTER(ImageBufferStruct),
                                                             ctypes.POINTER(ctypes.c_ssize_t), ctypes.POINTER(ImageBufferStruct),
                                                             ctypes.POINTER(ctypes.c_ssize_t), vp, vp]
        L.zimg_filter_graph_autotune_device.restype = ctypes.c_int
        L.zimg_filter_graph_autotune_device.argtypes = [vp, ctypes.c_uint, ctypes.POINTER(ImageBufferStruct),
                                                        ctypes.POINTER(ctypes.c_ssize_t), ctypes.POINTER(ImageBufferStruct),
                                                        ctypes.POINTER(ctypes.c_ssize_t), vp, vp, ctypes.POINTER(ctypes.c_uint)]
        L.zimg_filter_graph_process_device_batch_noalloc.restype = ctypes.c_int
        L.zimg_filter_graph_process_device_batch_noalloc.argtypes = [vp, ctypes.c_uint, ctypes.POINTER(ImageBufferStruct),
                                                                     ctypes.POINTER(ctypes.c_ssize_t), ctypes.POINTER(ImageBufferStruct),
                                                                     ctypes.POINTER(ctypes.c_ssize_t), vp]
        L.zimg_filter_graph_get_device_packed_tmp_size.restype = ctypes.c_int
        L.zimg_filter_graph_get_device_packed_tmp_size.argtypes = [vp, ctypes.c_uint, ctypes.c_int, ctypes.c_int, ctypes.POINTER(sz)]
        L.zimg_filter_graph_process_device_packed.restype = ctypes.c_int
        L.zimg_filter_graph_process_device_packed.argtypes = [vp, ctypes.c_uint, ctypes.POINTER(PackedBufferStruct),
                                                              ctypes.POINTER(PackedBufferStruct), vp, vp]
        L.zimg_filter_graph_describe.restype = sz
        L.zimg_filter_graph_describe.argtypes = [vp, ctypes.c_char_p, sz]
        L.zimg_filter_graph_process_device_profiled.restype = ctypes.c_int
        L.zimg_filter_graph_process_device_profiled.argtypes = [vp, ctypes.c_uint, ctypes.POINTER(ImageBufferStruct),
                                                                ctypes.POINTER(ctypes.c_ssize_t), ctypes.POINTER(ImageBufferStruct),
                                                                ctypes.POINTER(ctypes.c_ssize_t), vp, vp, ctypes.c_uint,
                                                                ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double),
                                                                ctypes.POINTER(ctypes.c_uint)]
        L.zimg_filter_graph_set_resize_precision.restype = ctypes.c_int
        L.zimg_filter_graph_set_resize_precision.argtypes = [vp, ctypes.c_int]
        L.zimg_filter_graph_set_transfer_surrogates.restype = ctypes.c_int
        L.zimg_filter_graph_set_transfer_surrogates.argtypes = [vp, ctypes.POINTER(vp), ctypes.c_uint]
        L.zimg_device_jobs_get_tmp_size.restype = ctypes.c_int
        L.zimg_device_jobs_get_tmp_size.argtypes = [ctypes.POINTER(vp), ctypes.c_uint, ctypes.POINTER(sz)]
        L.zimg_device_jobs_create.restype = vp
        L.zimg_device_jobs_create.argtypes = [ctypes.POINTER(vp), ctypes.c_uint, ctypes.POINTER(ImageBufferStruct),
                                              ctypes.POINTER(ImageBufferStruct), vp, ctypes.c_uint, vp]
        L.zimg_device_jobs_run.restype = ctypes.c_int
        L.zimg_device_jobs_run.argtypes = [vp, vp]
        L.zimg_device_jobs_describe.restype = sz
        L.zimg_device_jobs_describe.argtypes = [vp, ctypes.c_char_p, sz]
        L.zimg_device_jobs_free.restype = None
        L.zimg_device_jobs_free.argtypes = [vp]
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        L.zimg_transfer_surrogate_train.restype = vp
        L.zimg_transfer_surrogate_train.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_uint, ctypes.c_uint, vp]
        L.zimg_transfer_surrogate_apply.restype = ctypes.c_int
        L.zimg_transfer_surrogate_apply.argtypes = [vp, vp, sz, ctypes.c_float, ctypes.c_float, vp]
        L.zimg_transfer_exact_apply.restype = ctypes.c_int
        L.zimg_transfer_exact_apply.argtypes = [ctypes.c_int, ctypes.c_int, vp, sz, ctypes.c_float, ctypes.c_float, vp]
        L.zimg_transfer_surrogate_info.restype = ctypes.c_int
        L.zimg_transfer_surrogate_info.argtypes = [vp, dp, dp, dp]
        L.zimg_transfer_surrogate_param_count.restype = ctypes.c_uint
        L.zimg_transfer_surrogate_param_count.argtypes = []
        L.zimg_transfer_surrogate_get_params.restype = ctypes.c_int
        L.zimg_transfer_surrogate_get_params.argtypes = [vp, fp]
        L.zimg_transfer_surrogate_gradients.restype = ctypes.c_int
        L.zimg_transfer_surrogate_gradients.argtypes = [fp, fp, fp, ctypes.c_uint, fp]
        L.zimg_transfer_surrogate_eval.restype = ctypes.c_int
        L.zimg_transfer_surrogate_eval.argtypes = [fp, fp, fp, ctypes.c_uint]
        L.zimg_transfer_surrogate_free.restype = None
        L.zimg_transfer_surrogate_free.argtypes = [vp]
        _lib = L
    return _lib


class ZimgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"zimg error {code}: {msg}")
        self.code = code


def last_error():
    buf = ctypes.create_string_buffer(1024)
    code = lib().zimg_get_last_error(buf, 1024)
    return code, buf.value.decode(errors="replace")


def check(code):
    if code != ERROR_SUCCESS:
        c, msg = last_error()
        raise ZimgError(code, msg)


def image_format(version=API_VERSION, **kw):
    f = ImageFormatStruct()
    lib().zimg_image_format_default(ctypes.byref(f), version)
    for k, v in kw.items():
        if k in ("active_left", "active_top", "active_width", "active_height"):
            setattr(f.active_region, k[len("active_"):], v)
        else:
            setattr(f, k, v)
    return f


def builder_params(version=API_VERSION, **kw):
    p = GraphBuilderParamsStruct()
    lib().zimg_graph_builder_params_default(ctypes.byref(p), version)
    for k, v in kw.items():
        if k in ("allow_approximate_gamma", "scene_referred", "chromatic_adaptation"):
            v = bytes([int(v)])
        setattr(p, k, v)
    return p


def image_buffer(planes, version=API_VERSION):
    """planes: list of up to 4 (ptr, stride_bytes[, mask]) or None."""
    b = ImageBufferStruct()
    b.version = version
    for i, pl in enumerate(planes):
        if pl is None:
            continue
        b.plane[i].data = pl[0]
        b.plane[i].stride = pl[1]
        b.plane[i].mask = pl[2] if len(pl) > 2 else BUFFER_MAX
    return b


class FilterGraph:
    """Owns a zimg_filter_graph handle (zimg++.hpp:143-252)."""

    def __init__(self, src_format, dst_format, params=None):
        L = lib()
        self._h = L.zimg_filter_graph_build(ctypes.byref(src_format), ctypes.byref(dst_format),
                                            ctypes.byref(params) if params is not None else None)
        if not self._h:
            code, msg = last_error()
            raise ZimgError(code, msg)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.zimg_filter_graph_free(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def tmp_size(self):
        out = ctypes.c_size_t()
        check(lib().zimg_filter_graph_get_tmp_size(self._h, ctypes.byref(out)))
        return out.value

    def device_tmp_size(self, count=1):
        out = ctypes.c_size_t()
        check(lib().zimg_filter_graph_get_device_tmp_size(self._h, count, ctypes.byref(out)))
        return out.value

    def input_buffering(self):
        out = ctypes.c_uint()
        check(lib().zimg_filter_graph_get_input_buffering(self._h, ctypes.byref(out)))
        return out.value

    def output_buffering(self):
        out = ctypes.c_uint()
        check(lib().zimg_filter_graph_get_output_buffering(self._h, ctypes.byref(out)))
        return out.value

    def process_device_profiled(self, count, src, src_fs, dst, dst_fs, tmp=None, stream=None):
        """process_device_batch with per-stage HIP events: [(ms, bytes per frame)] per device stage."""
        sfs = (ctypes.c_ssize_t * 4)(*src_fs)
        dfs = (ctypes.c_ssize_t * 4)(*dst_fs)
        ms = (ctypes.c_float * 64)()
        by = (ctypes.c_double * 64)()
        n = ctypes.c_uint()
        check(lib().zimg_filter_graph_process_device_profiled(self._h, count, ctypes.byref(src), sfs, ctypes.byref(dst), dfs, tmp, stream,
                                                              64, ms, by, ctypes.byref(n)))
        return [(ms[k], by[k]) for k in range(min(n.value, 64))]

    def device_packed_tmp_size(self, count, src_packing, dst_packing):
        out = ctypes.c_size_t()
        check(lib().zimg_filter_graph_get_device_packed_tmp_size(self._h, count, src_packing, dst_packing, ctypes.byref(out)))
        return out.value

    def process_device_packed(self, count, src, dst, tmp=0, stream=0):
        check(lib().zimg_filter_graph_process_device_packed(self._h, count, ctypes.byref(src), ctypes.byref(dst), tmp, stream))

    def set_resize_precision(self, mode):
        check(lib().zimg_filter_graph_set_resize_precision(self._h, mode))

    def set_transfer_surrogates(self, surrogates):
        """zimg_filter_graph_set_transfer_surrogates; the graph keeps the
        TransferSurrogate objects alive while it uses them ([] restores the
        exact transfers)."""
        surrogates = list(surrogates)
        arr = (ctypes.c_void_p * max(len(surrogates), 1))(*[s._h for s in surrogates])
        check(lib().zimg_filter_graph_set_transfer_surrogates(self._h, arr, len(surrogates)))
        self._surrogates = surrogates

    def describe(self):
        n = lib().zimg_filter_graph_describe(self._h, None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        lib().zimg_filter_graph_describe(self._h, buf, n + 1)
        return buf.value.decode()

    def device_stages(self):
        """One line per device launch of the lowered graph (indented detail
        lines and the host staging summary left out)."""
        dev = self.describe().split("device:\n", 1)[1]
        return [l for l in dev.splitlines() if l and not l[0].isspace() and not l.startswith("host staging sets:")]

    def host_staging_sets(self):
        """Staging sets the host-buffer path (process) has created so far."""
        for line in self.describe().splitlines():
            if line.startswith("host staging sets:"):
                return int(line.split(":")[1])
        return 0

    def plan_trace(self):
        text = self.describe()
        plan = text.split("device:\n")[0].split("plan:\n", 1)[1]
        return [line for line in plan.splitlines() if line]

    def process(self, src, dst, tmp=None, unpack_cb=None, unpack_user=None, pack_cb=None, pack_user=None):
        """Host buffers (zimg_filter_graph_process)."""
        ucb = unpack_cb if unpack_cb is not None else CALLBACK()
        pcb = pack_cb if pack_cb is not None else CALLBACK()
        check(lib().zimg_filter_graph_process(self._h, ctypes.byref(src), ctypes.byref(dst), tmp, ucb, unpack_user, pcb, pack_user))

    def process_device(self, src, dst, tmp=0, stream=0):
        check(lib().zimg_filter_graph_process_device(self._h, ctypes.byref(src), ctypes.byref(dst), tmp, stream))

    def process_device_batch(self, count, src, src_frame_stride, dst, dst_frame_stride, tmp=0, stream=0):
        ss = (ctypes.c_ssize_t * 4)(*src_frame_stride)
        ds = (ctypes.c_ssize_t * 4)(*dst_frame_stride)
        check(lib().zimg_filter_graph_process_device_batch(self._h, count, ctypes.byref(src), ss, ctypes.byref(dst), ds, tmp, stream))

    def autotune_device(self, count, src, src_frame_stride, dst, dst_frame_stride, tmp=0, stream=0):
        """zimg_filter_graph_autotune_device: time the tiled resize steps' candidate
        tile shapes on these buffers (dst is overwritten) and keep the fastest;
        returns how many steps changed shape."""
        ss = (ctypes.c_ssize_t * 4)(*src_frame_stride)
        ds = (ctypes.c_ssize_t * 4)(*dst_frame_stride)
        n = ctypes.c_uint()
        check(lib().zimg_filter_graph_autotune_device(self._h, count, ctypes.byref(src), ss, ctypes.byref(dst), ds, tmp, stream,
                                                      ctypes.byref(n)))
        return n.value


NOALLOC_SCRATCH_BYTES = 128 << 20


def process_device_batch_noalloc(graph, count, src, src_frame_stride, dst, dst_frame_stride, stream=0):
    """zimg_filter_graph_process_device_batch_noalloc: no tmp, no allocation."""
    ss = (ctypes.c_ssize_t * 4)(*src_frame_stride)
    ds = (ctypes.c_ssize_t * 4)(*dst_frame_stride)
    check(lib().zimg_filter_graph_process_device_batch_noalloc(graph.handle, count, ctypes.byref(src), ss, ctypes.byref(dst), ds, stream))


def device_jobs_tmp_size(graphs):
    arr = (ctypes.c_void_p * len(graphs))(*[g.handle for g in graphs])
    out = ctypes.c_size_t()
    check(lib().zimg_device_jobs_get_tmp_size(arr, len(graphs), ctypes.byref(out)))
    return out.value


class DeviceJobs:
    """Batched jobs (zimg_device_jobs_*): graphs[k] on srcs[k] -> dsts[k]
    (image_buffer structs of device pointers), fixed at creation."""

    def __init__(self, graphs, srcs, dsts, tmp=None, flags=0, stream=None):
        self._graphs = list(graphs)  # keep the graphs alive
        n = len(self._graphs)
        arr = (ctypes.c_void_p * n)(*[g.handle for g in self._graphs])
        self._src = (ImageBufferStruct * n)(*srcs)
        self._dst = (ImageBufferStruct * n)(*dsts)
        self._h = lib().zimg_device_jobs_create(arr, n, self._src, self._dst, tmp, flags, stream)
        if not self._h:
            code, msg = last_error()
            raise ZimgError(code, msg)

    def run(self, stream=None):
        check(lib().zimg_device_jobs_run(self._h, stream))

    def describe(self):
        n = lib().zimg_device_jobs_describe(self._h, None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        lib().zimg_device_jobs_describe(self._h, buf, n + 1)
        return buf.value.decode()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.zimg_device_jobs_free(h)
            self._h = None


def _fptr(a):
    import numpy as _np
    assert a.dtype == _np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class TransferSurrogate:
    """zimg_transfer_surrogate_*: a trained MLP for one transfer function."""

    def __init__(self, transfer, to_linear=True, tolerance=1e-4, relative=False, steps=4000, seed=1, stream=None):
        self._h = lib().zimg_transfer_surrogate_train(transfer, int(to_linear), tolerance, int(relative), steps, seed, stream)
        if not self._h:
            code, msg = last_error()
            raise ZimgError(code, msg)

    def apply(self, ptr, n, pre=1.0, post=1.0, stream=None):
        check(lib().zimg_transfer_surrogate_apply(self._h, ptr, n, pre, post, stream))

    def info(self):
        c, e, l = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(lib().zimg_transfer_surrogate_info(self._h, ctypes.byref(c), ctypes.byref(e), ctypes.byref(l)))
        return {"coverage": c.value, "max_validated_error": e.value, "train_loss": l.value}

    def params(self):
        import numpy as _np
        out = _np.zeros(surrogate_param_count(), _np.float32)
        check(lib().zimg_transfer_surrogate_get_params(self._h, _fptr(out)))
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.zimg_transfer_surrogate_free(h)
            self._h = None


def pq_to_hlg_surrogate_tolerances(bits=12, lsb=0.5):
    """Per-bin tolerances of the two surrogates of an RGB PQ -> HLG conversion
    (PQ EOTF, relative; HLG OETF, absolute) from an output error budget of
    `lsb` codes at `bits` bits, split evenly between the two and halved again
    for the values between a bin's validation points.

    The HLG stage is E' = OETF(E * Y^((1 - g) / g)), g = 1.2, Y a positive
    combination of E's channels: a relative error e in E moves the scaled
    value by at most e (1 + (g - 1) / g) = 7e/6 relative, and OETF moves by
    at most 0.25 per unit of log E (sqrt(3E) / 2 below 1/12, where it reaches
    0.25; 12aE / (12E - b) above, falling from 0.25 to 0.183 at E = 1). So a
    relative PQ error e costs at most 0.2917 e (2^bits - 1) codes, and an
    absolute OETF error d costs d (2^bits - 1) codes."""
    codes = float((1 << bits) - 1)
    share = lsb / 2.0 / 2.0
    return share / (0.25 * 7.0 / 6.0 * codes), share / codes


def transfer_exact_apply(transfer, to_linear, ptr, n, pre=1.0, post=1.0, stream=None):
    check(lib().zimg_transfer_exact_apply(transfer, int(to_linear), ptr, n, pre, post, stream))


def surrogate_param_count():
    return lib().zimg_transfer_surrogate_param_count()


def surrogate_gradients(params, u, t):
    import numpy as _np
    grad = _np.zeros(surrogate_param_count() + 1, _np.float32)
    check(lib().zimg_transfer_surrogate_gradients(_fptr(params), _fptr(u), _fptr(t), len(u), _fptr(grad)))
    return grad


def surrogate_eval(params, u):
    import numpy as _np
    v = _np.zeros(len(u), _np.float32)
    check(lib().zimg_transfer_surrogate_eval(_fptr(params), _fptr(u), _fptr(v), len(u)))
    return v
