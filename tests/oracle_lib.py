"""ctypes bindings for the test oracle (oracle/liboracle.so) and, where it was
built, the reference bridge (oracle/_ref/libzimgref.so).

Test infrastructure only: the product library never imports this module.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libzimgref.so")

c_float_p = ctypes.POINTER(ctypes.c_float)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_u16_p = ctypes.POINTER(ctypes.c_uint16)
c_i16_p = ctypes.POINTER(ctypes.c_int16)
c_u32_p = ctypes.POINTER(ctypes.c_uint32)
c_u8_p = ctypes.POINTER(ctypes.c_uint8)

# zimg pixel types
BYTE, WORD, HALF, FLOAT = 0, 1, 2, 3
NP_TYPE = {BYTE: np.uint8, WORD: np.uint16, HALF: np.uint16, FLOAT: np.float32}

# resampling filters (zimg_resample_filter_e numbering)
POINT, BILINEAR, BICUBIC, SPLINE16, SPLINE36, LANCZOS, SPLINE64 = 0, 1, 2, 3, 4, 5, 6


def ptr(a, ctype=ctypes.c_void_p):
    return a.ctypes.data_as(ctype)


class CspOps(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("kind", ctypes.c_int * 6),
        ("m", ctypes.c_float * 54),
        ("func", ctypes.c_int * 6),
        ("pre", ctypes.c_float * 6),
        ("post", ctypes.c_float * 6),
        ("kr", ctypes.c_float * 6),
        ("kg", ctypes.c_float * 6),
        ("kb", ctypes.c_float * 6),
        ("nb", ctypes.c_float * 6),
        ("pb", ctypes.c_float * 6),
        ("nr", ctypes.c_float * 6),
        ("pr", ctypes.c_float * 6),
        ("aux_func", ctypes.c_int * 6),
    ]


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        lib = ctypes.CDLL(ORACLE_SO)
        lib.orc_powf.restype = ctypes.c_float
        lib.orc_powf.argtypes = [ctypes.c_float, ctypes.c_float]
        lib.orc_expf.restype = ctypes.c_float
        lib.orc_expf.argtypes = [ctypes.c_float]
        lib.orc_logf.restype = ctypes.c_float
        lib.orc_logf.argtypes = [ctypes.c_float]
        lib.orc_filter_eval.restype = ctypes.c_double
        lib.orc_filter_eval.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        lib.orc_filter_support.restype = ctypes.c_uint
        lib.orc_filter_support.argtypes = [ctypes.c_int, ctypes.c_double]
        lib.orc_compute_filter.restype = ctypes.c_int
        lib.orc_compute_filter.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_uint, ctypes.c_uint,
                                           ctypes.c_double, ctypes.c_double, c_u32_p, c_float_p, c_i16_p, c_u32_p]
        for name in ("orc_resize_h_u16", "orc_resize_v_u16"):
            f = getattr(lib, name)
            f.restype = None
            f.argtypes = [c_i16_p, c_u32_p, ctypes.c_uint, ctypes.c_uint, c_u16_p, ctypes.c_ssize_t, c_u16_p,
                          ctypes.c_ssize_t, ctypes.c_uint, ctypes.c_uint]
        for name in ("orc_resize_h_f32", "orc_resize_v_f32"):
            f = getattr(lib, name)
            f.restype = None
            f.argtypes = [c_float_p, c_u32_p, ctypes.c_uint, ctypes.c_uint, c_float_p, ctypes.c_ssize_t, c_float_p,
                          ctypes.c_ssize_t, ctypes.c_uint]
        lib.orc_scale_offset.restype = None
        lib.orc_scale_offset.argtypes = [ctypes.c_int, ctypes.c_uint, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_uint, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         c_float_p, c_float_p]
        lib.orc_half_to_float.restype = ctypes.c_float
        lib.orc_half_to_float.argtypes = [ctypes.c_uint16]
        lib.orc_float_to_half.restype = ctypes.c_uint16
        lib.orc_float_to_half.argtypes = [ctypes.c_float]
        lib.orc_compute_filter64.restype = ctypes.c_int
        lib.orc_compute_filter64.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_uint, ctypes.c_uint,
                                             ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_uint32), c_float_p,
                                             c_i16_p, c_double_p, c_u32_p]
        lib.orc_convert_to_half.restype = None
        lib.orc_convert_to_half.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_ssize_t, ctypes.c_void_p,
                                            ctypes.c_ssize_t, ctypes.c_uint, ctypes.c_uint, ctypes.c_float, ctypes.c_float]
        lib.orc_convert_to_float.restype = None
        lib.orc_convert_to_float.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_ssize_t, c_float_p, ctypes.c_ssize_t,
                                             ctypes.c_uint, ctypes.c_uint, ctypes.c_float, ctypes.c_float]
        lib.orc_left_shift.restype = None
        lib.orc_left_shift.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_ssize_t, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_ssize_t, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint]
        lib.orc_dither_ordered.restype = None
        lib.orc_dither_ordered.argtypes = [ctypes.c_int, c_u8_p, ctypes.c_uint, ctypes.c_int, ctypes.c_void_p,
                                           ctypes.c_ssize_t, ctypes.c_int, ctypes.c_void_p, ctypes.c_ssize_t,
                                           ctypes.c_uint, ctypes.c_uint, ctypes.c_float, ctypes.c_float, ctypes.c_uint]
        lib.orc_dither_ed.restype = None
        lib.orc_dither_ed.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_ssize_t, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_ssize_t, ctypes.c_uint, ctypes.c_uint, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_uint]
        lib.orc_bayer_table.restype = None
        lib.orc_bayer_table.argtypes = [c_u8_p]
        lib.orc_transfer_eval.restype = ctypes.c_int
        lib.orc_transfer_eval.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, c_float_p]
        lib.orc_transfer_scales.restype = None
        lib.orc_transfer_scales.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, c_float_p, c_float_p]
        lib.orc_colorspace_build.restype = ctypes.c_int
        lib.orc_colorspace_build.argtypes = [ctypes.c_int] * 6 + [ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                                  ctypes.POINTER(CspOps), ctypes.c_char_p, ctypes.c_size_t]
        lib.orc_colorspace_build_approx.restype = ctypes.c_int
        lib.orc_colorspace_build_approx.argtypes = lib.orc_colorspace_build.argtypes
        lib.orc_colorspace_apply.restype = None
        lib.orc_colorspace_apply.argtypes = [ctypes.POINTER(CspOps), c_float_p, c_float_p, c_float_p, ctypes.c_size_t]
        lib.orc_unresize_context.restype = ctypes.c_int
        lib.orc_unresize_context.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_double, c_u32_p, c_float_p, c_u32_p, c_float_p,
                                             c_float_p, c_float_p]
        for name in ("orc_unresize_h_f32", "orc_unresize_v_f32"):
            f = getattr(lib, name)
            f.restype = None
            f.argtypes = [c_float_p, c_u32_p, ctypes.c_uint, c_float_p, c_float_p, c_float_p, ctypes.c_uint, c_float_p,
                          ctypes.c_ssize_t, c_float_p, ctypes.c_ssize_t, ctypes.c_uint]
        lib.orc_ncl_rgb_to_yuv.restype = None
        lib.orc_ncl_rgb_to_yuv.argtypes = [ctypes.c_int, c_double_p]
        lib.orc_gamut_matrix.restype = None
        lib.orc_gamut_matrix.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, c_double_p]
        _oracle = lib
    return _oracle


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_SO)
        lib.ref_set_libm.argtypes = [ctypes.c_int]
        lib.ref_powf_n.argtypes = [c_float_p, c_float_p, c_float_p, ctypes.c_size_t]
        lib.ref_expf_n.argtypes = [c_float_p, c_float_p, ctypes.c_size_t]
        lib.ref_logf_n.argtypes = [c_float_p, c_float_p, ctypes.c_size_t]
        lib.ref_filter_eval.restype = ctypes.c_double
        lib.ref_filter_eval.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        lib.ref_compute_filter.restype = ctypes.c_int
        lib.ref_compute_filter.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_uint, ctypes.c_uint,
                                           ctypes.c_double, ctypes.c_double, c_u32_p, c_float_p, c_i16_p, c_u32_p]
        lib.ref_transfer.restype = ctypes.c_float
        lib.ref_transfer.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_float]
        lib.ref_transfer_scales.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, c_float_p, c_float_p]
        lib.ref_colorspace_apply.restype = ctypes.c_int
        lib.ref_colorspace_apply.argtypes = [ctypes.c_int] * 6 + [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                                  c_float_p, c_float_p, c_float_p, ctypes.c_size_t]
        lib.ref_colorspace_apply_auto.restype = ctypes.c_int
        lib.ref_colorspace_apply_auto.argtypes = lib.ref_colorspace_apply.argtypes
        lib.ref_unresize_context.restype = ctypes.c_int
        lib.ref_unresize_context.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_double, c_u32_p, c_float_p, c_u32_p, c_float_p,
                                             c_float_p, c_float_p]
        lib.ref_ncl_rgb_to_yuv.argtypes = [ctypes.c_int, c_double_p]
        lib.ref_half_to_float.restype = ctypes.c_float
        lib.ref_half_to_float.argtypes = [ctypes.c_uint16]
        lib.ref_float_to_half.restype = ctypes.c_uint16
        lib.ref_float_to_half.argtypes = [ctypes.c_float]
        lib.ref_blue_noise.argtypes = [c_u8_p]
        lib.ref_scale_offset.argtypes = [ctypes.c_int, ctypes.c_uint, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_uint, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         c_float_p, c_float_p]
        _ref = lib
    return _ref


# ---------------------------------------------------------------------------
# numpy-level helpers over the oracle


def compute_filter(lib, prefix, ftype, a, b, src, dst, shift, width):
    fn = getattr(lib, prefix + "compute_filter")
    fw = ctypes.c_uint32(0)
    rc = fn(ftype, a, b, src, dst, shift, width, ctypes.byref(fw), None, None, None)
    assert rc == 0
    fw = fw.value
    data = np.zeros(dst * fw, np.float32)
    data_i16 = np.zeros(dst * fw, np.int16)
    left = np.zeros(dst, np.uint32)
    rc = fn(ftype, a, b, src, dst, shift, width, ctypes.byref(ctypes.c_uint32(0)),
            ptr(data, c_float_p), ptr(data_i16, c_i16_p), ptr(left, c_u32_p))
    assert rc == 0
    return fw, data.reshape(dst, fw), data_i16.reshape(dst, fw), left


def oracle_filter64(ftype, a, b, src, dst, shift=0.0, width=None):
    """The unrounded FP64 resampling matrix (filter.cpp:249-306) as a dense
    (dst x src) float64 array."""
    lib = oracle()
    width = float(src if width is None else width)
    fw = ctypes.c_uint32(0)
    assert lib.orc_compute_filter64(ftype, a, b, src, dst, shift, width, ctypes.byref(fw), None, None, None, None) == 0
    fw = fw.value
    data = np.zeros(dst * fw, np.float32)
    data_i16 = np.zeros(dst * fw, np.int16)
    data64 = np.zeros(dst * fw, np.float64)
    left = np.zeros(dst, np.uint32)
    assert lib.orc_compute_filter64(ftype, a, b, src, dst, shift, width, ctypes.byref(ctypes.c_uint32(0)), ptr(data, c_float_p),
                                    ptr(data_i16, c_i16_p), ptr(data64, c_double_p), ptr(left, c_u32_p)) == 0
    m = np.zeros((dst, src), np.float64)
    rows = np.repeat(np.arange(dst), fw)
    cols = (left[:, None].astype(np.int64) + np.arange(fw)[None, :]).ravel()
    np.add.at(m, (rows, cols), data64)
    return m


def resize_plane_f64(img, ftype, a, b, dst_w, dst_h):
    """FP64 oracle of a resize: the FP64 matrices applied in FP64 (centre siting, no shift)."""
    src_h, src_w = img.shape
    x = np.asarray(img, np.float64)
    mh = oracle_filter64(ftype, a, b, src_w, dst_w) if dst_w != src_w else None
    mv = oracle_filter64(ftype, a, b, src_h, dst_h) if dst_h != src_h else None
    if mh is not None:
        x = x @ mh.T
    if mv is not None:
        x = mv @ x
    return x


def oracle_filter(ftype, a, b, src, dst, shift=0.0, width=None):
    return compute_filter(oracle(), "orc_", ftype, a, b, src, dst, shift, float(src if width is None else width))


def ref_filter(ftype, a, b, src, dst, shift=0.0, width=None):
    return compute_filter(ref(), "ref_", ftype, a, b, src, dst, shift, float(src if width is None else width))


def resize_h(img, filt, depth=16):
    fw, data, data_i16, left = filt
    h = img.shape[0]
    dst_w = left.shape[0]
    img = np.ascontiguousarray(img)
    if img.dtype == np.uint16:
        out = np.zeros((h, dst_w), np.uint16)
        oracle().orc_resize_h_u16(ptr(np.ascontiguousarray(data_i16), c_i16_p), ptr(left, c_u32_p), fw, dst_w,
                                  ptr(img, c_u16_p), img.strides[0], ptr(out, c_u16_p), out.strides[0], h, depth)
    else:
        out = np.zeros((h, dst_w), np.float32)
        oracle().orc_resize_h_f32(ptr(np.ascontiguousarray(data), c_float_p), ptr(left, c_u32_p), fw, dst_w,
                                  ptr(img, c_float_p), img.strides[0], ptr(out, c_float_p), out.strides[0], h)
    return out


def resize_v(img, filt, depth=16):
    fw, data, data_i16, left = filt
    w = img.shape[1]
    dst_h = left.shape[0]
    img = np.ascontiguousarray(img)
    if img.dtype == np.uint16:
        out = np.zeros((dst_h, w), np.uint16)
        oracle().orc_resize_v_u16(ptr(np.ascontiguousarray(data_i16), c_i16_p), ptr(left, c_u32_p), fw, dst_h,
                                  ptr(img, c_u16_p), img.strides[0], ptr(out, c_u16_p), out.strides[0], w, depth)
    else:
        out = np.zeros((dst_h, w), np.float32)
        oracle().orc_resize_v_f32(ptr(np.ascontiguousarray(data), c_float_p), ptr(left, c_u32_p), fw, dst_h,
                                  ptr(img, c_float_p), img.strides[0], ptr(out, c_float_p), out.strides[0], w)
    return out


def resize_h_first(xscale, yscale):
    """resize.cpp:12-18"""
    h_cost = max(xscale, 1.0) * 2.0 + xscale * max(yscale, 1.0)
    v_cost = max(yscale, 1.0) + yscale * max(xscale, 1.0) * 2.0
    return h_cost < v_cost


def resize_plane(img, ftype, a, b, dst_w, dst_h, shift_w=0.0, shift_h=0.0, subwidth=None, subheight=None, depth=16):
    """ResizeConversion::create (resize.cpp:38-102) over the oracle's kernels."""
    src_h, src_w = img.shape
    subwidth = float(src_w) if subwidth is None else subwidth
    subheight = float(src_h) if subheight is None else subheight
    skip_h = src_w == dst_w and shift_w == 0 and subwidth == src_w
    skip_v = src_h == dst_h and shift_h == 0 and subheight == src_h
    if skip_h and skip_v:
        return img.copy()
    if skip_h:
        return resize_v(img, oracle_filter(ftype, a, b, src_h, dst_h, shift_h, subheight), depth)
    if skip_v:
        return resize_h(img, oracle_filter(ftype, a, b, src_w, dst_w, shift_w, subwidth), depth)
    if resize_h_first(dst_w / subwidth, dst_h / subheight):
        t = resize_h(img, oracle_filter(ftype, a, b, src_w, dst_w, shift_w, subwidth), depth)
        return resize_v(t, oracle_filter(ftype, a, b, src_h, dst_h, shift_h, subheight), depth)
    t = resize_v(img, oracle_filter(ftype, a, b, src_h, dst_h, shift_h, subheight), depth)
    return resize_h(t, oracle_filter(ftype, a, b, src_w, dst_w, shift_w, subwidth), depth)


def scale_offset(fin, fout):
    """fin/fout: (type, depth, full, chroma, ycgco)"""
    s = ctypes.c_float()
    o = ctypes.c_float()
    oracle().orc_scale_offset(*fin, *fout, ctypes.byref(s), ctypes.byref(o))
    return s.value, o.value


def to_float(img, type_in, scale, offset):
    img = np.ascontiguousarray(img)
    out = np.zeros(img.shape, np.float32)
    oracle().orc_convert_to_float(type_in, img.ctypes.data, img.strides[0], ptr(out, c_float_p), out.strides[0],
                                  img.shape[1], img.shape[0], scale, offset)
    return out


def to_half(img, type_in, scale=1.0, offset=0.0):
    """HALF storage bits (uint16) of an integer or FLOAT plane."""
    img = np.ascontiguousarray(img)
    out = np.zeros(img.shape, np.uint16)
    oracle().orc_convert_to_half(type_in, img.ctypes.data, img.strides[0], out.ctypes.data, out.strides[0],
                                 img.shape[1], img.shape[0], scale, offset)
    return out


def half_to_float(bits):
    """FLOAT values of HALF storage bits (uint16 array), as the reference widens them."""
    return to_float(np.ascontiguousarray(bits, np.uint16), HALF, 1.0, 0.0)


def blue_noise():
    """The 64x64 RANDOM-dither threshold map (fixture taken from the reference table)."""
    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "blue_noise_64x64.npy"))


def dither(img, type_in, type_out, scale, offset, bits, dither_type=0, plane=0, blue_noise=None):
    img = np.ascontiguousarray(img)
    out = np.zeros(img.shape, NP_TYPE[type_out])
    bn = None if blue_noise is None else ptr(np.ascontiguousarray(blue_noise, np.uint8), c_u8_p)
    if dither_type == 3:
        oracle().orc_dither_ed(type_in, img.ctypes.data, img.strides[0], type_out, out.ctypes.data, out.strides[0],
                               img.shape[1], img.shape[0], scale, offset, bits)
    else:
        oracle().orc_dither_ordered(dither_type, bn, plane, type_in, img.ctypes.data, img.strides[0], type_out,
                                    out.ctypes.data, out.strides[0], img.shape[1], img.shape[0], scale, offset, bits)
    return out


def left_shift(img, type_in, type_out, shift):
    img = np.ascontiguousarray(img)
    out = np.zeros(img.shape, NP_TYPE[type_out])
    oracle().orc_left_shift(type_in, img.ctypes.data, img.strides[0], type_out, out.ctypes.data, out.strides[0],
                            img.shape[1], img.shape[0], shift)
    return out


def colorspace_ops(csp_in, csp_out, peak=100.0, scene=0, adapt=0, approx=0):
    """Operation chain; approx=1 gives the chain of the reference's AVX2 path with
    allow_approximate_gamma (FMA matrices, lookup-table gamma)."""
    ops = CspOps()
    buf = ctypes.create_string_buffer(512)
    build = oracle().orc_colorspace_build_approx if approx else oracle().orc_colorspace_build
    n = build(*csp_in, *csp_out, peak, scene, adapt, ctypes.byref(ops), buf, 512)
    return n, ops, buf.value.decode()


def colorspace_apply(ops, p0, p1, p2):
    p0 = np.ascontiguousarray(p0, np.float32).copy()
    p1 = np.ascontiguousarray(p1, np.float32).copy()
    p2 = np.ascontiguousarray(p2, np.float32).copy()
    oracle().orc_colorspace_apply(ctypes.byref(ops), ptr(p0, c_float_p), ptr(p1, c_float_p), ptr(p2, c_float_p), p0.size)
    return p0, p1, p2


def _aligned_copy(a, align=64):
    buf = np.empty(a.size + align // a.itemsize, a.dtype)
    off = (-buf.ctypes.data % align) // a.itemsize
    out = buf[off:off + a.size].reshape(a.shape)
    out[...] = a
    return out


def ref_colorspace_apply(csp_in, csp_out, p0, p1, p2, peak=100.0, scene=0, approx=0, adapt=0, cpu_auto=False):
    """The compiled reference's chain: its C path, or with cpu_auto its x86
    dispatch (AVX2 operations; -2 when the host lacks AVX2/FMA/F16C)."""
    # rows 64-byte aligned, as the reference's executor provides them (its SIMD loads require it)
    p0, p1, p2 = (_aligned_copy(np.ascontiguousarray(p, np.float32)) for p in (p0, p1, p2))
    fn = ref().ref_colorspace_apply_auto if cpu_auto else ref().ref_colorspace_apply
    n = fn(*csp_in, *csp_out, peak, scene, approx, adapt, ptr(p0, c_float_p), ptr(p1, c_float_p), ptr(p2, c_float_p), p0.size)
    return n, p0, p1, p2


# ---------------------------------------------------------------------------
# unresize


def unresize_context(up_dim, orig_dim, shift=0.0, use_ref=False):
    """(row_size, coef[orig, row_size], offset, lu_c, lu_l, lu_u) of unresizing
    up_dim samples to orig_dim (None when the reference refuses it)."""
    fn = ref().ref_unresize_context if use_ref else oracle().orc_unresize_context
    rs = ctypes.c_uint32(0)
    if fn(up_dim, orig_dim, shift, ctypes.byref(rs), None, None, None, None, None) != 0:
        return None
    rs = rs.value
    coef = np.zeros(orig_dim * rs, np.float32)
    off = np.zeros(orig_dim, np.uint32)
    c, l, u = (np.zeros(orig_dim, np.float32) for _ in range(3))
    assert fn(up_dim, orig_dim, shift, ctypes.byref(ctypes.c_uint32(0)), ptr(coef, c_float_p), ptr(off, c_u32_p), ptr(c, c_float_p),
              ptr(l, c_float_p), ptr(u, c_float_p)) == 0
    return rs, coef.reshape(orig_dim, rs), off, c, l, u


def unresize_h(img, ctx):
    rs, coef, off, c, l, u = ctx
    img = np.ascontiguousarray(img, np.float32)
    out = np.zeros((img.shape[0], off.shape[0]), np.float32)
    oracle().orc_unresize_h_f32(ptr(np.ascontiguousarray(coef), c_float_p), ptr(off, c_u32_p), rs, ptr(c, c_float_p), ptr(l, c_float_p),
                                ptr(u, c_float_p), off.shape[0], ptr(img, c_float_p), img.strides[0], ptr(out, c_float_p),
                                out.strides[0], img.shape[0])
    return out


def unresize_v(img, ctx):
    rs, coef, off, c, l, u = ctx
    img = np.ascontiguousarray(img, np.float32)
    out = np.zeros((off.shape[0], img.shape[1]), np.float32)
    oracle().orc_unresize_v_f32(ptr(np.ascontiguousarray(coef), c_float_p), ptr(off, c_u32_p), rs, ptr(c, c_float_p), ptr(l, c_float_p),
                                ptr(u, c_float_p), off.shape[0], ptr(img, c_float_p), img.strides[0], ptr(out, c_float_p),
                                out.strides[0], img.shape[1])
    return out


def unresize_plane(img, dst_w, dst_h, shift_w=0.0, shift_h=0.0):
    """UnresizeConversion::create (unresize/unresize.cpp:36-88) over the oracle's kernels."""
    up_h, up_w = img.shape
    skip_h = up_w == dst_w and shift_w == 0
    skip_v = up_h == dst_h and shift_h == 0
    if skip_h and skip_v:
        return np.array(img, np.float32)
    if skip_h:
        return unresize_v(img, unresize_context(up_h, dst_h, shift_h))
    if skip_v:
        return unresize_h(img, unresize_context(up_w, dst_w, shift_w))
    xs, ys = dst_w / up_w, dst_h / up_h
    if max(xs, 1.0) * 2.0 + xs * max(ys, 1.0) < max(ys, 1.0) + ys * max(xs, 1.0) * 2.0:
        t = unresize_h(img, unresize_context(up_w, dst_w, shift_w))
        return unresize_v(t, unresize_context(up_h, dst_h, shift_h))
    t = unresize_v(img, unresize_context(up_h, dst_h, shift_h))
    return unresize_h(t, unresize_context(up_w, dst_w, shift_w))
