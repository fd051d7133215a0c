"""GPU parity: the HIP path through the C ABI against the oracle, bit for bit.

Integer outputs must be identical; float outputs are compared bit for bit too
(the kernels evaluate the reference's C expressions in the same order without
contraction), which is stricter than the 1e-5 relative tolerance."""
import ctypes

import numpy as np
import pytest

import pipelines as P
import zimg_amd as z
from oracle_lib import (BICUBIC, BILINEAR, BYTE, FLOAT, LANCZOS, POINT, SPLINE16, SPLINE36, SPLINE64, WORD, blue_noise, dither,
                        left_shift, resize_plane, scale_offset)
from zimg_amd import workloads as W

pytestmark = pytest.mark.gpu

gpu_util = pytest.importorskip("gpu_util")


def _word_graph(sw, sh, dw, dh, ftype=z.RESIZE_LANCZOS, a=3.0, b=float("nan"), depth=16):
    src = z.image_format(width=sw, height=sh, pixel_type=z.PIXEL_WORD, depth=depth)
    dst = z.image_format(width=dw, height=dh, pixel_type=z.PIXEL_WORD, depth=depth)
    return z.FilterGraph(src, dst, z.builder_params(resample_filter=ftype, filter_param_a=a, filter_param_b=b))


def _assert_equal(got, want, what):
    assert got.shape == want.shape, what
    if got.dtype == np.float32:
        bad = np.count_nonzero(got.view(np.uint32) != want.view(np.uint32))
    else:
        bad = np.count_nonzero(got != want)
    if bad:
        idx = np.argwhere(got != want)[:5]
        pytest.fail(f"{what}: {bad} mismatches, first at {idx.tolist()}: got {got[tuple(idx[0])]} want {want[tuple(idx[0])]}")


def test_uhd_lanczos_word_bit_exact():
    """1920x1080 WORD -> 3840x2160 Lanczos3 (fused H+V launch) == reference C path."""
    print("case: WORD 1920x1080->3840x2160 lanczos3", flush=True)
    src = W.word_plane(1920, 1080, seed=1)
    g = _word_graph(1920, 1080, 3840, 2160)
    (out, *_) = gpu_util.run_graph(g, [src, None, None, None], [(2160, 3840, np.uint16), None, None, None])
    _assert_equal(out, P.word_resize(src, 3840, 2160), "uhd lanczos")


WORD_CASES = [
    (640, 480, 1280, 960, LANCZOS, 3.0, 0.0, 16),
    (640, 480, 320, 240, LANCZOS, 3.0, 0.0, 16),
    (641, 479, 1000, 333, BICUBIC, 0.0, 0.5, 16),
    (333, 1000, 641, 479, SPLINE36, 0.0, 0.0, 10),
    (500, 300, 500, 719, BILINEAR, 0.0, 0.0, 12),   # vertical only
    (500, 300, 1217, 300, SPLINE16, 0.0, 0.0, 16),  # horizontal only
    (96, 64, 1920, 1080, SPLINE64, 0.0, 0.0, 16),   # large upscale
    (1920, 1080, 200, 120, LANCZOS, 4.0, 0.0, 16),  # ~10x downscale, wide filters
    (64, 48, 97, 31, POINT, 0.0, 0.0, 16),
]


@pytest.mark.parametrize("sw,sh,dw,dh,ft,a,b,depth", WORD_CASES)
def test_word_resize_grid(sw, sh, dw, dh, ft, a, b, depth):
    print(f"case: WORD {sw}x{sh}->{dw}x{dh} filter {ft} depth {depth}", flush=True)
    src = (W.word_plane(sw, sh, seed=sw * 7 + dh) >> (16 - depth)).astype(np.uint16)
    g = _word_graph(sw, sh, dw, dh, ft, a, b if ft == BICUBIC else float("nan"), depth)
    for pad in (0, 100):
        (out, *_) = gpu_util.run_graph(g, [src, None, None, None], [(dh, dw, np.uint16), None, None, None], pad_bytes=pad)
        _assert_equal(out, P.word_resize(src, dw, dh, ft, a, b, depth), f"word resize pad={pad}")


IMMA_CASES = [
    # (sw, sh, dw, dh, filter, a, b, depth, frames)
    (641, 359, 1283, 719, LANCZOS, 3.0, 0.0, 10, 1),   # ragged: partial column blocks and row tiles
    (1000, 700, 1500, 1050, BICUBIC, 0.0, 0.5, 16, 2),  # 1.5x: widest windows that fit
    (500, 400, 1999, 1601, SPLINE36, 0.0, 0.0, 12, 1),
    (96, 64, 1920, 1080, SPLINE64, 0.0, 0.0, 16, 1),
    (17, 9, 40, 33, BILINEAR, 0.0, 0.0, 16, 3),         # smaller than one tile
]


@pytest.mark.parametrize("sw,sh,dw,dh,ft,a,b,depth,frames", IMMA_CASES)
def test_word_resize_integer_matrix_cores(sw, sh, dw, dh, ft, a, b, depth, frames):
    """Upscales whose 16-output windows fit run on the int8 matrix cores; bit-exact."""
    print(f"case: WORD int8 matrix cores {sw}x{sh}->{dw}x{dh} filter {ft} depth {depth} frames {frames}", flush=True)
    g = _word_graph(sw, sh, dw, dh, ft, a, b if ft == BICUBIC else float("nan"), depth)
    assert "integer matrix cores" in g.describe()
    src = np.stack([(W.word_plane(sw, sh, seed=sw + 13 * k) >> (16 - depth)).astype(np.uint16) for k in range(frames)])
    out, *_ = gpu_util.run_graph(g, [src, None, None, None], [(dh, dw, np.uint16), None, None, None], pad_bytes=32, pitch_align=16)
    out = out if frames > 1 else out[None]
    for k in range(frames):
        _assert_equal(out[k], P.word_resize(src[k], dw, dh, ft, a, b, depth), f"frame {k}")


@pytest.mark.parametrize("depth", [16, 12])
def test_integer_matrix_cores_clamp_at_both_ends(depth):
    """Full-range step edges (blocks of 0 and the largest code, single-sample
    spikes) make Lanczos3 ring past both ends of the range in the horizontal
    intermediate and in the output: 16-bit takes the saturating int16 pack,
    12-bit the clamp to [0, pixel_max]; bit-exact against the oracle."""
    sw, sh, dw, dh = 300, 170, 600, 340
    print(f"case: WORD int8 matrix cores clamp {sw}x{sh}->{dw}x{dh} depth {depth}", flush=True)
    top = (1 << depth) - 1
    yy, xx = np.mgrid[0:sh, 0:sw]
    src = np.where(((yy // 7) + (xx // 5)) % 2 == 0, top, 0).astype(np.uint16)
    rng = np.random.default_rng(depth)
    spikes = rng.random((sh, sw)) < 0.02
    src[spikes] = top - src[spikes]
    g = _word_graph(sw, sh, dw, dh, LANCZOS, 3.0, float("nan"), depth)
    assert "integer matrix cores" in g.describe()
    out, *_ = gpu_util.run_graph(g, [src, None, None, None], [(dh, dw, np.uint16), None, None, None], pitch_align=16)
    want = P.word_resize(src, dw, dh, LANCZOS, 3.0, 0.0, depth)
    assert (want == 0).any() and (want == top).any(), "the pattern reaches both clamps"
    _assert_equal(out, want, f"clamp depth {depth}")


@pytest.mark.parametrize("depth", [16, 10])
def test_tiled_word_downscale_clamps_at_both_ends(depth):
    """The same full-range step edges through a 2x Lanczos3 downscale on the
    tiled dot2 kernel (vertical pass first, row pairs): the ringing reaches 0
    and the largest code in the intermediate and the output."""
    sw, sh, dw, dh = 602, 342, 301, 171
    print(f"case: WORD tiled clamp {sw}x{sh}->{dw}x{dh} depth {depth}", flush=True)
    top = (1 << depth) - 1
    yy, xx = np.mgrid[0:sh, 0:sw]
    src = np.where(((yy // 9) + (xx // 7)) % 2 == 0, top, 0).astype(np.uint16)
    rng = np.random.default_rng(depth + 1)
    spikes = rng.random((sh, sw)) < 0.02
    src[spikes] = top - src[spikes]
    g = _word_graph(sw, sh, dw, dh, LANCZOS, 3.0, float("nan"), depth)
    assert "integer matrix cores" not in g.describe()
    out, *_ = gpu_util.run_graph(g, [src, None, None, None], [(dh, dw, np.uint16), None, None, None], pitch_align=16)
    want = P.word_resize(src, dw, dh, LANCZOS, 3.0, 0.0, depth)
    assert (want == 0).any() and (want == top).any(), "the pattern reaches both clamps"
    _assert_equal(out, want, f"clamp depth {depth}")


@pytest.mark.parametrize("sw,sh,dw,dh,frames", [(333, 200, 700, 419, 2), (16, 20, 61, 47, 1)])
def test_byte_to_word_upscale_integer_matrix_cores(sw, sh, dw, dh, frames):
    """BYTE 8-bit promoted to WORD 16-bit by the left shift folded into the
    matrix-core kernel's staging (depth_convert.cpp:23-30), then resized."""
    print(f"case: BYTE->WORD int8 matrix cores {sw}x{sh}->{dw}x{dh} frames {frames}", flush=True)
    g = z.FilterGraph(z.image_format(width=sw, height=sh, pixel_type=z.PIXEL_BYTE, depth=8),
                      z.image_format(width=dw, height=dh, pixel_type=z.PIXEL_WORD, depth=16),
                      z.builder_params(resample_filter=z.RESIZE_SPLINE36))
    assert "integer matrix cores" in g.describe()
    r = np.random.default_rng(sw)
    src = r.integers(0, 256, (frames, sh, sw)).astype(np.uint8)
    out, *_ = gpu_util.run_graph(g, [src, None, None, None], [(dh, dw, np.uint16), None, None, None], pitch_align=16)
    out = out if frames > 1 else out[None]
    for k in range(frames):
        want = P.word_resize(left_shift(src[k], BYTE, WORD, 8), dw, dh, SPLINE36, 0.0, 0.0, 16)
        _assert_equal(out[k], want, f"frame {k}")


def _alpha_fmt(w, h, alpha):
    return z.image_format(width=w, height=h, pixel_type=z.PIXEL_FLOAT, color_family=z.COLOR_RGB,
                          matrix_coefficients=z.MATRIX_RGB, transfer_characteristics=z.TRANSFER_BT709,
                          color_primaries=z.PRIMARIES_BT709, alpha=alpha)


def test_straight_alpha_resize_premultiplies_bit_exact():
    """RGB + straight alpha resized: premultiply (x * alpha, simple_filters.cpp:76-88), resize of all
    four planes, unpremultiply (x / clamp(alpha, 0, 1), 0 where alpha is 0, :98-110)."""
    print("case: FLOAT RGBA straight alpha 200x150->333x277 bicubic", flush=True)
    r = np.random.default_rng(5)
    w, h, dw, dh = 200, 150, 333, 277
    rgb = [r.random((h, w), dtype=np.float32) for _ in range(3)]
    alpha = r.random((h, w), dtype=np.float32)
    alpha[r.random((h, w)) < 0.2] = 0.0                   # transparent pixels
    alpha[:, :20] = 0.0                                    # a transparent band: resized alpha exactly 0
    alpha[r.random((h, w)) < 0.05] = np.float32(1.25)      # out of range: premultiply does not clamp
    g = z.FilterGraph(_alpha_fmt(w, h, z.ALPHA_STRAIGHT), _alpha_fmt(dw, dh, z.ALPHA_STRAIGHT))
    outs = gpu_util.run_graph(g, rgb + [alpha], [(dh, dw, np.float32)] * 4)
    prem = [(c * alpha).astype(np.float32) for c in rgb]
    res = [resize_plane(p, BICUBIC, 0.0, 0.5, dw, dh) for p in prem]
    ra = resize_plane(alpha, BICUBIC, 0.0, 0.5, dw, dh)
    ca = np.clip(ra, np.float32(0.0), np.float32(1.0))
    with np.errstate(divide="ignore", invalid="ignore"):
        want = [np.where(ca == 0, np.float32(0.0), (p / np.where(ca == 0, np.float32(1.0), ca))).astype(np.float32) for p in res]
    for k in range(3):
        _assert_equal(outs[k], want[k], f"plane {k}")
    _assert_equal(outs[3], ra, "alpha")
    assert np.count_nonzero(ca == 0) > 0 and np.count_nonzero(ra > 1) > 0


def test_premultiplied_to_straight_alpha_bit_exact():
    print("case: FLOAT RGBA premultiplied -> straight 96x64", flush=True)
    r = np.random.default_rng(6)
    w, h = 96, 64
    alpha = r.random((h, w), dtype=np.float32)
    alpha[r.random((h, w)) < 0.25] = 0.0
    rgb = [(r.random((h, w), dtype=np.float32) * alpha).astype(np.float32) for _ in range(3)]
    g = z.FilterGraph(_alpha_fmt(w, h, z.ALPHA_PREMULTIPLIED), _alpha_fmt(w, h, z.ALPHA_STRAIGHT))
    outs = gpu_util.run_graph(g, rgb + [alpha], [(h, w, np.float32)] * 4)
    ca = np.clip(alpha, np.float32(0.0), np.float32(1.0))
    for k in range(3):
        want = np.where(ca == 0, np.float32(0.0), rgb[k] / np.where(ca == 0, np.float32(1.0), ca)).astype(np.float32)
        _assert_equal(outs[k], want, f"plane {k}")
    _assert_equal(outs[3], alpha, "alpha")


FLOAT_CASES = [
    (3840, 2160, 1920, 1080, LANCZOS, 3.0, 0.0),
    (1280, 720, 3840, 2160, BICUBIC, 0.0, 0.5),
    (777, 555, 1111, 333, SPLINE36, 0.0, 0.0),
]


@pytest.mark.parametrize("sw,sh,dw,dh,ft,a,b", FLOAT_CASES)
def test_float_resize_bit_exact(sw, sh, dw, dh, ft, a, b):
    print(f"case: FLOAT {sw}x{sh}->{dw}x{dh} filter {ft}", flush=True)
    src = W.fp32_plane(sw, sh, seed=3)
    s, d, p = W.fp32_formats(sw, sh, dw, dh, ft, a, b if ft == BICUBIC else float("nan"))
    g = z.FilterGraph(s, d, p)
    (out, *_) = gpu_util.run_graph(g, [src, None, None, None], [(dh, dw, np.float32), None, None, None])
    _assert_equal(out, P.float_resize(src, dw, dh, ft, a, b), "float resize")


@pytest.mark.parametrize("w,h,frames", [(1920, 1080, 1), (98, 66, 1), (322, 180, 3), (1280, 720, 2)])
def test_fhd_709_to_2020_bit_exact(w, h, frames):
    """4:2:0 10-bit BT.709 -> BT.2020 NCL == reference chain, per frame."""
    print(f"case: 4:2:0 10-bit 709->2020 {w}x{h} x{frames}", flush=True)
    planes = [W.yuv420_10bit(w, h, seed=2 + k) for k in range(frames)]
    s, d, p = W.fhd_709_to_2020_formats(w, h)
    g = z.FilterGraph(s, d, p)
    stack = [np.stack([pl[c] for pl in planes]) if frames > 1 else planes[0][c] for c in range(3)]
    outs = gpu_util.run_graph(g, stack + [None], [(h, w, np.uint16), (h // 2, w // 2, np.uint16), (h // 2, w // 2, np.uint16), None])
    for k in range(frames):
        want = P.yuv420_10_709_to_2020(*planes[k])
        for got, ref, name in zip(outs[:3], want, "YUV"):
            _assert_equal(got[k] if frames > 1 else got, ref, f"709->2020 frame {k} plane {name}")


def test_709_to_2020_precision_form_within_one_code_value():
    """With the matrix-core precision form the chroma resizes run on MFMA:
    every plane within 1 code value of the exact chain (luma mixes in the
    chroma)."""
    w, h = 322, 180
    print(f"case: 709->2020 {w}x{h} precision form", flush=True)
    Y, U, V = W.yuv420_10bit(w, h, seed=9)
    s, d, p = W.fhd_709_to_2020_formats(w, h)
    g = z.FilterGraph(s, d, p)
    g.set_resize_precision(z.RESIZE_PRECISION_MFMA_FP16)
    outs = gpu_util.run_graph(g, [Y, U, V, None], [(h, w, np.uint16), (h // 2, w // 2, np.uint16), (h // 2, w // 2, np.uint16), None])
    want = P.yuv420_10_709_to_2020(Y, U, V)
    for got, ref, name in zip(outs[:3], want, "YUV"):
        assert np.max(np.abs(got.astype(np.int32) - ref.astype(np.int32))) <= 1, name


@pytest.mark.parametrize("w,h", [(640, 360), (3840, 2160)])
def test_pq_to_hlg_error_diffusion_bit_exact(w, h):
    """RGB FP32 PQ -> HLG -> 12-bit Floyd-Steinberg (row-skewed wavefront) == sequential reference."""
    print(f"case: PQ->HLG ED {w}x{h}", flush=True)
    R, G, B = W.pq_rgb(w, h, seed=4)
    s, d, p = W.uhd_pq_to_hlg_formats(w, h)
    g = z.FilterGraph(s, d, p)
    outs = gpu_util.run_graph(g, [R, G, B, None], [(h, w, np.uint16)] * 3 + [None])
    want = P.rgb_pq_to_hlg_12bit_ed(R, G, B)
    for got, ref, name in zip(outs[:3], want, "RGB"):
        _assert_equal(got, ref, f"pq->hlg plane {name}")


def test_batch_equals_single_frames():
    print("case: batch of 5 frames vs single", flush=True)
    frames = np.stack([W.word_plane(480, 270, seed=k) for k in range(5)])
    g = _word_graph(480, 270, 960, 540)
    batch, *_ = gpu_util.run_graph(g, [frames, None, None, None], [(540, 960, np.uint16), None, None, None], pad_bytes=64)
    for k in range(5):
        _assert_equal(batch[k], P.word_resize(frames[k], 960, 540), f"frame {k}")


def _yuv420_byte(w, h, seed):
    r = np.random.default_rng(seed)
    Y = r.integers(16, 236, (h, w)).astype(np.uint8)
    U = r.integers(16, 241, (h // 2, w // 2)).astype(np.uint8)
    V = r.integers(16, 241, (h // 2, w // 2)).astype(np.uint8)
    return Y, U, V


@pytest.mark.parametrize("src_type,dither_type", [("byte", 1), ("word10", 1), ("word10", 2)])
def test_yuv420_resize_ordered_dither_to_byte(src_type, dither_type):
    """4:2:0 u8 or 10-bit -> resized -> 8-bit with Bayer (1) or blue-noise (2)
    dither folded into the resize store (batched workload chain)."""
    print(f"case: 4:2:0 {src_type} 1280x720->1920x1080 spline36 dither {dither_type}", flush=True)
    sw, sh, dw, dh = 1280, 720, 1920, 1080
    Y8, U8, V8 = _yuv420_byte(sw, sh, 5)
    if src_type == "byte":
        planes, fmt_t, depth = (Y8, U8, V8), z.PIXEL_BYTE, 8
    else:
        planes, fmt_t, depth = tuple((p.astype(np.uint16) << 2) | 1 for p in (Y8, U8, V8)), z.PIXEL_WORD, 10
    common = dict(subsample_w=1, subsample_h=1, color_family=z.COLOR_YUV, matrix_coefficients=z.MATRIX_BT709,
                  transfer_characteristics=z.TRANSFER_BT709, color_primaries=z.PRIMARIES_BT709)
    s = z.image_format(width=sw, height=sh, pixel_type=fmt_t, depth=depth, **common)
    d = z.image_format(width=dw, height=dh, pixel_type=z.PIXEL_BYTE, depth=8, **common)
    p = z.builder_params(resample_filter=z.RESIZE_SPLINE36, resample_filter_uv=z.RESIZE_SPLINE36, dither_type=dither_type)
    g = z.FilterGraph(s, d, p)
    outs = gpu_util.run_graph(g, list(planes) + [None], [(dh, dw, np.uint8), (dh // 2, dw // 2, np.uint8), (dh // 2, dw // 2, np.uint8), None])

    # Oracle chain (graphbuilder.cpp:721-751 promotes limited BYTE to WORD16 by left shift).
    for q, (plane, got) in enumerate(zip(planes, outs[:3])):
        chroma = q > 0
        if src_type == "byte":
            x = left_shift(plane, BYTE, WORD, 8)
            wdepth = 16
        else:
            x = plane
            wdepth = 10
        cw_s, ch_s = plane.shape[1], plane.shape[0]
        cw_d, ch_d = (dw // 2, dh // 2) if chroma else (dw, dh)
        if chroma:
            scale_w = cw_d / cw_s
            shift_w = 0.25 - 0.25 / scale_w
        else:
            shift_w = 0.0
        r = resize_plane(x, SPLINE36, 0.0, 0.0, cw_d, ch_d, shift_w=shift_w, shift_h=0.0, subwidth=float(cw_s),
                         subheight=float(ch_s), depth=wdepth)
        sc, of = scale_offset((WORD, wdepth, 0, int(chroma), 0), (BYTE, 8, 0, int(chroma), 0))
        want = dither(r, WORD, BYTE, sc, of, 8, dither_type=dither_type, plane=q, blue_noise=blue_noise())
        _assert_equal(got, want, f"plane {q}")


def test_host_process_with_ring_buffers_and_callbacks():
    """zimg_filter_graph_process: host buffers, ring-buffer masks and callbacks (zimg.h:390-484)."""
    print("case: host process with callbacks", flush=True)
    sw, sh, dw, dh = 320, 240, 480, 360
    Y, U, V = W.yuv420_10bit(sw, sh, seed=6)
    s, d, p = W.fhd_709_to_2020_formats(sw, sh)
    d.width, d.height = dw, dh
    g = z.FilterGraph(s, d, p)
    in_lines, out_lines = g.input_buffering(), g.output_buffering()
    in_mask = z.lib().zimg_select_buffer_mask(in_lines)
    out_mask = z.lib().zimg_select_buffer_mask(out_lines)
    ring_in = [np.zeros((in_mask + 1, pl.shape[1]), np.uint16) for pl in (Y, U, V)]
    ring_out = [np.zeros((out_mask + 1, w_), np.uint16) for w_ in (dw, dw // 2, dw // 2)]
    result = [np.zeros((dh, dw), np.uint16), np.zeros((dh // 2, dw // 2), np.uint16), np.zeros((dh // 2, dw // 2), np.uint16)]
    calls = {"unpack": 0, "pack": 0}

    def unpack(user, i, left, right):
        calls["unpack"] += 1
        for q, pl in enumerate((Y, U, V)):
            sub = 1 if q else 0
            for r in range(i >> sub, (i + 2) >> sub):
                ring_in[q][r & in_mask] = pl[r]
        return 0

    def pack(user, i, left, right):
        calls["pack"] += 1
        for q in range(3):
            sub = 1 if q else 0
            for r in range(i >> sub, (i + 2) >> sub):
                result[q][r] = ring_out[q][r & out_mask]
        return 0

    ucb, pcb = z.CALLBACK(unpack), z.CALLBACK(pack)
    src = z.image_buffer([(r.ctypes.data, r.strides[0], in_mask) for r in ring_in])
    dst = z.image_buffer([(r.ctypes.data, r.strides[0], out_mask) for r in ring_out])
    tmp = np.zeros(g.tmp_size(), np.uint8)
    g.process(src, dst, tmp.ctypes.data, ucb, None, pcb, None)
    assert calls["unpack"] == sh // 2 and calls["pack"] == dh // 2

    # Same graph through the device entry point must agree.
    outs = gpu_util.run_graph(g, [Y, U, V, None], [(dh, dw, np.uint16), (dh // 2, dw // 2, np.uint16), (dh // 2, dw // 2, np.uint16), None])
    for q in range(3):
        _assert_equal(result[q], outs[q], f"host vs device plane {q}")


def test_failing_callback_reports_error():
    print("case: failing callback", flush=True)
    g = _word_graph(64, 48, 128, 96)
    src_img = W.word_plane(64, 48)
    out = np.zeros((96, 128), np.uint16)
    src = z.image_buffer([(src_img.ctypes.data, src_img.strides[0]), None, None, None])
    dst = z.image_buffer([(out.ctypes.data, out.strides[0]), None, None, None])
    bad = z.CALLBACK(lambda user, i, l, r: 1)
    with pytest.raises(z.ZimgError) as ei:
        g.process(src, dst, None, bad, None)
    assert ei.value.code == z.ERROR_USER_CALLBACK_FAILED


def test_host_process_concurrent_callers():
    """zimg_filter_graph_process is const and reentrant (zimg.h:495-499; the
    frame-parallel pattern of graphapp.cpp:355-377): 4 host threads x 3 calls
    on one graph, each with its own buffers and no tmp, give the device
    path's outputs bit for bit. Calls after the first round allocate nothing:
    the graph's staging-set count stays where the first round left it."""
    from concurrent.futures import ThreadPoolExecutor
    w, h = 320, 180
    print(f"case: host process, 4 threads x 3 calls, 4:2:0 709->2020 {w}x{h}", flush=True)
    s, d, p = W.fhd_709_to_2020_formats(w, h)
    g = z.FilterGraph(s, d, p)
    frames = [W.yuv420_10bit(w, h, seed=40 + k) for k in range(12)]
    specs = [(h, w, np.uint16), (h // 2, w // 2, np.uint16), (h // 2, w // 2, np.uint16), None]
    want = [gpu_util.run_graph(g, list(f) + [None], specs)[:3] for f in frames]

    def one(k):
        Y, U, V = frames[k]
        outs = [np.zeros((h, w), np.uint16), np.zeros((h // 2, w // 2), np.uint16), np.zeros((h // 2, w // 2), np.uint16)]
        src = z.image_buffer([(a.ctypes.data, a.strides[0]) for a in (Y, U, V)] + [None])
        dst = z.image_buffer([(a.ctypes.data, a.strides[0]) for a in outs] + [None])
        g.process(src, dst, None)
        return outs

    with ThreadPoolExecutor(max_workers=4) as ex:
        got = list(ex.map(one, range(12)))
    sets = g.host_staging_sets()
    assert 1 <= sets <= 4, sets
    with ThreadPoolExecutor(max_workers=4) as ex:
        got2 = list(ex.map(one, range(12)))
    for k in range(12):
        for q in range(3):
            _assert_equal(got[k][q], want[k][q], f"frame {k} plane {q}")
            _assert_equal(got2[k][q], want[k][q], f"frame {k} plane {q} (second round)")
    assert g.host_staging_sets() <= 4
