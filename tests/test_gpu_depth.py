"""GPU parity of the depth rows: HALF storage (both directions, every half
value, NaN payloads, ties, denormal results), integer <-> HALF, RANDOM
(blue-noise) dither, and HALF planes through resize and colourspace, against
the oracle, bit for bit."""
import numpy as np
import pytest

import zimg_amd as z
from oracle_lib import (BICUBIC, BYTE, FLOAT, HALF, WORD, blue_noise, colorspace_apply, colorspace_ops, dither,
                        half_to_float, resize_plane, scale_offset, to_half)

pytestmark = pytest.mark.gpu

gpu_util = pytest.importorskip("gpu_util")

M_709, M_RGB, T_709, P_709 = 3, 1, 4, 4


def _assert_bits(got, want, what):
    assert got.shape == want.shape, what
    g = got.view(np.uint32) if got.dtype == np.float32 else got
    w = want.view(np.uint32) if want.dtype == np.float32 else want
    bad = np.argwhere(g != w)
    if bad.size:
        i = tuple(bad[0])
        pytest.fail(f"{what}: {len(bad)} mismatches, first at {list(i)}: got {g[i]:#x} want {w[i]:#x}")


def _grey(w, h, t, depth, **kw):
    return z.image_format(width=w, height=h, pixel_type=t, depth=depth, **kw)


def test_half_to_float_every_value():
    print("case: GREY HALF 256x256 (all 65536 values) -> FLOAT", flush=True)
    h = np.arange(65536, dtype=np.uint16).reshape(256, 256)
    g = z.FilterGraph(_grey(256, 256, z.PIXEL_HALF, 16), _grey(256, 256, z.PIXEL_FLOAT, 32), z.builder_params())
    (out, *_) = gpu_util.run_graph(g, [h, None, None, None], [(256, 256, np.float32), None, None, None])
    _assert_bits(out, half_to_float(h), "half -> float")


def _float_patterns(n, seed):
    rng = np.random.default_rng(seed)
    exps = rng.integers(0, 256, n).astype(np.uint32)
    mants = rng.integers(0, 1 << 23, n).astype(np.uint32)
    near = rng.integers(96, 146, n).astype(np.uint32)  # exponents around the half range
    ties = (rng.integers(0, 1 << 10, n).astype(np.uint32) << 13) | 0x1000
    special = np.array([0, 0x80000000, 0x7F800000, 0xFF800000, 0x7F800001, 0x7FA00000, 0xFFC00001, 0x7FFFFFFF,
                        0x477FEFFF, 0x477FF000, 0x477FF001, 0x33000000, 0x33000001, 0x387FC000, 0x00000001, 0x00800000],
                       np.uint32)
    bits = np.concatenate([(exps << 23) | mants, (near << 23) | mants, (near << 23) | ties, special])
    bits |= rng.integers(0, 2, bits.size).astype(np.uint32) << 31
    return bits


def test_float_to_half_patterns():
    bits = _float_patterns(40000, 12)
    w = 512
    h = (bits.size + w - 1) // w
    bits = np.resize(bits, w * h).reshape(h, w)
    print(f"case: GREY FLOAT {w}x{h} bit patterns -> HALF", flush=True)
    src = bits.view(np.float32)
    g = z.FilterGraph(_grey(w, h, z.PIXEL_FLOAT, 32), _grey(w, h, z.PIXEL_HALF, 16), z.builder_params())
    (out, *_) = gpu_util.run_graph(g, [src, None, None, None], [(h, w, np.uint16), None, None, None])
    _assert_bits(out, to_half(src, FLOAT), "float -> half")


@pytest.mark.parametrize("dither_type", [z.DITHER_NONE, z.DITHER_ORDERED, z.DITHER_RANDOM, z.DITHER_ERROR_DIFFUSION])
def test_word_to_half_to_byte(dither_type):
    """WORD 10-bit -> HALF (integer_to_half), then HALF -> BYTE 8-bit full range
    (dither reads the half through half_to_float), three YUV 4:4:4 planes."""
    print(f"case: YUV444 WORD10 333x131 -> HALF -> BYTE dither {dither_type}", flush=True)
    w, h = 333, 131
    rng = np.random.default_rng(21)
    planes = [rng.integers(64, 941, (h, w)).astype(np.uint16) for _ in range(3)]
    common = dict(color_family=z.COLOR_YUV, matrix_coefficients=z.MATRIX_BT709, transfer_characteristics=z.TRANSFER_BT709,
                  color_primaries=z.PRIMARIES_BT709)
    s = _grey(w, h, z.PIXEL_WORD, 10, **common)
    m = _grey(w, h, z.PIXEL_HALF, 16, **common)
    d = _grey(w, h, z.PIXEL_BYTE, 8, pixel_range=z.RANGE_FULL, **common)
    g1 = z.FilterGraph(s, m, z.builder_params())
    halves = gpu_util.run_graph(g1, planes + [None], [(h, w, np.uint16)] * 3 + [None])
    g2 = z.FilterGraph(m, d, z.builder_params(dither_type=dither_type))
    outs = gpu_util.run_graph(g2, halves[:3] + [None], [(h, w, np.uint8)] * 3 + [None])
    oracle_dither = {z.DITHER_NONE: 0, z.DITHER_ORDERED: 1, z.DITHER_RANDOM: 2, z.DITHER_ERROR_DIFFUSION: 3}[dither_type]
    for q in range(3):
        chroma = int(q > 0)
        sc, of = scale_offset((WORD, 10, 0, chroma, 0), (HALF, 16, 0, chroma, 0))
        want_h = to_half(planes[q], WORD, sc, of)
        _assert_bits(halves[q], want_h, f"word -> half plane {q}")
        sc, of = scale_offset((HALF, 16, 0, chroma, 0), (BYTE, 8, 1, chroma, 0))
        want = dither(want_h, HALF, BYTE, sc, of, 8, dither_type=oracle_dither, plane=q, blue_noise=blue_noise())
        _assert_bits(outs[q], want, f"half -> byte plane {q}")


def test_random_dither_float_to_word_batch():
    """FLOAT RGB -> WORD 10-bit full range, blue-noise dither, every plane variant,
    several frames in one launch (the map repeats every 64 rows/columns)."""
    print("case: RGB FLOAT 200x150 x3 frames -> WORD10 random dither", flush=True)
    w, h, F = 200, 150, 3
    rng = np.random.default_rng(22)
    planes = [rng.random((F, h, w), dtype=np.float32) * 1.1 - 0.05 for _ in range(3)]
    common = dict(color_family=z.COLOR_RGB, matrix_coefficients=z.MATRIX_RGB, pixel_range=z.RANGE_FULL)
    g = z.FilterGraph(_grey(w, h, z.PIXEL_FLOAT, 32, **common), _grey(w, h, z.PIXEL_WORD, 10, **common),
                      z.builder_params(dither_type=z.DITHER_RANDOM))
    outs = gpu_util.run_graph(g, planes + [None], [(h, w, np.uint16)] * 3 + [None])
    sc, of = scale_offset((FLOAT, 32, 1, 0, 0), (WORD, 10, 1, 0, 0))
    for q in range(3):
        for f in range(F):
            want = dither(planes[q][f], FLOAT, WORD, sc, of, 10, dither_type=2, plane=q, blue_noise=blue_noise())
            _assert_bits(outs[q][f], want, f"plane {q} frame {f}")


def test_half_resize_to_half():
    """HALF -> FLOAT (folded into the resize load) -> bicubic 333x200 -> 700x411
    -> HALF (folded into the resize store): one launch."""
    print("case: GREY HALF 333x200 -> 700x411 bicubic -> HALF", flush=True)
    sw, sh, dw, dh = 333, 200, 700, 411
    rng = np.random.default_rng(23)
    src = (rng.random((sh, sw), dtype=np.float32) * 2 - 0.5).astype(np.float16).view(np.uint16)
    g = z.FilterGraph(_grey(sw, sh, z.PIXEL_HALF, 16), _grey(dw, dh, z.PIXEL_HALF, 16),
                      z.builder_params(resample_filter=z.RESIZE_BICUBIC, filter_param_a=0.0, filter_param_b=0.5))
    launches = g.device_stages()
    assert len(launches) == 1, launches
    (out, *_) = gpu_util.run_graph(g, [src, None, None, None], [(dh, dw, np.uint16), None, None, None], pad_bytes=6)
    want = to_half(resize_plane(half_to_float(src), BICUBIC, 0.0, 0.5, dw, dh), FLOAT)
    _assert_bits(out, want, "half resize")


def test_half_yuv_to_half_rgb():
    """HALF YUV 4:4:4 BT.709 -> HALF RGB: widening folded into the colourspace
    load, the half rounding into its store."""
    print("case: HALF YUV444 709 320x97 -> HALF RGB", flush=True)
    w, h = 320, 97
    rng = np.random.default_rng(24)
    Y = rng.random((h, w), dtype=np.float32)
    U = rng.random((h, w), dtype=np.float32) - 0.5
    V = rng.random((h, w), dtype=np.float32) - 0.5
    planes = [p.astype(np.float16).view(np.uint16) for p in (Y, U, V)]
    yuv = dict(color_family=z.COLOR_YUV, matrix_coefficients=z.MATRIX_BT709, transfer_characteristics=z.TRANSFER_BT709,
               color_primaries=z.PRIMARIES_BT709)
    rgb = dict(yuv, color_family=z.COLOR_RGB, matrix_coefficients=z.MATRIX_RGB)
    g = z.FilterGraph(_grey(w, h, z.PIXEL_HALF, 16, **yuv), _grey(w, h, z.PIXEL_HALF, 16, **rgb), z.builder_params())
    launches = g.device_stages()
    assert len(launches) == 1, launches
    outs = gpu_util.run_graph(g, planes + [None], [(h, w, np.uint16)] * 3 + [None])
    n, ops, _ = colorspace_ops((M_709, T_709, P_709), (M_RGB, T_709, P_709))
    assert n == 1
    r, gg, b = colorspace_apply(ops, *(half_to_float(p) for p in planes))
    for q, want in enumerate((r, gg, b)):
        _assert_bits(outs[q], to_half(want, FLOAT), f"plane {q}")


@pytest.mark.parametrize("w,out", [(640, "byte"), (641, "byte"), (640, "word"), (643, "word")])
def test_error_diffusion_frames_widths(w, out):
    """FLOAT -> BYTE 8-bit / WORD 12-bit error diffusion at even and odd
    widths (odd: the last column pair of a row is half outside it), values
    past both ends of the range; 200 rows end in a strip of 8 rows, and two
    frames share one launch; every frame against the sequential oracle."""
    h, F = 200, 2
    print(f"case: GREY FLOAT {w}x{h} x{F} frames -> {out} error diffusion", flush=True)
    rng = np.random.default_rng(29)
    src = rng.random((F, h, w), dtype=np.float32) * 1.2 - 0.1
    ptype, depth, dt, otype = (z.PIXEL_BYTE, 8, np.uint8, BYTE) if out == "byte" else (z.PIXEL_WORD, 12, np.uint16, WORD)
    g = z.FilterGraph(_grey(w, h, z.PIXEL_FLOAT, 32), _grey(w, h, ptype, depth, pixel_range=z.RANGE_FULL),
                      z.builder_params(dither_type=z.DITHER_ERROR_DIFFUSION))
    (o, *_) = gpu_util.run_graph(g, [src, None, None, None], [(h, w, dt), None, None, None])
    sc, of = scale_offset((FLOAT, 32, 1, 0, 0), (otype, depth, 1, 0, 0))
    for f in range(F):
        want = dither(src[f], FLOAT, otype, sc, of, depth, dither_type=3, plane=0)
        _assert_bits(o[f], want, f"error diffusion {w} wide, frame {f}")


@pytest.mark.parametrize("w", [20000, 32639])
def test_error_diffusion_wide_planes(w):
    """Planes wider than a 1024-entry LDS ring per wave (up to the 32639-sample
    limit: 2048 entries, 128 KiB of LDS) through the error-diffusion
    wavefront, three strips so the rows cross between waves through the
    ring, against the sequential oracle."""
    h = 150
    print(f"case: GREY FLOAT {w}x{h} -> BYTE error diffusion", flush=True)
    rng = np.random.default_rng(23)
    src = rng.random((h, w), dtype=np.float32)
    g = z.FilterGraph(_grey(w, h, z.PIXEL_FLOAT, 32), _grey(w, h, z.PIXEL_BYTE, 8, pixel_range=z.RANGE_FULL),
                      z.builder_params(dither_type=z.DITHER_ERROR_DIFFUSION))
    (out, *_) = gpu_util.run_graph(g, [src, None, None, None], [(h, w, np.uint8), None, None, None])
    sc, of = scale_offset((FLOAT, 32, 1, 0, 0), (BYTE, 8, 1, 0, 0))
    want = dither(src, FLOAT, BYTE, sc, of, 8, dither_type=3, plane=0)
    _assert_bits(out, want, f"error diffusion {w} wide")
