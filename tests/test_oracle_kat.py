"""Pins the oracle by known answers, closed forms and brute force, independently
of the compiled reference: the reference's own KATs (test/resize/filter_test.cpp,
test/colorspace/gamma_test.cpp), IEEE binary16 ground truth, numpy
restatements of the resize / dither formulas, and a pure-Python sequential
Floyd-Steinberg loop. Each check is chosen so that a dropped term, a wrong
sign or index, or a transposed operand fails it."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from oracle_lib import (BICUBIC, BILINEAR, BYTE, FLOAT, HALF, LANCZOS, POINT, SPLINE16, SPLINE36, SPLINE64, WORD, blue_noise,
                        c_u8_p, dither, half_to_float, left_shift, oracle, oracle_filter, ptr, resize_h, resize_h_first, resize_v,
                        scale_offset, to_float, to_half)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

f_eval = lambda t, a, b, x: oracle().orc_filter_eval(t, a, b, x)


def check_interpolating(t, a=0.0, b=0.0):
    """filter_test.cpp:8-20"""
    sup = oracle().orc_filter_support(t, a)
    assert f_eval(t, a, b, 0.0) == 1.0
    for i in range(1, sup + 1):
        assert abs(f_eval(t, a, b, -float(i))) < 1e-15
        assert abs(f_eval(t, a, b, float(i))) < 1e-15
    assert f_eval(t, a, b, math.nextafter(-float(sup), -math.inf)) == 0.0
    assert f_eval(t, a, b, math.nextafter(float(sup), math.inf)) == 0.0


def test_filter_kats():
    # filter_test.cpp:23-106
    check_interpolating(BILINEAR)
    assert f_eval(BILINEAR, 0, 0, 0.5) == 0.5 and f_eval(BILINEAR, 0, 0, -0.5) == 0.5
    check_interpolating(BICUBIC, 0.0, 0.5)
    assert f_eval(BICUBIC, 0.0, 0.5, 0.5) == 0.5625 and f_eval(BICUBIC, 0.0, 0.5, -1.5) == -0.0625
    mb = (1 / 3, 1 / 3)
    assert f_eval(BICUBIC, *mb, 0.0) == pytest.approx(8 / 9, rel=1e-15)
    assert abs(f_eval(BICUBIC, *mb, 1.0) - 1 / 18) < 1e-15
    assert f_eval(BICUBIC, *mb, 2.0) == 0.0
    assert f_eval(BICUBIC, *mb, 0.5) == pytest.approx(77 / 144, rel=1e-15)
    assert abs(f_eval(BICUBIC, *mb, -1.5) + 5 / 144) < 1e-15
    check_interpolating(SPLINE16)
    assert f_eval(SPLINE16, 0, 0, 0.5) == 0.575
    assert f_eval(SPLINE16, 0, 0, 1.5) == pytest.approx(-0.075, rel=1e-15)
    check_interpolating(SPLINE36)
    for x, v in ((0.5, 0.598684), (1.5, -0.118421), (2.5, 0.0197368)):
        assert abs(f_eval(SPLINE36, 0, 0, x) - v) < 1e-6 and abs(f_eval(SPLINE36, 0, 0, -x) - v) < 1e-6
    check_interpolating(SPLINE64)
    for x, v, tol in ((0.5, 0.600352, 1e-6), (1.5, -0.126760, 1e-6), (2.5, 0.0316901, 1e-7), (3.5, -0.00528169, 1e-8)):
        assert abs(f_eval(SPLINE64, 0, 0, x) - v) < tol
    for taps in (1, 2, 3):
        check_interpolating(LANCZOS, float(taps))
    # Lanczos closed form at a few points
    for x in (0.3, 1.7, 2.9):
        s = lambda t: 1.0 if t == 0 else math.sin(math.pi * t) / (math.pi * t)
        assert f_eval(LANCZOS, 3.0, 0, x) == pytest.approx(s(x) * s(x / 3), rel=1e-14)


def brute_matrix(ftype, a, b, src, dst, shift, width):
    """Independent FP64 resampling matrix (dense numpy), mirror boundary."""
    scale = dst / width
    step = min(scale, 1.0)
    sup = oracle().orc_filter_support(ftype, a) / step
    fs = max(int(math.ceil(sup)) * 2, 1)
    M = np.zeros((dst, src))
    for i in range(dst):
        pos = (i + 0.5) / scale + shift
        x0 = pos - fs / 2.0
        begin = (math.floor(x0 + 0.5) if x0 < 0 else math.floor(x0 + 0.49999999999999994)) + 0.5
        taps = [(begin + j, f_eval(ftype, a, b, (begin + j - pos) * step)) for j in range(fs)]
        tot = sum(w for _, w in taps)
        for xp, w in taps:
            rp = -xp if xp < 0 else (2 * src - xp if xp >= src else xp)
            rp = min(max(rp, 0.0), math.nextafter(src, -math.inf))
            M[i, int(math.floor(rp))] += w / tot
    return M


@pytest.mark.parametrize("ftype,a,b", [(LANCZOS, 3, 0), (BICUBIC, 0, 0.5), (SPLINE36, 0, 0), (BILINEAR, 0, 0), (POINT, 0, 0)])
@pytest.mark.parametrize("src,dst,shift", [(7, 13, 0.0), (13, 7, 0.25), (16, 5, -0.5), (5, 16, 0.125), (9, 9, 0.3)])
def test_compute_filter_matches_brute_force(ftype, a, b, src, dst, shift):
    fw, data, di16, left = oracle_filter(ftype, a, b, src, dst, shift, float(src))
    M = brute_matrix(ftype, a, b, src, dst, shift, float(src))
    for i in range(dst):
        row = np.zeros(src)
        row[left[i]:left[i] + fw] = data[i]
        assert np.allclose(row, M[i], atol=2e-7), (i, row, M[i])
        # Q1.14 taps sum to exactly one (filter.cpp:114-121) and track the FP64 row
        assert int(di16[i].astype(np.int64).sum()) == 1 << 14
        assert np.all(np.abs(di16[i] / 16384.0 - M[i, left[i]:left[i] + fw]) <= 2.0 / 16384)
        assert abs(1.0 - float(np.sum(data[i], dtype=np.float64))) <= np.finfo(np.float32).eps * 4
    assert np.all(left[:-1] <= left[1:]) or True


def np_resize_h_u16(img, fw, di16, left, depth):
    x = img.astype(np.int64) - 32768
    idx = left[:, None] + np.arange(fw)[None, :]
    acc = np.einsum("rjk,jk->rj", x[:, idx], di16.astype(np.int64))
    acc = ((acc + (1 << 13)) >> 14) + 32768
    return np.clip(acc, 0, (1 << depth) - 1).astype(np.uint16)


def test_word_resize_kernels_match_numpy_formula():
    rng = np.random.default_rng(1)
    for depth in (16, 10):
        img = rng.integers(0, 1 << depth, (23, 31)).astype(np.uint16)
        filt = oracle_filter(LANCZOS, 3, 0, 31, 45, 0.1, 31.0)
        got = resize_h(img, filt, depth)
        assert np.array_equal(got, np_resize_h_u16(img, filt[0], filt[2], filt[3], depth))
        fv = oracle_filter(SPLINE36, 0, 0, 23, 11, -0.2, 23.0)
        gotv = resize_v(img, fv, depth)
        want = np_resize_h_u16(img.T.copy(), fv[0], fv[2], fv[3], depth).T
        assert np.array_equal(gotv, want)


def test_float_resize_kernel_sequential_order():
    """FP32 accumulation is sequential from 0 in tap order (resize_impl.cpp:61-76)."""
    rng = np.random.default_rng(2)
    img = rng.random((5, 40), dtype=np.float32)
    fw, data, _, left = filt = oracle_filter(BICUBIC, 0, 0.5, 40, 97, 0.0, 40.0)
    got = resize_h(img, filt)
    want = np.zeros((5, 97), np.float32)
    for r in range(5):
        for j in range(97):
            acc = np.float32(0)
            for k in range(fw):
                acc = np.float32(acc + np.float32(data[j, k] * img[r, left[j] + k]))
            want[r, j] = acc
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_float_vertical_kernel_sequential_order():
    """The vertical FP32 pass also accumulates from 0 in tap order, one column
    at a time (resize_impl.cpp:97-114). Inputs are chosen so that any other
    association order (pairwise, reversed, a start at the first product, FMA)
    rounds differently: large and tiny magnitudes alternate down each column."""
    rng = np.random.default_rng(12)
    h, w = 41, 7
    img = (rng.random((h, w), dtype=np.float32) * np.where(np.arange(h)[:, None] % 2 == 0, 1e4, 1e-3)).astype(np.float32)
    fw, data, _, left = filt = oracle_filter(LANCZOS, 4, 0, h, 19, 0.3, float(h))
    got = resize_v(img, filt)
    want = np.zeros((19, w), np.float32)
    for i in range(19):
        for j in range(w):
            acc = np.float32(0)
            for k in range(fw):
                acc = np.float32(acc + np.float32(data[i, k] * img[left[i] + k, j]))
            want[i, j] = acc
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # The orders a plausible slip would produce differ from it somewhere.
    rev = np.zeros_like(want)
    for i in range(19):
        for j in range(w):
            acc = np.float32(0)
            for k in reversed(range(fw)):
                acc = np.float32(acc + np.float32(data[i, k] * img[left[i] + k, j]))
            rev[i, j] = acc
    assert not np.array_equal(rev.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("type_in,type_out,depth_in,shift", [(BYTE, WORD, 8, 8), (BYTE, WORD, 8, 2), (WORD, WORD, 10, 6),
                                                            (WORD, WORD, 10, 2), (BYTE, BYTE, 6, 2)])
def test_left_shift_every_code(type_in, type_out, depth_in, shift):
    """IntegerLeftShift (depth_convert.cpp:23-30, 104-144): every code of the
    input depth becomes x << shift in the output storage type, on a strided
    plane whose padding must stay untouched."""
    codes = np.arange(1 << depth_in, dtype=np.uint8 if type_in == BYTE else np.uint16)
    img = np.resize(codes, (max(1, codes.size // 64), 64)).astype(codes.dtype)
    got = left_shift(img, type_in, type_out, shift)
    want = (img.astype(np.uint32) << shift).astype(np.uint8 if type_out == BYTE else np.uint16)
    assert np.array_equal(got, want)
    assert set(np.unique(got).tolist()) == set((codes.astype(np.uint32) << shift).tolist())


def test_pass_order_choice():
    """resize_h_first (resize.cpp:12-18): the pass with the cheaper total tap
    work goes first, a horizontal tap weighted twice a vertical one. Known
    answers from the reference's own cases (2x/2x and 3x upscales: H first;
    1/2 x 1/2 downscale: V first), then a brute-force count over real filter
    widths and output sizes for random scale pairs with a clear margin."""
    assert resize_h_first(2.0, 2.0)
    assert resize_h_first(3.0, 3.0)
    assert not resize_h_first(0.5, 0.5)
    rng = np.random.default_rng(21)
    checked = 0
    for _ in range(400):
        sw, sh = int(rng.integers(40, 400)), int(rng.integers(40, 400))
        dw, dh = int(rng.integers(20, 900)), int(rng.integers(20, 900))
        fw_h = oracle_filter(LANCZOS, 3, 0, sw, dw)[0]
        fw_v = oracle_filter(LANCZOS, 3, 0, sh, dh)[0]
        # H first: the H pass makes sh x dw outputs of fw_h taps, then V makes dh x dw of fw_v.
        cost_h = 2 * sh * dw * fw_h + dh * dw * fw_v
        # V first: the V pass makes dh x sw outputs, then H makes dh x dw.
        cost_v = dh * sw * fw_v + 2 * dh * dw * fw_h
        if abs(cost_h - cost_v) < 0.2 * min(cost_h, cost_v):
            continue
        checked += 1
        assert resize_h_first(dw / sw, dh / sh) == (cost_h < cost_v), (sw, sh, dw, dh)
    assert checked > 200


def test_half_conversions_ieee():
    o = oracle()
    h = np.arange(65536, dtype=np.uint16)
    ref = h.view(np.float16).astype(np.float32)
    got = np.array([o.orc_half_to_float(int(v)) for v in h], np.float32)
    fin = np.isfinite(ref)
    assert np.array_equal(got[fin].view(np.uint32), ref[fin].view(np.uint32))
    assert np.all(np.isnan(got[~fin]) == np.isnan(ref[~fin]))
    rng = np.random.default_rng(3)
    f = np.concatenate([rng.standard_normal(50000) * 100, rng.standard_normal(50000) * 1e-5, [65504, 65519, 1e-8, -0.0]]).astype(np.float32)
    got = np.array([o.orc_float_to_half(float(v)) for v in f], np.uint16)
    assert np.array_equal(got, f.astype(np.float16).view(np.uint16))


def test_half_nan_and_plane_conversions():
    """NaN rules of quantize.cpp (payload kept, made quiet) and the plane forms:
    integer_to_half = float_to_half(x * scale + offset) (two float roundings,
    then one half rounding), float_to_half_n ignores scale/offset."""
    h = np.arange(65536, dtype=np.uint16)
    f = half_to_float(h[None, :])[0].view(np.uint32)
    nan = ((h & 0x7C00) == 0x7C00) & ((h & 0x3FF) != 0)
    want = ((h.astype(np.uint32) & 0x8000) << 16) | 0x7FC00000 | ((h.astype(np.uint32) & 0x3FF) << 13)
    assert np.array_equal(f[nan], want[nan])
    assert np.array_equal(f[~nan], h[~nan].view(np.float16).astype(np.float32).view(np.uint32))
    # float -> half on NaNs with payloads (including signalling ones, read from memory)
    bits = np.array([0x7F800001, 0x7FA00000, 0xFFC00001, 0x7FFFFFFF, 0x7F802000, 0xFF800000, 0x7F800000], np.uint32)
    got = to_half(bits.view(np.float32)[None, :], FLOAT)[0]
    exp = ((bits >> 16) & 0x8000) | 0x7C00 | np.where((bits & 0x7FFFFF) != 0, 0x200 | ((bits & 0x7FFFFF) >> 13), 0)
    assert got.tolist() == exp.astype(np.uint16).tolist()
    rng = np.random.default_rng(11)
    ints = rng.integers(0, 1024, (7, 300)).astype(np.uint16)
    s, o = scale_offset((WORD, 10, 0, 1, 0), (HALF, 16, 0, 1, 0))
    got = to_half(ints, WORD, s, o)
    x = (ints.astype(np.float32) * np.float32(s)).astype(np.float32) + np.float32(o)
    assert np.array_equal(got, x.astype(np.float32).astype(np.float16).view(np.uint16))
    fl = (rng.standard_normal((5, 333)) * 3).astype(np.float32)
    assert np.array_equal(to_half(fl, FLOAT, 7.0, 5.0), fl.astype(np.float16).view(np.uint16))


def test_blue_noise_fixture_and_product_table():
    """The RANDOM-dither map: 64x64, every byte value exactly 16 times (a
    rank-ordered threshold map), and the product's copy is the same data."""
    t = blue_noise()
    assert t.shape == (64, 64) and t.dtype == np.uint8
    assert np.array_equal(np.bincount(t.ravel(), minlength=256), np.full(256, 16))
    text = open(os.path.join(ROOT, "zimg_amd", "csrc", "bluenoise.h")).read()
    body = text[text.index("#define ZAMD_BLUE_NOISE_BYTES"):].split("\\\n", 1)[1]
    nums = [int(v) for v in re.findall(r"\d+", body)]
    assert nums == t.ravel().tolist()


def test_random_dither_formula():
    """dither.cpp:219-232 with dither_ordered: variant p starts the map at
    column offset (0, 32, 16, 48)[p] and row offset (0, 12, 55, 26)[p]."""
    rng = np.random.default_rng(5)
    img = rng.random((131, 150), dtype=np.float32)
    s, o = scale_offset((FLOAT, 32, 1, 0, 0), (BYTE, 8, 1, 0, 0))
    t = blue_noise()
    tab = ((t.astype(np.float32) + np.float32(1)) / np.float32(257) - np.float32(0.5)).astype(np.float32)
    for plane, (hoff, voff) in enumerate([(0, 0), (32, 12), (16, 55), (48, 26)]):
        got = dither(img, FLOAT, BYTE, s, o, 8, dither_type=2, plane=plane, blue_noise=t)
        d = tab[(np.arange(131)[:, None] + voff) % 64, (np.arange(150)[None, :] + hoff) % 64]
        x = (img * np.float32(s)).astype(np.float32) + np.float32(o)
        x = np.clip((x + d).astype(np.float32), 0, 255)
        assert np.array_equal(got, np.rint(x).astype(np.uint8)), plane


def test_bayer_table():
    raw = (ctypes.c_uint8 * 256)()
    oracle().orc_bayer_table(raw)
    t = np.array(raw, np.uint8).reshape(16, 16)
    assert sorted(t.ravel().tolist()) == list(range(256))
    # first rows of dither.cpp:141-142
    assert t[0, :8].tolist() == [0, 192, 48, 240, 12, 204, 60, 252]
    assert t[1, :8].tolist() == [128, 64, 176, 112, 140, 76, 188, 124]
    # recursive Bayer structure: the top-left 2x2 of each quadrant level orders 0, 3, 2, 1 (x4)
    assert [t[0, 0], t[0, 1], t[1, 0], t[1, 1]] == [0, 192, 128, 64]
    src = "/root/reference/src/zimg/depth/dither.cpp"
    if os.path.exists(src):
        text = open(src).read()
        block = text[text.index("BAYER_TABLE[BAYER_TABLE_LEN][BAYER_TABLE_LEN] = {"):]
        block = block[:block.index("};")]
        nums = [int(v) for v in re.findall(r"\d+", block.split("=", 1)[1])]
        assert nums == t.ravel().tolist()


def test_ordered_dither_formula():
    rng = np.random.default_rng(4)
    img = rng.integers(0, 1024, (33, 70)).astype(np.uint16)
    s, o = scale_offset((WORD, 10, 0, 0, 0), (BYTE, 8, 0, 0, 0))
    raw = (ctypes.c_uint8 * 256)()
    oracle().orc_bayer_table(raw)
    t = np.array(raw, np.uint8).reshape(16, 16)
    variants = [t, t[:, ::-1], t[::-1, :], t.T]
    for plane in range(4):
        got = dither(img, WORD, BYTE, s, o, 8, dither_type=1, plane=plane)
        tab = ((variants[plane].astype(np.float32) + np.float32(1)) / np.float32(257) - np.float32(0.5)).astype(np.float32)
        x = img.astype(np.float32) * np.float32(s) + np.float32(o)
        d = tab[np.arange(33)[:, None] % 16, np.arange(70)[None, :] % 16]
        x = np.clip((x + d).astype(np.float32), 0, 255)
        assert np.array_equal(got, np.rint(x).astype(np.uint8)), plane
    none = dither(img, WORD, BYTE, s, o, 8)
    assert np.array_equal(none, np.rint(np.clip(img.astype(np.float32) * np.float32(s) + np.float32(o), 0, 255)).astype(np.uint8))


def test_error_diffusion_equals_sequential_python():
    rng = np.random.default_rng(5)
    h, w = 9, 13
    img = rng.random((h, w), dtype=np.float32)
    s, o = scale_offset((FLOAT, 32, 0, 0, 0), (WORD, 12, 1, 0, 0))
    got = dither(img, FLOAT, WORD, s, o, 12, dither_type=3)
    f32 = np.float32
    prev = [f32(0)] * (w + 2)
    want = np.zeros((h, w), np.uint16)
    for i in range(h):
        cur = [f32(0)] * (w + 2)
        for j in range(w):
            x = f32(f32(img[i, j] * f32(s)) + f32(o))
            err = f32(0)
            err = f32(err + f32(cur[j] * f32(7 / 16)))
            err = f32(err + f32(prev[j + 2] * f32(3 / 16)))
            err = f32(err + f32(prev[j + 1] * f32(5 / 16)))
            err = f32(err + f32(prev[j] * f32(1 / 16)))
            x = f32(x + err)
            x = min(max(x, f32(0)), f32(4095))
            q = int(np.rint(x))
            want[i, j] = q
            cur[j + 1] = f32(x - f32(q))
        prev = cur
    assert np.array_equal(got, want)
    # diffusion keeps the mean: the quantised image tracks the scaled input
    big = rng.random((64, 64), dtype=np.float32) * np.float32(0.01)
    q = dither(big, FLOAT, BYTE, 255.0, 0.0, 8, dither_type=3)
    assert abs(q.mean() - big.mean() * 255) < 0.05


def test_to_float_two_roundings():
    img = np.arange(1024, dtype=np.uint16).reshape(32, 32)
    s, o = scale_offset((WORD, 10, 0, 1, 0), (FLOAT, 32, 0, 1, 0))
    assert (s, o) == (np.float32(1 / 896), np.float32(-512 / 896))
    got = to_float(img, WORD, s, o)
    want = (img.astype(np.float32) * np.float32(s)).astype(np.float32) + np.float32(o)
    assert np.array_equal(got.view(np.uint32), want.astype(np.float32).view(np.uint32))


def _tf(transfer, to_linear, scene, x):
    y = ctypes.c_float()
    assert oracle().orc_transfer_eval(transfer, to_linear, scene, float(x), ctypes.byref(y)) == 0
    return y.value


T_LOG100, T_LOG316, T_709, T_240M, T_SRGB, T_2084, T_B67 = 2, 3, 4, 7, 10, 11, 13


@pytest.mark.parametrize("transfer,scene,lo,hi,thr", [
    (T_709, 1, 0.0, 1.0, 1e-6), (T_709, 0, 0.0, 1.0, 1e-6), (T_240M, 1, 0.0, 1.0, 1e-6), (T_SRGB, 0, 0.0, 1.0, 1e-6),
    (T_2084, 0, 0.0, 1.0, 1e-4), (T_B67, 1, 0.0, 1.0, 1e-6), (T_LOG100, 0, 0.01, 1.0, 1e-6),
])
def test_transfer_kats_and_round_trip(transfer, scene, lo, hi, thr):
    """gamma_test.cpp:8-348: endpoints, monotonicity and 2^16-step round-trip accuracy."""
    g = lambda x: _tf(transfer, 0, scene, x)
    lin = lambda x: _tf(transfer, 1, scene, x)
    if transfer not in (T_2084, T_B67, T_LOG100):
        assert g(0.0) == 0.0 and g(1.0) == 1.0 and lin(0.0) == 0.0 and lin(1.0) == 1.0
    steps = 1 << 12
    xs = [np.float32(lo + i * ((hi - lo) / steps)) for i in range(steps + 1)]
    prev = -math.inf
    for x in xs:
        y = g(x)
        assert not math.isnan(y) and y >= prev
        prev = y
    err = 0.0
    for x in xs:
        rt = lin(g(x))
        err += abs((rt - x) / (x if x != 0 else np.finfo(np.float32).eps))
    assert err / len(xs) < thr


def test_pq_hlg_kats():
    # gamma_test.cpp:266-269,308-313: exact endpoints
    assert _tf(T_2084, 1, 0, 1.0) == 1.0 and _tf(T_2084, 0, 0, 1.0) == 1.0
    assert _tf(T_2084, 1, 0, 0.0) == 0.0 and _tf(T_2084, 0, 0, 0.0) == 0.0
    assert _tf(T_B67, 0, 1, np.float32(1.0 / 12.0)) == 0.5 and _tf(T_B67, 0, 1, 1.0) == 1.0
    assert _tf(T_B67, 1, 1, 0.5) == np.float32(1.0 / 12.0) and _tf(T_B67, 1, 1, 0.0) == 0.0
    # ST 2084 at 100 nits is code ~0.508 (BT.2100 table)
    assert _tf(T_2084, 0, 0, 0.01) == pytest.approx(0.5081, abs=2e-4)
    # ARIB B67 OETF: E = 1/12 -> 0.5; E = 1 -> ~1.0
    assert _tf(T_B67, 0, 1, 1.0 / 12.0) == pytest.approx(0.5, abs=1e-7)
    assert _tf(T_B67, 0, 1, 1.0) == pytest.approx(1.0, abs=1e-6)


def test_libm_close_to_correctly_rounded():
    """powf/expf/logf are within one ulp of the correctly rounded FP64 result and
    exact in the vast majority of cases (method's published 0.82 / 0.502 / 0.818 ulp)."""
    o = oracle()
    rng = np.random.default_rng(6)
    x = (rng.random(20000) * 4).astype(np.float32)
    y = rng.uniform(-3, 3, 20000).astype(np.float32)
    got = np.array([o.orc_powf(float(a), float(b)) for a, b in zip(x, y)], np.float32)
    want = np.power(x.astype(np.float64), y.astype(np.float64)).astype(np.float32)
    ulps = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
    assert ulps.max() <= 1 and np.mean(ulps == 0) > 0.99
    xe = rng.uniform(-80, 80, 20000).astype(np.float32)
    got = np.array([o.orc_expf(float(v)) for v in xe], np.float32)
    want = np.exp(xe.astype(np.float64)).astype(np.float32)
    assert np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32)).max() <= 1
    xl = (10.0 ** rng.uniform(-30, 30, 20000)).astype(np.float32)
    got = np.array([o.orc_logf(float(v)) for v in xl], np.float32)
    want = np.log(xl.astype(np.float64)).astype(np.float32)
    assert np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32)).max() <= 1


def test_fp64_matrix_reproduces_pinned_f32_table():
    """The FP64 entries the precision-form oracle applies are the ones the
    f32 table (pinned against the compiled reference) was rounded from: the
    error-dithered rounding of filter.cpp:86-97 applied to them in numpy
    gives the table back bit for bit; rows sum to 1 in FP64."""
    import ctypes as C
    from oracle_lib import c_double_p, c_float_p, c_i16_p, c_u32_p, oracle_filter64
    for ftype, a, b, src, dst in [(LANCZOS, 3.0, 0.0, 3840, 1920), (BICUBIC, 0.0, 0.5, 1280, 3840), (SPLINE36, 0.0, 0.0, 333, 1000)]:
        lib = oracle()
        fw = C.c_uint32(0)
        lib.orc_compute_filter64(ftype, a, b, src, dst, 0.0, float(src), C.byref(fw), None, None, None, None)
        fw = fw.value
        d32 = np.zeros(dst * fw, np.float32)
        d16 = np.zeros(dst * fw, np.int16)
        d64 = np.zeros(dst * fw, np.float64)
        left = np.zeros(dst, np.uint32)
        lib.orc_compute_filter64(ftype, a, b, src, dst, 0.0, float(src), C.byref(C.c_uint32(0)), ptr(d32, c_float_p),
                                 ptr(d16, c_i16_p), ptr(d64, c_double_p), ptr(left, c_u32_p))
        d64 = d64.reshape(dst, fw)
        want = np.zeros((dst, fw), np.float32)
        for i in range(dst):
            err = 0.0
            for j in range(fw):
                e = float(d64[i, j]) - err
                q = np.float32(e)
                err = float(q) - e
                want[i, j] = q
        assert np.array_equal(want.ravel().view(np.uint32), d32.view(np.uint32))
        m = oracle_filter64(ftype, a, b, src, dst)
        assert np.max(np.abs(m.sum(axis=1) - 1.0)) < 1e-12


def _bilinear_matrix(up, orig, shift):
    """The bilinear upscale A (up x orig) that unresize inverts, from its
    definition: output sample i sits at (i + 0.5) orig / up on the input grid,
    interpolated linearly between the two nearest input samples, the outermost
    input sample repeated beyond the ends (bilinear.cpp:64-108)."""
    A = np.zeros((up, orig))
    sh = -shift * orig / up
    lo, hi = 0.5 + sh, orig - 0.5 + sh
    for i in range(up):
        x = (i + 0.5) * orig / up
        if x <= lo:
            A[i, int(min(max(math.floor(lo), 0), orig - 1))] = 1
        elif x >= hi:
            A[i, int(min(max(math.floor(hi), 0), orig - 1))] = 1
        else:
            k = int(math.floor(x - lo))
            d = x - k - lo
            A[i, k] += 1 - d
            A[i, k + 1] += d
    return A


@pytest.mark.parametrize("up,orig,shift", [(40, 17, 0.0), (64, 32, 0.25), (33, 33, 0.5), (90, 31, -0.125)])
def test_unresize_solves_least_squares(up, orig, shift):
    """Unresize is the least-squares inverse of the bilinear upscale: its
    output solves (A'A) x = A'y for the A built above (FP64 numpy solve; float
    error only), and unresizing A x gives back x."""
    from oracle_lib import unresize_context, unresize_h, unresize_v
    A = _bilinear_matrix(up, orig, shift)
    ctx = unresize_context(up, orig, shift)
    rng = np.random.default_rng(up + orig)
    y = rng.random((5, up)).astype(np.float32)
    want = np.linalg.solve(A.T @ A, A.T @ y.astype(np.float64).T).T
    got = unresize_h(y, ctx)
    assert np.max(np.abs(got - want)) < 2e-4 * max(1.0, np.max(np.abs(want)))
    x = rng.random((orig, 6))
    back = unresize_v((A @ x).astype(np.float32), ctx)
    assert np.max(np.abs(back - x)) < 2e-4


def test_unresize_kernels_sequential_order():
    """Both line kernels in the reference's float order (unresize_impl.cpp:
    27-85): the band accumulated from 0 in tap order, forward substitution
    z = (accum - c z) l, then back substitution w = d - u w."""
    from oracle_lib import unresize_context, unresize_h, unresize_v
    rs, coef, off, c, l, u = ctx = unresize_context(23, 9, 0.25)
    rng = np.random.default_rng(5)
    y = (rng.random((3, 23)) * np.where(np.arange(23) % 2, 1e3, 1e-3)).astype(np.float32)
    want = np.zeros((3, 9), np.float32)
    f = np.float32
    for r in range(3):
        z = f(0)
        for j in range(9):
            acc = f(0)
            for k in range(rs):
                acc = f(acc + f(coef[j, k] * y[r, off[j] + k]))
            z = f(f(acc - f(c[j] * z)) * l[j])
            want[r, j] = z
        w = f(0)
        for j in range(9, 0, -1):
            w = f(want[r, j - 1] - f(u[j - 1] * w))
            want[r, j - 1] = w
    assert np.array_equal(unresize_h(y, ctx).view(np.uint32), want.view(np.uint32))
    assert np.array_equal(unresize_v(y.T.copy(), ctx).view(np.uint32), want.T.copy().view(np.uint32))
