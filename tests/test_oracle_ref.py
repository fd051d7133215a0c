"""Pins the oracle's restatement against the reference's own compiled leaf
sources (oracle/_ref, built from resize/filter.cpp, colorspace/*.cpp,
depth/quantize.cpp, depth/blue.cpp and the vendored musl libm), bit for bit."""
import ctypes

import numpy as np
import pytest

from oracle_lib import (unresize_context, BICUBIC, BILINEAR, LANCZOS, POINT, SPLINE16, SPLINE36, SPLINE64, blue_noise, c_float_p, c_u8_p, colorspace_apply,
                        colorspace_ops, oracle, oracle_filter, ptr, ref, ref_colorspace_apply, ref_filter)

pytestmark = pytest.mark.ref

FILTERS = [(LANCZOS, 3, 0), (BICUBIC, 0, 0.5), (BICUBIC, 1 / 3, 1 / 3), (BILINEAR, 0, 0), (POINT, 0, 0),
           (SPLINE16, 0, 0), (SPLINE36, 0, 0), (SPLINE64, 0, 0), (LANCZOS, 4, 0)]
DIMS = [(1920, 3840), (1080, 2160), (3840, 1920), (2160, 1080), (1280, 3840), (720, 2160), (960, 1920),
        (1920, 960), (17, 5), (5, 17), (100, 100), (640, 480), (1, 7), (7, 1), (3840, 1280), (1440, 720)]
SHIFTS = [0.0, 0.25, -0.5, 0.125, 1 / 3]


@pytest.mark.parametrize("ftype,a,b", FILTERS)
def test_filter_tables_bit_exact(ftype, a, b):
    """compute_filter + matrix_to_filter (filter.cpp:43-306): f32 and Q1.14 tables and left offsets."""
    for src, dst in DIMS:
        for shift in SHIFTS:
            for width in (float(src), src - 0.5):
                A = oracle_filter(ftype, a, b, src, dst, shift, width)
                B = ref_filter(ftype, a, b, src, dst, shift, width)
                assert A[0] == B[0], (src, dst, shift, width)
                assert np.array_equal(A[1].view(np.uint32), B[1].view(np.uint32))
                assert np.array_equal(A[2], B[2])
                assert np.array_equal(A[3], B[3])


@pytest.mark.parametrize("ftype,a,b", FILTERS)
def test_filter_eval_bit_exact(ftype, a, b):
    o, r = oracle(), ref()
    for x in np.linspace(-5, 5, 2001):
        assert o.orc_filter_eval(ftype, a, b, x) == r.ref_filter_eval(ftype, a, b, x)


def _ref_vec(fn, *arrs):
    out = np.zeros_like(arrs[0])
    fn(*[ptr(a, c_float_p) for a in arrs], ptr(out, c_float_p), arrs[0].size)
    return out


def test_libm_bit_exact_vs_vendored_musl():
    """powf/expf/logf/log10f restatements == the reference's test libm (test/main.cpp:10-18)."""
    o, r = oracle(), ref()
    r.ref_set_libm(1)
    rng = np.random.default_rng(11)
    n = 100_000
    x = np.concatenate([rng.random(n, dtype=np.float32) * 2, (10.0 ** rng.uniform(-30, 30, n)).astype(np.float32)])
    y = np.concatenate([rng.uniform(-3, 3, n), rng.uniform(-3, 3, n)]).astype(np.float32)
    y[:16] = [2.4, 1 / 2.4, 0.45, 1 / 0.45, 1 / 78.84375, 1 / 0.1593017578125, 78.84375, 0.1593017578125,
              1.2, 1 / 1.2, 2.2, 2.8, 2.6, -1 / 6, 0.2, 0.0]
    ref_p = _ref_vec(r.ref_powf_n, x, y)
    orc_p = np.array([o.orc_powf(float(a), float(b)) for a, b in zip(x, y)], np.float32)
    assert np.array_equal(ref_p.view(np.uint32), orc_p.view(np.uint32))
    xe = rng.uniform(-100, 100, n).astype(np.float32)
    assert np.array_equal(_ref_vec(r.ref_expf_n, xe).view(np.uint32),
                          np.array([o.orc_expf(float(v)) for v in xe], np.float32).view(np.uint32))
    xl = (10.0 ** rng.uniform(-38, 38, n)).astype(np.float32)
    assert np.array_equal(_ref_vec(r.ref_logf_n, xl).view(np.uint32),
                          np.array([o.orc_logf(float(v)) for v in xl], np.float32).view(np.uint32))


def test_glibc_powf_equals_musl_powf():
    """The library default (glibc powf) and the test libm agree on this host, so
    device parity with the musl method is parity with the default build too."""
    r = ref()
    rng = np.random.default_rng(5)
    x = (rng.random(500_000, dtype=np.float32) * 4).astype(np.float32)
    y = rng.uniform(-3, 3, 500_000).astype(np.float32)
    r.ref_set_libm(0)
    g = _ref_vec(r.ref_powf_n, x, y)
    r.ref_set_libm(1)
    m = _ref_vec(r.ref_powf_n, x, y)
    assert np.array_equal(g.view(np.uint32), m.view(np.uint32))


def test_colorspace_chains_bit_exact():
    """Operation chains (colorspace_graph.cpp BFS + operation_impl.cpp ops) on random
    float triples, over a seeded sample of all (matrix, transfer, primaries) pairs."""
    ref().ref_set_libm(1)
    rng = np.random.default_rng(2)
    pl = [rng.uniform(-0.2, 1.2, 512).astype(np.float32) for _ in range(3)]
    csps = [(m, t, p) for m in range(13) for t in range(14) for p in range(11)]
    pairs = [(csps[i], csps[j]) for i, j in np.random.default_rng(3).integers(0, len(csps), size=(2500, 2))]
    pairs += [((3, 4, 4), (7, 4, 6)), ((1, 11, 6), (1, 13, 6)), ((12, 11, 6), (1, 1, 6)), ((8, 4, 6), (1, 1, 6))]
    checked = 0
    for cin, cout in pairs:
        for peak, scene, adapt in [(100.0, 0, 0), (203.0, 0, 1), (1000.0, 1, 0)]:
            n_o, ops, _ = colorspace_ops(cin, cout, peak, scene, adapt)
            n_r, a0, a1, a2 = ref_colorspace_apply(cin, cout, *pl, peak, scene, 0, adapt)
            assert (n_o < 0) == (n_r < 0), (cin, cout)
            if n_o < 0:
                continue
            assert n_o == n_r
            b = colorspace_apply(ops, *pl)
            for x, y in zip((a0, a1, a2), b):
                assert np.array_equal(x, y, equal_nan=True), (cin, cout, peak, scene, adapt)
            checked += 1
    assert checked > 1000


def test_approximate_gamma_chains_bit_exact():
    """allow_approximate_gamma: the oracle's restatement of the AVX2 path
    (FMA matrices, 65537-entry to-linear and 65536-entry to-gamma tables,
    scene-referred HLG) against the reference's own AVX2 sources compiled
    here and run through its cpu AUTO dispatch, over every transfer pair and
    a seeded sample of whole colourspace pairs, on values that reach every
    table edge: the domain ends, out-of-range and huge values, infinities,
    NaN, denormals and exact table points."""
    ref().ref_set_libm(1)
    rng = np.random.default_rng(41)
    edge = np.array([-0.5, 1.5, -0.51, 1.51, 0.0, -0.0, 1.0, 1e-40, -1e-40, 3e9, -3e9, 1e30, np.inf, -np.inf, np.nan,
                     2.0 ** -14, 65504.0, 65520.0, 0.25, 0.5 + 1.0 / 65536], np.float32)
    grid = (np.arange(0, 65537, 97, dtype=np.float32) / 65536 * 2 - 0.5).astype(np.float32)
    base = np.concatenate([edge, grid, rng.uniform(-0.7, 1.7, 3000).astype(np.float32),
                           rng.uniform(0, 1, 1000).astype(np.float32) ** 4])
    pl = [np.roll(base, 7 * k) for k in range(3)]
    csps = [(m, t, p) for m in range(13) for t in range(14) for p in range(11)]
    pairs = [((1, ti, 4), (1, to, 4)) for ti in range(1, 14) for to in range(1, 14) if ti != to]
    pairs += [(csps[i], csps[j]) for i, j in np.random.default_rng(43).integers(0, len(csps), size=(600, 2))]
    pairs += [((3, 4, 4), (7, 4, 6)), ((1, 11, 6), (1, 13, 6)), ((12, 11, 6), (1, 1, 6)), ((8, 4, 6), (1, 1, 6))]
    checked = 0
    for cin, cout in pairs:
        for peak, scene, adapt in [(100.0, 0, 0), (203.0, 0, 1), (1000.0, 1, 0)]:
            n_o, ops, _ = colorspace_ops(cin, cout, peak, scene, adapt, approx=1)
            n_r, a0, a1, a2 = ref_colorspace_apply(cin, cout, *pl, peak, scene, 1, adapt, cpu_auto=True)
            if n_r == -2:
                pytest.skip("host CPU lacks AVX2/FMA/F16C")
            assert (n_o < 0) == (n_r < 0), (cin, cout)
            if n_o < 0:
                continue
            assert n_o == n_r
            b = colorspace_apply(ops, *pl)
            for x, y in zip((a0, a1, a2), b):
                assert np.array_equal(x, y, equal_nan=True), (cin, cout, peak, scene, adapt)
            checked += 1
    assert checked > 400
    # the mode changes results: the tables are approximations of the exact chain
    n_e, ops_e, _ = colorspace_ops((3, 4, 4), (7, 4, 6))
    n_a, ops_a, _ = colorspace_ops((3, 4, 4), (7, 4, 6), approx=1)
    assert n_e == n_a == 5
    ex, ap = colorspace_apply(ops_e, *pl), colorspace_apply(ops_a, *pl)
    assert not np.array_equal(ex[0], ap[0], equal_nan=True)


def test_half_conversions_bit_exact():
    """Every half value, and floats over the whole exponent range incl. ties,
    denormal results, overflow, infinities and quiet NaNs with payloads; bits
    compared, NaN payloads included."""
    o, r = oracle(), ref()
    for h in range(65536):
        a, b = np.float32(o.orc_half_to_float(h)), np.float32(r.ref_half_to_float(h))
        assert a.view(np.uint32) == b.view(np.uint32), hex(h)
    rng = np.random.default_rng(9)
    exps = rng.integers(0, 256, 40000).astype(np.uint32)
    mants = rng.integers(0, 1 << 23, 40000).astype(np.uint32)
    ties = (rng.integers(0, 1 << 10, 2000).astype(np.uint32) << 13) | 0x1000
    tie_exps = rng.integers(100, 145, 2000).astype(np.uint32)
    bits = np.concatenate([(exps << 23) | mants, (tie_exps << 23) | ties,
                           [0x477FEFFF, 0x477FF000, 0x477FF001, 0x33000000, 0x33000001, 0x387FC000, 0x7FC01234, 0xFFC00000]])
    bits = bits | (rng.integers(0, 2, bits.size).astype(np.uint32) << 31)
    for f in bits.astype(np.uint32).view(np.float32):
        if np.isnan(f) and not (f.view(np.uint32) & 0x400000):
            continue  # a signalling NaN is quietened by the by-value call itself
        assert o.orc_float_to_half(float(f)) == r.ref_float_to_half(float(f)), hex(int(f.view(np.uint32)))


def test_blue_noise_fixture_equals_reference_table():
    raw = np.zeros((64, 64), np.uint8)
    ref().ref_blue_noise(ptr(raw, c_u8_p))
    assert np.array_equal(raw, blue_noise())


def test_scale_offset_bit_exact():
    o, r = oracle(), ref()
    fmts = [(0, 8, 0, 0, 0), (0, 8, 1, 0, 0), (0, 8, 0, 1, 0), (1, 10, 0, 0, 0), (1, 10, 0, 1, 0), (1, 12, 1, 0, 0),
            (1, 16, 0, 0, 0), (1, 16, 0, 1, 1), (3, 32, 0, 0, 0), (3, 32, 0, 1, 0), (2, 16, 0, 0, 0)]
    for fi in fmts:
        for fo in fmts:
            so = [ctypes.c_float(), ctypes.c_float()]
            sr = [ctypes.c_float(), ctypes.c_float()]
            o.orc_scale_offset(*fi, *fo, ctypes.byref(so[0]), ctypes.byref(so[1]))
            r.ref_scale_offset(*fi, *fo, ctypes.byref(sr[0]), ctypes.byref(sr[1]))
            assert (so[0].value, so[1].value) == (sr[0].value, sr[1].value)


@pytest.mark.parametrize("up,orig,shift", [(13, 7, 0.0), (100, 37, 0.25), (3840, 1920, 0.0), (2160, 1080, -0.25), (64, 64, 0.5),
                                           (17, 1, 0.0), (1999, 1000, 0.125), (9, 8, -1.5)])
def test_unresize_context_bit_exact(up, orig, shift):
    """create_bilinear_context (unresize/bilinear.cpp:112-172), compiled from
    the reference: the A' band, its offsets and the LU factors of A'A, bit for bit."""
    a = unresize_context(up, orig, shift)
    b = unresize_context(up, orig, shift, use_ref=True)
    assert b is not None and a is not None
    assert a[0] == b[0]
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32))
    # upscaling by unresize is refused by both
    assert unresize_context(7, 13) is None and unresize_context(7, 13, use_ref=True) is None
