"""GPU parity of the wider graph lowering against oracle compositions:
every colourspace operation family (all transfer functions both ways, the
matrix families including constant luminance and ICtCp, gamut changes with and
without chromatic adaptation, scene- and display-referred), fractional active
regions, integer subrectangles (CopyRect), interlaced field shifts, and the
grey <-> colour conversions.

The oracle side derives each chain from the reference's rules (plane geometry
of graphbuilder.cpp in tests/pipelines.py, pinned against the reference's
traces by test_api_cpu.py; operation chains from the oracle's own BFS), never
from this library's planner."""
import numpy as np
import pytest

import pipelines as P
import zimg_amd as z
from oracle_lib import (BICUBIC, BILINEAR, FLOAT, LANCZOS, SPLINE36, WORD, colorspace_apply, colorspace_ops, dither,
                        scale_offset, to_float)

pytestmark = pytest.mark.gpu

gpu_util = pytest.importorskip("gpu_util")

API_2_5 = (2 << 8) | 5

# C API code -> the reference's internal enumeration (zimg.cpp:126-238).
MATRIX = {z.MATRIX_RGB: 1, z.MATRIX_BT709: 3, z.MATRIX_FCC: 4, z.MATRIX_BT470_BG: 2, z.MATRIX_ST170_M: 2,
          z.MATRIX_ST240_M: 5, z.MATRIX_YCGCO: 6, z.MATRIX_BT2020_NCL: 7, z.MATRIX_BT2020_CL: 8,
          z.MATRIX_CHROMATICITY_DERIVED_NCL: 9, z.MATRIX_CHROMATICITY_DERIVED_CL: 10, z.MATRIX_BT2100_ICTCP: 12}
TRANSFER = {z.TRANSFER_BT709: 4, z.TRANSFER_BT601: 4, z.TRANSFER_BT2020_10: 4, z.TRANSFER_BT2020_12: 4,
            z.TRANSFER_ST240_M: 7, z.TRANSFER_BT470_M: 5, z.TRANSFER_BT470_BG: 6, z.TRANSFER_IEC_61966_2_4: 8,
            z.TRANSFER_BT1361: 9, z.TRANSFER_IEC_61966_2_1: 10, z.TRANSFER_LINEAR: 1, z.TRANSFER_LOG_100: 2,
            z.TRANSFER_LOG_316: 3, z.TRANSFER_ST2084: 11, z.TRANSFER_ST428: 12, z.TRANSFER_ARIB_B67: 13}
PRIMARIES = {z.PRIMARIES_BT470_M: 1, z.PRIMARIES_BT470_BG: 2, z.PRIMARIES_BT709: 4, z.PRIMARIES_ST170_M: 3,
             z.PRIMARIES_ST240_M: 3, z.PRIMARIES_FILM: 5, z.PRIMARIES_BT2020: 6, z.PRIMARIES_ST428: 7,
             z.PRIMARIES_ST431_2: 8, z.PRIMARIES_ST432_1: 9, z.PRIMARIES_EBU3213_E: 10}
# one API code per distinct internal transfer / primaries
TRANSFERS = [z.TRANSFER_BT709, z.TRANSFER_ST240_M, z.TRANSFER_BT470_M, z.TRANSFER_BT470_BG, z.TRANSFER_IEC_61966_2_4,
             z.TRANSFER_BT1361, z.TRANSFER_IEC_61966_2_1, z.TRANSFER_LINEAR, z.TRANSFER_LOG_100, z.TRANSFER_LOG_316,
             z.TRANSFER_ST2084, z.TRANSFER_ST428, z.TRANSFER_ARIB_B67]
PRIMS = [z.PRIMARIES_BT470_M, z.PRIMARIES_BT470_BG, z.PRIMARIES_BT709, z.PRIMARIES_ST170_M, z.PRIMARIES_FILM,
         z.PRIMARIES_BT2020, z.PRIMARIES_ST428, z.PRIMARIES_ST431_2, z.PRIMARIES_ST432_1, z.PRIMARIES_EBU3213_E]


def _assert_bits(got, want, what):
    assert got.shape == want.shape, what
    if got.dtype == np.float32:
        g, w = got.view(np.uint32), want.view(np.uint32)
        bad = (g != w) & ~(np.isnan(got) & np.isnan(want))
    else:
        bad = got != want
    n = int(np.count_nonzero(bad))
    if n:
        idx = np.argwhere(bad)[0]
        pytest.fail(f"{what}: {n} mismatches, first at {idx.tolist()}: got {got[tuple(idx)]!r} want {want[tuple(idx)]!r}")


def _float_fmt(w, h, m, t, p, **kw):
    fam = z.COLOR_RGB if m == z.MATRIX_RGB else z.COLOR_YUV
    return z.image_format(width=w, height=h, pixel_type=z.PIXEL_FLOAT, color_family=fam, matrix_coefficients=m,
                          transfer_characteristics=t, color_primaries=p, **kw)


def _csp_pairs():
    pairs = []
    # every transfer to every transfer, through linear light (RGB, BT.709 primaries)
    for ti in TRANSFERS:
        for to in TRANSFERS:
            if ti != to:
                pairs.append(((z.MATRIX_RGB, ti, z.PRIMARIES_BT709), (z.MATRIX_RGB, to, z.PRIMARIES_BT709)))
    # every matrix to every other, with the transfers constant luminance and ICtCp need
    mats = [m for m in MATRIX if m not in (z.MATRIX_ST170_M,)]
    for mi in mats:
        for mo in mats:
            if mi == mo:
                continue
            for t in (z.TRANSFER_BT709, z.TRANSFER_ST2084, z.TRANSFER_ARIB_B67):
                pairs.append(((mi, t, z.PRIMARIES_BT2020), (mo, t, z.PRIMARIES_BT2020)))
    # every gamut change in linear light and through a gamma
    for pi in PRIMS:
        for po in PRIMS:
            if pi != po:
                pairs.append(((z.MATRIX_RGB, z.TRANSFER_LINEAR, pi), (z.MATRIX_RGB, z.TRANSFER_LINEAR, po)))
                pairs.append(((z.MATRIX_BT709, z.TRANSFER_IEC_61966_2_1, pi), (z.MATRIX_BT2020_NCL, z.TRANSFER_ST2084, po)))
    # seeded sample of whole triples
    r = np.random.default_rng(31)
    mats_l, trs_l, prs_l = list(MATRIX), list(TRANSFER), list(PRIMARIES)
    for _ in range(300):
        a = (mats_l[r.integers(len(mats_l))], trs_l[r.integers(len(trs_l))], prs_l[r.integers(len(prs_l))])
        b = (mats_l[r.integers(len(mats_l))], trs_l[r.integers(len(trs_l))], prs_l[r.integers(len(prs_l))])
        if a != b:
            pairs.append((a, b))
    return pairs


PARAM_SETS = [(100.0, 0, 0), (203.0, 0, 1), (1000.0, 1, 0)]


@pytest.mark.parametrize("peak,scene,adapt", PARAM_SETS)
def test_colorspace_every_family_on_device(peak, scene, adapt):
    """k_colorspace against the oracle's operation chains, bit for bit, for
    every transfer pair (so every device libm routine: powf, expf, logf,
    log10f), every matrix pair (NCL, CL, YCgCo, ICtCp, chromaticity-derived),
    every gamut pair, and 300 random triples. Unreachable pairs must fail to
    build with NO_COLORSPACE_CONVERSION exactly where the oracle finds no path."""
    print(f"case: colourspace families peak {peak} scene {scene} adapt {adapt}", flush=True)
    w, h = 67, 13
    r = np.random.default_rng(7)
    planes = [r.uniform(-0.2, 1.2, (h, w)).astype(np.float32) for _ in range(3)]
    planes[0][0, :8] = [0.0, 1.0, 1e-6, 1e-30, 0.018, 0.5, 1e-3, 2.0]  # exact endpoints and tiny values
    params = z.builder_params(version=API_2_5, nominal_peak_luminance=peak, scene_referred=scene, chromatic_adaptation=adapt)
    built = nopath = 0
    for cin, cout in _csp_pairs():
        ii = (MATRIX[cin[0]], TRANSFER[cin[1]], PRIMARIES[cin[2]])
        oo = (MATRIX[cout[0]], TRANSFER[cout[1]], PRIMARIES[cout[2]])
        n, ops, _ = colorspace_ops(ii, oo, peak, scene, adapt)
        try:
            g = z.FilterGraph(_float_fmt(w, h, *cin), _float_fmt(w, h, *cout), params)
        except z.ZimgError as e:
            if e.code == z.ERROR_UNKNOWN and z.PRIMARIES_ST428 in (cin[2], cout[2]):
                # luma coefficients derived from XYZ "primaries" (chromaticity-derived
                # matrices, the HLG OOTF): the reference raises InternalError there
                # too (colorspace_param.cpp:70-101), and the oracle finds no chain
                assert n < 0, (cin, cout)
                continue
            assert n < 0 and e.code == z.ERROR_NO_COLORSPACE_CONVERSION, (cin, cout, e)
            nopath += 1
            continue
        assert n >= 0, (cin, cout)
        outs = gpu_util.run_graph(g, planes + [None], [(h, w, np.float32)] * 3 + [None])
        want = colorspace_apply(ops, *planes) if n else planes
        for k in range(3):
            _assert_bits(outs[k], want[k], f"{cin}->{cout} plane {k}")
        built += 1
    assert built > 700 and nopath > 10, (built, nopath)


@pytest.mark.parametrize("peak,scene,adapt", PARAM_SETS)
def test_colorspace_approximate_gamma_on_device(peak, scene, adapt):
    """allow_approximate_gamma: the device runs the reference's AVX2 chain
    (FMA matrices, lookup-table gamma, scene-referred HLG), bit for bit
    against the oracle's restatement of it (itself pinned against the
    compiled AVX2 sources), over every transfer pair, the matrix and gamut
    pairs and the random triples, on values that reach the table edges."""
    print(f"case: approximate gamma peak {peak} scene {scene} adapt {adapt}", flush=True)
    w, h = 67, 13
    r = np.random.default_rng(8)
    planes = [r.uniform(-0.7, 1.7, (h, w)).astype(np.float32) for _ in range(3)]
    planes[0][0, :16] = [-0.5, 1.5, -0.51, 1.51, 0.0, 1.0, 1e-40, 3e9, -3e9, np.inf, -np.inf, np.nan, 65504.0, 65520.0, 0.25, 2.0 ** -14]
    params = z.builder_params(version=API_2_5, nominal_peak_luminance=peak, scene_referred=scene, chromatic_adaptation=adapt,
                              allow_approximate_gamma=1)
    built = 0
    for cin, cout in _csp_pairs():
        ii = (MATRIX[cin[0]], TRANSFER[cin[1]], PRIMARIES[cin[2]])
        oo = (MATRIX[cout[0]], TRANSFER[cout[1]], PRIMARIES[cout[2]])
        n, ops, _ = colorspace_ops(ii, oo, peak, scene, adapt, approx=1)
        try:
            g = z.FilterGraph(_float_fmt(w, h, *cin), _float_fmt(w, h, *cout), params)
        except z.ZimgError as e:
            assert n < 0 and e.code in (z.ERROR_NO_COLORSPACE_CONVERSION, z.ERROR_UNKNOWN), (cin, cout, e)
            continue
        assert n >= 0, (cin, cout)
        outs = gpu_util.run_graph(g, planes + [None], [(h, w, np.float32)] * 3 + [None])
        want = colorspace_apply(ops, *planes) if n else planes
        for k in range(3):
            _assert_bits(outs[k], want[k], f"{cin}->{cout} plane {k}")
        built += 1
    assert built > 700, built


def test_fhd_709_to_2020_approximate_gamma():
    """The 4:2:0 10-bit BT.709 -> BT.2020 chain with allow_approximate_gamma:
    lookup-table gamma in the colourspace stage, bit for bit."""
    w, h = 640, 360
    print(f"case: 4:2:0 10-bit 709->2020 approximate gamma {w}x{h}", flush=True)
    from zimg_amd import workloads as W
    Y, U, V = W.yuv420_10bit(w, h, seed=21)
    s, d, p = W.fhd_709_to_2020_formats(w, h)
    p.allow_approximate_gamma = b"\x01"
    g = z.FilterGraph(s, d, p)
    assert "lut_gamma" in g.describe() and "lut_linear" in g.describe()
    outs = _run420(g, [Y, U, V], w, h)
    for got, ref, name in zip(outs[:3], P.yuv420_10_709_to_2020(Y, U, V, approx=1), "YUV"):
        _assert_bits(got, ref, f"plane {name}")


def _word420_fmt(w, h, depth=10, **kw):
    d = dict(width=w, height=h, pixel_type=z.PIXEL_WORD, depth=depth, subsample_w=1, subsample_h=1, color_family=z.COLOR_YUV,
             matrix_coefficients=z.MATRIX_BT709, transfer_characteristics=z.TRANSFER_BT709, color_primaries=z.PRIMARIES_BT709)
    d.update(kw)
    return z.image_format(**d)


def _yuv420_planes(w, h, seed, depth=10):
    r = np.random.default_rng(seed)
    top = (1 << depth) - 1
    return [r.integers(0, top + 1, (h, w)).astype(np.uint16), r.integers(0, top + 1, (h // 2, w // 2)).astype(np.uint16),
            r.integers(0, top + 1, (h // 2, w // 2)).astype(np.uint16)]


def _run420(g, planes, w, h):
    return gpu_util.run_graph(g, planes + [None], [(h, w, np.uint16), (h // 2, w // 2, np.uint16), (h // 2, w // 2, np.uint16), None])


@pytest.mark.parametrize("parity_in,parity_out,loc,dh", [(1, 0, z.CHROMA_LEFT, 240), (2, 0, z.CHROMA_CENTER, 240),
                                                         (1, 1, z.CHROMA_BOTTOM_LEFT, 480), (0, 2, z.CHROMA_TOP, 120)])
def test_interlaced_field_shifts(parity_in, parity_out, loc, dh):
    """Field parity moves luma by a quarter line and chroma by the siting-adjusted
    offset (graphbuilder.cpp:113-171, 348-366): WORD 4:2:0 10-bit, bicubic."""
    w, h = 320, 240
    print(f"case: interlaced {parity_in}->{parity_out} loc {loc} {w}x{h}->{w}x{dh}", flush=True)
    planes = _yuv420_planes(w, h, 11 + parity_in)
    src = _word420_fmt(w, h, field_parity=parity_in, chroma_location=loc)
    dst = _word420_fmt(w, dh, field_parity=parity_out, chroma_location=loc)
    g = z.FilterGraph(src, dst, z.builder_params(resample_filter=z.RESIZE_BICUBIC, resample_filter_uv=z.RESIZE_BICUBIC))
    outs = gpu_util.run_graph(g, planes + [None], [(dh, w, np.uint16), (dh // 2, w // 2, np.uint16), (dh // 2, w // 2, np.uint16), None])
    gs = P.plane_geometry(w, h, ssw=1, ssh=1, parity=parity_in, chroma_location=loc)
    gd = P.plane_geometry(w, dh, ssw=1, ssh=1, parity=parity_out, chroma_location=loc)
    for q in range(3):
        want = P.resize_geom(planes[q], gs[q], gd[q], BICUBIC, 0.0, 0.5, depth=10)
        _assert_bits(outs[q], want, f"plane {q}")


def test_active_region_fractional_resize():
    """A fractional source window (zimg_image_format.active_region) on a 4:2:0
    WORD image with LEFT chroma siting, Lanczos3 luma and spline36 chroma."""
    w, h, dw, dh = 300, 200, 412, 258
    act = (10.5, 7.25, 250.3, 160.7)
    print(f"case: active region {act} {w}x{h}->{dw}x{dh}", flush=True)
    planes = _yuv420_planes(w, h, 5)
    src = _word420_fmt(w, h, active_left=act[0], active_top=act[1], active_width=act[2], active_height=act[3])
    dst = _word420_fmt(dw, dh)
    g = z.FilterGraph(src, dst, z.builder_params(resample_filter=z.RESIZE_LANCZOS, filter_param_a=3.0,
                                                 resample_filter_uv=z.RESIZE_SPLINE36))
    outs = _run420(g, planes, dw, dh)
    gs = P.plane_geometry(w, h, ssw=1, ssh=1, active=act)
    gd = P.plane_geometry(dw, dh, ssw=1, ssh=1)
    for q in range(3):
        f = (LANCZOS, 3.0, 0.0) if q == 0 else (SPLINE36, 0.0, 0.0)
        want = P.resize_geom(planes[q], gs[q], gd[q], *f, depth=10)
        _assert_bits(outs[q], want, f"plane {q}")


def test_active_region_float_grey():
    w, h, dw, dh = 257, 131, 100, 333
    act = (-3.5, 2.75, 240.0, 120.5)
    print(f"case: FLOAT grey active region {act}", flush=True)
    img = np.random.default_rng(8).random((h, w), dtype=np.float32)
    src = z.image_format(width=w, height=h, pixel_type=z.PIXEL_FLOAT, active_left=act[0], active_top=act[1], active_width=act[2],
                         active_height=act[3])
    dst = z.image_format(width=dw, height=dh, pixel_type=z.PIXEL_FLOAT)
    g = z.FilterGraph(src, dst, z.builder_params(resample_filter=z.RESIZE_SPLINE36))
    (out, *_) = gpu_util.run_graph(g, [img, None, None, None], [(dh, dw, np.float32), None, None, None])
    want = P.resize_geom(img, P.plane_geometry(w, h, grey=True, active=act)[0], P.plane_geometry(dw, dh, grey=True)[0],
                         SPLINE36, 0.0, 0.0)
    _assert_bits(out, want, "grey")


def test_integer_subrectangle_is_a_copy():
    """A pixel-aligned window equal to the output size lowers to CopyRect
    (graphbuilder.cpp:540-571, 920-937): the output is the window itself."""
    w, h, dw, dh, left, top = 300, 200, 200, 150, 32, 24
    print(f"case: subrectangle ({left},{top}) {dw}x{dh} of {w}x{h}", flush=True)
    planes = _yuv420_planes(w, h, 6)
    src = _word420_fmt(w, h, active_left=float(left), active_top=float(top), active_width=float(dw), active_height=float(dh))
    g = z.FilterGraph(src, _word420_fmt(dw, dh))
    assert any("subrectangle" in t for t in g.plan_trace())
    outs = _run420(g, planes, dw, dh)
    _assert_bits(outs[0], planes[0][top:top + dh, left:left + dw], "luma")
    for q in (1, 2):
        _assert_bits(outs[q], planes[q][top // 2:(top + dh) // 2, left // 2:(left + dw) // 2], f"chroma {q}")


def test_grey_to_yuv_and_rgb():
    """grey -> YUV fills chroma with the neutral code (graphbuilder.cpp:629-658);
    grey -> RGB repeats luma in all three planes (:615-627)."""
    w, h = 130, 66
    print("case: grey -> YUV 4:2:0 WORD, grey -> RGB FLOAT", flush=True)
    y = np.random.default_rng(9).integers(0, 1024, (h, w)).astype(np.uint16)
    src = z.image_format(width=w, height=h, pixel_type=z.PIXEL_WORD, depth=10, transfer_characteristics=z.TRANSFER_BT709,
                         color_primaries=z.PRIMARIES_BT709)
    g = z.FilterGraph(src, _word420_fmt(w, h, matrix_coefficients=z.MATRIX_BT709))
    assert g.plan_trace() == ["grey_to_yuv"]
    outs = _run420(g, [y, None, None], w, h)
    _assert_bits(outs[0], y, "luma")
    for q in (1, 2):
        _assert_bits(outs[q], np.full((h // 2, w // 2), 512, np.uint16), f"chroma {q}")
    yf = np.random.default_rng(10).random((h, w), dtype=np.float32)
    g = z.FilterGraph(z.image_format(width=w, height=h, pixel_type=z.PIXEL_FLOAT),
                      _float_fmt(w, h, z.MATRIX_RGB, z.TRANSFER_UNSPECIFIED, z.PRIMARIES_UNSPECIFIED))
    outs = gpu_util.run_graph(g, [yf, None, None, None], [(h, w, np.float32)] * 3 + [None])
    for q in range(3):
        _assert_bits(outs[q], yf, f"rgb {q}")


def test_grey_colorspace_round_trips():
    """grey_to_grey_colorspace / yuv_to_grey_colorspace (graphbuilder_test.cpp
    traces): grey is expanded to RGB with the RGB matrix, converted, and the
    luma of the result kept."""
    w, h = 96, 40
    print("case: grey and YUV to grey through a gamut change", flush=True)
    r = np.random.default_rng(12)
    yf = r.random((h, w), dtype=np.float32)
    src = z.image_format(width=w, height=h, pixel_type=z.PIXEL_FLOAT, matrix_coefficients=z.MATRIX_BT709,
                         transfer_characteristics=z.TRANSFER_BT709, color_primaries=z.PRIMARIES_ST431_2)
    dst = z.image_format(width=w, height=h, pixel_type=z.PIXEL_FLOAT, matrix_coefficients=z.MATRIX_BT470_BG,
                         transfer_characteristics=z.TRANSFER_BT709, color_primaries=z.PRIMARIES_BT2020)
    g = z.FilterGraph(src, dst)
    assert g.plan_trace() == ["grey_to_rgb", g.plan_trace()[1], "yuv_to_grey"]
    (out, *_) = gpu_util.run_graph(g, [yf, None, None, None], [(h, w, np.float32), None, None, None])
    n, ops, _ = colorspace_ops((1, 4, 8), (2, 4, 6))
    _assert_bits(out, colorspace_apply(ops, yf, yf, yf)[0], "grey->grey")

    planes = [r.random((h, w), dtype=np.float32), r.uniform(-0.5, 0.5, (h, w)).astype(np.float32),
              r.uniform(-0.5, 0.5, (h, w)).astype(np.float32)]
    g = z.FilterGraph(_float_fmt(w, h, z.MATRIX_BT709, z.TRANSFER_BT709, z.PRIMARIES_ST431_2), dst)
    (out, *_) = gpu_util.run_graph(g, planes + [None], [(h, w, np.float32), None, None, None])
    n, ops, _ = colorspace_ops((3, 4, 8), (2, 4, 6))
    _assert_bits(out, colorspace_apply(ops, *planes)[0], "yuv->grey")


@pytest.mark.parametrize("matrix,transfer", [(z.MATRIX_BT2020_CL, z.TRANSFER_BT2020_10), (z.MATRIX_BT2100_ICTCP, z.TRANSFER_ST2084),
                                             (z.MATRIX_BT2100_ICTCP, z.TRANSFER_ARIB_B67)])
def test_cl_and_ictcp_420_to_rgb(matrix, transfer):
    """4:2:0 10-bit constant-luminance / ICtCp to RGB FLOAT 4:4:4 BT.709: depth
    to float, chroma upsample (bilinear, the API default), the operation chain."""
    w, h = 192, 108
    print(f"case: 4:2:0 10-bit matrix {matrix} transfer {transfer} -> RGB FLOAT", flush=True)
    planes = _yuv420_planes(w, h, 13)
    planes[0] = np.clip(planes[0], 64, 940)
    planes[1] = np.clip(planes[1], 64, 960)
    planes[2] = np.clip(planes[2], 64, 960)
    src = _word420_fmt(w, h, matrix_coefficients=matrix, transfer_characteristics=transfer, color_primaries=z.PRIMARIES_BT2020)
    dst = _float_fmt(w, h, z.MATRIX_RGB, z.TRANSFER_BT709, z.PRIMARIES_BT709)
    g = z.FilterGraph(src, dst)
    outs = gpu_util.run_graph(g, planes + [None], [(h, w, np.float32)] * 3 + [None])
    s, o = scale_offset((WORD, 10, 0, 0, 0), (FLOAT, 32, 0, 0, 0))
    Yf = to_float(planes[0], WORD, s, o)
    s, o = scale_offset((WORD, 10, 0, 1, 0), (FLOAT, 32, 0, 1, 0))
    gs = P.plane_geometry(w, h, ssw=1, ssh=1)
    g444 = P.plane_geometry(w, h)
    Uf, Vf = (P.resize_geom(to_float(planes[q], WORD, s, o), gs[q], g444[q], BILINEAR, 0.0, 0.0) for q in (1, 2))
    n, ops, _ = colorspace_ops((MATRIX[matrix], TRANSFER[transfer], 6), (1, 4, 4))
    assert n > 0
    want = colorspace_apply(ops, Yf, Uf, Vf)
    for q in range(3):
        _assert_bits(outs[q], want[q], f"plane {q}")


@pytest.mark.parametrize("sw,sh,dw,dh", [(640, 360, 333, 200), (500, 300, 500, 143), (777, 101, 300, 101), (64, 64, 64, 64)])
def test_unresize_float_bit_exact(sw, sh, dw, dh):
    """Unresize (resample_filter ZIMG_RESIZE_UNRESIZE): FLOAT planes, both
    passes or one, against the oracle's restatement of unresize_impl.cpp."""
    from oracle_lib import unresize_plane
    print(f"case: unresize FLOAT {sw}x{sh}->{dw}x{dh}", flush=True)
    img = np.random.default_rng(sw + dh).random((2, sh, sw), dtype=np.float32)
    g = z.FilterGraph(z.image_format(width=sw, height=sh, pixel_type=z.PIXEL_FLOAT),
                      z.image_format(width=dw, height=dh, pixel_type=z.PIXEL_FLOAT), z.builder_params(resample_filter=-1))
    (out, *_) = gpu_util.run_graph(g, [img, None, None, None], [(dh, dw, np.float32), None, None, None], pad_bytes=12)
    for k in range(2):
        _assert_bits(out[k], unresize_plane(img[k], dw, dh), f"frame {k}")


def test_unresize_yuv420_word_chain():
    """4:2:0 10-bit WORD unresized horizontally: to FLOAT, unresize with the
    LEFT-siting chroma shift, back to WORD (graphbuilder.cpp:721-724,761-796)."""
    from oracle_lib import unresize_plane, dither as odither
    w, h, dw = 640, 360, 320
    print(f"case: unresize 4:2:0 WORD {w}x{h}->{dw}x{h}", flush=True)
    planes = _yuv420_planes(w, h, 31)
    g = z.FilterGraph(_word420_fmt(w, h), _word420_fmt(dw, h), z.builder_params(resample_filter=-1))
    outs = gpu_util.run_graph(g, planes + [None], [(h, dw, np.uint16), (h // 2, dw // 2, np.uint16), (h // 2, dw // 2, np.uint16), None])
    gs, gd = P.plane_geometry(w, h, ssw=1, ssh=1), P.plane_geometry(dw, h, ssw=1, ssh=1)
    for q in range(3):
        chroma = q > 0
        s_, o_ = scale_offset((WORD, 10, 0, int(chroma), 0), (FLOAT, 32, 0, 0, 0))
        f = to_float(planes[q], WORD, s_, o_)
        shift_w = P.resize_params(gs[q], gd[q])[0]
        r = unresize_plane(f, gd[q][0], gd[q][1], shift_w=shift_w)
        s2, o2 = scale_offset((FLOAT, 32, 0, 0, 0), (WORD, 10, 0, int(chroma), 0))
        _assert_bits(outs[q], odither(r, FLOAT, WORD, s2, o2, 10, plane=q), f"plane {q}")
