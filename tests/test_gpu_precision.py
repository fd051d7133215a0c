"""GPU: the matrix-core precision form of FLOAT resize (banded contraction on
FP16 / BF16 MFMA with hi+lo split operands and FP32 accumulation), judged
against the FP64 oracle (the FP64 resampling matrices applied in FP64) and
against the exact FP32 path. Bounds: FP16 hi+lo within 4e-6 absolute of FP64
on [0, 1] data, BF16 hi+lo within 6e-5; the exact FP32 path itself sits
within 2e-6."""
import json

import numpy as np
import pytest

import pipelines as P
import zimg_amd as z
from oracle_lib import BICUBIC, LANCZOS, resize_plane_f64
from zimg_amd import workloads as W

pytestmark = pytest.mark.gpu

gpu_util = pytest.importorskip("gpu_util")

CASES = [
    # (sw, sh, dw, dh, filter, a, b)  -- the two FP32 precision-form workloads, then small odd shapes
    (3840, 2160, 1920, 1080, LANCZOS, 3.0, 0.0),
    (1280, 720, 3840, 2160, BICUBIC, 0.0, 0.5),
    (333, 217, 1001, 77, LANCZOS, 4.0, 0.0),
    (97, 400, 31, 999, BICUBIC, 0.0, 0.5),
    # two launches through an HBM intermediate (120-tap horizontal pass), and one pass only
    (2000, 64, 100, 500, LANCZOS, 3.0, 0.0),
    (500, 300, 700, 300, LANCZOS, 3.0, 0.0),
]

BOUND = {z.RESIZE_PRECISION_MFMA_FP16: 4e-6, z.RESIZE_PRECISION_MFMA_BF16: 6e-5}


def _run(sw, sh, dw, dh, ft, a, b, mode, src):
    s, d, p = W.fp32_formats(sw, sh, dw, dh, ft, a, b)
    g = z.FilterGraph(s, d, p)
    g.set_resize_precision(mode)
    (out, *_) = gpu_util.run_graph(g, [src, None, None, None], [(dh, dw, np.float32), None, None, None])
    return out


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}-{c[2]}x{c[3]}-f{c[4]}")
def test_mfma_precision_form_error(case):
    sw, sh, dw, dh, ft, a, b = case
    print(f"case: FP32 {sw}x{sh}->{dw}x{dh} filter {ft} precision forms", flush=True)
    src = W.fp32_plane(sw, sh, seed=31)
    exact = _run(sw, sh, dw, dh, ft, a, b, z.RESIZE_PRECISION_EXACT, src)
    np.testing.assert_array_equal(exact.view(np.uint32), P.float_resize(src, dw, dh, ft, a, b).view(np.uint32))
    f64 = resize_plane_f64(src, ft, a, b, dw, dh)
    report = {"case": f"{sw}x{sh}->{dw}x{dh}", "fp32_vs_fp64_max": float(np.max(np.abs(exact - f64)))}
    assert report["fp32_vs_fp64_max"] < 2e-6
    for mode, name in ((z.RESIZE_PRECISION_MFMA_FP16, "fp16"), (z.RESIZE_PRECISION_MFMA_BF16, "bf16")):
        out = _run(sw, sh, dw, dh, ft, a, b, mode, src)
        e64 = out.astype(np.float64) - f64
        e32 = out - exact
        report[name] = {"max_vs_fp64": float(np.max(np.abs(e64))), "rms_vs_fp64": float(np.sqrt(np.mean(e64 ** 2))),
                        "max_vs_fp32": float(np.max(np.abs(e32))), "rms_vs_fp32": float(np.sqrt(np.mean(e32.astype(np.float64) ** 2)))}
        assert report[name]["max_vs_fp64"] < BOUND[mode], (name, report[name])
    print(json.dumps(report), flush=True)


@pytest.mark.parametrize("src_half,dst_half,pad", [(True, False, 0), (False, True, 0), (False, False, 4)],
                         ids=["half-in", "half-out", "unaligned-rows"])
def test_mfma_one_launch_general_loads_and_stores(src_half, dst_half, pad):
    """The one-launch matrix-core form with conversions in its load or store,
    or with rows not 16-byte aligned (its per-sample paths): within the FP16
    hi+lo bound of the FP64 oracle applied to the same source values."""
    sw, sh, dw, dh = 701, 333, 350, 505
    print(f"case: precision form {sw}x{sh}->{dw}x{dh} half_in={src_half} half_out={dst_half} pad={pad}", flush=True)
    src = W.fp32_plane(sw, sh, seed=12)
    if src_half:
        src = src.astype(np.float16)
    s = z.image_format(width=sw, height=sh, pixel_type=z.PIXEL_HALF if src_half else z.PIXEL_FLOAT)
    d = z.image_format(width=dw, height=dh, pixel_type=z.PIXEL_HALF if dst_half else z.PIXEL_FLOAT)
    g = z.FilterGraph(s, d, z.builder_params(resample_filter=z.RESIZE_LANCZOS, filter_param_a=3.0))
    g.set_resize_precision(z.RESIZE_PRECISION_MFMA_FP16)
    assert "one launch" in g.describe()
    (out, *_) = gpu_util.run_graph(g, [src, None, None, None], [(dh, dw, np.float16 if dst_half else np.float32), None, None, None],
                                   pad_bytes=pad)
    f64 = resize_plane_f64(src.astype(np.float64), LANCZOS, 3.0, 0.0, dw, dh)
    err = np.max(np.abs(out.astype(np.float64) - f64))
    # HALF output adds its own rounding: half an ulp, 2^-11 for values in [1, 2)
    # (Lanczos overshoot reaches past 1).
    bound = 5.0e-4 if dst_half else BOUND[z.RESIZE_PRECISION_MFMA_FP16]
    assert err < bound, err


def test_mfma_form_keeps_exact_integer_and_dithered_paths():
    """WORD resizes stay on the exact integer path whatever the setting."""
    src = W.word_plane(640, 360, seed=4)
    s = z.image_format(width=640, height=360, pixel_type=z.PIXEL_WORD, depth=16)
    d = z.image_format(width=1280, height=720, pixel_type=z.PIXEL_WORD, depth=16)
    g = z.FilterGraph(s, d, z.builder_params(resample_filter=z.RESIZE_LANCZOS, filter_param_a=3.0))
    g.set_resize_precision(z.RESIZE_PRECISION_MFMA_BF16)
    (out, *_) = gpu_util.run_graph(g, [src, None, None, None], [(720, 1280, np.uint16), None, None, None])
    np.testing.assert_array_equal(out, P.word_resize(src, 1280, 720))
