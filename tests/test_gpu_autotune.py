"""GPU: zimg_filter_graph_autotune_device times the candidate tile shapes of a
graph's tiled resize steps on the caller's buffers and keeps the fastest; the
outputs stay the oracle's, bit for bit, whichever shape it keeps."""
import numpy as np
import pytest

import pipelines as P
import zimg_amd as z
from zimg_amd import workloads as W

pytestmark = pytest.mark.gpu

gpu_util = pytest.importorskip("gpu_util")


@pytest.mark.parametrize("case", ["up3", "down2"])
def test_autotuned_graph_equals_oracle(case):
    if case == "up3":
        sw, sh, dw, dh, ft, a, b = 1280, 720, 3840, 2160, z.RESIZE_BICUBIC, 0.0, 0.5
    else:
        sw, sh, dw, dh, ft, a, b = 1920, 1080, 960, 540, z.RESIZE_LANCZOS, 3.0, 0.0
    print(f"case: FLOAT {sw}x{sh} -> {dw}x{dh} filter {ft}, 2 frames, autotuned", flush=True)
    frames = np.stack([W.fp32_plane(sw, sh, seed=31), W.fp32_plane(sw, sh, seed=32)])
    g = z.FilterGraph(*W.fp32_formats(sw, sh, dw, dh, ft, a, b))
    before = g.describe()
    (out, *_) = gpu_util.run_graph(g, [frames, None, None, None], [(dh, dw, np.float32), None, None, None], autotune=True)
    print("tile before:", [l for l in before.splitlines() if "tiled=" in l], "after:", [l for l in g.describe().splitlines() if "tiled=" in l], flush=True)
    for f in range(2):
        want = P.float_resize(frames[f], dw, dh, ft, a, b)
        np.testing.assert_array_equal(out[f].view(np.uint32), want.view(np.uint32), err_msg=f"frame {f}")
    # Processing again with the kept shape gives the same bits.
    (again, *_) = gpu_util.run_graph(g, [frames, None, None, None], [(dh, dw, np.float32), None, None, None])
    np.testing.assert_array_equal(again.view(np.uint32), out.view(np.uint32))
