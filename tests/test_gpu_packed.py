"""GPU: packed layouts in HBM (zimg_filter_graph_process_device_packed):
NV12, P010 / P016 and YUY2 / UYVY sources and destinations give exactly what
the planar device path gives on the same samples, and the packing itself is
the layout's definition (the YUY2 byte order of src/testapp/frame.cpp:
334-385, 506-550; NV12 / P010 semi-planar chroma)."""
import numpy as np
import pytest

import zimg_amd as z

pytestmark = pytest.mark.gpu

gpu_util = pytest.importorskip("gpu_util")


def _fmt(w, h, t, depth, ssw, ssh):
    return z.image_format(width=w, height=h, pixel_type=t, depth=depth, subsample_w=ssw, subsample_h=ssh, color_family=z.COLOR_YUV,
                          matrix_coefficients=z.MATRIX_BT709, transfer_characteristics=z.TRANSFER_BT709, color_primaries=z.PRIMARIES_BT709)


def _nv12(Y, U, V, shift):
    uv = np.empty((U.shape[0], U.shape[1] * 2), Y.dtype)
    uv[:, 0::2], uv[:, 1::2] = U, V
    return [(Y.astype(np.uint32) << shift).astype(Y.dtype), (uv.astype(np.uint32) << shift).astype(Y.dtype)]


def _yuy2(Y, U, V, uyvy=False):
    h, w = Y.shape
    p = np.empty((h, 2 * w), np.uint8)
    if uyvy:
        p[:, 1::2], p[:, 0::4], p[:, 2::4] = Y, U, V
    else:
        p[:, 0::2], p[:, 1::4], p[:, 3::4] = Y, U, V
    return [p]


def _dev(torch, arrs, frames=1):
    keep, planes = [], []
    for a in arrs:
        t = torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1)).to("cuda")
        keep.append(t)
        planes.append((t.data_ptr(), a.shape[-1] * a.itemsize, a.shape[-2] * a.shape[-1] * a.itemsize))
    return keep, planes


def _out(torch, shapes):
    keep, planes = [], []
    for (h, w, dt) in shapes:
        t = torch.full((h * w * np.dtype(dt).itemsize,), 0x77, dtype=torch.uint8, device="cuda")
        keep.append((t, h, w, dt))
        planes.append((t.data_ptr(), w * np.dtype(dt).itemsize, h * w * np.dtype(dt).itemsize))
    return keep, planes


def _read(torch, keep):
    torch.cuda.synchronize()
    return [t.cpu().numpy().view(dt).reshape(h, w) for (t, h, w, dt) in keep]


@pytest.mark.parametrize("src_pack,dst_pack,word,shift", [
    (z.PACKING_NV12, z.PACKING_NV12, False, 0),   # NV12 -> NV12
    (z.PACKING_NV12, z.PACKING_NV12, True, 6),    # P010 -> P010
    (z.PACKING_NV12, z.PACKING_PLANAR, True, 0),  # P016 -> planar
    (z.PACKING_PLANAR, z.PACKING_NV12, False, 0), # planar -> NV12
])
def test_nv12_p010_resize_equals_planar(src_pack, dst_pack, word, shift):
    import torch
    w, h, dw, dh = 322, 180, 640, 360
    depth = 10 if (word and shift == 6) else (16 if word else 8)
    t = z.PIXEL_WORD if word else z.PIXEL_BYTE
    dt = np.uint16 if word else np.uint8
    print(f"case: packed {src_pack}->{dst_pack} word {word} shift {shift} {w}x{h}->{dw}x{dh}", flush=True)
    r = np.random.default_rng(w + shift)
    top = 1 << depth
    Y, U, V = r.integers(0, top, (h, w)).astype(dt), r.integers(0, top, (h // 2, w // 2)).astype(dt), \
        r.integers(0, top, (h // 2, w // 2)).astype(dt)
    g = z.FilterGraph(_fmt(w, h, t, depth, 1, 1), _fmt(dw, dh, t, depth, 1, 1), z.builder_params(resample_filter=z.RESIZE_SPLINE36,
                                                                                                  resample_filter_uv=z.RESIZE_SPLINE36))
    specs = [(dh, dw, dt), (dh // 2, dw // 2, dt), (dh // 2, dw // 2, dt)]
    want = gpu_util.run_graph(g, [Y, U, V, None], specs + [None])[:3]
    src_arrs = _nv12(Y, U, V, shift) if src_pack == z.PACKING_NV12 else [Y, U, V]
    ks, sp = _dev(torch, src_arrs)
    if dst_pack == z.PACKING_NV12:
        kd, dp = _out(torch, [(dh, dw, dt), (dh // 2, dw, dt)])
    else:
        kd, dp = _out(torch, specs)
    src = z.packed_buffer(src_pack, sp, shift if src_pack == z.PACKING_NV12 else 0)
    dst = z.packed_buffer(dst_pack, dp, shift if dst_pack == z.PACKING_NV12 else 0)
    tmp = torch.empty(max(g.device_packed_tmp_size(1, src_pack, dst_pack), 1), dtype=torch.uint8, device="cuda")
    g.process_device_packed(1, src, dst, tmp.data_ptr(), 0)
    got = _read(torch, kd)
    if dst_pack == z.PACKING_NV12:
        expect = _nv12(*want, shift)
    else:
        expect = want
    for q, (a, b) in enumerate(zip(got, expect)):
        np.testing.assert_array_equal(a, b, err_msg=f"plane {q}")


@pytest.mark.parametrize("uyvy", [False, True])
def test_yuy2_to_i420_and_back(uyvy):
    """YUY2 / UYVY 4:2:2 BYTE: to 4:2:0 planar, and 4:2:0 planar to the packed
    layout, both equal to the planar path on the unpacked samples."""
    import torch
    w, h = 320, 180
    pack = z.PACKING_UYVY if uyvy else z.PACKING_YUY2
    print(f"case: {'UYVY' if uyvy else 'YUY2'} {w}x{h} <-> 4:2:0", flush=True)
    r = np.random.default_rng(9 + uyvy)
    Y, U, V = r.integers(16, 236, (h, w)).astype(np.uint8), r.integers(16, 241, (h, w // 2)).astype(np.uint8), \
        r.integers(16, 241, (h, w // 2)).astype(np.uint8)
    g = z.FilterGraph(_fmt(w, h, z.PIXEL_BYTE, 8, 1, 0), _fmt(w, h, z.PIXEL_BYTE, 8, 1, 1))
    specs = [(h, w, np.uint8), (h // 2, w // 2, np.uint8), (h // 2, w // 2, np.uint8)]
    want = gpu_util.run_graph(g, [Y, U, V, None], specs + [None])[:3]
    ks, sp = _dev(torch, _yuy2(Y, U, V, uyvy))
    kd, dp = _out(torch, specs)
    tmp = torch.empty(max(g.device_packed_tmp_size(1, pack, z.PACKING_PLANAR), 1), dtype=torch.uint8, device="cuda")
    g.process_device_packed(1, z.packed_buffer(pack, sp), z.packed_buffer(z.PACKING_PLANAR, dp), tmp.data_ptr(), 0)
    for q, (a, b) in enumerate(zip(_read(torch, kd), want)):
        np.testing.assert_array_equal(a, b, err_msg=f"unpack plane {q}")

    # and back: 4:2:0 -> 4:2:2 packed
    g2 = z.FilterGraph(_fmt(w, h, z.PIXEL_BYTE, 8, 1, 1), _fmt(w, h, z.PIXEL_BYTE, 8, 1, 0))
    Y2, U2, V2 = want
    want2 = gpu_util.run_graph(g2, [Y2, U2, V2, None], [(h, w, np.uint8), (h, w // 2, np.uint8), (h, w // 2, np.uint8), None])[:3]
    ks2, sp2 = _dev(torch, [Y2, U2, V2])
    kd2, dp2 = _out(torch, [(h, 2 * w, np.uint8)])
    tmp2 = torch.empty(max(g2.device_packed_tmp_size(1, z.PACKING_PLANAR, pack), 1), dtype=torch.uint8, device="cuda")
    g2.process_device_packed(1, z.packed_buffer(z.PACKING_PLANAR, sp2), z.packed_buffer(pack, dp2), tmp2.data_ptr(), 0)
    np.testing.assert_array_equal(_read(torch, kd2)[0], _yuy2(*want2, uyvy)[0])
