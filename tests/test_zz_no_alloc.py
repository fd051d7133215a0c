"""GPU: the allocation-free entry point
(zimg_filter_graph_process_device_batch_noalloc) gives exactly what
zimg_filter_graph_process_device_batch gives with a caller tmp, on two calls
in a row (the second on a smaller input, over the scratch the first left
behind), and rejects a batch whose intermediates exceed its scratch."""
import numpy as np
import pytest

import zimg_amd as z
from zimg_amd import workloads as W

pytestmark = pytest.mark.gpu


def _buffers(torch, planes, specs, frames):
    keep, sb, sfs, db, dfs, outs = [], [None] * 4, [0] * 4, [None] * 4, [0] * 4, []
    for p, a in enumerate(planes):
        if a is None:
            continue
        t = torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
        keep.append(t)
        F, h, w = a.shape
        sb[p] = (t.data_ptr(), w * a.itemsize)
        sfs[p] = h * w * a.itemsize
    for p, spec in enumerate(specs):
        if spec is None:
            continue
        h, w, dt = spec
        isz = np.dtype(dt).itemsize
        t = torch.full((frames * h * w * isz,), 0x5A, dtype=torch.uint8, device="cuda")
        keep.append(t)
        outs.append((t, h, w, dt))
        db[p] = (t.data_ptr(), w * isz)
        dfs[p] = h * w * isz
    return keep, z.image_buffer(sb), sfs, z.image_buffer(db), dfs, outs


def _read(torch, outs, frames):
    torch.cuda.synchronize()
    return [t.cpu().numpy().view(dt).reshape(frames, h, w).copy() for (t, h, w, dt) in outs]


def _c2(w, h, frames, seed):
    planes = [W.yuv420_10bit(w, h, seed=seed + k) for k in range(frames)]
    stacks = [np.stack([pl[c] for pl in planes]) for c in range(3)] + [None]
    specs = [(h, w, np.uint16), (h // 2, w // 2, np.uint16), (h // 2, w // 2, np.uint16), None]
    return z.FilterGraph(*W.fhd_709_to_2020_formats(w, h)), stacks, specs


def _both(torch, g, stacks, specs, frames):
    keep, src, sfs, dst, dfs, outs = _buffers(torch, stacks, specs, frames)
    z.process_device_batch_noalloc(g, frames, src, sfs, dst, dfs, 0)
    got = _read(torch, outs, frames)
    for (t, *_r) in outs:
        t.fill_(0x5A)
    tmp = torch.empty(max(g.device_tmp_size(frames), 1), dtype=torch.uint8, device="cuda")
    g.process_device_batch(frames, src, sfs, dst, dfs, tmp.data_ptr(), 0)
    want = _read(torch, outs, frames)
    return got, want


def test_noalloc_equals_tmp_entry_point_twice_then_smaller():
    import torch
    print("case: noalloc 4:2:0 709->2020 640x360 x3, then 320x180 x2", flush=True)
    g, stacks, specs = _c2(640, 360, 3, seed=70)
    assert 0 < g.device_tmp_size(3) <= z.NOALLOC_SCRATCH_BYTES
    got, want = _both(torch, g, stacks, specs, 3)
    for q in range(3):
        np.testing.assert_array_equal(got[q], want[q], err_msg=f"first call plane {q}")
    g2, stacks2, specs2 = _c2(320, 180, 2, seed=80)
    got, want = _both(torch, g2, stacks2, specs2, 2)
    for q in range(3):
        np.testing.assert_array_equal(got[q], want[q], err_msg=f"second (smaller) call plane {q}")
    # and a graph with no intermediates at all
    gw = z.FilterGraph(*W.uhd_lanczos_formats(320, 180, 640, 360))
    assert gw.device_tmp_size(2) == 0
    src = np.stack([W.word_plane(320, 180, seed=s) for s in (1, 2)])
    got, want = _both(torch, gw, [src, None, None, None], [(360, 640, np.uint16), None, None, None], 2)
    np.testing.assert_array_equal(got[0], want[0])


def test_noalloc_rejects_batches_beyond_its_scratch():
    """FHD 4:2:0 709->2020 keeps FP32 4:4:4 chroma intermediates (about 32 MiB
    a frame): a batch whose intermediates exceed the 128 MiB scratch is
    refused before anything is launched."""
    import torch
    w, h = 1920, 1080
    g, _, specs = _c2(w, h, 1, seed=1)
    per = g.device_tmp_size(1)
    if per == 0:
        pytest.skip("this graph keeps no intermediate planes")
    frames = z.NOALLOC_SCRATCH_BYTES // per + 1
    print(f"case: noalloc size limit (FHD 709->2020, {frames} frames)", flush=True)
    assert g.device_tmp_size(frames) > z.NOALLOC_SCRATCH_BYTES
    # whole-size planes, so nothing could be out of bounds even if a launch happened
    stacks = [np.zeros((frames, h, w), np.uint16), np.zeros((frames, h // 2, w // 2), np.uint16),
              np.zeros((frames, h // 2, w // 2), np.uint16), None]
    keep, src, sfs, dst, dfs, outs = _buffers(torch, stacks, specs, frames)
    with pytest.raises(z.ZimgError) as ei:
        z.process_device_batch_noalloc(g, frames, src, sfs, dst, dfs, 0)
    assert ei.value.code == z.ERROR_ILLEGAL_ARGUMENT
