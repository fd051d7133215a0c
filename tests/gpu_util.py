"""Device-memory plumbing for the GPU tests: numpy planes -> HBM (torch byte
tensors), one call through the C ABI, HBM -> numpy. Row pitch padding is
optional so strided layouts are exercised too."""
import numpy as np

import zimg_amd as z


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible")
    return torch


def _upload(torch, arr, pad_bytes, pitch_align=4):
    """arr: (F, h, w) numpy -> (tensor, pitch, frame_stride)."""
    F, h, w = arr.shape
    rowbytes = w * arr.itemsize
    pitch = (rowbytes + pad_bytes + pitch_align - 1) // pitch_align * pitch_align
    host = np.zeros((F, h, pitch), np.uint8)
    host[:, :, :rowbytes] = arr.view(np.uint8).reshape(F, h, rowbytes)
    t = torch.from_numpy(host.reshape(-1)).to("cuda")
    return t, pitch, h * pitch


def run_graph(graph, src_planes, dst_specs, frames=None, pad_bytes=0, stream=0, pitch_align=4, autotune=False):
    """src_planes: list of 4 (numpy (h,w) or (F,h,w), or None).
    dst_specs: list of 4 ((h, w, dtype) or None). Row pitches are rounded up to
    pitch_align bytes. autotune: time the graph's tile shapes on these buffers
    first (zimg_filter_graph_autotune_device). Returns list of numpy outputs."""
    torch = _torch()
    keep = []
    src_bufs, src_fs = [None] * 4, [0] * 4
    F = 1
    for p, a in enumerate(src_planes):
        if a is None:
            continue
        a3 = a if a.ndim == 3 else a[None]
        F = a3.shape[0]
        t, pitch, fs = _upload(torch, np.ascontiguousarray(a3), pad_bytes, pitch_align)
        keep.append(t)
        src_bufs[p] = (t.data_ptr(), pitch)
        src_fs[p] = fs
    if frames is None:
        frames = F
    dst_bufs, dst_fs, dst_t = [None] * 4, [0] * 4, [None] * 4
    for p, spec in enumerate(dst_specs):
        if spec is None:
            continue
        h, w, dt = spec
        rowbytes = w * np.dtype(dt).itemsize
        pitch = (rowbytes + pad_bytes + pitch_align - 1) // pitch_align * pitch_align
        t = torch.full((frames * h * pitch,), 0xAB, dtype=torch.uint8, device="cuda")
        dst_t[p] = (t, h, w, dt, pitch)
        dst_bufs[p] = (t.data_ptr(), pitch)
        dst_fs[p] = h * pitch
    tmp_size = graph.device_tmp_size(frames)
    tmp = torch.empty(max(tmp_size, 1), dtype=torch.uint8, device="cuda")
    src = z.image_buffer(src_bufs)
    dst = z.image_buffer(dst_bufs)
    if autotune:
        graph.autotune_device(frames, src, src_fs, dst, dst_fs, tmp.data_ptr(), stream)
    if frames == 1:
        graph.process_device(src, dst, tmp.data_ptr(), stream)
    else:
        graph.process_device_batch(frames, src, src_fs, dst, dst_fs, tmp.data_ptr(), stream)
    torch.cuda.synchronize()
    outs = []
    for p in range(4):
        if dst_t[p] is None:
            outs.append(None)
            continue
        t, h, w, dt, pitch = dst_t[p]
        host = t.cpu().numpy().reshape(frames, h, pitch)
        rowbytes = w * np.dtype(dt).itemsize
        a = np.ascontiguousarray(host[:, :, :rowbytes]).view(dt).reshape(frames, h, w)
        outs.append(a if frames > 1 else a[0])
    return outs
