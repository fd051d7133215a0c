"""GPU: batched uneven jobs in one persistent launch (captured into a HIP
graph, with and without the tile autotuner) produce exactly what each job's
own launch sequence produces, and that equals the oracle chain."""
import numpy as np
import pytest

import zimg_amd as z
from oracle_lib import BYTE, WORD, dither, left_shift, resize_plane, scale_offset
from zimg_amd import workloads as W

pytestmark = pytest.mark.gpu

gpu_util = pytest.importorskip("gpu_util")

SMALL = [(320, 180), (480, 270), (640, 360), (160, 90)]


def _setup(jobs):
    import torch
    graphs, srcs, dsts, keep, planes, outs = [], [], [], [], [], []
    for k, j in enumerate(jobs):
        g = z.FilterGraph(*W.batch_job_graph(j))
        graphs.append(g)
        pl = W.batch_job_planes(j, seed=500 + k)
        planes.append(pl)
        sb, db, ob = [], [], []
        (dw, dh) = j["dst"]
        for p, a in enumerate(pl):
            t = torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1)).to("cuda")
            keep.append(t)
            sb.append((t.data_ptr(), a.shape[1] * a.itemsize))
            w, h = (dw, dh) if p == 0 else (dw // 2, dh // 2)
            pitch = (w + 63) // 64 * 64
            o = torch.full((h * pitch,), 0xCD, dtype=torch.uint8, device="cuda")
            keep.append(o)
            db.append((o.data_ptr(), pitch))
            ob.append((o, h, w, pitch))
        srcs.append(z.image_buffer(sb + [None]))
        dsts.append(z.image_buffer(db + [None]))
        outs.append(ob)
    return graphs, srcs, dsts, keep, planes, outs


def _read(ob):
    import torch
    torch.cuda.synchronize()
    return [o.cpu().numpy().reshape(h, pitch)[:, :w].copy() for (o, h, w, pitch) in ob]


def _oracle(job, pl):
    (sw, sh), (dw, dh), bits = job["src"], job["dst"], job["bits"]
    filt = {"bilinear": (1, 0.0, 0.0), "bicubic": (2, 0.0, 0.5), "spline36": (4, 0.0, 0.0), "lanczos3": (5, 3.0, 0.0)}[job["filter"]]
    res = []
    for q, plane in enumerate(pl):
        chroma = q > 0
        if bits == 8:
            x, wd = left_shift(plane, BYTE, WORD, 8), 16
        else:
            x, wd = plane, 10
        cw_s, ch_s = plane.shape[1], plane.shape[0]
        cw_d, ch_d = (dw // 2, dh // 2) if chroma else (dw, dh)
        shift_w = (0.25 - 0.25 / (cw_d / cw_s)) if chroma else 0.0
        r = resize_plane(x, filt[0], filt[1], filt[2], cw_d, ch_d, shift_w=shift_w, shift_h=0.0, subwidth=float(cw_s),
                         subheight=float(ch_s), depth=wd)
        sc, of = scale_offset((WORD, wd, 0, int(chroma), 0), (BYTE, 8, 0, int(chroma), 0))
        res.append(dither(r, WORD, BYTE, sc, of, 8, dither_type=1, plane=q))
    return res


@pytest.mark.parametrize("flags", [0, z.DEVICE_JOBS_AUTOTUNE, z.DEVICE_JOBS_NO_GRAPH])
def test_persistent_jobs_equal_per_job_launches(flags):
    print(f"case: 10 uneven small jobs, flags {flags}", flush=True)
    jobs = W.batch_jobs(10, seed=77, sizes=SMALL)
    graphs, srcs, dsts, keep, planes, outs = _setup(jobs)
    for g, s, d in zip(graphs, srcs, dsts):
        g.process_device(s, d, None, None)
    want = [_read(o) for o in outs]
    for o in outs:
        for (t, *_rest) in o:
            t.fill_(0xCD)
    jb = z.DeviceJobs(graphs, srcs, dsts, None, flags, None)
    print(jb.describe(), flush=True)
    # jobs whose taps exceed the persistent kernel's capacity run their own launches first
    pers = int(jb.describe().split("persistent=")[1].split()[0])
    assert 8 <= pers <= 10
    for rep in range(3):  # graph replays
        jb.run(None)
        got = [_read(o) for o in outs]
        for k in range(len(jobs)):
            for p in range(3):
                np.testing.assert_array_equal(got[k][p], want[k][p], err_msg=f"job {k} plane {p} replay {rep}")
    for k in (0, 3):
        ref = _oracle(jobs[k], planes[k])
        for p in range(3):
            np.testing.assert_array_equal(want[k][p], ref[p], err_msg=f"oracle job {k} plane {p}")
