"""GPU: batched uneven jobs in one persistent launch (captured into a HIP
graph, with and without the tile autotuner) produce exactly what each job's
own launch sequence produces, and that equals the oracle chain."""
import numpy as np
import pytest

import pipelines as P
import zimg_amd as z
from oracle_lib import BYTE, WORD, dither, resize_plane, scale_offset
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
    return P.batch_job(job, pl)


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


@pytest.fixture(scope="module")
def batch64():
    """The committed 64-job list (720p..UHD, u8 / 10-bit 4:2:0, four filters,
    Bayer dither to 8-bit), one seeded plane set per job, and the oracle chain
    of every job and plane (computed on host threads)."""
    import json
    import os
    from concurrent.futures import ThreadPoolExecutor
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "batch64_jobs.json")))
    jobs = fx["jobs"]
    assert jobs == W.batch_jobs(64, fx["seed"])
    planes = [W.batch_job_planes(j, seed=900 + k) for k, j in enumerate(jobs)]
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 4)) as ex:
        want = list(ex.map(lambda k: _oracle(jobs[k], planes[k]), range(len(jobs))))
    return jobs, planes, want


def _setup_planes(jobs, planes):
    import torch
    graphs, srcs, dsts, keep, outs = [], [], [], [], []
    for j, pl in zip(jobs, planes):
        graphs.append(z.FilterGraph(*W.batch_job_graph(j)))
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
    return graphs, srcs, dsts, keep, outs


@pytest.mark.parametrize("flags", [0, z.DEVICE_JOBS_AUTOTUNE, z.DEVICE_JOBS_NO_GRAPH])
def test_batch64_every_job_equals_oracle(batch64, flags):
    """The full batched workload: 64 uneven jobs (192 planes, 720p..UHD) in one
    persistent launch, every plane of every job against the oracle chain,
    under graph replay (twice), the tile autotuner, and direct launches."""
    jobs, planes, want = batch64
    print(f"case: batch64 full list, flags {flags}", flush=True)
    graphs, srcs, dsts, keep, outs = _setup_planes(jobs, planes)
    jb = z.DeviceJobs(graphs, srcs, dsts, None, flags, None)
    print(jb.describe(), flush=True)
    assert int(jb.describe().split("persistent=")[1].split()[0]) == 64
    for rep in range(2 if flags != z.DEVICE_JOBS_NO_GRAPH else 1):
        for o in outs:
            for (t, *_rest) in o:
                t.fill_(0xCD)
        jb.run(None)
        got = [_read(o) for o in outs]
        for k in range(len(jobs)):
            for p in range(3):
                np.testing.assert_array_equal(got[k][p], want[k][p], err_msg=f"job {k} {jobs[k]} plane {p} replay {rep}")


def test_jobs_from_fresh_graphs_with_lazy_tables():
    """DeviceJobs built from graphs that never ran before: jobs outside the
    persistent launch (a FLOAT resize on the FP16 matrix cores, a WORD upscale
    on the integer matrix cores followed by error diffusion) are captured into
    the HIP graph with every table they use already resident."""
    import torch
    print("case: fresh graphs, precision-form FLOAT job + WORD upscale with error diffusion", flush=True)
    sw, sh, dw, dh = 320, 180, 640, 360
    fsrc = W.fp32_plane(sw, sh, seed=4)
    s, d, p = W.fp32_formats(sw, sh, dw, dh, z.RESIZE_BICUBIC, 0.0, 0.5)
    gf = z.FilterGraph(s, d, p)
    gf.set_resize_precision(z.RESIZE_PRECISION_MFMA_FP16)
    wsrc = W.word_plane(sw, sh, seed=5)
    gw = z.FilterGraph(z.image_format(width=sw, height=sh, pixel_type=z.PIXEL_WORD, depth=16),
                       z.image_format(width=dw, height=dh, pixel_type=z.PIXEL_BYTE, depth=8, pixel_range=z.RANGE_FULL),
                       z.builder_params(resample_filter=z.RESIZE_LANCZOS, filter_param_a=3.0, dither_type=z.DITHER_ERROR_DIFFUSION))
    assert "integer matrix cores" in gw.describe() and "error_diffusion" in gw.describe()
    ft = torch.from_numpy(fsrc.reshape(-1)).to("cuda")
    wt = torch.from_numpy(wsrc.view(np.uint8).reshape(-1)).to("cuda")
    fo = torch.zeros((dh * dw,), dtype=torch.float32, device="cuda")
    wo = torch.zeros((dh * dw,), dtype=torch.uint8, device="cuda")
    graphs = [gf, gw]
    srcs = [z.image_buffer([(ft.data_ptr(), sw * 4), None, None, None]), z.image_buffer([(wt.data_ptr(), sw * 2), None, None, None])]
    dsts = [z.image_buffer([(fo.data_ptr(), dw * 4), None, None, None]), z.image_buffer([(wo.data_ptr(), dw), None, None, None])]
    tmp = torch.empty(max(z.device_jobs_tmp_size(graphs), 1), dtype=torch.uint8, device="cuda")
    jb = z.DeviceJobs(graphs, srcs, dsts, tmp.data_ptr(), 0, None)
    assert "graph=1" in jb.describe()
    jb.run(None)
    torch.cuda.synchronize()
    got_f = fo.cpu().numpy().reshape(dh, dw)
    exact = P.float_resize(fsrc, dw, dh, z.RESIZE_BICUBIC, 0.0, 0.5)
    assert np.max(np.abs(got_f - exact)) < 2e-6
    r = resize_plane(wsrc, 5, 3.0, 0.0, dw, dh, depth=16)
    sc, of = scale_offset((1, 16, 0, 0, 0), (0, 8, 1, 0, 0))
    want_w = dither(r, WORD, BYTE, sc, of, 8, dither_type=3)
    np.testing.assert_array_equal(wo.cpu().numpy().reshape(dh, dw), want_w)
