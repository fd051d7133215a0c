"""GPU: the learned transfer surrogate. The training kernels' gradients and
the inference layers' outputs against the FP64 network oracle; training on
the GPU; per-bin validation with exact fallback; the PQ -> HLG workload's
error in 12-bit LSB against the FP64 chain (bound 0.5 LSB); and what the
trained network gains or loses in speed against the exact transfer."""
import json
import time

import numpy as np
import pytest

import mlp_oracle as M
import pipelines as P
import zimg_amd as z
from oracle_lib import colorspace_ops
from zimg_amd import workloads as W

pytestmark = pytest.mark.gpu

gpu_util = pytest.importorskip("gpu_util")


def test_training_gradients_and_inference_match_fp64_oracle():
    print("case: MLP gradients n=256 vs FP64 oracle", flush=True)
    p = M.init_params(11).astype(np.float32)
    r = np.random.default_rng(12)
    u = r.uniform(-1, 1, 256).astype(np.float32)
    t = (np.sin(2.5 * u) * 0.7).astype(np.float32)
    g = z.surrogate_gradients(p, u, t)
    want, loss = M.grad(p.astype(np.float64), u.astype(np.float64), t.astype(np.float64))
    assert abs(g[-1] - loss) <= 1e-4 * loss + 1e-7
    for name, lo, hi in [("w1", 0, 32), ("b1", 32, 64), ("W2", 64, 1088), ("b2", 1088, 1120), ("w3", 1120, 1152), ("b3", 1152, 1153)]:
        scale = np.max(np.abs(want[lo:hi])) + 1e-12
        err = np.max(np.abs(g[lo:hi] - want[lo:hi])) / scale
        assert err < 2e-4, (name, err)
    v = z.surrogate_eval(p, u)
    vw, _ = M.forward(p.astype(np.float64), u.astype(np.float64))
    assert np.max(np.abs(v - vw)) < 2e-5 * (np.max(np.abs(vw)) + 1)


def _device(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda")


@pytest.mark.parametrize("transfer,to_linear,tol,rel", [(z.TRANSFER_BT709, True, 1e-4, False), (z.TRANSFER_ST2084, True, 4e-5, True),
                                                       (z.TRANSFER_ARIB_B67, False, 3e-5, False)])
def test_trained_surrogate_within_tolerance_on_fresh_samples(transfer, to_linear, tol, rel):
    print(f"case: train surrogate transfer {transfer} to_linear {to_linear}", flush=True)
    import torch
    s = z.TransferSurrogate(transfer, to_linear, tol, rel, steps=20000, seed=7)
    info = s.info()
    print(json.dumps(info), flush=True)
    r = np.random.default_rng(13)
    x = r.uniform(0, 1, 200000).astype(np.float32)
    if transfer == z.TRANSFER_ARIB_B67:
        x = (x * 1.2).astype(np.float32)
    xd = _device(x)
    s.apply(xd.data_ptr(), x.size)
    got = xd.cpu().numpy().astype(np.float64)
    func = {z.TRANSFER_BT709: lambda v: np.power(v, 2.4), z.TRANSFER_ST2084: M.pq_eotf, z.TRANSFER_ARIB_B67: M.hlg_oetf}[transfer]
    want = func(x.astype(np.float64))
    err = np.abs(got - want)
    if rel:
        err = err / np.maximum(np.abs(want), 1e-30)
    # every value is either inside a validated bin (tolerance, with a margin for
    # points between the 64 validated per bin) or evaluated exactly (FP32 rounding)
    exact_err = 2e-6 if not rel else 2e-5
    assert np.max(err) <= max(2 * tol, exact_err), np.max(err)


@pytest.fixture(scope="module")
def chain_surrogates():
    """The PQ EOTF (relative) and HLG OETF (absolute) surrogates of the PQ ->
    HLG workload, each validated per bin against the tolerance that the
    12-bit output budget of 0.5 LSB allows it."""
    tol_pq, tol_hlg = z.pq_to_hlg_surrogate_tolerances(bits=12, lsb=0.5)
    pq = z.TransferSurrogate(z.TRANSFER_ST2084, True, tol_pq, True, steps=20000, seed=7)
    hlg = z.TransferSurrogate(z.TRANSFER_ARIB_B67, False, tol_hlg, False, steps=20000, seed=7)
    print(json.dumps({"tol_pq_rel": tol_pq, "tol_hlg_abs": tol_hlg, "pq": pq.info(), "hlg": hlg.info()}), flush=True)
    return pq, hlg


def _float_formats(w, h):
    s, d, p = W.uhd_pq_to_hlg_formats(w, h)
    d.pixel_type = z.PIXEL_FLOAT
    d.depth = 32
    return s, d, p


def _lsb12(outs, want):
    return max(float(np.max(np.abs(np.clip(o.astype(np.float64), 0, 1) - np.clip(wv, 0, 1)))) * 4095 for o, wv in zip(outs, want))


def test_uhd_pq_to_hlg_learned_through_error_diffusion(chain_surrogates):
    """3840x2160 PQ -> HLG -> 12-bit error diffusion with both transfers on the
    network: the values entering error diffusion are within 0.5 LSB (12-bit)
    of the FP64 chain, the diffused codes are exactly the oracle's error
    diffusion of those values, and so within 1 + that of the FP64 chain."""
    print("case: PQ->HLG learned 3840x2160 through ED", flush=True)
    w, h = 3840, 2160
    pq, hlg = chain_surrogates
    R, G, B = W.pq_rgb(w, h, seed=9)
    n, ops, _ = colorspace_ops((P.M_RGB, P.T_2084, P.P_2020), (P.M_RGB, P.T_B67, P.P_2020), 203.0, 0, 0)
    want = M.pq_to_hlg_f64(R, G, B, float(ops.kr[1]), float(ops.kg[1]), float(ops.kb[1]))
    gf = z.FilterGraph(*_float_formats(w, h))
    gf.set_transfer_surrogates([pq, hlg])
    desc = gf.describe()
    assert "surrogate(f" in desc and "hlg_ootf_scale" in desc, desc
    lin = gpu_util.run_graph(gf, [R, G, B, None], [(h, w, np.float32)] * 3 + [None])
    lsb = _lsb12(lin[:3], want)
    ge = z.FilterGraph(*W.uhd_pq_to_hlg_formats(w, h))
    ge.set_transfer_surrogates([pq, hlg])
    codes = gpu_util.run_graph(ge, [R, G, B, None], [(h, w, np.uint16)] * 3 + [None])
    exact = gpu_util.run_graph(z.FilterGraph(*_float_formats(w, h)), [R, G, B, None], [(h, w, np.float32)] * 3 + [None])
    report = {"max_lsb_learned": lsb, "max_lsb_exact": _lsb12(exact[:3], want),
              "changed_values": float(np.mean([np.mean(a != b) for a, b in zip(lin[:3], exact[:3])]))}
    ed_err = 0.0
    for c, f, wv in zip(codes[:3], lin[:3], want):
        _assert_equal_codes(c, P.dither(f, P.FLOAT, P.WORD, 4095.0, 0.0, 12, dither_type=3))
        ed_err = max(ed_err, float(np.max(np.abs(c.astype(np.float64) - 4095.0 * np.clip(wv, 0, 1)))))
    report["max_code_error_after_ed"] = ed_err
    print(json.dumps(report), flush=True)
    assert report["changed_values"] > 0.05, report  # the network, not the fallback, produced most values
    assert lsb <= 0.5, report
    assert ed_err <= 1.0 + lsb + 1e-3, report


def _assert_equal_codes(got, want):
    d = np.argwhere(got != want)
    assert d.size == 0, (len(d), d[:5].tolist())


def test_learned_segments_around_integer_input_and_output(chain_surrogates):
    """A surrogate as the first operation behind an integer input and as the
    last in front of a quantised output: the conversion launches added around
    it equal the oracle's conversions of the learned FLOAT result."""
    print("case: learned PQ EOTF, WORD in / WORD out 640x360", flush=True)
    pq, _ = chain_surrogates
    w, h = 640, 360
    r = np.random.default_rng(5)
    planes = [r.integers(0, 1024, (h, w)).astype(np.uint16) for _ in range(3)]
    common = dict(color_family=z.COLOR_RGB, matrix_coefficients=z.MATRIX_RGB, color_primaries=z.PRIMARIES_BT2020,
                  pixel_range=z.RANGE_FULL)
    src = z.image_format(width=w, height=h, pixel_type=z.PIXEL_WORD, depth=10, transfer_characteristics=z.TRANSFER_ST2084, **common)
    dst_f = z.image_format(width=w, height=h, pixel_type=z.PIXEL_FLOAT, transfer_characteristics=z.TRANSFER_LINEAR, **common)
    dst_w = z.image_format(width=w, height=h, pixel_type=z.PIXEL_WORD, depth=16, transfer_characteristics=z.TRANSFER_LINEAR, **common)
    prm = z.builder_params(nominal_peak_luminance=10000.0)
    gf = z.FilterGraph(src, dst_f, prm)
    gf.set_transfer_surrogates([pq])
    lin = gpu_util.run_graph(gf, planes + [None], [(h, w, np.float32)] * 3 + [None])
    gw = z.FilterGraph(src, dst_w, prm)
    gw.set_transfer_surrogates([pq])
    assert gw.device_tmp_size(1) >= 3 * w * h * 4
    codes = gpu_util.run_graph(gw, planes + [None], [(h, w, np.uint16)] * 3 + [None])
    # the exact graph's input conversion, then the surrogate: compare with the exact FLOAT result
    ref = gpu_util.run_graph(z.FilterGraph(src, dst_f, prm), planes + [None], [(h, w, np.float32)] * 3 + [None])
    for l, c, e in zip(lin[:3], codes[:3], ref[:3]):
        _assert_equal_codes(c, P.dither(l, P.FLOAT, P.WORD, 65535.0, 0.0, 16))
        nz = e > 1e-4
        assert np.max(np.abs(l[nz] / e[nz] - 1.0)) < 4e-4
    gf.set_transfer_surrogates([])
    assert "surrogate" not in gf.describe()
    back = gpu_util.run_graph(gf, planes + [None], [(h, w, np.float32)] * 3 + [None])
    for b, e in zip(back[:3], ref[:3]):
        _assert_equal_codes(b.view(np.uint32), e.view(np.uint32))


def test_pq_to_hlg_learned_vs_exact_speed(chain_surrogates):
    """Frames/s of the PQ -> HLG colourspace stage (3840x2160, 8 frames) with
    the learned transfers and with the exact ones, timed with HIP events on
    the launch stream; reported, with the per-value instruction counts that
    explain the ratio."""
    import torch
    pq, hlg = chain_surrogates
    w, h, frames = 3840, 2160, 8
    R, G, B = W.pq_rgb(w, h, seed=9)
    times = {}
    for name, sur in [("exact", []), ("learned", [pq, hlg])]:
        g = z.FilterGraph(*_float_formats(w, h))
        g.set_transfer_surrogates(sur)
        src = [torch.from_numpy(np.stack([c] * frames)).to("cuda") for c in (R, G, B)]
        dst = [torch.empty((frames, h, w), dtype=torch.float32, device="cuda") for _ in range(3)]
        tmp = torch.empty(max(g.device_tmp_size(frames), 1), dtype=torch.uint8, device="cuda")
        sb = z.image_buffer([(t.data_ptr(), w * 4) for t in src] + [None])
        db = z.image_buffer([(t.data_ptr(), w * 4) for t in dst] + [None])
        fs = [w * h * 4] * 3 + [0]
        stream = torch.cuda.current_stream()
        run = lambda: g.process_device_batch(frames, sb, fs, db, fs, tmp.data_ptr(), stream.cuda_stream)
        run()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(5):
            run()
        e1.record(stream)
        torch.cuda.synchronize()
        times[name] = e0.elapsed_time(e1) / 5
    report = {"fps_exact": frames / times["exact"] * 1e3, "fps_learned": frames / times["learned"] * 1e3,
              "gain": times["exact"] / times["learned"]}
    print(json.dumps(report), flush=True)
    assert times["exact"] > 0 and times["learned"] > 0


def _pq_formats(w, h):
    common = dict(color_family=z.COLOR_RGB, matrix_coefficients=z.MATRIX_RGB, color_primaries=z.PRIMARIES_BT2020,
                  pixel_range=z.RANGE_FULL)
    src = z.image_format(width=w, height=h, pixel_type=z.PIXEL_FLOAT, transfer_characteristics=z.TRANSFER_ST2084, **common)
    dst = z.image_format(width=w, height=h, pixel_type=z.PIXEL_FLOAT, transfer_characteristics=z.TRANSFER_ARIB_B67, **common)
    return src, dst, z.builder_params(nominal_peak_luminance=203.0)
