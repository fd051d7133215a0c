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


def test_pq_to_hlg_learned_error_and_gain():
    """The learned-form workload at 960x540: PQ EOTF by the surrogate (exact
    fallback outside validated bins), then the exact display-referred HLG stage
    on the device; 12-bit LSB error of the FLOAT result against FP64."""
    print("case: PQ->HLG learned form 960x540", flush=True)
    import torch
    w, h = 960, 540
    R, G, B = W.pq_rgb(w, h, seed=21)
    s = z.TransferSurrogate(z.TRANSFER_ST2084, True, 4e-5, True, steps=20000, seed=7)
    planes = [_device(c.ravel()) for c in (R, G, B)]
    for t in planes:
        s.apply(t.data_ptr(), t.numel(), 1.0, 10000.0 / 203.0)
    lin = [t.cpu().numpy().reshape(h, w) for t in planes]
    common = dict(color_family=z.COLOR_RGB, matrix_coefficients=z.MATRIX_RGB, color_primaries=z.PRIMARIES_BT2020,
                  pixel_range=z.RANGE_FULL)
    src = z.image_format(width=w, height=h, pixel_type=z.PIXEL_FLOAT, transfer_characteristics=z.TRANSFER_LINEAR, **common)
    dst = z.image_format(width=w, height=h, pixel_type=z.PIXEL_FLOAT, transfer_characteristics=z.TRANSFER_ARIB_B67, **common)
    g = z.FilterGraph(src, dst, z.builder_params(nominal_peak_luminance=203.0))
    outs = gpu_util.run_graph(g, lin + [None], [(h, w, np.float32)] * 3 + [None])
    n, ops, _ = colorspace_ops((P.M_RGB, P.T_2084, P.P_2020), (P.M_RGB, P.T_B67, P.P_2020), 203.0, 0, 0)
    want = M.pq_to_hlg_f64(R, G, B, float(ops.kr[1]), float(ops.kg[1]), float(ops.kb[1]))
    lsb = max(float(np.max(np.abs(np.clip(o, 0, 1) - np.clip(wv, 0, 1)))) * 4095 for o, wv in zip(outs, want))
    # exact path for comparison
    ref = gpu_util.run_graph(z.FilterGraph(*_pq_formats(w, h)), [R, G, B, None], [(h, w, np.float32)] * 3 + [None])
    lsb_exact = max(float(np.max(np.abs(np.clip(o, 0, 1) - np.clip(wv, 0, 1)))) * 4095 for o, wv in zip(ref, want))
    # throughput of the transfer stage alone, surrogate vs exact, 3 x UHD values
    x = torch.from_numpy(np.tile(R.ravel(), 16)[: 3 * 3840 * 2160].astype(np.float32)).to("cuda")
    y = x.clone()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / 5

    t_s = timed(lambda: (y.copy_(x), s.apply(y.data_ptr(), y.numel(), 1.0, 10000.0 / 203.0)))
    t_e = timed(lambda: (y.copy_(x), z.transfer_exact_apply(z.TRANSFER_ST2084, True, y.data_ptr(), y.numel(), 1.0, 10000.0 / 203.0)))
    report = {"max_lsb_learned": lsb, "max_lsb_exact": lsb_exact, "coverage": s.info()["coverage"],
              "transfer_ms_learned": t_s * 1e3, "transfer_ms_exact": t_e * 1e3, "gain": t_e / t_s}
    print(json.dumps(report), flush=True)
    assert lsb <= 0.5, report
    assert lsb_exact <= 0.05, report


def _pq_formats(w, h):
    common = dict(color_family=z.COLOR_RGB, matrix_coefficients=z.MATRIX_RGB, color_primaries=z.PRIMARIES_BT2020,
                  pixel_range=z.RANGE_FULL)
    src = z.image_format(width=w, height=h, pixel_type=z.PIXEL_FLOAT, transfer_characteristics=z.TRANSFER_ST2084, **common)
    dst = z.image_format(width=w, height=h, pixel_type=z.PIXEL_FLOAT, transfer_characteristics=z.TRANSFER_ARIB_B67, **common)
    return src, dst, z.builder_params(nominal_peak_luminance=203.0)
