"""CPU pins of the learned form's test oracles: the FP64 network gradient
against central finite differences, and the FP64 PQ -> HLG chain against the
oracle's FP32 colourspace chain (which is pinned to the compiled reference)."""
import numpy as np

import mlp_oracle as M
import pipelines as P
from oracle_lib import colorspace_apply, colorspace_ops
from zimg_amd import workloads as W


def test_network_gradient_matches_finite_differences():
    p = M.init_params(3)
    r = np.random.default_rng(4)
    u = r.uniform(-1, 1, 64)
    t = np.sin(2.5 * u) * 0.7
    g, _ = M.grad(p, u, t)
    idx = list(range(0, 32, 5)) + list(range(32, 64, 7)) + list(range(64, 1088, 97)) + list(range(1088, 1153, 9)) + [1152]
    for k in idx:
        h = 1e-6 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (M.grad(pp, u, t)[1] - M.grad(pm, u, t)[1]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(g[k]), (k, fd, g[k])


def test_fp64_pq_to_hlg_chain_matches_fp32_oracle_chain():
    R, G, B = W.pq_rgb(256, 128, seed=5)
    n, ops, _ = colorspace_ops((P.M_RGB, P.T_2084, P.P_2020), (P.M_RGB, P.T_B67, P.P_2020), 203.0, 0, 0)
    assert n == 2
    want = colorspace_apply(ops, R, G, B)
    kr, kg, kb = (float(ops.kr[1]), float(ops.kg[1]), float(ops.kb[1]))
    got = M.pq_to_hlg_f64(R, G, B, kr, kg, kb)
    # the FP32 chain's own rounding (powf with exponents up to 1/m1 = 6.3) is
    # ~1e-5 at the brightest codes, 0.04 LSB of a 12-bit output
    for a, b in zip(got, want):
        assert np.max(np.abs(a - b)) < 3e-5
