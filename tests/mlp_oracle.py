"""FP64 restatement of the 1-32-32-1 SiLU network the learned transfer
surrogate uses (test infrastructure only): forward pass and the gradient of
mean((v - t)^2), in the parameter layout w1[32], b1[32], W2[32][32] (row =
output unit), b2[32], w3[32], b3. The network has no counterpart in the
reference; these formulas are pinned by central finite differences."""
import numpy as np

H = 32
OFF = {"w1": 0, "b1": 32, "W2": 64, "b2": 64 + 1024, "w3": 64 + 1024 + 32, "b3": 64 + 1024 + 64}
NPARAM = 64 + 1024 + 64 + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[0:32], p[32:64], p[64:64 + 1024].reshape(32, 32), p[OFF["b2"]:OFF["b2"] + 32], p[OFF["w3"]:OFF["w3"] + 32],
            p[OFF["b3"]])


def sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(p, u):
    w1, b1, W2, b2, w3, b3 = unpack(p)
    u = np.asarray(u, np.float64)
    z1 = np.outer(w1, u) + b1[:, None]          # [32, n]
    h1 = z1 * sig(z1)
    z2 = W2 @ h1 + b2[:, None]
    h2 = z2 * sig(z2)
    v = w3 @ h2 + b3
    return v, (u, z1, h1, z2, h2)


def grad(p, u, t):
    """(gradient vector, mean loss) of mean((v - t)^2)."""
    w1, b1, W2, b2, w3, b3 = unpack(p)
    v, (u, z1, h1, z2, h2) = forward(p, u)
    n = len(u)
    d = v - np.asarray(t, np.float64)
    dv = 2.0 * d / n
    g = np.zeros(NPARAM)
    g[OFF["w3"]:OFF["w3"] + 32] = h2 @ dv
    g[OFF["b3"]] = dv.sum()
    s2 = sig(z2)
    dz2 = (w3[:, None] * dv[None, :]) * (s2 * (1 + z2 * (1 - s2)))
    g[OFF["b2"]:OFF["b2"] + 32] = dz2.sum(axis=1)
    g[64:64 + 1024] = (dz2 @ h1.T).ravel()
    dh1 = W2.T @ dz2
    s1 = sig(z1)
    dz1 = dh1 * (s1 * (1 + z1 * (1 - s1)))
    g[0:32] = dz1 @ u
    g[32:64] = dz1.sum(axis=1)
    return g, float(np.mean(d * d))


def init_params(seed):
    r = np.random.default_rng(seed)
    p = np.zeros(NPARAM)
    p[0:32] = r.uniform(-3, 3, 32)
    p[32:64] = r.uniform(-3, 3, 32)
    p[64:64 + 1024] = r.uniform(-1 / np.sqrt(32), 1 / np.sqrt(32), 1024)
    p[OFF["b2"]:OFF["b2"] + 32] = r.uniform(-0.2, 0.2, 32)
    p[OFF["w3"]:OFF["w3"] + 32] = r.uniform(-0.2, 0.2, 32)
    p[OFF["b3"]] = 0.1
    return p


# ---- FP64 chain of the learned-form workload: PQ codes -> linear (1.0 = the
# reference white of `peak` nits) -> display-referred HLG (gamma.cpp:177-188,
# 271-284; operation_impl.cpp:79-115) ----
PQ_M1, PQ_M2, PQ_C1, PQ_C2, PQ_C3 = 0.1593017578125, 78.84375, 0.8359375, 18.8515625, 18.6875
HLG_A, HLG_B, HLG_C = 0.17883277, 0.28466892, 0.55991073


def pq_eotf(x):
    x = np.asarray(x, np.float64)
    xp = np.power(np.maximum(x, 0.0), 1.0 / PQ_M2)
    y = np.power(np.maximum(xp - PQ_C1, 0.0) / (PQ_C2 - PQ_C3 * xp), 1.0 / PQ_M1)
    return np.where(x > 0, y, 0.0)


def hlg_oetf(x):
    x = np.maximum(np.asarray(x, np.float64), 0.0)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(x <= 1.0 / 12.0, np.sqrt(3.0 * x), HLG_A * np.log(np.maximum(12.0 * x - HLG_B, 1e-300)) + HLG_C)


def pq_to_hlg_f64(R, G, B, kr, kg, kb, peak=203.0):
    lin = [pq_eotf(c) * (10000.0 / peak) for c in (R, G, B)]
    r, g, b = (c * (peak / 1000.0) for c in lin)
    yd = np.maximum(kr * r + kg * g + kb * b, 1.1754943508222875e-38)
    ys = np.power(yd, (1.0 - 1.2) / 1.2)
    return [hlg_oetf(c * ys) for c in (r, g, b)]
