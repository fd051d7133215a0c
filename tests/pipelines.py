"""Oracle compositions of whole graphs. Each follows the node chain the
reference's GraphBuilder produces for that format pair (graphbuilder.cpp:
892-1072, traced in tests/golden/graphbuilder_traces.json), calling the oracle
restatement for every stage. They are written from the reference's rules, not
from this library's planner, so the two derivations check each other."""
import numpy as np

from oracle_lib import (BICUBIC, FLOAT, LANCZOS, WORD, colorspace_apply, colorspace_ops, dither, resize_plane,
                        scale_offset, to_float)

# internal colourspace enums (colorspace.h)
M_709, M_2020 = 3, 7
T_709, T_2084, T_B67 = 4, 11, 13
P_709, P_2020 = 4, 6
M_RGB = 1


def word_resize(src, dst_w, dst_h, ftype=LANCZOS, a=3.0, b=0.0, depth=16):
    """Single WORD plane, GREY: one ResizeConversion (H/V order by cost)."""
    return resize_plane(src, ftype, a, b, dst_w, dst_h, depth=depth)


def float_resize(src, dst_w, dst_h, ftype, a, b):
    return resize_plane(src, ftype, a, b, dst_w, dst_h)


def yuv420_10_709_to_2020(Y, U, V):
    """FHD 4:2:0 10-bit limited BT.709 -> BT.2020 NCL (transfer BT.709), chroma
    siting LEFT (API default), bicubic b=0 c=0.5 chroma filter.

    Chain (graphbuilder.cpp:967-1019): luma WORD->FLOAT; chroma WORD->FLOAT at
    half size, resize to 4:4:4 (shift_w +0.25); colourspace; luma FLOAT->WORD;
    chroma resize to 4:2:0 (shift_w -0.5) then FLOAT->WORD (dither NONE)."""
    h, w = Y.shape
    luma_in = (WORD, 10, 0, 0, 0)
    chroma_in = (WORD, 10, 0, 1, 0)
    luma_f = (FLOAT, 32, 0, 0, 0)
    chroma_f = (FLOAT, 32, 0, 1, 0)
    s, o = scale_offset(luma_in, luma_f)
    Yf = to_float(Y, WORD, s, o)
    s, o = scale_offset(chroma_in, chroma_f)
    Uf = to_float(U, WORD, s, o)
    Vf = to_float(V, WORD, s, o)
    up = lambda c: resize_plane(c, BICUBIC, 0.0, 0.5, w, h, shift_w=0.25, shift_h=0.0, subwidth=w / 2, subheight=h / 2)
    Uf, Vf = up(Uf), up(Vf)
    n, ops, _ = colorspace_ops((M_709, T_709, P_709), (M_2020, T_709, P_2020), 100.0, 0, 0)
    assert n == 5
    Yc, Uc, Vc = colorspace_apply(ops, Yf, Uf, Vf)
    s, o = scale_offset(luma_f, luma_in)
    Yo = dither(Yc, FLOAT, WORD, s, o, 10)
    down = lambda c: resize_plane(c, BICUBIC, 0.0, 0.5, w // 2, h // 2, shift_w=-0.5, shift_h=0.0, subwidth=float(w), subheight=float(h))
    s, o = scale_offset(chroma_f, chroma_in)
    Uo = dither(down(Uc), FLOAT, WORD, s, o, 10, plane=1)
    Vo = dither(down(Vc), FLOAT, WORD, s, o, 10, plane=2)
    return Yo, Uo, Vo


def rgb_pq_to_hlg_12bit_ed(R, G, B, peak=203.0):
    """RGB FP32 BT.2020 PQ -> HLG (display-referred, peak nits) -> WORD 12-bit full
    range with Floyd-Steinberg error diffusion, per plane."""
    n, ops, _ = colorspace_ops((M_RGB, T_2084, P_2020), (M_RGB, T_B67, P_2020), peak, 0, 0)
    assert n == 2
    Rc, Gc, Bc = colorspace_apply(ops, R, G, B)
    s, o = scale_offset((FLOAT, 32, 0, 0, 0), (WORD, 12, 1, 0, 0))
    return [dither(p, FLOAT, WORD, s, o, 12, dither_type=3) for p in (Rc, Gc, Bc)]
