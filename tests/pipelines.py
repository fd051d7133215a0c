"""Oracle compositions of whole graphs. Each follows the node chain the
reference's GraphBuilder produces for that format pair (graphbuilder.cpp:
892-1072, traced in tests/golden/graphbuilder_traces.json), calling the oracle
restatement for every stage. They are written from the reference's rules, not
from this library's planner, so the two derivations check each other."""
import numpy as np

from oracle_lib import (BICUBIC, BYTE, FLOAT, LANCZOS, WORD, colorspace_apply, colorspace_ops, dither, left_shift,
                        resize_plane, scale_offset, to_float)

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


def batch_job(job, planes):
    """One job of the batched form (zimg_amd.workloads.batch_job_graph): each
    4:2:0 plane (8-bit promoted to WORD by a left shift, or 10-bit WORD)
    resized as WORD with the job's filter (chroma sited LEFT, so shifted
    horizontally by 1/4 - 1/4 x scale), then ordered (Bayer) dither to 8 bits,
    as the reference's graph for that format pair."""
    (sw, sh), (dw, dh), bits = job["src"], job["dst"], job["bits"]
    filt = {"bilinear": (1, 0.0, 0.0), "bicubic": (2, 0.0, 0.5), "spline36": (4, 0.0, 0.0), "lanczos3": (5, 3.0, 0.0)}[job["filter"]]
    res = []
    for q, plane in enumerate(planes):
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


def yuv420_10_709_to_2020(Y, U, V, approx=0):
    """FHD 4:2:0 10-bit limited BT.709 -> BT.2020 NCL (transfer BT.709), chroma
    siting LEFT (API default), bicubic b=0 c=0.5 chroma filter. approx=1: the
    colourspace chain of allow_approximate_gamma (the reference's AVX2 path).

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
    n, ops, _ = colorspace_ops((M_709, T_709, P_709), (M_2020, T_709, P_2020), 100.0, 0, 0, approx=approx)
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


# ---- plane geometry of the reference's graph builder ---------------------------
# internal_state (graphbuilder.cpp:368-392): every plane carries its size and an
# active window in its own sample units; chroma planes are scaled from luma
# (chroma_from_luma, :331-346) and shifted by chroma siting and field parity
# (apply_pixel_siting, :348-366, with the offsets of :64-171). A resize between
# two planes maps the destination window onto the source (resize_plane,
# :769-776).

CHROMA_LOC = {0: ("left", "center"), 1: ("center", "center"), 2: ("left", "top"), 3: ("center", "top"),
              4: ("left", "bottom"), 5: ("center", "bottom")}


def _chroma_offset_w(loc, sub):
    return -0.5 + 0.5 * sub if loc == "left" else 0.0


def _chroma_offset_h(loc, sub):
    if loc == "top":
        return -0.5 + 0.5 * sub
    if loc == "bottom":
        return 0.5 - 0.5 * sub
    return 0.0


def _parity_adjust(parity, off):
    return -0.25 + off / 2 if parity == 1 else 0.25 + off / 2 if parity == 2 else off


def _luma_parity(parity):
    return -0.25 if parity == 1 else 0.25 if parity == 2 else 0.0


def plane_geometry(width, height, grey=False, ssw=0, ssh=0, parity=0, chroma_location=0, active=None):
    """[(width, height, left, top, active_width, active_height)] for Y, then U and V."""
    al, at, aw, ah = active if active is not None else (0.0, 0.0, float(width), float(height))
    planes = [[width, height, al, at, aw, ah]]
    if not grey:
        sw, sh = 1.0 / (1 << ssw), 1.0 / (1 << ssh)
        u = [width >> ssw, height >> ssh, al * sw, at * sh, aw * sw, ah * sh]
        planes += [u, list(u)]
    planes[0][3] -= _luma_parity(parity)
    if not grey:
        loc_w, loc_h = CHROMA_LOC[chroma_location]
        sub_w = planes[1][0] / planes[0][0]
        sub_h = planes[1][1] / planes[0][1]
        ow = _chroma_offset_w(loc_w, sub_w)
        oh = _parity_adjust(parity, _chroma_offset_h(loc_h, sub_h))
        for q in (1, 2):
            planes[q][2] -= ow
            planes[q][3] -= oh
    return [tuple(p) for p in planes]


def resize_params(src_plane, dst_plane):
    """(shift_w, shift_h, subwidth, subheight) of a resize between two planes, or
    None when nothing about the plane changes (graphbuilder.cpp:769-776)."""
    if tuple(src_plane) == tuple(dst_plane):
        return None
    _, _, sl, st, saw, sah = src_plane
    dw, dh, dl, dt, daw, dah = dst_plane
    scale_w = daw / saw
    scale_h = dah / sah
    return (sl - dl / scale_w, st - dt / scale_h, saw * (dw / daw), sah * (dh / dah))


def resize_geom(img, src_plane, dst_plane, ftype, a, b, depth=16):
    """One plane through resize_plane with the builder's geometry."""
    rp = resize_params(src_plane, dst_plane)
    if rp is None:
        return img.copy()
    sw_, sh_, subw, subh = rp
    return resize_plane(img, ftype, a, b, dst_plane[0], dst_plane[1], shift_w=sw_, shift_h=sh_, subwidth=subw, subheight=subh,
                        depth=depth)
