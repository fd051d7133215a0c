"""Seeded synthetic inputs and formats of the benchmark workloads.

  uhd_lanczos      one WORD plane 1920x1080 -> 3840x2160, Lanczos (3 taps)
  fhd_709_to_2020  FHD YUV 4:2:0 10-bit limited, BT.709 -> BT.2020 NCL, bicubic chroma
  fp32_down/up     FP32 plane 3840x2160 -> 1920x1080 Lanczos3, 1280x720 -> 3840x2160 bicubic
  uhd_pq_to_hlg    3840x2160 RGB FP32 PQ -> HLG (203-nit peak) -> 12-bit with error diffusion

Used by the tests and bench.py; nothing here is on the processing path.
"""
import numpy as np

from . import (COLOR_RGB, COLOR_YUV, DITHER_ERROR_DIFFUSION, DITHER_ORDERED, MATRIX_BT709, MATRIX_BT2020_NCL, MATRIX_RGB,
               PIXEL_BYTE, PIXEL_FLOAT, PIXEL_WORD, PRIMARIES_BT709, PRIMARIES_BT2020, RANGE_FULL, RANGE_LIMITED, RESIZE_BICUBIC,
               RESIZE_BILINEAR, RESIZE_LANCZOS, RESIZE_SPLINE36, TRANSFER_ARIB_B67, TRANSFER_BT709, TRANSFER_ST2084, builder_params,
               image_format)


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


# ---- WORD Lanczos upscale ------------------------------------------------------

def word_plane(w, h, seed=0):
    """clamp(round_half_even(32768 + 16000 sin(2 pi x/97) cos(2 pi y/61) + N(0, 2000)), 0, 65535)."""
    r = rng(seed)
    x = np.arange(w, dtype=np.float64)[None, :]
    y = np.arange(h, dtype=np.float64)[:, None]
    v = 32768.0 + 16000.0 * np.sin(2 * np.pi * x / 97) * np.cos(2 * np.pi * y / 61) + r.normal(0.0, 2000.0, (h, w))
    return np.clip(np.rint(v), 0, 65535).astype(np.uint16)


def uhd_lanczos_formats(sw=1920, sh=1080, dw=3840, dh=2160, depth=16, taps=3):
    src = image_format(width=sw, height=sh, pixel_type=PIXEL_WORD, depth=depth)
    dst = image_format(width=dw, height=dh, pixel_type=PIXEL_WORD, depth=depth)
    return src, dst, builder_params(resample_filter=RESIZE_LANCZOS, filter_param_a=taps)


# ---- FHD BT.709 -> BT.2020 ----------------------------------------------------

def yuv420_10bit(w, h, seed=0):
    """Smooth RGB gradient field + N(0, 8 code values) converted to BT.709 Y'CbCr,
    10-bit limited range (Y 64-940, CbCr 64-960), chroma 2x2 box-averaged."""
    r = rng(seed)
    x = np.linspace(0.0, 1.0, w)[None, :]
    y = np.linspace(0.0, 1.0, h)[:, None]
    R = 0.5 + 0.45 * np.sin(2 * np.pi * (x * 1.3 + y * 0.4))
    G = 0.5 + 0.45 * np.cos(2 * np.pi * (x * 0.7 - y * 1.1))
    B = 0.5 + 0.45 * np.sin(2 * np.pi * (x * 0.5 + y * 0.9) + 1.0)
    kr, kb = 0.2126, 0.0722
    kg = 1 - kr - kb
    Y = kr * R + kg * G + kb * B
    Cb = (B - Y) / (2 - 2 * kb)
    Cr = (R - Y) / (2 - 2 * kr)
    Yq = 64 + 876 * Y + r.normal(0, 8, (h, w))
    Cbq = 512 + 896 * Cb + r.normal(0, 8, (h, w))
    Crq = 512 + 896 * Cr + r.normal(0, 8, (h, w))
    sub = lambda c: 0.25 * (c[0::2, 0::2] + c[1::2, 0::2] + c[0::2, 1::2] + c[1::2, 1::2])
    Yo = np.clip(np.rint(Yq), 64, 940).astype(np.uint16)
    Uo = np.clip(np.rint(sub(Cbq)), 64, 960).astype(np.uint16)
    Vo = np.clip(np.rint(sub(Crq)), 64, 960).astype(np.uint16)
    return Yo, Uo, Vo


def fhd_709_to_2020_formats(w=1920, h=1080):
    common = dict(width=w, height=h, pixel_type=PIXEL_WORD, depth=10, subsample_w=1, subsample_h=1, color_family=COLOR_YUV,
                  pixel_range=RANGE_LIMITED)
    src = image_format(**common, matrix_coefficients=MATRIX_BT709, transfer_characteristics=TRANSFER_BT709,
                       color_primaries=PRIMARIES_BT709)
    dst = image_format(**common, matrix_coefficients=MATRIX_BT2020_NCL, transfer_characteristics=TRANSFER_BT709,
                       color_primaries=PRIMARIES_BT2020)
    # The API default chroma filter is bilinear; this workload uses bicubic b=0, c=0.5.
    return src, dst, builder_params(resample_filter_uv=RESIZE_BICUBIC, filter_param_a_uv=0.0, filter_param_b_uv=0.5)


# ---- FP32 precision planes ----------------------------------------------------

def fp32_plane(w, h, seed=0):
    """U[0,1] mixed with vertical step edges every 64 px (alternating 0 / 1 bands)."""
    r = rng(seed)
    v = r.random((h, w), dtype=np.float32)
    band = ((np.arange(w) // 64) % 2).astype(np.float32)[None, :]
    edge_cols = (np.arange(w) % 64) < 8
    v[:, edge_cols] = np.broadcast_to(band, (h, w))[:, edge_cols]
    return v


def fp32_formats(sw, sh, dw, dh, filt, a=float("nan"), b=float("nan")):
    src = image_format(width=sw, height=sh, pixel_type=PIXEL_FLOAT)
    dst = image_format(width=dw, height=dh, pixel_type=PIXEL_FLOAT)
    return src, dst, builder_params(resample_filter=filt, filter_param_a=a, filter_param_b=b)


# ---- UHD PQ -> HLG with error diffusion ---------------------------------------

def pq_rgb(w, h, seed=0):
    """PQ code values U[0,1] with 10% of pixels drawn from [0.9, 1] (highlights)."""
    r = rng(seed)
    planes = []
    for _ in range(3):
        v = r.random((h, w), dtype=np.float32)
        hi = r.random((h, w)) < 0.10
        v[hi] = (0.9 + 0.1 * r.random(int(hi.sum()))).astype(np.float32)
        planes.append(v)
    return planes


def uhd_pq_to_hlg_formats(w=3840, h=2160, peak=203.0):
    src = image_format(width=w, height=h, pixel_type=PIXEL_FLOAT, color_family=COLOR_RGB, matrix_coefficients=MATRIX_RGB,
                       transfer_characteristics=TRANSFER_ST2084, color_primaries=PRIMARIES_BT2020)
    dst = image_format(width=w, height=h, pixel_type=PIXEL_WORD, depth=12, pixel_range=RANGE_FULL, color_family=COLOR_RGB,
                       matrix_coefficients=MATRIX_RGB, transfer_characteristics=TRANSFER_ARIB_B67,
                       color_primaries=PRIMARIES_BT2020)
    return src, dst, builder_params(dither_type=DITHER_ERROR_DIFFUSION, nominal_peak_luminance=peak)


# Batched uneven jobs: YUV 4:2:0, BYTE 8-bit or WORD 10-bit limited range, each
# resized to another size of the set with its own filter, then ordered-dithered
# to 8-bit. Sizes and filters are drawn from a seeded generator.
BATCH_SIZES = [(1280, 720), (1920, 1080), (2560, 1440), (3840, 2160)]
BATCH_FILTERS = [("bilinear", RESIZE_BILINEAR), ("bicubic", RESIZE_BICUBIC), ("spline36", RESIZE_SPLINE36), ("lanczos3", RESIZE_LANCZOS)]


def batch_jobs(n=64, seed=2024, sizes=BATCH_SIZES):
    r = rng(seed)
    jobs = []
    for k in range(n):
        si = int(r.integers(len(sizes)))
        di = int(r.integers(len(sizes) - 1))
        di = di + 1 if di >= si else di
        fname, _ = BATCH_FILTERS[int(r.integers(len(BATCH_FILTERS)))]
        bits = 8 if r.random() < 0.5 else 10
        jobs.append({"src": list(sizes[si]), "dst": list(sizes[di]), "filter": fname, "bits": bits})
    return jobs


def batch_job_graph(job):
    """(src format, dst format, params) of one batched job."""
    common = dict(subsample_w=1, subsample_h=1, color_family=COLOR_YUV, matrix_coefficients=MATRIX_BT709,
                  transfer_characteristics=TRANSFER_BT709, color_primaries=PRIMARIES_BT709)
    (sw, sh), (dw, dh) = job["src"], job["dst"]
    bits = job["bits"]
    src = image_format(width=sw, height=sh, pixel_type=PIXEL_BYTE if bits == 8 else PIXEL_WORD, depth=bits, **common)
    dst = image_format(width=dw, height=dh, pixel_type=PIXEL_BYTE, depth=8, **common)
    filt = dict(BATCH_FILTERS)[job["filter"]]
    a = 3.0 if filt == RESIZE_LANCZOS else float("nan")
    b = float("nan")
    if filt == RESIZE_BICUBIC:
        a, b = 0.0, 0.5
    params = builder_params(resample_filter=filt, resample_filter_uv=filt, filter_param_a=a, filter_param_b=b,
                            filter_param_a_uv=a, filter_param_b_uv=b, dither_type=DITHER_ORDERED)
    return src, dst, params


def batch_job_planes(job, seed):
    """Seeded Y, U, V planes of a job: smooth gradients plus noise, limited range."""
    (w, h), bits = job["src"], job["bits"]
    r = rng(seed)
    lo, hi = (16, 235) if bits == 8 else (64, 940)
    clo, chi = (16, 240) if bits == 8 else (64, 960)
    dt = np.uint8 if bits == 8 else np.uint16
    yy, xx = np.mgrid[0:h, 0:w]
    Y = lo + (hi - lo) * (0.5 + 0.4 * np.sin(xx / 53.0) * np.cos(yy / 37.0)) + r.normal(0, 4 if bits == 8 else 16, (h, w))
    cy, cx = np.mgrid[0:h // 2, 0:w // 2]
    U = (clo + chi) / 2 + (chi - clo) * 0.3 * np.sin(cx / 29.0 + cy / 41.0) + r.normal(0, 3, (h // 2, w // 2))
    V = (clo + chi) / 2 + (chi - clo) * 0.3 * np.cos(cx / 31.0 - cy / 23.0) + r.normal(0, 3, (h // 2, w // 2))
    return [np.clip(np.rint(P), lo if P is Y else clo, hi if P is Y else chi).astype(dt) for P in (Y, U, V)]
