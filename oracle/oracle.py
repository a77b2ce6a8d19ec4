"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by the tests, the smoke check and the bench's CPU-baseline leg as the checker,
never by the product.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "liboracle.so")
REF_STATS = os.path.join(HERE, "_ref", "libref_stats.so")
HIST_BINS = 20001

if not os.path.exists(PATH):
    raise ImportError(f"{PATH} is not built: run make -C oracle")
lib = C.CDLL(PATH)

_i, _ll, _dp, _vp = C.c_int, C.c_longlong, C.c_double, C.c_void_p


class Params(C.Structure):
    _fields_ = [("w", _i), ("h", _i), ("depth", _i), ("window", _i), ("pad", _i), ("mask_index", _i),
                ("vlt", _i), ("band_base", _i), ("band_size", _i), ("tvi", _i * 4), ("weight", _i * 4),
                ("sw", _i * 5), ("sh", _i * 5), ("n", _ll * 5), ("k", _ll * 5), ("lut_size", _i)]

    def offsets(self):
        off = [0]
        for s in range(5):
            off.append(off[-1] + self.n[s])
        return off


class Result(C.Structure):
    _fields_ = [("sum_fx", C.c_ulonglong * 5), ("count", C.c_ulonglong * 5), ("k", C.c_ulonglong * 5),
                ("thr_bits", C.c_uint * 5), ("above", C.c_ulonglong * 5), ("above_fx", C.c_ulonglong * 5),
                ("pooled", C.c_double * 5), ("score", C.c_double)]


class Hist(C.Structure):
    _fields_ = [("bins", C.c_uint64 * HIST_BINS), ("n", C.c_uint64)]


class Accum(C.Structure):
    _fields_ = [("frames", C.c_uint64), ("cambi", _dp), ("cambi_source", _dp),
                ("all_r", _dp * 3), ("all_d", _dp * 3), ("st_r", _dp * 3), ("st_d", _dp * 3),
                ("sh_all_r", _dp * 2), ("sh_all_x", _dp * 2), ("sh_st_r", _dp * 2), ("sh_st_x", _dp * 2),
                ("own", _dp), ("mine", _dp), ("amount", Hist * 3), ("matching", Hist * 2), ("lasting", Hist),
                ("psnr", _dp * 3), ("color_r", _dp * 3), ("color_d", _dp * 3), ("vmaf", _dp), ("vmaf_min", _dp)]


class Frame(C.Structure):
    _fields_ = [("y", _vp), ("u", _vp), ("v", _vp), ("ypitch", _i), ("cpitch", _i), ("semi", _i), ("bytes", _i),
                ("shift", _i), ("depth", _i)]


class Parts(C.Structure):
    _fields_ = [("fidelity", _dp), ("texture", _dp), ("banding", _dp), ("grain", _dp), ("score", _dp)]


class RunConfig(C.Structure):
    _fields_ = [("w", _i), ("h", _i), ("depth", _i), ("up", _i), ("win", _i), ("per_segment", _i),
                ("window_base", _i), ("threads", _i)]


def _sig(name, args, res=None):
    f = getattr(lib, name)
    f.argtypes = args
    f.restype = res
    return f


_sig("orc_bt1886_luminance", [_i], _dp)
_sig("orc_tvi_for_diff", [_i, _dp], _i)
_sig("orc_vlt_luma", [_dp], _i)
_sig("orc_cambi_params_init", [C.POINTER(Params), _i, _i, _i, _i], _i)
_sig("orc_recip_lut", [_vp, _i])
_sig("orc_from8", [_vp, _i, _i, _i, _vp])
_sig("orc_spatial_mask", [_vp, _i, _i, _i, _vp])
_sig("orc_decimate16", [_vp, _i, _i, _i, _vp])
_sig("orc_decimate8", [_vp, _i, _i, _i, _vp])
_sig("orc_mode_filter", [_vp, _i, _i, _vp])
_sig("orc_pyramid", [C.POINTER(Params), _vp, _i, _vp, _vp])
_sig("orc_cvalues", [C.POINTER(Params), _vp, _vp, _i, _i, _vp, _vp, _i, _i])
_sig("orc_fixed24", [C.c_float], C.c_ulonglong)
_sig("orc_pool", [_vp, _ll, _ll, C.POINTER(Result), _i])
_sig("orc_cambi_score_of", [C.POINTER(Params), C.POINTER(Result)], _dp)
_sig("orc_cambi", [C.POINTER(Params), _vp, _i, _vp, C.POINTER(Result), _i, _i], _i)
_sig("orc_window_stats", [_vp, _i, _vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp])
_sig("orc_window_values", [_vp, _vp])
_sig("orc_persist", [_vp, _vp, _i, _i, _i, _vp])
_sig("orc_persist_values", [_vp, _vp])
_sig("orc_quantile_linear", [_vp, _ll, _dp], _dp)
_sig("orc_quantile_f32", [_vp, _ll, _dp], C.c_float)
_sig("orc_select_rank_u64", [_vp, _ll, _ll], C.c_ulonglong)
_sig("orc_tree_sum256", [_vp, _ll], _dp)
_sig("orc_hist_bin", [_dp], _i)
_sig("orc_hist_add", [C.POINTER(Hist), _dp])
_sig("orc_hist_quantile", [C.POINTER(Hist), _dp], _dp)
_sig("orc_accum_clear", [C.POINTER(Accum)])
_sig("orc_accum_merge", [C.POINTER(Accum), C.POINTER(Accum)])
_sig("orc_accum_score", [C.POINTER(Accum), _dp, C.POINTER(Parts)])
_sig("orc_collect_frame", [C.POINTER(Accum), _vp, _i, _vp])
_sig("orc_collect_lasting", [C.POINTER(Accum), _vp, _vp, _i])
_sig("orc_run_create", [C.POINTER(RunConfig)], _vp)
_sig("orc_run_destroy", [_vp])
_sig("orc_run_frame", [_vp, _vp, _i, _vp, _i, C.POINTER(_dp), C.POINTER(_dp)], _i)
_sig("orc_run_frame_yuv", [_vp, _vp, _vp, C.POINTER(_dp), C.POINTER(_dp), _vp], _i)
_sig("orc_run_segments", [_vp], _i)
_sig("orc_frame_planes", [C.POINTER(Frame), _i, _i, _i, _vp, _vp, _vp, _i, _i])
_sig("orc_plane_sse", [_vp, _vp, _ll], C.c_ulonglong)
_sig("orc_plane_sse_f", [_vp, _vp, _ll], C.c_ulonglong)
_sig("orc_strength_fixed", [C.c_uint], C.c_ulonglong)
_sig("orc_color_sums", [_vp, _vp, _ll, _i, _vp])
_sig("orc_psnr", [C.c_ulonglong, _ll, _i], _dp)
_sig("orc_run_segment", [_vp, _i], C.POINTER(Accum))


def params(w, h, depth=10, window_base=65):
    p = Params()
    assert lib.orc_cambi_params_init(C.byref(p), w, h, depth, window_base) == 0
    return p


def recip_lut(n):
    out = np.empty(n, np.float32)
    lib.orc_recip_lut(out.ctypes.data, n)
    return out


def spatial_mask(img, index):
    img = np.ascontiguousarray(img, np.uint16)
    h, w = img.shape
    out = np.empty((h, w), np.uint8)
    lib.orc_spatial_mask(img.ctypes.data, w, h, index, out.ctypes.data)
    return out


def mode_filter(img):
    img = np.ascontiguousarray(img, np.uint16)
    h, w = img.shape
    out = np.empty((h, w), np.uint16)
    lib.orc_mode_filter(img.ctypes.data, w, h, out.ctypes.data)
    return out


def from8(img):
    img = np.ascontiguousarray(img, np.uint16)
    h, w = img.shape
    out = np.empty((h, w), np.uint16)
    lib.orc_from8(img.ctypes.data, w, w, h, out.ctypes.data)
    return out


def pyramid(p, img):
    """Images and masks of the 5 scales, as lists of 2-D arrays."""
    img = np.ascontiguousarray(img, np.uint16)
    off = p.offsets()
    images = np.empty(off[5], np.uint16)
    masks = np.empty(off[5], np.uint8)
    lib.orc_pyramid(C.byref(p), img.ctypes.data, img.shape[1], images.ctypes.data, masks.ctypes.data)
    ims = [images[off[s]:off[s + 1]].reshape(p.sh[s], p.sw[s]) for s in range(5)]
    mks = [masks[off[s]:off[s + 1]].reshape(p.sh[s], p.sw[s]) for s in range(5)]
    return ims, mks


def cvalues(p, img, mask, brute=False, threads=0):
    img = np.ascontiguousarray(img, np.uint16)
    mask = np.ascontiguousarray(mask, np.uint8)
    h, w = img.shape
    lut = recip_lut(p.lut_size)
    out = np.empty((h, w), np.float32)
    lib.orc_cvalues(C.byref(p), img.ctypes.data, mask.ctypes.data, w, h, lut.ctypes.data, out.ctypes.data, int(brute),
                    threads)
    return out


def cambi(p, img, brute=False, threads=0, want_cvalues=False):
    img = np.ascontiguousarray(img, np.uint16)
    r = Result()
    cv = np.empty(p.offsets()[5], np.float32) if want_cvalues else None
    assert lib.orc_cambi(C.byref(p), img.ctypes.data, img.shape[1], None if cv is None else cv.ctypes.data,
                         C.byref(r), int(brute), threads) == 0
    if want_cvalues:
        off = p.offsets()
        return r, [cv[off[s]:off[s + 1]].reshape(p.sh[s], p.sw[s]) for s in range(5)]
    return r


def pool(c, k):
    c = np.ascontiguousarray(c, np.float32).ravel()
    r = Result()
    lib.orc_pool(c.ctypes.data, c.size, k, C.byref(r), 0)
    return r


def window_stats(yr, yd, up, win, want_mid=True):
    yr = np.ascontiguousarray(yr, np.uint16)
    yd = np.ascontiguousarray(yd, np.uint16)
    h, w = yr.shape
    windows = (w // win) * (h // win)
    sums = np.empty((windows, 12), np.int64)
    stats = np.empty((windows, 8), np.float64)
    mid_r = np.zeros((h, w), np.int32) if want_mid else None
    mid_d = np.zeros((h, w), np.int32) if want_mid else None
    lib.orc_window_stats(yr.ctypes.data, w, yd.ctypes.data, w, up, w, h, win,
                         None if mid_r is None else mid_r.ctypes.data, None if mid_d is None else mid_d.ctypes.data,
                         sums.ctypes.data, stats.ctypes.data)
    return sums, stats, mid_r, mid_d


def persist(ring, mid_d, win):
    """ring: 13 int32 planes at offsets -6..6."""
    planes = [np.ascontiguousarray(r, np.int32) for r in ring]
    h, w = planes[0].shape
    ptrs = (_vp * 13)(*[r.ctypes.data for r in planes])
    mid_d = np.ascontiguousarray(mid_d, np.int32)
    sums = np.empty(((w // win) * (h // win), 5), np.int64)
    lib.orc_persist(ptrs, mid_d.ctypes.data, w, h, win, sums.ctypes.data)
    return sums


def quantile_linear(v, q):
    v = np.ascontiguousarray(v, np.float64)
    return lib.orc_quantile_linear(v.ctypes.data, v.size, q)


def quantile_f32(v, q):
    """The structure threshold as the reference evaluates it (src/main.cpp:279-289) on float keys."""
    v = np.ascontiguousarray(v, np.float32)
    return lib.orc_quantile_f32(v.ctypes.data, v.size, q)


def tree_sum256(v):
    v = np.ascontiguousarray(v, np.float64)
    return lib.orc_tree_sum256(v.ctypes.data, v.size)


def hist_of(pcts):
    h = Hist()
    for p in pcts:
        lib.orc_hist_add(C.byref(h), float(p))
    return h


def hist_quantile(h, q):
    return lib.orc_hist_quantile(C.byref(h), q)


# Formats of decoded frames: (semi-planar, bytes per sample, right shift, depth), as view_of's table
# (src/main.cpp:135-147).
FORMATS = {"yuv420p": (0, 1, 0, 8), "yuv420p10": (0, 2, 0, 10), "nv12": (1, 1, 0, 8), "p010": (1, 2, 6, 10)}


def frame_planes(fmt, y, u, v, x0, y0, w, h, pd):
    """frame_planes of a frame held in numpy planes: y (rows x pitch samples); u and v for a planar format, or u the
    interleaved (rows x 2 * cw) plane of a semi-planar one.  Returns packed y16, u, v."""
    semi, nbytes, shift, depth = FORMATS[fmt]
    dt = np.uint8 if nbytes == 1 else np.uint16
    y = np.ascontiguousarray(y, dt)
    u = np.ascontiguousarray(u, dt)
    v = None if semi else np.ascontiguousarray(v, dt)
    f = Frame(y.ctypes.data, u.ctypes.data, 0 if v is None else v.ctypes.data, y.shape[1] * nbytes,
              u.shape[1] * nbytes, semi, nbytes, shift, depth)
    oy = np.empty((h, w), np.uint16)
    ou = np.empty((h // 2, w // 2), np.uint16)
    ov = np.empty((h // 2, w // 2), np.uint16)
    lib.orc_frame_planes(C.byref(f), x0, y0, pd, oy.ctypes.data, ou.ctypes.data, ov.ctypes.data, w, h)
    return oy, ou, ov


def plane_sse(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.dtype == np.float32:
        return lib.orc_plane_sse_f(a.ctypes.data, b.ctypes.data, a.size)
    a = a.astype(np.uint16, copy=False)
    b = b.astype(np.uint16, copy=False)
    return lib.orc_plane_sse(a.ctypes.data, b.ctypes.data, a.size)


def strength_fixed(s):
    return lib.orc_strength_fixed(int(s))


def color_sums(u, v, up):
    u = np.ascontiguousarray(u, np.uint16)
    v = np.ascontiguousarray(v, np.uint16)
    out = np.zeros(3, np.uint64)
    lib.orc_color_sums(u.ctypes.data, v.ctypes.data, u.size, up, out.ctypes.data)
    return out


def psnr(sse, n, depth):
    return lib.orc_psnr(int(sse), int(n), depth)


class Run:
    """The whole run of one (source, encode) pair (src/main.cpp's queue/collect) on the CPU."""

    def __init__(self, w, h, depth=10, up=0, win=16, per_segment=1440, window_base=65, threads=0):
        self.cfg = RunConfig(w, h, depth, up, win, per_segment, window_base, threads)
        self.ptr = lib.orc_run_create(C.byref(self.cfg))
        assert self.ptr

    def frame(self, src, enc):
        src = np.ascontiguousarray(src, np.uint16)
        enc = np.ascontiguousarray(enc, np.uint16)
        a, b = _dp(), _dp()
        lib.orc_run_frame(self.ptr, src.ctypes.data, src.shape[1], enc.ctypes.data, enc.shape[1], C.byref(a),
                          C.byref(b))
        return a.value, b.value

    def frame_yuv(self, src, enc):
        """src, enc: (y, u, v) planes at the pipeline depth.  Returns (cambi, cambi_source, psnr[3])."""
        keep = [np.ascontiguousarray(p, np.uint16) for p in (*src, *enc)]
        sp = (_vp * 3)(*[p.ctypes.data for p in keep[:3]])
        ep = (_vp * 3)(*[p.ctypes.data for p in keep[3:]])
        a, b = _dp(), _dp()
        ps = np.zeros(3)
        assert lib.orc_run_frame_yuv(self.ptr, sp, ep, C.byref(a), C.byref(b), ps.ctypes.data) == 0
        return a.value, b.value, ps

    def segments(self):
        return lib.orc_run_segments(self.ptr)

    def segment(self, s):
        return lib.orc_run_segment(self.ptr, s).contents

    def merged(self):
        a = Accum()
        lib.orc_accum_clear(C.byref(a))
        for s in range(self.segments()):
            lib.orc_accum_merge(C.byref(a), C.byref(self.segment(s)))
        return a

    def __del__(self):
        try:
            lib.orc_run_destroy(self.ptr)
        except Exception:
            pass


def accum_summary(a, vmaf_total=float("nan")):
    """The figures es_scorer_segment reports, from an oracle accumulator."""
    parts = Parts()
    lib.orc_accum_score(C.byref(a), vmaf_total, C.byref(parts))
    out = {"frames": a.frames, "cambi": a.cambi, "cambi_source": a.cambi_source,
           "all_r": list(a.all_r), "all_d": list(a.all_d), "st_r": list(a.st_r), "st_d": list(a.st_d),
           "sh_all_r": list(a.sh_all_r), "sh_all_x": list(a.sh_all_x), "sh_st_r": list(a.sh_st_r),
           "sh_st_x": list(a.sh_st_x), "own": a.own, "mine": a.mine,
           "n_amount": [a.amount[k].n for k in range(3)], "n_matching": [a.matching[k].n for k in range(2)],
           "n_lasting": a.lasting.n,
           "median_amount": [hist_quantile(a.amount[k], 0.5) for k in range(3)],
           "tenth_amount": [hist_quantile(a.amount[k], 0.1) for k in range(3)],
           "median_matching": [hist_quantile(a.matching[k], 0.5) for k in range(2)],
           "tenth_matching": [hist_quantile(a.matching[k], 0.1) for k in range(2)],
           "median_lasting": hist_quantile(a.lasting, 0.5), "tenth_lasting": hist_quantile(a.lasting, 0.1),
           "texture": parts.texture, "banding": parts.banding, "grain": parts.grain, "score": parts.score,
           "psnr": list(a.psnr), "color_source": list(a.color_r), "color_encode": list(a.color_d)}
    hists = np.zeros((6, HIST_BINS), np.uint64)
    for k in range(3):
        hists[k] = np.frombuffer(a.amount[k].bins, np.uint64)
    for k in range(2):
        hists[3 + k] = np.frombuffer(a.matching[k].bins, np.uint64)
    hists[5] = np.frombuffer(a.lasting.bins, np.uint64)
    return out, hists


# ---- the source-CAMBI cache's file naming and records (src/cache.cpp:14-59) ----

FNV_OFFSET, FNV_PRIME = 0xcbf29ce484222325, 0x100000001b3


def fnv1a64(data, h=FNV_OFFSET):
    """64-bit FNV-1a (Fowler, Noll, Vo): for each byte, xor then multiply by the FNV prime, modulo 2^64."""
    for b in bytes(data):
        h = ((h ^ b) * FNV_PRIME) & 0xFFFFFFFFFFFFFFFF
    return h


def cambi_cache_name(path, version, x0, y0, w, h, depth, sample=1 << 20):
    """The cache file's name: FNV-1a of "cambi <version> <size> <x0> <y0> <w> <h> <depth>", then of the source's first
    and last `sample` bytes (the whole file twice when it is shorter)."""
    data = open(path, "rb").read()
    size = len(data)
    hsh = fnv1a64(f"cambi {version} {size} {x0} {y0} {w} {h} {depth}".encode())
    hsh = fnv1a64(data[:sample], hsh)
    hsh = fnv1a64(data[max(size - sample, 0):][:sample], hsh)
    return f"{hsh:016x}.cambi"


def cambi_cache_records(file):
    """The whole 16-byte {int64 pts, double score} records of a cache file, as (pts, score) pairs."""
    raw = open(file, "rb").read()
    n = len(raw) // 16
    rec = np.frombuffer(raw[:n * 16], dtype=[("pts", "<i8"), ("score", "<f8")])
    return [(int(a), float(b)) for a, b in rec]


def ref_stats():
    """The reference's own stats.hpp behind oracle/_ref's shim, or None when it was not built."""
    if not os.path.exists(REF_STATS):
        return None
    r = C.CDLL(REF_STATS)
    r.ref_hist.argtypes = [_vp, C.c_long, _vp, _vp, _vp, _i, _vp]
    r.ref_hist.restype = None
    r.ref_score.argtypes = [C.c_uint64, _dp, _dp, _dp, _dp, _dp, _vp, C.c_long, _vp]
    r.ref_score.restype = None
    return r
