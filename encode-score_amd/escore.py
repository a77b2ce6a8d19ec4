"""ctypes bindings of libencscore_hip (include/encscore.h).

Device memory is the library's own (es_malloc); host data are numpy arrays.  Every call checks the status and raises
EsError with the library's message.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(HERE, "libencscore_hip.so")

ES_PLAIN = 0
ES_BATCHED = 1
HIST_BINS = 20001

if not os.path.exists(library_path):
    raise ImportError(f"{library_path} is not built: run __graft_entry__.build() or make -C encode-score_amd")

lib = C.CDLL(library_path)


class EsError(RuntimeError):
    pass


_vp, _i, _sz, _ll, _u64, _dp = C.c_void_p, C.c_int, C.c_size_t, C.c_longlong, C.c_uint64, C.POINTER(C.c_double)


class CambiInfo(C.Structure):
    _fields_ = [("w", _i), ("h", _i), ("depth", _i), ("window", _i), ("pad", _i), ("mask_index", _i),
                ("vlt", _i), ("band_base", _i), ("band_size", _i), ("tvi", _i * 4), ("weight", _i * 4),
                ("sw", _i * 5), ("sh", _i * 5), ("n", _ll * 5), ("k", _ll * 5)]


class CambiScale(C.Structure):
    _fields_ = [("sum_fx", C.c_ulonglong), ("count", C.c_ulonglong), ("k", C.c_ulonglong), ("thr_bits", C.c_uint),
                ("above", C.c_ulonglong), ("above_fx", C.c_ulonglong), ("pooled", C.c_double)]


class EncodeDesc(C.Structure):
    _fields_ = [("w", _i), ("h", _i), ("length", _i), ("pitch", _i), ("bytes", _i), ("shift", _i)]


class ScorerConfig(C.Structure):
    _fields_ = [("mode", _i), ("src_w", _i), ("src_h", _i), ("src_pitch", _i), ("src_bytes", _i), ("src_shift", _i),
                ("depth", _i), ("win", _i), ("per_segment", _i), ("window_base", _i), ("batch", _i), ("tile_h", _i),
                ("use_graph", _i), ("n_encodes", _i), ("encodes", C.POINTER(EncodeDesc)), ("profile", _i),
                ("cv_bmax", _i), ("cv_waves", _i), ("serial_stages", _i), ("cv_tile_blocks", _i), ("early_energies", _i),
                ("cv_mfma", _i), ("cv_queues", _i), ("tile_h0", _i)]


class TuneResult(C.Structure):
    _fields_ = [("batch", _i), ("tile_h", _i), ("cv_bmax", _i), ("cv_waves", _i), ("ms_per_step", C.c_double),
                ("trials", _i), ("cv_tile_blocks", _i), ("early_energies", _i), ("cv_mfma", _i), ("cv_queues", _i),
                ("tile_h0", _i)]


class Segment(C.Structure):
    _fields_ = [("frames", _u64), ("cambi", C.c_double), ("cambi_source", C.c_double),
                ("all_r", C.c_double * 3), ("all_d", C.c_double * 3), ("st_r", C.c_double * 3), ("st_d", C.c_double * 3),
                ("sh_all_r", C.c_double * 2), ("sh_all_x", C.c_double * 2), ("sh_st_r", C.c_double * 2),
                ("sh_st_x", C.c_double * 2), ("own", C.c_double), ("mine", C.c_double),
                ("n_amount", _u64 * 3), ("n_matching", _u64 * 2), ("n_lasting", _u64),
                ("median_amount", C.c_double * 3), ("tenth_amount", C.c_double * 3),
                ("median_matching", C.c_double * 2), ("tenth_matching", C.c_double * 2),
                ("median_lasting", C.c_double), ("tenth_lasting", C.c_double),
                ("texture", C.c_double), ("banding", C.c_double), ("grain", C.c_double), ("score", C.c_double),
                ("psnr", C.c_double * 3), ("color_source", C.c_double * 3), ("color_encode", C.c_double * 3),
                ("chroma_frames", _u64), ("vmaf", C.c_double), ("vmaf_min", C.c_double), ("vmaf_frames", _u64),
                ("fidelity", C.c_double)]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


ES_FMT_YUV420P, ES_FMT_YUV420P10, ES_FMT_NV12, ES_FMT_P010 = 0, 1, 2, 3
FORMATS = {"yuv420p": ES_FMT_YUV420P, "yuv420p10": ES_FMT_YUV420P10, "nv12": ES_FMT_NV12, "p010": ES_FMT_P010}


class Frame(C.Structure):
    """es_frame: a decoded 4:2:0 frame in device memory; pitches in bytes."""
    _fields_ = [("y", _vp), ("u", _vp), ("v", _vp), ("ypitch", _i), ("cpitch", _i), ("semi", _i), ("bytes", _i),
                ("shift", _i), ("depth", _i)]


class RunInfo(C.Structure):
    _fields_ = [("source", C.c_char_p), ("paths", C.POINTER(C.c_char_p)), ("model", C.c_char_p), ("start", C.c_double),
                ("offset", C.c_double), ("fps", C.c_double), ("segment_seconds", C.c_double), ("width", _i),
                ("height", _i)]


class SynthDesc(C.Structure):
    _fields_ = [("kind", _i), ("seed", _u64), ("salt", _i), ("full_w", _i), ("full_h", _i), ("crop_x", _i),
                ("crop_y", _i), ("w", _i), ("h", _i), ("depth", _i), ("grain_q8", _i), ("blur_q8", _i),
                ("quant_step", _i), ("quant_where", _i), ("block", _i), ("block_pct", _i), ("box3", _i)]


def _sig(name, args, res=_i):
    f = getattr(lib, name)
    f.argtypes = args
    f.restype = res
    return f


_sig("es_last_error", [], C.c_char_p)
_sig("es_cambi_cache_file", [_vp], C.c_char_p)
for _n, _a in [
    ("es_init", [_i]), ("es_device_count", [C.POINTER(_i)]), ("es_device_synchronize", []),
    ("es_stream_create", [C.POINTER(_vp)]), ("es_stream_destroy", [_vp]), ("es_stream_synchronize", [_vp]),
    ("es_event_create", [C.POINTER(_vp)]), ("es_event_destroy", [_vp]), ("es_event_record", [_vp, _vp]),
    ("es_stream_wait_event", [_vp, _vp]), ("es_event_synchronize", [_vp]),
    ("es_event_elapsed_ms", [_vp, _vp, C.POINTER(C.c_float)]),
    ("es_malloc", [C.POINTER(_vp), _sz]), ("es_free", [_vp]), ("es_host_alloc", [C.POINTER(_vp), _sz]),
    ("es_host_free", [_vp]), ("es_memcpy_htod", [_vp, _vp, _sz, _vp]), ("es_memcpy_dtoh", [_vp, _vp, _sz, _vp]),
    ("es_memcpy_dtod", [_vp, _vp, _sz, _vp]), ("es_memset", [_vp, _i, _sz, _vp]),
    ("es_cambi_params", [_i, _i, _i, _i, C.POINTER(CambiInfo)]),
    ("es_cambi_create", [_i, _i, _i, C.POINTER(_vp)]), ("es_cambi_create_ex", [_i, _i, _i, _i, _i, C.POINTER(_vp)]),
    ("es_cambi_destroy", [_vp]), ("es_cambi_launch", [_vp, _vp, _i, _vp]),
    ("es_cambi_launch_ex", [_vp, _vp, _i, _i, _i, _vp]), ("es_cambi_finish", [_vp, _dp]),
    ("es_cambi_scales", [_vp, C.POINTER(CambiScale)]), ("es_cambi_plane", [_vp, _i, _vp]),
    ("es_cambi_cvalues", [_vp, _i, _vp]), ("es_cambi_set_method", [_vp, _i, _i]),
    ("es_window_stats", [_vp, _vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp]),
    ("es_window_values", [_vp, _i, _vp]),
    ("es_persist_stats", [_vp, _sz, _i, _vp, _i, _i, _i, _vp, _vp]),
    ("es_window_stats_f32", [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp]),
    ("es_persist_stats_f32", [_vp, _sz, _i, _vp, _i, _i, _vp, _vp]),
    ("es_select_rank_u32", [_vp, _sz, _sz, C.POINTER(C.c_uint32), _vp]),
    ("es_select_rank_u64", [_vp, _sz, _sz, C.POINTER(C.c_uint64), _vp]),
    ("es_hist_quantiles", [_vp, _i, _vp, _i, _vp, _vp]), ("es_hist_quantiles_noalloc", [_vp, _i, _vp, _i, _vp, _vp, _vp]),
    ("es_scorer_create", [C.POINTER(ScorerConfig), C.POINTER(_vp)]), ("es_scorer_destroy", [_vp]),
    ("es_scorer_autotune", [C.POINTER(ScorerConfig), _i, _vp, _vp, C.POINTER(TuneResult)]),
    ("es_scorer_submit", [_vp, _i, _vp, _vp, _vp]), ("es_scorer_sync", [_vp]),
    ("es_scorer_frames", [_vp, C.POINTER(C.c_long)]), ("es_scorer_segments", [_vp, _i, C.POINTER(_i)]),
    ("es_scorer_segment", [_vp, _i, _i, C.POINTER(Segment), _vp]),
    ("es_scorer_cambi", [_vp, _i, C.c_long, _vp, _vp]), ("es_scorer_thresholds", [_vp, _i, C.c_long, _vp]),
    ("es_scorer_stage_ms", [_vp, _vp, C.POINTER(C.c_long)]), ("es_copy_kernel", [_vp, _vp, _sz, _vp]),
    ("es_synth_frame", [C.POINTER(SynthDesc), _i, _vp, _i, _vp]),
    ("es_frame_view", [_i, _vp, _vp, _vp, _i, _i, C.POINTER(Frame)]),
    ("es_frame_planes", [C.POINTER(Frame), _i, _i, _i, _vp, _vp, _vp, _i, _i, _vp]),
    ("es_plane_sse", [_vp, _vp, _vp, _vp, _sz, _vp, _vp]),
    ("es_color_sums", [_vp, _vp, _sz, _i, _vp, _vp]),
    ("es_color_values", [_vp, _sz, _vp]),
    ("es_psnr", [_u64, _sz, _i, _dp]),
    ("es_scorer_submit_frames", [_vp, _i, C.POINTER(Frame), C.POINTER(Frame), _vp]),
    ("es_scorer_psnr", [_vp, _i, C.c_long, _vp]),
    ("es_scorer_set_vmaf", [_vp, _i, C.c_long, _vp]),
    ("es_scorer_write_json", [_vp, C.POINTER(RunInfo), C.c_char_p]),
    ("es_cambi_cache_open", [C.c_char_p, _i, _i, _i, _i, _i, C.c_char_p, C.POINTER(_vp)]),
    ("es_cambi_cache_find", [_vp, C.c_int64, _dp, C.POINTER(_i)]),
    ("es_cambi_cache_add", [_vp, C.c_int64, C.c_double]), ("es_cambi_cache_close", [_vp]),
    ("es_scorer_use_cache", [_vp, C.c_char_p, C.c_char_p]),
    ("es_scorer_submit_pts", [_vp, _i, C.POINTER(Frame), C.POINTER(Frame), _vp, _vp]),
    ("es_scorer_cache_hits", [_vp, C.POINTER(C.c_long)]),
    ("es_scorer_cache_file", [_vp, _i, C.POINTER(C.c_char_p)]),
    ("es_synth_frame_host", [C.POINTER(SynthDesc), _i, _vp, _i, _i]),
]:
    _sig(_n, _a)


def check(rc):
    if rc != 0:
        raise EsError(f"status {rc}: {lib.es_last_error().decode()}")


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    if isinstance(a, (Stream, Event)):
        return a.ptr
    return a


_initialised = False


def init(device=0):
    global _initialised
    check(lib.es_init(device))
    _initialised = True


def device_count():
    n = _i(0)
    check(lib.es_device_count(C.byref(n)))
    return n.value


def synchronize():
    check(lib.es_device_synchronize())


class DeviceArray:
    """A block of device memory with a numpy dtype and shape."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if hasattr(shape, "__len__") else (shape,)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        check(lib.es_malloc(C.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a, stream)
        return d

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib.es_memcpy_htod(self.ptr, a.ctypes.data, self.nbytes, _ptr(stream)))

    def host(self, stream=None):
        out = np.empty(self.shape, self.dtype)
        check(lib.es_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes, _ptr(stream)))
        return out

    def zero(self):
        check(lib.es_memset(self.ptr, 0, self.nbytes, None))

    def free(self):
        if self.ptr:
            lib.es_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = _vp()
        check(lib.es_stream_create(C.byref(p)))
        self.ptr = p.value

    def synchronize(self):
        check(lib.es_stream_synchronize(self.ptr))

    def __del__(self):
        try:
            lib.es_stream_destroy(self.ptr)
        except Exception:
            pass


class Event:
    def __init__(self):
        p = _vp()
        check(lib.es_event_create(C.byref(p)))
        self.ptr = p.value

    def record(self, stream):
        check(lib.es_event_record(self.ptr, _ptr(stream)))

    def synchronize(self):
        check(lib.es_event_synchronize(self.ptr))

    def elapsed_ms(self, end):
        ms = C.c_float()
        check(lib.es_event_elapsed_ms(self.ptr, end.ptr, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            lib.es_event_destroy(self.ptr)
        except Exception:
            pass


def cambi_params(w, h, depth=10, window_base=65):
    info = CambiInfo()
    check(lib.es_cambi_params(w, h, depth, window_base, C.byref(info)))
    return info


class Cambi:
    """The CAMBI object of src/cambi.hpp: launch() then finish() per picture."""

    def __init__(self, w, h, depth=10, window_base=65, keep_cvalues=False, cv_mfma=True, cv_bmax=32):
        p = _vp()
        check(lib.es_cambi_create_ex(w, h, depth, window_base, int(keep_cvalues), C.byref(p)))
        self.ptr = p.value
        self.info = cambi_params(w, h, depth, window_base)
        if not cv_mfma or cv_bmax != 32:
            check(lib.es_cambi_set_method(self.ptr, int(cv_mfma), cv_bmax))

    def launch(self, img, pitch, stream=None, bytes=2, shift=0):
        check(lib.es_cambi_launch_ex(self.ptr, _ptr(img), pitch, bytes, shift, _ptr(stream)))

    def finish(self):
        s = C.c_double()
        check(lib.es_cambi_finish(self.ptr, C.byref(s)))
        return s.value

    def scales(self):
        out = (CambiScale * 5)()
        check(lib.es_cambi_scales(self.ptr, out))
        return [{n: getattr(o, n) for n, _ in CambiScale._fields_} for o in out]

    def plane(self, s):
        out = np.empty((self.info.sh[s], self.info.sw[s]), np.uint16)
        check(lib.es_cambi_plane(self.ptr, s, out.ctypes.data))
        return out

    def cvalues(self, s):
        out = np.empty((self.info.sh[s], self.info.sw[s]), np.float32)
        check(lib.es_cambi_cvalues(self.ptr, s, out.ctypes.data))
        return out

    def __del__(self):
        try:
            lib.es_cambi_destroy(self.ptr)
        except Exception:
            pass


def window_stats(yr, yd, up, w, h, win, mid_r, mid_d, sums, fp64=False, stream=None):
    check(lib.es_window_stats(_ptr(yr), _ptr(yd), up, w, h, win, _ptr(mid_r), _ptr(mid_d), _ptr(sums), int(fp64),
                              _ptr(stream)))


def window_values(sums):
    sums = np.ascontiguousarray(sums, np.int64)
    n = sums.size // 12
    out = np.empty((n, 8), np.float64)
    check(lib.es_window_values(sums.ctypes.data, n, out.ctypes.data))
    return out


def window_stats_f32(yr, yd, up, w, h, mid_r, mid_d, out, stream=None):
    """The reference's window_stats outputs: float mid planes and 8 floats per 16 x 16 window."""
    check(lib.es_window_stats_f32(_ptr(yr), _ptr(yd), up, w, h, _ptr(mid_r), _ptr(mid_d), _ptr(out), _ptr(stream)))


def persist_stats_f32(ring, plane, centre, mid_d, w, h, out, stream=None):
    """The reference's persist_stats on 13 float mid planes: 5 floats per 16 x 16 window."""
    check(lib.es_persist_stats_f32(_ptr(ring), plane, centre, _ptr(mid_d), w, h, _ptr(out), _ptr(stream)))


def persist_stats(ring, plane, centre, mid_d, w, h, win, sums, stream=None):
    check(lib.es_persist_stats(_ptr(ring), plane, centre, _ptr(mid_d), w, h, win, _ptr(sums), _ptr(stream)))


def select_rank(keys, n, rank, stream=None):
    if keys.dtype == np.uint32:
        out = C.c_uint32()
        check(lib.es_select_rank_u32(keys.ptr, n, rank, C.byref(out), _ptr(stream)))
    else:
        out = C.c_uint64()
        check(lib.es_select_rank_u64(keys.ptr, n, rank, C.byref(out), _ptr(stream)))
    return out.value


def hist_quantiles_noalloc(hists, nhist, qs, stream=None):
    """hist_quantiles through es_hist_quantiles_noalloc (results staged in the library's static device array)."""
    qs = np.ascontiguousarray(qs, np.float64)
    out = np.empty((nhist, qs.size), np.float64)
    counts = np.empty(nhist, np.uint64)
    check(lib.es_hist_quantiles_noalloc(_ptr(hists), nhist, qs.ctypes.data, qs.size, out.ctypes.data,
                                        counts.ctypes.data, _ptr(stream)))
    return out, counts


def hist_quantiles(hists, nhist, qs):
    qs = np.ascontiguousarray(qs, np.float64)
    out = np.empty((nhist, qs.size), np.float64)
    counts = np.empty(nhist, np.uint64)
    check(lib.es_hist_quantiles(_ptr(hists), nhist, qs.ctypes.data, qs.size, out.ctypes.data, counts.ctypes.data))
    return out, counts


class Scorer:
    """The measure path of one source against its encodes (src/main.cpp's queue/collect)."""

    def __init__(self, mode, src_w, src_h, encodes, depth=10, win=16, per_segment=1440, window_base=65, batch=16,
                 tile_h=0, use_graph=False, src_pitch=None, src_bytes=2, src_shift=0, profile=False, cv_bmax=0,
                 cv_waves=0, serial_stages=False, cv_tile_blocks=False, early_energies=False, cv_mfma=False,
                 cv_queues=0, tile_h0=0):
        self._encodes = (EncodeDesc * len(encodes))(*encodes)
        cfg = ScorerConfig(mode, src_w, src_h, src_pitch or src_w, src_bytes, src_shift, depth, win, per_segment,
                           window_base, batch, tile_h, int(use_graph), len(encodes), self._encodes, int(profile),
                           cv_bmax, cv_waves, int(serial_stages), int(cv_tile_blocks), int(early_energies),
                           int(cv_mfma), int(cv_queues), tile_h0)
        p = _vp()
        check(lib.es_scorer_create(C.byref(cfg), C.byref(p)))
        self.ptr = p.value
        self.n_encodes = len(encodes)

    def submit_frames(self, src_frames, enc_frames, stream=None, pts=None):
        """src_frames: n Frame structures; enc_frames: n rows of n_encodes Frame structures; pts: the source frames'
        timestamps for the source-CAMBI cache (use_cache), or None."""
        n = len(src_frames)
        src = (Frame * n)(*src_frames)
        flat = [f for row in enc_frames for f in row]
        enc = (Frame * max(1, len(flat)))(*flat)
        if pts is None:
            check(lib.es_scorer_submit_frames(self.ptr, n, src, enc, _ptr(stream)))
        else:
            p = np.ascontiguousarray(pts, np.int64)
            check(lib.es_scorer_submit_pts(self.ptr, n, src, enc, p.ctypes.data, _ptr(stream)))

    def use_cache(self, source_path, directory=None):
        check(lib.es_scorer_use_cache(self.ptr, str(source_path).encode(),
                                      None if directory is None else str(directory).encode()))

    def cache_hits(self):
        n = C.c_long()
        check(lib.es_scorer_cache_hits(self.ptr, C.byref(n)))
        return n.value

    def cache_file(self, crop=0):
        f = C.c_char_p()
        check(lib.es_scorer_cache_file(self.ptr, crop, C.byref(f)))
        return f.value.decode()

    def psnr(self, e, n):
        out = np.empty((n, 3))
        check(lib.es_scorer_psnr(self.ptr, e, n, out.ctypes.data))
        return out

    def set_vmaf(self, e, vmaf):
        v = np.ascontiguousarray(vmaf, np.float64)
        check(lib.es_scorer_set_vmaf(self.ptr, e, v.size, v.ctypes.data))

    def _info(self, info):
        paths = [p.encode() for p in info["paths"]]
        self._paths = (C.c_char_p * max(1, len(paths)))(*paths)
        return RunInfo(info["source"].encode(), self._paths, info.get("model", "vmaf_v0.6.1").encode(),
                       info.get("start", 0.0), info.get("offset", 0.0), info["fps"], info["segment_seconds"],
                       info["width"], info["height"])

    def write_json(self, info, path):
        """info: dict of source, paths, model, start, offset, fps, segment_seconds, width, height."""
        ri = self._info(info)
        check(lib.es_scorer_write_json(self.ptr, C.byref(ri), str(path).encode()))

    def submit(self, src_ptrs, enc_ptrs, stream=None):
        """src_ptrs: device pointers of n source frames; enc_ptrs: n x n_encodes device pointers."""
        n = len(src_ptrs)
        src = (_vp * n)(*src_ptrs)
        flat = [p for row in enc_ptrs for p in row]
        enc = (_vp * len(flat))(*flat)
        check(lib.es_scorer_submit(self.ptr, n, src, enc, _ptr(stream)))

    def sync(self):
        check(lib.es_scorer_sync(self.ptr))

    def frames(self):
        f = C.c_long()
        check(lib.es_scorer_frames(self.ptr, C.byref(f)))
        return f.value

    def segments(self, e):
        n = _i()
        check(lib.es_scorer_segments(self.ptr, e, C.byref(n)))
        return n.value

    def segment(self, e, seg=-1, with_hist=False):
        s = Segment()
        hist = np.zeros((6, HIST_BINS), np.uint64) if with_hist else None
        check(lib.es_scorer_segment(self.ptr, e, seg, C.byref(s), _ptr(hist)))
        return (s.as_dict(), hist) if with_hist else s.as_dict()

    def cambi(self, e, n):
        a, b = np.empty(n), np.empty(n)
        check(lib.es_scorer_cambi(self.ptr, e, n, a.ctypes.data, b.ctypes.data))
        return a, b

    STAGES = ("upload", "cambi_base", "cambi_scale", "cambi_cvalues", "cambi_topk", "cambi_finish", "window_stats",
              "frame_sums", "persist", "collect", "results")

    def stage_ms(self):
        ms = np.zeros(len(self.STAGES))
        rounds = C.c_long()
        check(lib.es_scorer_stage_ms(self.ptr, ms.ctypes.data, C.byref(rounds)))
        return dict(zip(self.STAGES, ms)), rounds.value

    def thresholds(self, e, n):
        q = np.empty(n)
        check(lib.es_scorer_thresholds(self.ptr, e, n, q.ctypes.data))
        return q

    def __del__(self):
        try:
            lib.es_scorer_destroy(self.ptr)
        except Exception:
            pass


def frame_view(fmt, y, u, v, ypitch, cpitch):
    """es_frame_view: the Frame of device planes y, u (and v when planar) of `fmt` (a name of FORMATS or its code)."""
    f = Frame()
    check(lib.es_frame_view(FORMATS.get(fmt, fmt), _ptr(y), _ptr(u), _ptr(v), ypitch, cpitch, C.byref(f)))
    return f


def frame_planes(frame, x0, y0, pd, y16, u, v, w, h, stream=None):
    check(lib.es_frame_planes(C.byref(frame), x0, y0, pd, _ptr(y16), _ptr(u), _ptr(v), w, h, _ptr(stream)))


def plane_sse(a, b, n, sse, floats=False, stream=None):
    """Adds the squared error of n samples of device planes a and b (uint16, or float32 with floats) to the device
    uint64 *sse."""
    if floats:
        check(lib.es_plane_sse(None, None, _ptr(a), _ptr(b), n, _ptr(sse), _ptr(stream)))
    else:
        check(lib.es_plane_sse(_ptr(a), _ptr(b), None, None, n, _ptr(sse), _ptr(stream)))


def color_sums(u, v, n, up, sums, stream=None):
    check(lib.es_color_sums(_ptr(u), _ptr(v), n, up, _ptr(sums), _ptr(stream)))


def color_values(sums, n):
    s = np.ascontiguousarray(sums, np.uint64)
    out = np.empty(3)
    check(lib.es_color_values(s.ctypes.data, n, out.ctypes.data))
    return out


def psnr(sse, n, depth):
    out = C.c_double()
    check(lib.es_psnr(int(sse), int(n), depth, C.byref(out)))
    return out.value


NOPTS = -(1 << 63)


class CambiCache:
    """The source-CAMBI cache (CambiCache, src/cache.hpp) over es_cambi_cache_*."""

    def __init__(self, path, x0, y0, w, h, depth, directory=None):
        p = _vp()
        check(lib.es_cambi_cache_open(str(path).encode(), x0, y0, w, h, depth,
                                      None if directory is None else str(directory).encode(), C.byref(p)))
        self.ptr = p.value

    def find(self, pts):
        v, found = C.c_double(), _i()
        check(lib.es_cambi_cache_find(self.ptr, pts, C.byref(v), C.byref(found)))
        return v.value if found.value else None

    def add(self, pts, score):
        check(lib.es_cambi_cache_add(self.ptr, pts, score))

    def file(self):
        return lib.es_cambi_cache_file(self.ptr).decode()

    def close(self):
        if self.ptr:
            check(lib.es_cambi_cache_close(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def copy_kernel(dst, src, nbytes, stream=None):
    check(lib.es_copy_kernel(_ptr(dst), _ptr(src), nbytes, _ptr(stream)))


def synth_frame(desc, frame, dev, bytes=2, stream=None):
    check(lib.es_synth_frame(C.byref(desc), frame, _ptr(dev), bytes, _ptr(stream)))


def synth_frame_host(desc, frame, bytes=2, threads=8):
    out = np.empty((desc.h, desc.w), np.uint8 if bytes == 1 else np.uint16)
    check(lib.es_synth_frame_host(C.byref(desc), frame, out.ctypes.data, bytes, threads))
    return out


def autotune(src_w, src_h, encodes, src_ptrs, enc_ptrs, depth=10, win=16, src_bytes=2, src_shift=0, cache=None,
             serial_stages=False):
    """The library's autotuner (es_scorer_autotune) on the given device frames (len(src_ptrs) steps; enc_ptrs[i] holds
    step i's encode planes): a dict of batch, tile_h, tile_h0, cv_bmax, cv_waves, the c-value method, queues and grid,
    when the energies start, ms_per_step and trials.  With `cache` (a JSON file path) a result is stored under the
    workload's shape and reused."""
    import json
    key = json.dumps({"src": [src_w, src_h, src_bytes, src_shift], "enc": [[e.w, e.h, e.bytes, e.shift] for e in encodes],
                      "depth": depth, "win": win, "serial": bool(serial_stages), "device": _device_key()}, sort_keys=True)
    store = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            store = json.load(f)
        if key in store:
            return dict(store[key], cached=True)
    descs = (EncodeDesc * len(encodes))(*encodes)
    cfg = ScorerConfig(ES_BATCHED, src_w, src_h, src_w, src_bytes, src_shift, depth, win, 1440, 65, 16, 0, 1,
                       len(encodes), descs, 0, 0, 0, int(serial_stages))
    n = len(src_ptrs)
    src = (_vp * n)(*src_ptrs)
    flat = [p for row in enc_ptrs for p in row]
    enc = (_vp * len(flat))(*flat)
    r = TuneResult()
    check(lib.es_scorer_autotune(C.byref(cfg), n, src, enc, C.byref(r)))
    out = {"batch": r.batch, "tile_h": r.tile_h, "tile_h0": r.tile_h0, "cv_bmax": r.cv_bmax, "cv_waves": r.cv_waves,
           "cv_tile_blocks": bool(r.cv_tile_blocks), "early_energies": bool(r.early_energies),
           "cv_mfma": bool(r.cv_mfma), "cv_queues": r.cv_queues,
           "ms_per_step": r.ms_per_step, "trials": r.trials}
    if cache:
        store[key] = out
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        with open(cache, "w") as f:
            json.dump(store, f, indent=1, sort_keys=True)
    return dict(out, cached=False)


def _device_key():
    n = C.c_int()
    check(lib.es_device_count(C.byref(n)))
    return f"gfx950 x{n.value}"
