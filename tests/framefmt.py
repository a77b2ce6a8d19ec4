"""Decoded frames for the ingest tests: numpy sample values packed into the storage layout of each 4:2:0 format
(planar or semi-planar, 1 or 2 bytes per sample, P010's high-bit samples), with row padding."""
import numpy as np

# (semi-planar, bytes per sample, right shift, depth), as view_of's table (src/main.cpp:135-147)
FORMATS = {"yuv420p": (0, 1, 0, 8), "yuv420p10": (0, 2, 0, 10), "nv12": (1, 1, 0, 8), "p010": (1, 2, 6, 10)}


def values(rng, fmt, w, h):
    """Random luma (h x w) and chroma ((h / 2) x (w / 2)) sample values of the format's depth."""
    depth = FORMATS[fmt][3]
    top = 1 << depth
    return (rng.integers(0, top, (h, w)), rng.integers(0, top, (h // 2, w // 2)),
            rng.integers(0, top, (h // 2, w // 2)))


def pack(fmt, y, u, v, pad=0, junk=None):
    """The stored planes (y, u, v or y, uv, None) of sample values y, u, v: rows padded by `pad` samples, P010's
    unused low bits filled from `junk` (an rng) when given."""
    semi, nbytes, shift, _ = FORMATS[fmt]
    dt = np.uint8 if nbytes == 1 else np.uint16

    def store(a, cols):
        out = np.zeros((a.shape[0], cols), dt)
        s = a.astype(np.uint32) << shift
        if junk is not None and shift:
            s |= junk.integers(0, 1 << shift, a.shape).astype(np.uint32)
        out[:, :a.shape[1]] = s.astype(dt)
        return out

    h, w = y.shape
    py = store(y, w + pad)
    if semi:
        uv = np.empty((u.shape[0], 2 * u.shape[1]), np.int64)
        uv[:, 0::2] = u
        uv[:, 1::2] = v
        return py, store(uv, 2 * u.shape[1] + 2 * pad), None
    return py, store(u, u.shape[1] + pad), store(v, v.shape[1] + pad)


def expected(fmt, y, u, v, x0, y0, w, h, pd):
    """frame_planes' result by numpy slicing: the crop at (x0, y0) and the chroma crop at (x0 // 2, y0 // 2)."""
    up = pd - FORMATS[fmt][3]
    cx, cy = x0 // 2, y0 // 2
    return ((y[y0:y0 + h, x0:x0 + w] << up).astype(np.uint16),
            (u[cy:cy + h // 2, cx:cx + w // 2] << up).astype(np.uint16),
            (v[cy:cy + h // 2, cx:cx + w // 2] << up).astype(np.uint16))
