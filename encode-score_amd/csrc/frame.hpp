// Frame ingest and the PSNR and colour reductions: descriptors and launch wrappers (frame_kernels.hip).
#pragma once
#include "encscore.h"
#include "es_internal.hpp"

namespace es {

// One chroma plane as the frame-sums kernel reads it: sample(cx, cy) = (plane[cy * pitch + cx * step] >> rshift)
// << shift, with step 2 for a semi-planar frame's interleaved u and v (v's pointer one sample after u's).
struct ChromaIn {
    const void *ptr;
    int pitch, step, bytes, rshift, shift;
};

// The figures of one (source, encode) frame pair beside the measures (src/main.cpp:543-556): both frames cropped to
// w x h (w, h even) at the pipeline depth.  out[0..2] += the luma, u and v squared errors; out[3..5] += the
// encode's colour sums (10-bit u, v and 2^23 times the strength, see strength_fixed); with src_color, out[6..8] +=
// the source's.  The luma area [0, lw) x [0, lh) is left out of out[0]: the window-stats kernel adds its squared
// error there (lw = lh = 0: none; lh a multiple of kFrameSumRows).  The sums are integers, so the order the blocks
// and kernels add them in does not change them.
struct FrameSumJob {
    PlaneIn ry, dy;
    ChromaIn ru, rv, du, dv;
    int w, h, up, src_color;
    int lw, lh;
    unsigned long long *out;
};
constexpr int kFrameSums = 9;
constexpr int kFrameSumRows = 16;  // luma rows per block (8 chroma rows)

void launch_frame_planes(const es_frame &f, int x0, int y0, int lshift, uint16_t *y16, uint16_t *u, uint16_t *v, int w,
                         int h, hipStream_t st);
void launch_plane_sse(const uint16_t *a, const uint16_t *b, const float *fa, const float *fb, size_t n,
                      unsigned long long *sse, hipStream_t st);
void launch_color_sums(const uint16_t *u, const uint16_t *v, size_t n, int up, unsigned long long *sums, hipStream_t st);
void launch_color_to_double(const unsigned long long *exact, double *sums, hipStream_t st);
// jobs: device array; max_h: the tallest job's luma height
void launch_frame_sums(const FrameSumJob *jobs, int njobs, int max_h, hipStream_t st);

// The luma plane of `f` cropped at (x0, y0) for a pipeline of depth pd, read in place.
inline PlaneIn frame_luma(const es_frame &f, int x0, int y0, int pd)
{
    const char *p = (const char *)f.y + (size_t)y0 * f.ypitch + (size_t)x0 * f.bytes;
    PlaneIn in{};
    in.ptr = p;
    in.pitch = f.ypitch / f.bytes;
    in.bytes = f.bytes;
    in.shift = pd - f.depth;
    in.rshift = f.shift;
    return in;
}

// The u (k = 0) or v (k = 1) plane of `f` cropped at chroma (x0 / 2, y0 / 2) for a pipeline of depth pd.
inline ChromaIn frame_chroma(const es_frame &f, int k, int x0, int y0, int pd)
{
    ChromaIn c{};
    const int cx = x0 >> 1, cy = y0 >> 1;
    c.pitch = f.cpitch / f.bytes;
    c.bytes = f.bytes;
    c.rshift = f.shift;
    c.shift = pd - f.depth;
    if (f.semi) {
        c.step = 2;
        c.ptr = (const char *)f.u + (size_t)cy * f.cpitch + (size_t)(2 * cx + k) * f.bytes;
    } else {
        c.step = 1;
        c.ptr = (const char *)(k ? f.v : f.u) + (size_t)cy * f.cpitch + (size_t)cx * f.bytes;
    }
    return c;
}

// Checks a frame description against the pipeline depth; with_chroma also checks its chroma planes.
void check_frame(const es_frame &f, int pd, bool with_chroma);

}  // namespace es
