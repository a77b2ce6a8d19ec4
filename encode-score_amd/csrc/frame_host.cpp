// es_frame_view / es_frame_planes (src/main.cpp:128-153, :496-505; src/kernels.cu:297-310), es_plane_sse /
// es_color_sums (src/kernels.cu:453-484, the launches of src/main.cpp:525, :543-554), es_color_values and es_psnr
// (src/main.cpp:273-277, :594, :603-604).
#include "frame.hpp"
#include "host_common.hpp"

#include <algorithm>
#include <cmath>

namespace es {

void check_frame(const es_frame &f, int pd, bool with_chroma)
{
    require(f.y != nullptr, "null luma plane");
    require(f.bytes == 1 || f.bytes == 2, "samples must be 1 or 2 bytes");
    require(f.depth >= 1 && f.depth <= pd, "the frame's depth must be at most the pipeline's");
    require(pd - f.depth <= 8, "the frame's depth is too far below the pipeline's");
    require(f.shift >= 0 && f.shift <= 8 && (f.bytes == 2 || f.shift == 0), "bad sample shift");
    require(f.ypitch > 0 && f.ypitch % f.bytes == 0, "luma pitch must be a positive multiple of the sample size");
    if (!with_chroma) return;
    require(f.u != nullptr && (f.semi || f.v != nullptr), "null chroma plane");
    require(f.cpitch > 0 && f.cpitch % f.bytes == 0, "chroma pitch must be a positive multiple of the sample size");
}

}  // namespace es

using namespace es;

extern "C" {

int es_frame_view(int format, const void *y, const void *u, const void *v, int ypitch, int cpitch, es_frame *out)
{
    return guard([&] {
        require(out != nullptr, "null argument");
        es_frame f{};
        f.y = y;
        f.u = u;
        f.ypitch = ypitch;
        f.cpitch = cpitch;
        switch (format) {
        case ES_FMT_YUV420P: f.semi = 0, f.bytes = 1, f.shift = 0, f.depth = 8; break;
        case ES_FMT_YUV420P10: f.semi = 0, f.bytes = 2, f.shift = 0, f.depth = 10; break;
        case ES_FMT_NV12: f.semi = 1, f.bytes = 1, f.shift = 0, f.depth = 8; break;
        case ES_FMT_P010: f.semi = 1, f.bytes = 2, f.shift = 6, f.depth = 10; break;
        default: throw Error(ES_EINVAL, "unsupported pixel format (4:2:0 at 8 or 10 bits only)");
        }
        f.v = f.semi ? nullptr : v;
        *out = f;
    });
}

int es_frame_planes(const es_frame *f, int x0, int y0, int pd, uint16_t *y16, uint16_t *u, uint16_t *v, int w, int h,
                    es_stream s)
{
    return guard([&] {
        require(f && y16 && u && v, "null pointer");
        require(pd == 8 || pd == 10, "pipeline depth must be 8 or 10");
        check_frame(*f, pd, true);
        require(w > 0 && h > 0 && w % 2 == 0 && h % 2 == 0, "the crop must have a positive even width and height");
        require(x0 >= 0 && y0 >= 0, "negative crop offset");
        require(f->ypitch / f->bytes >= x0 + w, "luma pitch below the crop");
        require(((uintptr_t)y16 & 3) == 0, "y16 must be 4-byte aligned");
        ensure_init();
        launch_frame_planes(*f, x0, y0, pd - f->depth, y16, u, v, w, h, hs(s));
    });
}

int es_plane_sse(const uint16_t *a, const uint16_t *b, const float *fa, const float *fb, size_t n, uint64_t *sse,
                 es_stream s)
{
    return guard([&] {
        require(sse != nullptr, "null output");
        require(a ? b != nullptr : (fa && fb), "null plane");
        ensure_init();
        launch_plane_sse(a, b, fa, fb, n, (unsigned long long *)sse, hs(s));
    });
}

int es_color_sums(const uint16_t *u, const uint16_t *v, size_t n, int up, uint64_t *sums, es_stream s)
{
    return guard([&] {
        require(u && v && sums, "null pointer");
        require(up >= 0 && up <= 2, "up must be 0 to 2");
        ensure_init();
        launch_color_sums(u, v, n, up, (unsigned long long *)sums, hs(s));
    });
}

int es_color_values(const uint64_t *sums, size_t n, double *out)
{
    return guard([&] {
        require(sums && out, "null pointer");
        const double d = (double)n;
        out[0] = (double)sums[0] / d;
        out[1] = (double)sums[1] / d;
        out[2] = std::ldexp((double)sums[2], -23) / d;
    });
}

int es_psnr(uint64_t sse, size_t n, int depth, double *out)
{
    return guard([&] {
        require(out != nullptr, "null output");
        require(depth >= 1 && depth <= 16, "bad depth");
        const double mse = (double)sse / (double)n, peak = (double)((1 << depth) - 1);
        *out = std::min(10.0 * std::log10(peak * peak / std::max(mse, 1e-16)), 6.0 * depth + 12);
    });
}

}  // extern "C"
