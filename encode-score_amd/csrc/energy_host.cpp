// es_window_stats / es_persist_stats (the launches of src/main.cpp:539-542, :558-566), es_window_values, and the
// select entry points.
#include "energy.hpp"
#include "host_common.hpp"

#include <cmath>

using namespace es;

namespace {

template <class T>
T *thread_scratch()
{
    static thread_local DevBuf *buf = nullptr;  // one small buffer per calling thread, kept until the process ends
    if (!buf) buf = new DevBuf(64);
    return buf->as<T>();
}

}  // namespace

extern "C" {

int es_window_stats(const uint16_t *yr, const uint16_t *yd, int up, int w, int h, int win, int32_t *mid_r,
                    int32_t *mid_d, int64_t *sums, int fp64, es_stream s)
{
    return guard([&] {
        ensure_init();
        require(yr && yd && sums, "null pointer");
        require(win == 16 || win == 32 || win == 64, "window size must be 16, 32 or 64");
        require(w >= win && h >= win, "picture smaller than one window");
        require(up >= 0 && up <= 2, "up must be 0 to 2");
        WsPair p{};
        p.r = PlaneIn{yr, w, 2, 0, 0};
        p.d = PlaneIn{yd, w, 2, 0, 0};
        p.w = w;
        p.h = h;
        p.up = up;
        p.mid_r = mid_r;
        p.mid_d = mid_d;
        p.sums = (long long *)sums;
        const auto *d = (const WsPair *)arena().stage(&p, sizeof p, hs(s));
        launch_window_stats(d, 1, w, h, win, fp64 != 0, hs(s));
        arena().done(hs(s));
    });
}

int es_window_stats_f32(const uint16_t *yr, const uint16_t *yd, int up, int w, int h, float *mid_r, float *mid_d,
                        float *out, es_stream s)
{
    return guard([&] {
        ensure_init();
        require(yr && yd && out, "null pointer");
        require(w >= 16 && h >= 16, "picture smaller than one window");
        require(up >= 0 && up <= 2, "up must be 0 to 2");
        WsPair p{};
        p.r = PlaneIn{yr, w, 2, 0, 0};
        p.d = PlaneIn{yd, w, 2, 0, 0};
        p.w = w;
        p.h = h;
        p.up = up;
        p.midf_r = mid_r;
        p.midf_d = mid_d;
        p.vals = out;
        const auto *d = (const WsPair *)arena().stage(&p, sizeof p, hs(s));
        launch_window_stats(d, 1, w, h, 16, true, hs(s));
        arena().done(hs(s));
    });
}

int es_persist_stats_f32(const float *ring, size_t plane, int centre, const float *mid_d, int w, int h, float *out,
                         es_stream s)
{
    return guard([&] {
        ensure_init();
        require(ring && mid_d && out, "null pointer");
        require(centre >= 0 && centre < kRing, "centre slot must be 0 to 12");
        require(plane >= (size_t)w * h && w >= 16 && h >= 16, "bad plane size");
        PersistF32Job j{};
        for (int o = -kRadius; o <= kRadius; o++) j.ring[o + kRadius] = ring + (size_t)((centre + o + kRing) % kRing) * plane;
        j.mid_d = mid_d;
        j.w = w;
        j.h = h;
        j.out = out;
        const auto *d = (const PersistF32Job *)arena().stage(&j, sizeof j, hs(s));
        launch_persist_f32(d, 1, w, h, hs(s));
        arena().done(hs(s));
    });
}

int es_window_values(const int64_t *sums, int windows, double *v)
{
    return guard([&] {
        require(sums && v && windows >= 0, "bad argument");
        for (int i = 0; i < windows; i++) {
            const int64_t *z = sums + (size_t)i * 12;
            double *o = v + (size_t)i * 8;
            o[0] = (double)z[0] / 81.0;
            o[1] = (double)z[1] / 6561.0;
            o[2] = (double)z[2] / 81.0;
            o[3] = (double)z[3] / 6561.0;
            o[4] = (double)z[4] / 81.0;
            o[5] = (double)z[5] / 6561.0;
            for (int p = 0; p < 2; p++) {
                const double d = (double)(z[6 + 3 * p] - z[7 + 3 * p]), x = (double)z[8 + 3 * p];
                o[6 + p] = std::sqrt(d * d + 4.0 * (x * x)) / 81.0;
            }
        }
    });
}

int es_persist_stats(const int32_t *ring, size_t plane, int centre, const int32_t *mid_d, int w, int h, int win,
                     int64_t *sums, es_stream s)
{
    return guard([&] {
        ensure_init();
        require(ring && mid_d && sums, "null pointer");
        require(win == 16 || win == 32 || win == 64, "window size must be 16, 32 or 64");
        require(centre >= 0 && centre < kRing, "centre slot must be 0 to 12");
        require(plane >= (size_t)w * h && w >= win && h >= win, "bad plane size");
        PersistJob j{};
        for (int o = -kRadius; o <= kRadius; o++) j.ring[o + kRadius] = ring + (size_t)((centre + o + kRing) % kRing) * plane;
        j.mid_d = mid_d;
        j.w = w;
        j.h = h;
        j.out = (long long *)sums;
        const auto *d = (const PersistJob *)arena().stage(&j, sizeof j, hs(s));
        launch_persist(d, 1, w, h, win, 0, hs(s));
        arena().done(hs(s));
    });
}

int es_copy_kernel(void *dst, const void *src, size_t bytes, es_stream s)
{
    return guard([&] {
        ensure_init();
        require(dst && src && bytes % 16 == 0, "bad argument");
        launch_copy(dst, src, bytes, hs(s));
    });
}

int es_select_rank_u32(const uint32_t *keys, size_t n, size_t rank, uint32_t *out, es_stream s)
{
    return guard([&] {
        ensure_init();
        require(keys && out && rank < n, "bad argument");
        unsigned *d = thread_scratch<unsigned>();
        launch_select_u32(keys, (long long)n, (long long)rank, d, hs(s));
        hip_check(hipMemcpyAsync(out, d, 4, hipMemcpyDeviceToHost, hs(s)), "reading the select");
        hip_check(hipStreamSynchronize(hs(s)), "select");
    });
}

int es_select_rank_u64(const uint64_t *keys, size_t n, size_t rank, uint64_t *out, es_stream s)
{
    return guard([&] {
        ensure_init();
        require(keys && out && rank < n, "bad argument");
        unsigned long long *d = thread_scratch<unsigned long long>();
        launch_select_u64((const unsigned long long *)keys, (long long)n, (long long)rank, d, hs(s));
        hip_check(hipMemcpyAsync(out, d, 8, hipMemcpyDeviceToHost, hs(s)), "reading the select");
        hip_check(hipStreamSynchronize(hs(s)), "select");
    });
}

int es_hist_quantiles(const uint64_t *hists, int nhist, const double *qs, int nq, double *out, uint64_t *counts)
{
    return guard([&] {
        ensure_init();
        require(hists && qs && out && nhist >= 0 && nq >= 0, "bad argument");
        if (!nhist || !nq) return;
        DevBuf dq((size_t)nq * 8), dout((size_t)nhist * nq * 8), dc((size_t)nhist * 8);
        hip_check(hipMemcpy(dq.p, qs, (size_t)nq * 8, hipMemcpyHostToDevice), "quantiles");
        launch_hist_quantiles((const unsigned long long *)hists, nhist, dq.as<double>(), nq, dout.as<double>(),
                              dc.as<unsigned long long>(), nullptr);
        hip_check(hipMemcpy(out, dout.p, dout.bytes, hipMemcpyDeviceToHost), "quantiles");
        if (counts) hip_check(hipMemcpy(counts, dc.p, dc.bytes, hipMemcpyDeviceToHost), "quantiles");
    });
}

}  // extern "C"
