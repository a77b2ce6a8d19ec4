// es_kernel_get / es_kernel_launch: the by-name kernels of the reference's measure path (Kernels::get and the launches
// of src/main.cpp:229-234, :513-566), with each call's arguments given as the reference kernel's argument array (one
// pointer per argument, as cuLaunchKernel's kernelParams).  The launch shape is this library's own: a launch names
// the picture it covers, w x h pixels (Kernels::launch's w and h, src/gpu.cpp:15-19).
#include "energy.hpp"
#include "frame.hpp"
#include "host_common.hpp"

#include <cstring>

using namespace es;

namespace {

enum class ByName { FramePlanes, WindowStats, PersistStats, PlaneSse, ColorSums };

struct NamedKernel {
    const char *name;
    ByName which;
};

const NamedKernel kNamed[] = {
    {"frame_planes", ByName::FramePlanes},   // src/kernels.cu:297-310
    {"window_stats", ByName::WindowStats},   // src/kernels.cu:346-347
    {"persist_stats", ByName::PersistStats}, // src/kernels.cu:435
    {"plane_sse", ByName::PlaneSse},         // src/kernels.cu:454-455
    {"color_sums", ByName::ColorSums},       // src/kernels.cu:470
};

// Rethrows a failed call of this library's C ABI with its message.
void check(int status)
{
    if (status != ES_OK) throw Error(status, es_last_error());
}

template <class T>
T arg(void **args, int i)
{
    require(args[i] != nullptr, "null kernel argument");
    T v;
    std::memcpy(&v, args[i], sizeof v);
    return v;
}

}  // namespace

struct es_kernel_s {
    NamedKernel k;
};

extern "C" {

int es_kernel_get(const char *name, es_kernel *out)
{
    return guard([&] {
        require(name && out, "null argument");
        for (const NamedKernel &k : kNamed) {
            if (std::strcmp(k.name, name) == 0) {
                static es_kernel_s handles[sizeof kNamed / sizeof kNamed[0]];
                es_kernel_s &h = handles[&k - kNamed];
                h.k = k;
                *out = &h;
                return;
            }
        }
        throw Error(ES_EINVAL, std::string(name) + ": no such kernel on the measure path");
    });
}

int es_kernel_launch(es_kernel k, unsigned w, unsigned h, void **args, es_stream s)
{
    return guard([&] {
        require(k && args, "null argument");
        switch (k->k.which) {
        case ByName::FramePlanes: {
            // frame_planes(Frame f, int x0, int y0, int pd, u16 *y16, u16 *u, u16 *v, int w, int h)
            const es_frame f = arg<es_frame>(args, 0);
            check(es_frame_planes(&f, arg<int>(args, 1), arg<int>(args, 2), arg<int>(args, 3), arg<uint16_t *>(args, 4),
                                  arg<uint16_t *>(args, 5), arg<uint16_t *>(args, 6), arg<int>(args, 7),
                                  arg<int>(args, 8), s));
            break;
        }
        case ByName::WindowStats:
            // window_stats(const u16 *yr, const u16 *yd, int up, int w, int h, float *mid_r, float *mid_d, float *out)
            check(es_window_stats_f32(arg<const uint16_t *>(args, 0), arg<const uint16_t *>(args, 1), arg<int>(args, 2),
                                      arg<int>(args, 3), arg<int>(args, 4), arg<float *>(args, 5), arg<float *>(args, 6),
                                      arg<float *>(args, 7), s));
            break;
        case ByName::PersistStats:
            // persist_stats(const float *ring, size_t plane, int centre, const float *mid_d, int w, float *out); the
            // picture's height is the launch's
            check(es_persist_stats_f32(arg<const float *>(args, 0), arg<size_t>(args, 1), arg<int>(args, 2),
                                       arg<const float *>(args, 3), arg<int>(args, 4), (int)h, arg<float *>(args, 5), s));
            break;
        case ByName::PlaneSse:
            // plane_sse(const u16 *a, const u16 *b, const float *fa, const float *fb, size_t n, u64 *sse)
            check(es_plane_sse(arg<const uint16_t *>(args, 0), arg<const uint16_t *>(args, 1), arg<const float *>(args, 2),
                               arg<const float *>(args, 3), arg<size_t>(args, 4), arg<uint64_t *>(args, 5), s));
            break;
        case ByName::ColorSums: {
            // color_sums(const u16 *u, const u16 *v, size_t n, int up, double *sums): the exact sums into a cleared
            // stream-ordered slot (the zeros are staged on the stream), then added to the caller's doubles
            const unsigned long long zeros[3] = {0, 0, 0};
            const int up = arg<int>(args, 3);
            require(arg<const uint16_t *>(args, 0) && arg<const uint16_t *>(args, 1) && arg<double *>(args, 4),
                    "null pointer");
            require(up >= 0 && up <= 2, "up must be 0 to 2");
            ensure_init();
            auto *exact = (unsigned long long *)arena().stage(zeros, sizeof zeros, hs(s));
            launch_color_sums(arg<const uint16_t *>(args, 0), arg<const uint16_t *>(args, 1), arg<size_t>(args, 2),
                              up, exact, hs(s));
            launch_color_to_double(exact, arg<double *>(args, 4), hs(s));
            arena().done(hs(s));
            break;
        }
        }
        (void)w;
    });
}

}  // extern "C"
