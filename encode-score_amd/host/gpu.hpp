// Drop-in for the reference's src/gpu.hpp (cu_check, class Kernels) over the C ABI of libencscore_hip: errors become
// std::runtime_error("what: message") as cu_check's do (src/gpu.hpp:7-13); Kernels::load sets up the device and its
// kernels (src/gpu.cpp:4-6); Kernels::get names a kernel of the measure path and Kernels::launch queues it with the
// reference kernel's argument array (src/gpu.hpp:17-20, src/gpu.cpp:8-19).  No CUDA types: streams are es_stream
// handles and kernels es_kernel handles.
#pragma once
#include "encscore.h"

#include <stdexcept>
#include <string>

inline void es_check(int status, const char *what)
{
    if (status == ES_OK) return;
    throw std::runtime_error(std::string(what) + ": " + es_last_error());
}

class Kernels {
public:
    // Makes `device` current and loads the kernels (the reference loads its embedded fatbin here).
    void load(int device = 0) { es_check(es_init(device), "loading kernels"); }
    // "frame_planes", "window_stats", "persist_stats", "plane_sse" or "color_sums" (src/main.cpp:229-234).
    es_kernel get(const char *name) const
    {
        es_kernel k = nullptr;
        es_check(es_kernel_get(name, &k), name);
        return k;
    }
    // Queues kernel `f` over a w x h picture with `args` as its argument array (one pointer per argument of the
    // reference kernel); the library chooses the launch shape.
    static void launch(es_kernel f, unsigned w, unsigned h, void **args, es_stream stream = nullptr)
    {
        es_check(es_kernel_launch(f, w, h, args, stream), "kernel launch");
    }
    // A non-blocking stream of the device (src/main.cpp:368).
    static es_stream stream()
    {
        es_stream s = nullptr;
        es_check(es_stream_create(&s), "creating a stream");
        return s;
    }
};
