// Drop-in for the reference's src/gpu.hpp (cu_check, class Kernels) over the C ABI of libencscore_hip: errors become
// std::runtime_error("what: message") as cu_check's do (src/gpu.hpp:7-13), and Kernels::load sets up the device
// and its kernels (src/gpu.cpp:4-6).  No CUDA types: streams are es_stream handles.
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
    // A non-blocking stream of the device (src/main.cpp:368).
    static es_stream stream()
    {
        es_stream s = nullptr;
        es_check(es_stream_create(&s), "creating a stream");
        return s;
    }
};
