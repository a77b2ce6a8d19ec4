// Host-side helpers of the C ABI: error capture, argument checks, and the descriptor arena that stages small kernel
// descriptors into device memory in stream order.
#pragma once
#include "encscore.h"
#include "es_internal.hpp"

#include <cstring>
#include <mutex>
#include <new>
#include <vector>

namespace es {

void set_error(const std::string &m);
void ensure_init();

template <class F>
int guard(F &&f)
{
    try {
        f();
        return ES_OK;
    } catch (const Error &e) {
        set_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error("out of host memory");
        return ES_ENOMEM;
    } catch (const std::exception &e) {
        set_error(e.what());
        return ES_EINVAL;
    }
}

inline void require(bool ok, const std::string &msg)
{
    if (!ok) throw Error(ES_EINVAL, msg);
}

inline hipStream_t hs(es_stream s) { return reinterpret_cast<hipStream_t>(s); }

// Device memory owned by an object; freed on destruction.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr, o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p, bytes = o.bytes;
            o.p = nullptr, o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n)
    {
        release();
        if (n == 0) return;
        const hipError_t r = hipMalloc(&p, n);
        if (r == hipErrorOutOfMemory) throw Error(ES_ENOMEM, "device memory: " + std::to_string(n) + " bytes");
        hip_check(r, "hipMalloc");
        bytes = n;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

// Pinned host memory owned by an object.
struct HostBuf {
    void *p = nullptr;
    size_t bytes = 0;
    HostBuf() = default;
    explicit HostBuf(size_t n) { alloc(n); }
    HostBuf(const HostBuf &) = delete;
    HostBuf &operator=(const HostBuf &) = delete;
    HostBuf(HostBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr, o.bytes = 0; }
    HostBuf &operator=(HostBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p, bytes = o.bytes;
            o.p = nullptr, o.bytes = 0;
        }
        return *this;
    }
    ~HostBuf() { release(); }
    void alloc(size_t n)
    {
        release();
        if (n == 0) return;
        hip_check(hipHostMalloc(&p, n, hipHostMallocDefault), "hipHostMalloc");
        bytes = n;
        std::memset(p, 0, n);
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

// A ring of descriptor slots: each call copies its descriptor bytes to a pinned slot and queues the copy into the
// matching device slot on the caller's stream; a slot is reused only after the event recorded with its last use.
class ArgArena {
public:
    static constexpr size_t kSlot = 4096;
    static constexpr int kSlots = 256;
    void *stage(const void *data, size_t bytes, hipStream_t st);
    void done(hipStream_t st);  // records the slot's event after the launch that reads it
private:
    std::mutex m_;
    HostBuf host_;
    DevBuf dev_;
    hipEvent_t ev_[kSlots] = {};
    bool used_[kSlots] = {};
    int next_ = 0, cur_ = -1;
};
ArgArena &arena();

// Host CAMBI parameters of a picture size (Cambi::Cambi, src/cambi.cpp:39-58).
struct CambiInfo {
    int w, h, depth, window, pad, mask_index, vlt, band_base, band_size;
    int tvi[kDiffs], weight[kDiffs];
    int sw[kScales], sh[kScales];
    long long n[kScales], k[kScales], off[kScales + 1];
};
CambiInfo cambi_info(int w, int h, int depth, int window_base);

}  // namespace es
