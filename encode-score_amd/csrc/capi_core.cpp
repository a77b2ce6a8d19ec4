// Runtime part of the C ABI: errors, device set-up, memory, streams and events.
#include "host_common.hpp"

#include <cmath>

namespace es {

static thread_local std::string g_error;
static int g_device = -1;

void set_error(const std::string &m) { g_error = m; }

void ensure_init()
{
    if (g_device < 0) throw Error(ES_ESTATE, "es_init has not been called");
    hip_check(hipSetDevice(g_device), "hipSetDevice");
}

void *ArgArena::stage(const void *data, size_t bytes, hipStream_t st)
{
    require(bytes <= kSlot, "descriptor too large for the argument arena");
    std::lock_guard<std::mutex> lock(m_);
    if (!host_.p) {
        host_.alloc(kSlot * kSlots);
        dev_.alloc(kSlot * kSlots);
        for (auto &e : ev_) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    }
    const int s = next_;
    next_ = (next_ + 1) % kSlots;
    if (used_[s]) hip_check(hipEventSynchronize(ev_[s]), "argument slot");
    char *h = host_.as<char>() + (size_t)s * kSlot;
    char *d = dev_.as<char>() + (size_t)s * kSlot;
    std::memcpy(h, data, bytes);
    hip_check(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st), "staging a descriptor");
    cur_ = s;
    return d;
}

void ArgArena::done(hipStream_t st)
{
    std::lock_guard<std::mutex> lock(m_);
    if (cur_ < 0) return;
    hip_check(hipEventRecord(ev_[cur_], st), "hipEventRecord");
    used_[cur_] = true;
    cur_ = -1;
}

ArgArena &arena()
{
    static ArgArena *a = new ArgArena();  // lives until the process ends
    return *a;
}

// BT.1886 on the limited 10-bit range (libvmaf's luminance functions, called at src/cambi.cpp:41-53): white 300
// cd/m^2, black 0.01 cd/m^2, gamma 2.4, V = (clip(s, 64, 940) - 64) / 876.
static double luminance(int sample)
{
    const int s = sample < 64 ? 64 : sample > 940 ? 940 : sample;
    const double v = (double)(s - 64) / 876.0;
    const double lw = std::pow(300.0, 1.0 / 2.4), lb = std::pow(0.01, 1.0 / 2.4);
    const double a = std::pow(lw - lb, 2.4), b = lb / (lw - lb);
    return a * std::pow(v + b > 0 ? v + b : 0.0, 2.4);
}

// get_tvi_for_diff: the last code of the range whose `diff`-code step raises the luminance by more than 1.9%.
static int tvi_for_diff(int diff)
{
    int last = 64;
    for (int s = 64; s <= 940; s++) {
        const double l = luminance(s);
        if (!(luminance(s + diff) - l > 0.019 * l)) break;
        last = s;
    }
    return last;
}

// get_vlt_luma with threshold 0: the first code from 0 whose luminance reaches it.
static int vlt_luma(double thr)
{
    for (int s = 0; s < 1024; s++)
        if (luminance(s) >= thr) return s;
    return 1023;
}

CambiInfo cambi_info(int w, int h, int depth, int window_base)
{
    require(w >= 1 && h >= 1, "picture size must be positive");
    require(depth == 8 || depth == 10, "depth must be 8 or 10");
    require(window_base >= 1, "window base must be positive");
    CambiInfo c{};
    c.w = w;
    c.h = h;
    c.depth = depth;
    c.window = ((window_base * (w + h) / 375) >> 4) | 1;  // cambi.c's adjust_window_size (src/cambi.cpp:39)
    c.pad = c.window >> 1;
    require(c.pad <= kMaxPad, "CAMBI window above 65 (picture too large for the window base)");
    const uint32_t blocks = (uint32_t)((w >> 6) * (h >> 6));
    int lg = 0;
    for (uint32_t t = blocks ? blocks - 1 : 0; t; t >>= 1) lg++;
    c.mask_index = (49 + 3 * (lg - 11) - 1) >> 1;  // src/cambi.cpp:40
    static const int tvi[kDiffs] = {tvi_for_diff(1), tvi_for_diff(2), tvi_for_diff(3), tvi_for_diff(4)};
    static const int vlt = vlt_luma(0.0);
    for (int d = 0; d < kDiffs; d++) {
        c.tvi[d] = tvi[d] + kDiffs;
        c.weight[d] = d + 1;
    }
    c.vlt = vlt;
    c.band_base = std::max(c.vlt - 3 * kDiffs + 1, 0);
    c.band_size = c.tvi[kDiffs - 1] + 1 - c.band_base;
    require(c.band_size + 2 * kDiffs <= kMaxBand + 2 * kDiffs, "band too wide");
    c.sw[0] = w;
    c.sh[0] = h;
    for (int s = 1; s < kScales; s++) {
        c.sw[s] = (c.sw[s - 1] + 1) >> 1;
        c.sh[s] = (c.sh[s - 1] + 1) >> 1;
    }
    c.off[0] = 0;
    for (int s = 0; s < kScales; s++) {
        c.n[s] = (long long)c.sw[s] * c.sh[s];
        long long k = (long long)(int)(0.6 * (double)c.n[s]);  // topk_count (src/cambi.cpp:29)
        c.k[s] = std::min(std::max(k, 1LL), c.n[s]);
        c.off[s + 1] = c.off[s] + c.n[s];
    }
    return c;
}

}  // namespace es

using namespace es;

extern "C" {

const char *es_last_error(void) { return g_error.c_str(); }

int es_init(int device)
{
    return guard([&] {
        int n = 0;
        hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
        require(device >= 0 && device < n, "no such device");
        hip_check(hipSetDevice(device), "hipSetDevice");
        hip_check(hipFree(nullptr), "initialising the device");
        g_device = device;
    });
}

int es_device_count(int *count)
{
    return guard([&] {
        require(count != nullptr, "null argument");
        hip_check(hipGetDeviceCount(count), "hipGetDeviceCount");
    });
}

int es_device_synchronize(void)
{
    return guard([&] {
        ensure_init();
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    });
}

int es_stream_create(es_stream *out)
{
    return guard([&] {
        ensure_init();
        require(out != nullptr, "null argument");
        hipStream_t s;
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        *out = reinterpret_cast<es_stream>(s);
    });
}

int es_stream_destroy(es_stream s)
{
    return guard([&] {
        if (s) hip_check(hipStreamDestroy(hs(s)), "hipStreamDestroy");
    });
}

int es_stream_synchronize(es_stream s)
{
    return guard([&] { hip_check(hipStreamSynchronize(hs(s)), "hipStreamSynchronize"); });
}

int es_event_create(es_event *out)
{
    return guard([&] {
        ensure_init();
        require(out != nullptr, "null argument");
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        *out = reinterpret_cast<es_event>(e);
    });
}

int es_event_destroy(es_event e)
{
    return guard([&] {
        if (e) hip_check(hipEventDestroy(reinterpret_cast<hipEvent_t>(e)), "hipEventDestroy");
    });
}

int es_event_record(es_event e, es_stream s)
{
    return guard([&] { hip_check(hipEventRecord(reinterpret_cast<hipEvent_t>(e), hs(s)), "hipEventRecord"); });
}

int es_stream_wait_event(es_stream s, es_event e)
{
    return guard([&] { hip_check(hipStreamWaitEvent(hs(s), reinterpret_cast<hipEvent_t>(e), 0), "hipStreamWaitEvent"); });
}

int es_event_synchronize(es_event e)
{
    return guard([&] { hip_check(hipEventSynchronize(reinterpret_cast<hipEvent_t>(e)), "hipEventSynchronize"); });
}

int es_event_elapsed_ms(es_event start, es_event end, float *ms)
{
    return guard([&] {
        require(ms != nullptr, "null argument");
        hip_check(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(end)),
                  "hipEventElapsedTime");
    });
}

int es_malloc(void **ptr, size_t bytes)
{
    return guard([&] {
        ensure_init();
        require(ptr != nullptr, "null argument");
        const hipError_t r = hipMalloc(ptr, bytes);
        if (r == hipErrorOutOfMemory) throw Error(ES_ENOMEM, "device memory: " + std::to_string(bytes) + " bytes");
        hip_check(r, "hipMalloc");
    });
}

int es_free(void *ptr)
{
    return guard([&] {
        if (ptr) hip_check(hipFree(ptr), "hipFree");
    });
}

int es_host_alloc(void **ptr, size_t bytes)
{
    return guard([&] {
        ensure_init();
        require(ptr != nullptr, "null argument");
        hip_check(hipHostMalloc(ptr, bytes, hipHostMallocDefault), "hipHostMalloc");
    });
}

int es_host_free(void *ptr)
{
    return guard([&] {
        if (ptr) hip_check(hipHostFree(ptr), "hipHostFree");
    });
}

static int copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, es_stream s)
{
    return guard([&] {
        ensure_init();
        if (!bytes) return;
        require(dst && src, "null pointer");
        if (s) hip_check(hipMemcpyAsync(dst, src, bytes, kind, hs(s)), "hipMemcpyAsync");
        else hip_check(hipMemcpy(dst, src, bytes, kind), "hipMemcpy");
    });
}

int es_memcpy_htod(void *dst, const void *src, size_t bytes, es_stream s) { return copy(dst, src, bytes, hipMemcpyHostToDevice, s); }
int es_memcpy_dtoh(void *dst, const void *src, size_t bytes, es_stream s) { return copy(dst, src, bytes, hipMemcpyDeviceToHost, s); }
int es_memcpy_dtod(void *dst, const void *src, size_t bytes, es_stream s) { return copy(dst, src, bytes, hipMemcpyDeviceToDevice, s); }

int es_memset(void *dst, int value, size_t bytes, es_stream s)
{
    return guard([&] {
        ensure_init();
        if (!bytes) return;
        require(dst != nullptr, "null pointer");
        if (s) hip_check(hipMemsetAsync(dst, value, bytes, hs(s)), "hipMemsetAsync");
        else hip_check(hipMemset(dst, value, bytes), "hipMemset");
    });
}

int es_cambi_params(int w, int h, int depth, int window_base, es_cambi_info *out)
{
    return guard([&] {
        require(out != nullptr, "null argument");
        const CambiInfo c = cambi_info(w, h, depth, window_base);
        out->w = c.w;
        out->h = c.h;
        out->depth = c.depth;
        out->window = c.window;
        out->pad = c.pad;
        out->mask_index = c.mask_index;
        out->vlt = c.vlt;
        out->band_base = c.band_base;
        out->band_size = c.band_size;
        for (int d = 0; d < 4; d++) {
            out->tvi[d] = c.tvi[d];
            out->weight[d] = c.weight[d];
        }
        for (int s = 0; s < 5; s++) {
            out->sw[s] = c.sw[s];
            out->sh[s] = c.sh[s];
            out->n[s] = c.n[s];
            out->k[s] = c.k[s];
        }
    });
}

}  // extern "C"
