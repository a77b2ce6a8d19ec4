// es_hist_quantiles_noalloc: Histogram::quantile (src/stats.hpp:21-30) of device histograms for callers that cannot
// allocate.  The per-histogram results are staged in `scratch`, a static device array of this file (one copy per GPU,
// held until the process ends); the quantiles travel in the kernel's arguments.  No device or pinned memory is
// allocated.  Every call writes every entry it reads back, so nothing needs clearing; calls that overlap on one GPU
// race on the array.
#include "energy.hpp"
#include "host_common.hpp"

#include <mutex>

namespace {

constexpr int kMaxHists = ES_NOALLOC_MAX_HISTS, kMaxQuantiles = ES_NOALLOC_MAX_QUANTILES;
constexpr size_t kOutBytes = (size_t)kMaxHists * kMaxQuantiles * sizeof(double);
constexpr size_t kCountBytes = (size_t)kMaxHists * sizeof(unsigned long long);
static_assert(kMaxQuantiles <= es::kMaxListQuantiles, "the quantiles must fit the kernel's argument list");

// [0, kOutBytes): out[h * nq + j]; then counts[h]
static __device__ __align__(256) unsigned char scratch[kOutBytes + kCountBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
unsigned char *g_scratch[kMaxDevices] = {};  // the array's address on each device, looked up once per device

// The array's device address on the current device.  The array's name in host code is a host-side stand-in, not a
// device address, so the address is looked up (with the device current) and only this pointer is used.
unsigned char *device_scratch()
{
    int dev = 0;
    es::hip_check(hipGetDevice(&dev), "hipGetDevice");
    es::require(dev >= 0 && dev < kMaxDevices, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        es::hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "looking up the scratch array");
        es::require(p != nullptr, "the scratch array has no device address");
        g_scratch[dev] = static_cast<unsigned char *>(p);
    }
    return g_scratch[dev];
}

}  // namespace

extern "C" int es_hist_quantiles_noalloc(const uint64_t *hists, int nhist, const double *qs, int nq, double *out,
                                         uint64_t *counts, es_stream s)
{
    using namespace es;
    return guard([&] {
        ensure_init();
        require(hists && qs && out && nhist >= 0 && nq >= 0, "bad argument");
        require(nhist <= kMaxHists, "more histograms than the no-allocation form holds (ES_NOALLOC_MAX_HISTS)");
        require(nq <= kMaxQuantiles, "more quantiles than the no-allocation form holds (ES_NOALLOC_MAX_QUANTILES)");
        if (!nhist || !nq) return;
        unsigned char *base = device_scratch();
        double *d_out = reinterpret_cast<double *>(base);
        unsigned long long *d_counts = reinterpret_cast<unsigned long long *>(base + kOutBytes);
        QuantileList list{};
        for (int j = 0; j < nq; j++) list.q[j] = qs[j];
        const hipStream_t st = hs(s);
        launch_hist_quantiles_list((const unsigned long long *)hists, nhist, list, nq, d_out, d_counts, st);
        hip_check(hipMemcpyAsync(out, d_out, (size_t)nhist * nq * sizeof(double), hipMemcpyDeviceToHost, st), "quantiles");
        if (counts)
            hip_check(hipMemcpyAsync(counts, d_counts, (size_t)nhist * sizeof(uint64_t), hipMemcpyDeviceToHost, st),
                      "quantiles");
        hip_check(hipStreamSynchronize(st), "quantiles");
    });
}
