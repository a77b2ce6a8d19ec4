// Device collect: the per-window part of the reference's host collect (src/main.cpp:576-643) for the batched form, and
// the order-statistic selects.
//
// collect: one block of 256 threads per (encode, frame).  It selects the structure threshold, numpy's linear 0.8
// quantile of the source's oriented energies (src/main.cpp:279-289, :606-608), with an in-block radix select (8
// passes of 8-bit digits over the FP64 bit patterns, which order like the non-negative values); then per window the
// structure mask, the histogram entries (src/stats.hpp:16) and the exact frame totals; and, for the centre frame six
// back, the lasting-texture windows (src/main.cpp:630-641).  Frame totals of the oriented energy use the canonical
// order: blocks of 256 consecutive windows added as a halving tree, the blocks in order.
//
// hist_quantiles: Histogram::quantile (src/stats.hpp:21-30) for many histograms, the bin where the running count
// first exceeds q (n - 1): a block scan over the 20001 bins.
//
// select_u32 / select_u64: the key of a given ascending rank of an array, by the same in-block radix select.
#include "device_util.hpp"
#include "energy.hpp"

#include <algorithm>

namespace es {

// Window figures from the exact sums (src/kernels.cu:421-425), as the oracle computes them.
__device__ __forceinline__ void window_values(const long long *z, double *v)
{
    v[0] = (double)z[0] / 81.0;
    v[1] = (double)z[1] / 6561.0;
    v[2] = (double)z[2] / 81.0;
    v[3] = (double)z[3] / 6561.0;
    v[4] = (double)z[4] / 81.0;
    v[5] = (double)z[5] / 6561.0;
#pragma unroll
    for (int p = 0; p < 2; p++) {
        const double d = (double)(z[6 + 3 * p] - z[7 + 3 * p]), x = (double)z[8 + 3 * p];
        v[6 + p] = sqrt(__dadd_rn(__dmul_rn(d, d), __dmul_rn(4.0, __dmul_rn(x, x)))) / 81.0;
    }
}

__device__ __forceinline__ double oriented_r(const long long *z)
{
    const double d = (double)(z[6] - z[7]), x = (double)z[8];
    return sqrt(__dadd_rn(__dmul_rn(d, d), __dmul_rn(4.0, __dmul_rn(x, x)))) / 81.0;
}

// The 256-bin digit whose running count (ascending) first exceeds `rank`; the first wave of the block scans the bins
// as 4 per lane.  Returns the digit and leaves the rank within it in *rank_out.  Callers synchronise around it.
__device__ int find_digit(const unsigned *hist, unsigned long long rank, unsigned long long *rank_out, int *digit_out)
{
    const int tid = threadIdx.x;
    if (tid < 64) {
        const unsigned c0 = hist[4 * tid], c1 = hist[4 * tid + 1], c2 = hist[4 * tid + 2], c3 = hist[4 * tid + 3];
        const unsigned s = c0 + c1 + c2 + c3;
        const unsigned incl = wave_scan_incl(s), excl = incl - s;
        const unsigned long long m = __ballot((unsigned long long)incl > rank);
        const int lane = m ? __ffsll((long long)m) - 1 : 63;
        if (tid == lane) {
            unsigned long long cum = excl;
            int d = 4 * tid;
            const unsigned c[4] = {c0, c1, c2, c3};
            int k = 0;
            for (; k < 3 && cum + c[k] <= rank; k++) cum += c[k];
            d += k;
            *digit_out = d;
            *rank_out = rank - cum;
        }
    }
    __syncthreads();
    return *digit_out;
}

// The key of ascending rank `rank` among keys[0 .. n), all threads of the block taking part; *eq_out receives how
// many keys equal it and *below_out how many of those come before rank.
template <typename K>
__device__ K block_select(const K *keys, long long n, unsigned long long rank, unsigned long long *eq_out,
                          unsigned long long *before_out)
{
    constexpr int kBits = sizeof(K) * 8;
    __shared__ unsigned hist[256];
    __shared__ unsigned long long rank_sh;
    __shared__ int digit_sh;
    const int tid = threadIdx.x;
    K prefix = 0;
    unsigned long long r = rank;
    unsigned eq = 0;
    for (int shift = kBits - 8; shift >= 0; shift -= 8) {
        for (int i = tid; i < 256; i += blockDim.x) hist[i] = 0;
        __syncthreads();
        for (long long i = tid; i < n; i += blockDim.x) {
            const K key = keys[i];
            const bool match = shift == kBits - 8 || (key >> (shift + 8)) == (prefix >> (shift + 8));
            if (match) atomicAdd(&hist[(unsigned)(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        const int d = find_digit(hist, r, &rank_sh, &digit_sh);
        r = rank_sh;
        eq = hist[d];
        prefix |= (K)d << shift;
        __syncthreads();
    }
    *eq_out = eq;
    *before_out = r;
    return prefix;
}

template <typename K>
__device__ K block_min_above(const K *keys, long long n, K a)
{
    __shared__ K part[16];
    K m = ~(K)0;
    for (long long i = threadIdx.x; i < n; i += blockDim.x) {
        const K key = keys[i];
        if (key > a && key < m) m = key;
    }
    for (int o = 32; o; o >>= 1) {
        const K other = __shfl_xor(m, o);
        m = other < m ? other : m;
    }
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    m = part[0];
    for (int wv = 1; wv < (int)(blockDim.x >> 6); wv++) m = part[wv] < m ? part[wv] : m;
    __syncthreads();
    return m;
}

constexpr int kCollectThreads = 256;

// The lasting-texture windows of the centre frame six back (src/main.cpp:630-641), with exact [own, mine, count].
__device__ void collect_lasting(const CollectJob &J)
{
    if (!J.psrc) return;
    const int n = J.windows, tid = threadIdx.x;
    long long lz[3] = {0, 0, 0};
    for (int i = tid; i < n; i += kCollectThreads) {
        long long i0, i1, iab, iaa, ibb;
        if (J.lasting_plain) {
            const long long *p = J.psrc + (size_t)i * 5;
            i0 = p[0];
            i1 = p[1];
            iab = p[2];
            iaa = p[3];
            ibb = p[4];
        } else {
            const long long *p = J.psrc + (size_t)i * 4;
            i0 = p[0];
            iab = p[1];
            iaa = p[2];
            ibb = p[3];
            i1 = J.penc[i];
        }
        const double v0 = (double)i0 / 78732.0, v1 = (double)i1 / 78732.0;
        const double v2 = (double)iab / 236196.0, v3 = (double)iaa / 236196.0, v4 = (double)ibb / 236196.0;
        const double agree = v2 / sqrt(fmax(__dmul_rn(v3, v4), 1e-9));
        if (J.cmask[i] && agree > 0.5 && v0 > 0) {
            lz[0] += i0;
            lz[1] += i1;
            lz[2] += 1;
            atomicAdd(J.lhist + (size_t)5 * kHistBins + hist_bin(100.0 * v1 / v0), 1ull);
        }
    }
    __shared__ long long part[3][kCollectThreads / 64];
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const long long s = wave_sum(lz[k]);
        if (lane == 0) part[k][wave] = s;
    }
    __syncthreads();
    if (tid == 0)
        for (int k = 0; k < 3; k++) {
            long long s = 0;
            for (int wv = 0; wv < kCollectThreads / 64; wv++) s += part[k][wv];
            J.tot_lasting[k] = s;
        }
}

__global__ __launch_bounds__(kCollectThreads) void collect_kernel(const CollectJob *jobs, double *scratch, int max_windows,
                                                                   int phase)
{
    const CollectJob &J = jobs[blockIdx.x];
    const int n = J.windows, tid = threadIdx.x;
    if (n <= 0) return;  // an inactive job (an encode that has ended)
    if (phase == 1) {
        collect_lasting(J);
        return;
    }
    unsigned long long *keys = (unsigned long long *)scratch + (size_t)blockIdx.x * max_windows;
    for (int i = tid; i < n; i += kCollectThreads) keys[i] = (unsigned long long)__double_as_longlong(oriented_r(J.sums + (size_t)i * 12));
    __threadfence_block();
    __syncthreads();
    // numpy's linear quantile of the oriented energies at 0.8
    const double pos = 0.8 * (double)(n - 1);
    const long long lo = (long long)pos;
    unsigned long long eq, before;
    const unsigned long long ka = block_select(keys, n, (unsigned long long)lo, &eq, &before);
    double q = __longlong_as_double((long long)ka);
    if (lo + 1 < n) {
        unsigned long long kb = ka;
        if (before + 1 >= eq) kb = block_min_above(keys, n, ka);
        const double a = q, b = __longlong_as_double((long long)kb);
        q = __dadd_rn(a, __dmul_rn(b - a, pos - (double)lo));
    }
    if (tid == 0 && J.quantile) *J.quantile = q;

    __shared__ double tree[4][kCollectThreads];
    __shared__ long long part[12][kCollectThreads / 64];
    long long zi[12] = {0};
    double tot[4] = {0, 0, 0, 0};
    for (int b = 0; b < n; b += kCollectThreads) {
        const int i = b + tid;
        double o[4] = {0, 0, 0, 0};
        if (i < n) {
            const long long *z = J.sums + (size_t)i * 12;
            double v[8];
            window_values(z, v);
            const bool m = v[6] >= q && v[6] > 0;
            J.mask[i] = (uint8_t)m;
#pragma unroll
            for (int k = 0; k < 6; k++) {
                zi[k] += z[k];
                if (m) zi[6 + k] += z[k];
            }
            o[0] = v[6];
            o[1] = v[7];
            o[2] = m ? v[6] : 0.0;
            o[3] = m ? v[7] : 0.0;
            if (m) {
                const double er[3] = {v[0], v[1], v[6]}, ed[3] = {v[2], v[3], v[7]};
#pragma unroll
                for (int k = 0; k < 3; k++)
                    atomicAdd(J.hist + (size_t)k * kHistBins + hist_bin(100.0 * sqrt(ed[k] / fmax(er[k], 1e-9))), 1ull);
#pragma unroll
                for (int k = 0; k < 2; k++)
                    atomicAdd(J.hist + (size_t)(3 + k) * kHistBins + hist_bin(100.0 * v[4 + k] / fmax(er[k], 1e-9)), 1ull);
            }
        }
#pragma unroll
        for (int k = 0; k < 4; k++) tree[k][tid] = o[k];
        __syncthreads();
        for (int s = kCollectThreads / 2; s > 0; s >>= 1) {
            if (tid < s)
#pragma unroll
                for (int k = 0; k < 4; k++) tree[k][tid] += tree[k][tid + s];
            __syncthreads();
        }
        if (tid == 0)
#pragma unroll
            for (int k = 0; k < 4; k++) tot[k] += tree[k][0];
        __syncthreads();
    }
    // block sums of the exact totals
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const long long s = wave_sum(zi[k]);
        if (lane == 0) part[k][wave] = s;
    }
    __syncthreads();
    if (tid == 0) {
        for (int k = 0; k < 12; k++) {
            long long s = 0;
            for (int wv = 0; wv < kCollectThreads / 64; wv++) s += part[k][wv];
            J.tot_int[k] = s;
        }
        for (int k = 0; k < 4; k++) J.tot_orient[k] = tot[k];
    }
}

void launch_collect(const CollectJob *jobs, int njobs, double *scratch, int max_windows, int phase, hipStream_t st)
{
    if (njobs <= 0) return;
    hipLaunchKernelGGL(collect_kernel, dim3(njobs), dim3(kCollectThreads), 0, st, jobs, scratch, max_windows, phase);
    hip_check(hipGetLastError(), "collect");
}

// Histogram::quantile for nhist histograms of kHistBins bins and nq quantiles each: out[h * nq + j], NaN when empty.
__global__ __launch_bounds__(256) void hist_quantile_kernel(const unsigned long long *hists, const double *qs, int nq,
                                                            double *out, unsigned long long *counts)
{
    constexpr int kChunk = (kHistBins + 255) / 256;
    __shared__ unsigned long long csum[256];
    __shared__ unsigned long long total_sh;
    const unsigned long long *h = hists + (size_t)blockIdx.x * kHistBins;
    const int tid = threadIdx.x, b0 = tid * kChunk, b1 = min(b0 + kChunk, kHistBins);
    unsigned long long s = 0;
    for (int b = b0; b < b1; b++) s += h[b];
    csum[tid] = s;
    __syncthreads();
    if (tid == 0) {
        unsigned long long run = 0;
        for (int i = 0; i < 256; i++) {
            const unsigned long long c = csum[i];
            csum[i] = run;  // exclusive
            run += c;
        }
        total_sh = run;
    }
    __syncthreads();
    const unsigned long long n = total_sh;
    if (tid == 0 && counts) counts[blockIdx.x] = n;
    for (int j = 0; j < nq; j++) {
        if (n == 0) {
            if (tid == 0) out[(size_t)blockIdx.x * nq + j] = __longlong_as_double(0x7FF8000000000000ll);
            continue;
        }
        const double target = qs[j] * (double)(n - 1);
        unsigned long long seen = csum[tid];
        // the first bin with seen > target lies in this thread's chunk iff the chunk's end passes it and its start
        // does not
        if ((double)seen <= target && (double)(seen + s) > target) {
            for (int b = b0; b < b1; b++) {
                seen += h[b];
                if ((double)seen > target) {
                    out[(size_t)blockIdx.x * nq + j] = -500.0 + (double)b * 0.05;
                    break;
                }
            }
        }
        if (tid == 0 && (double)n <= target) out[(size_t)blockIdx.x * nq + j] = 500.0;
    }
}

void launch_hist_quantiles(const unsigned long long *hists, int nhist, const double *qs, int nq, double *out,
                           unsigned long long *counts, hipStream_t st)
{
    if (nhist <= 0) return;
    hipLaunchKernelGGL(hist_quantile_kernel, dim3(nhist), dim3(256), 0, st, hists, qs, nq, out, counts);
    hip_check(hipGetLastError(), "histogram quantiles");
}

// A plain copy of 16-byte vectors, grid-strided, for calibrating HBM byte counters against a known size.
__global__ __launch_bounds__(256) void copy_kernel(uint4 *dst, const uint4 *src, size_t n)
{
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

void launch_copy(void *dst, const void *src, size_t bytes, hipStream_t st)
{
    const size_t n = bytes / 16;
    if (!n) return;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, st, (uint4 *)dst, (const uint4 *)src, n);
    hip_check(hipGetLastError(), "copy");
}

template <typename K>
__global__ __launch_bounds__(1024) void select_kernel(const K *keys, long long n, long long rank, K *out)
{
    unsigned long long eq, before;
    const K k = block_select(keys, n, (unsigned long long)rank, &eq, &before);
    if (threadIdx.x == 0) *out = k;
}

void launch_select_u64(const unsigned long long *keys, long long n, long long rank, unsigned long long *out,
                       hipStream_t st)
{
    hipLaunchKernelGGL(select_kernel<unsigned long long>, dim3(1), dim3(1024), 0, st, keys, n, rank, out);
    hip_check(hipGetLastError(), "select");
}

void launch_select_u32(const unsigned int *keys, long long n, long long rank, unsigned int *out, hipStream_t st)
{
    hipLaunchKernelGGL(select_kernel<unsigned int>, dim3(1), dim3(1024), 0, st, keys, n, rank, out);
    hip_check(hipGetLastError(), "select");
}

}  // namespace es
