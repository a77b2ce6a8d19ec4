// Device collect: the per-window part of the reference's host collect (src/main.cpp:576-643) for the batched form, and
// the order-statistic selects.
//
// collect, per (encode, frame) job: the structure threshold, numpy's linear 0.8 quantile of the source's oriented
// energies (src/main.cpp:279-289, :606-608), by a radix select in one block per job (8 passes of 8-bit digits over the
// FP64 bit patterns, which order like the non-negative values); then, one block per chunk of 256 windows, the
// structure mask, the histogram entries (src/stats.hpp:16) and the exact frame totals; and, for the centre frame six
// back, the lasting-texture windows (src/main.cpp:630-641).  Frame totals of the oriented energy use the canonical
// order: each chunk of 256 consecutive windows added as a halving tree, the chunks in order.
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

constexpr int kCollectThreads = 256;  // one chunk of 256 windows per block
constexpr uint16_t kNoBin = 0xFFFF;   // a window that adds nothing to a histogram
constexpr int kSelectThreads = 1024;

// Sums of a chunk's per-thread values over the block (exact integers), from the block's 4 waves.
template <int N, int P>
__device__ __forceinline__ void block_sum_ll(long long (&v)[N], long long (&part)[P][kCollectThreads / 64])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();  // an earlier call's readers are done with part
#pragma unroll
    for (int k = 0; k < N; k++) {
        const long long s = wave_sum(v[k]);
        if (lane == 0) part[k][wave] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = part[k][0] + part[k][1] + part[k][2] + part[k][3];
}

// Keys for the structure threshold: the bit patterns of the source's oriented energies rounded once to float (the
// reference's per-window figures are floats; non-negative floats order like their bit patterns).
__global__ __launch_bounds__(kCollectThreads) void collect_keys_kernel(const CollectJob *jobs, CollectWork w)
{
    const CollectJob &J = jobs[blockIdx.y];
    const int i = blockIdx.x * kCollectThreads + threadIdx.x;
    if (i < J.windows)
        w.keys[(size_t)blockIdx.y * w.max_windows + i] =
            (unsigned long long)__float_as_uint(__double2float_rn(oriented_r(J.sums + (size_t)i * 12)));
}

// numpy's linear quantile at 0.8 of a job's keys as the reference evaluates it (src/main.cpp:279-289): the key a of
// rank lo = floor(0.8 (n - 1)) by 4 passes of 8-bit digits, each pass counting into per-wave histograms, and the next
// larger key b for lo + 1; then b - a in float, a + (b - a)(pos - lo) in double, rounded to float.
__global__ __launch_bounds__(kSelectThreads) void collect_select_kernel(const CollectJob *jobs, CollectWork w)
{
    constexpr int kWaves = kSelectThreads / 64;
    __shared__ unsigned hist[kWaves][256];
    __shared__ unsigned long long rank_sh;
    __shared__ int digit_sh;
    const CollectJob &J = jobs[blockIdx.x];
    const long long n = J.windows;
    if (n <= 0) return;
    const unsigned long long *keys = w.keys + (size_t)blockIdx.x * w.max_windows;
    const int tid = threadIdx.x, wave = tid >> 6;
    const double pos = 0.8 * (double)(n - 1);
    const long long lo = (long long)pos;
    unsigned long long prefix = 0, r = (unsigned long long)lo;
    unsigned eq = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int i = tid & 63; i < 256; i += 64) hist[wave][i] = 0;
        __syncthreads();
#pragma unroll 4
        for (long long i = tid; i < n; i += kSelectThreads) {
            const unsigned long long key = keys[i];
            if (shift == 24 || (key >> (shift + 8)) == (prefix >> (shift + 8)))
                atomicAdd(&hist[wave][(unsigned)(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (tid < 256) {
            unsigned t = 0;
#pragma unroll
            for (int k = 0; k < kWaves; k++) t += hist[k][tid];
            hist[0][tid] = t;
        }
        __syncthreads();
        const int d = find_digit(hist[0], r, &rank_sh, &digit_sh);
        r = rank_sh;
        eq = hist[0][d];
        prefix |= (unsigned long long)d << shift;
        __syncthreads();
    }
    const float a = __uint_as_float((unsigned)prefix);
    float q = a;
    if (lo + 1 < n) {
        unsigned long long kb = prefix;
        if (r + 1 >= eq) kb = block_min_above(keys, n, prefix);  // no second copy of the key at rank lo + 1
        const float b = __uint_as_float((unsigned)kb);
        q = __double2float_rn(__dadd_rn((double)a, __dmul_rn((double)__fsub_rn(b, a), pos - (double)lo)));
    }
    if (tid == 0) *J.quantile = q;
}

// Per chunk of 256 windows: the window figures, the structure mask, the histogram entries, the chunk's exact
// integer totals and its oriented-energy sums as a halving tree; the job's last chunk block adds the chunks in order.
__global__ __launch_bounds__(kCollectThreads) void collect_apply_kernel(const CollectJob *jobs, CollectWork w)
{
    __shared__ double tree[4][kCollectThreads];
    __shared__ long long part[12][kCollectThreads / 64];
    const int job = blockIdx.y, chunk = blockIdx.x, tid = threadIdx.x;
    const CollectJob &J = jobs[job];
    const int n = J.windows, nchunks = (n + kCollectThreads - 1) / kCollectThreads;
    if (chunk >= nchunks) return;  // uniform over the block
    const float q = (float)*J.quantile;  // a float value held in a double
    const int i = chunk * kCollectThreads + tid;
    long long zi[12] = {0};
    double o[4] = {0, 0, 0, 0};
    bool m = false;
    int bins[5] = {0, 0, 0, 0, 0};
    if (i < n) {
        const long long *z = J.sums + (size_t)i * 12;
        long long zz[12];
#pragma unroll
        for (int k = 0; k < 12; k++) zz[k] = z[k];
        double v[8];
        window_values(zz, v);
        const float of = __double2float_rn(v[6]);  // compared as the reference's float figure
        m = of >= q && of > 0.0f;
        J.mask[i] = (uint8_t)m;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            zi[k] = zz[k];
            zi[6 + k] = m ? zz[k] : 0;
        }
        o[0] = v[6];
        o[1] = v[7];
        o[2] = m ? v[6] : 0.0;
        o[3] = m ? v[7] : 0.0;
        if (m) {
            const double er[3] = {v[0], v[1], v[6]}, ed[3] = {v[2], v[3], v[7]};
#pragma unroll
            for (int k = 0; k < 3; k++) bins[k] = hist_bin(100.0 * sqrt(ed[k] / fmax(er[k], 1e-9)));
#pragma unroll
            for (int k = 0; k < 2; k++) bins[3 + k] = hist_bin(100.0 * v[4 + k] / fmax(er[k], 1e-9));
        }
    }
    if (i < n) {
        uint16_t *bo = w.bins + ((size_t)job * w.max_windows + i) * 6;
#pragma unroll
        for (int k = 0; k < 5; k++) bo[k] = m ? (uint16_t)bins[k] : kNoBin;
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
    block_sum_ll(zi, part);
    const size_t c = (size_t)job * w.max_chunks + chunk;
    if (tid < 12) w.cint[c * 12 + tid] = zi[tid];
    if (tid < 4) w.cdbl[c * 4 + tid] = tree[tid][0];
}

// The lasting-texture windows of the centre frame six back (src/main.cpp:630-641), per chunk of 256 windows, with
// exact [own, mine, count] totals added up by the job's last chunk block.
__global__ __launch_bounds__(kCollectThreads) void collect_lasting_kernel(const CollectJob *jobs, CollectWork w)
{
    __shared__ long long part[3][kCollectThreads / 64];
    const int job = blockIdx.y, chunk = blockIdx.x, tid = threadIdx.x;
    const CollectJob &J = jobs[job];
    const int n = J.windows, nchunks = (n + kCollectThreads - 1) / kCollectThreads;
    if (!J.psrc || chunk >= nchunks) return;  // uniform over the block
    const int i = chunk * kCollectThreads + tid;
    long long lz[3] = {0, 0, 0};
    bool keep = false;
    int bin = 0;
    if (i < n) {
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
            lz[0] = i0;
            lz[1] = i1;
            lz[2] = 1;
            keep = true;
            bin = hist_bin(100.0 * v1 / v0);
        }
    }
    if (i < n) w.bins[((size_t)job * w.max_windows + i) * 6 + 5] = keep ? (uint16_t)bin : kNoBin;
    block_sum_ll(lz, part);
    const size_t c = (size_t)job * w.max_chunks + chunk;
    if (tid < 3) w.clast[c * 3 + tid] = lz[tid];
}

// A job's totals from its chunks' partial results (written by the launches before): the integer parts in any order,
// every thread taking some chunks; the oriented sums chunk by chunk in order, staged through LDS 256 chunks at a time
// and added by one thread.
__global__ __launch_bounds__(kCollectThreads) void collect_total_kernel(const CollectJob *jobs, CollectWork w)
{
    __shared__ double stage[4][kCollectThreads];
    __shared__ long long part[15][kCollectThreads / 64];
    const int job = blockIdx.x, tid = threadIdx.x;
    const CollectJob &J = jobs[job];
    const int n = J.windows, nchunks = (n + kCollectThreads - 1) / kCollectThreads;
    if (n <= 0) return;
    const long long *ci = w.cint + (size_t)job * w.max_chunks * 12;
    const long long *cl = w.clast + (size_t)job * w.max_chunks * 3;
    const double *cd = w.cdbl + (size_t)job * w.max_chunks * 4;
    const bool lasting = J.psrc != nullptr;
    long long t[15] = {0};
    for (int k = tid; k < nchunks; k += kCollectThreads) {
#pragma unroll
        for (int j = 0; j < 12; j++) t[j] += ci[(size_t)k * 12 + j];
        if (lasting)
#pragma unroll
            for (int j = 0; j < 3; j++) t[12 + j] += cl[(size_t)k * 3 + j];
    }
    block_sum_ll(t, part);
    if (tid < 12) J.tot_int[tid] = t[tid];
    if (lasting && tid >= 12 && tid < 15) J.tot_lasting[tid - 12] = t[tid];
    double tot[4] = {0, 0, 0, 0};
    for (int k0 = 0; k0 < nchunks; k0 += kCollectThreads) {
        __syncthreads();
        if (k0 + tid < nchunks)
#pragma unroll
            for (int j = 0; j < 4; j++) stage[j][tid] = cd[(size_t)(k0 + tid) * 4 + j];
        __syncthreads();
        if (tid == 0)
            for (int k = 0; k < kCollectThreads && k0 + k < nchunks; k++)
#pragma unroll
                for (int j = 0; j < 4; j++) tot[j] += stage[j][k];
    }
    if (tid == 0)
#pragma unroll
        for (int j = 0; j < 4; j++) J.tot_orient[j] = tot[j];
}

// The histogram entries of the round.  One block per (frame, encode, histogram) counts the frame's bins in LDS and
// writes the counts out; then one block per range of bins and (encode, histogram) adds the frames' counts, frame by
// frame in order, into the histogram of each frame's segment (a run of frames of one segment is added at once).  Each
// bin of each histogram has one writer, so plain read-modify-writes add the counts exactly.
constexpr int kHistThreads = 1024;
__global__ __launch_bounds__(kHistThreads) void collect_hist_kernel(const CollectJob *jobs, CollectWork w)
{
    __shared__ unsigned cnt[kHistBins];
    const int job = blockIdx.x / 6, k = blockIdx.x % 6, tid = threadIdx.x;
    const CollectJob &J = jobs[job];
    const int n = J.windows;
    if (n <= 0 || (k == 5 && !J.psrc)) return;  // uniform over the block: nothing to count
    for (int b = tid; b < kHistBins; b += kHistThreads) cnt[b] = 0;
    __syncthreads();
    const uint16_t *bi = w.bins + (size_t)job * w.max_windows * 6 + k;
    for (int i = tid; i < n; i += kHistThreads) {
        const unsigned b = bi[(size_t)i * 6];
        if (b != kNoBin) atomicAdd(&cnt[b], 1u);
    }
    __syncthreads();
    unsigned *out = w.fhist + ((size_t)job * 6 + k) * kHistBins;
    for (int b = tid; b < kHistBins; b += kHistThreads) out[b] = cnt[b];
}

__global__ __launch_bounds__(256) void collect_merge_kernel(const CollectJob *jobs, int nframes, int nenc, CollectWork w)
{
    const int e = blockIdx.y / 6, k = blockIdx.y % 6;
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= kHistBins) return;
    unsigned long long *cur = nullptr, acc = 0;
    for (int f = 0; f < nframes; f++) {
        const int job = f * nenc + e;
        const CollectJob &J = jobs[job];
        if (J.windows <= 0 || (k == 5 && !J.psrc)) continue;  // uniform: this frame has no counts here
        unsigned long long *target = (k < 5 ? J.hist : J.lhist) + (size_t)k * kHistBins;
        if (target != cur) {
            if (cur && acc) cur[b] += acc;
            cur = target;
            acc = 0;
        }
        acc += w.fhist[((size_t)job * 6 + k) * kHistBins + b];
    }
    if (cur && acc) cur[b] += acc;
}

void launch_collect(const CollectJob *jobs, int nframes, int nenc, const CollectWork &w, hipStream_t st)
{
    const int njobs = nframes * nenc;
    if (njobs <= 0) return;
    const dim3 grid(w.max_chunks, njobs);
    hipLaunchKernelGGL(collect_keys_kernel, grid, dim3(kCollectThreads), 0, st, jobs, w);
    hipLaunchKernelGGL(collect_select_kernel, dim3(njobs), dim3(kSelectThreads), 0, st, jobs, w);
    hipLaunchKernelGGL(collect_apply_kernel, grid, dim3(kCollectThreads), 0, st, jobs, w);
    hipLaunchKernelGGL(collect_lasting_kernel, grid, dim3(kCollectThreads), 0, st, jobs, w);
    hipLaunchKernelGGL(collect_total_kernel, dim3(njobs), dim3(kCollectThreads), 0, st, jobs, w);
    hipLaunchKernelGGL(collect_hist_kernel, dim3(njobs * 6), dim3(kHistThreads), 0, st, jobs, w);
    hipLaunchKernelGGL(collect_merge_kernel, dim3((kHistBins + 255) / 256, nenc * 6), dim3(256), 0, st, jobs, nframes,
                       nenc, w);
    hip_check(hipGetLastError(), "collect");
}

// Histogram::quantile for nhist histograms of kHistBins bins and nq quantiles each: out[h * nq + j], NaN when empty.
// Q is the quantiles' source: a device array, or a QuantileList passed by value.  Every call writes every entry of out
// and (when given) of counts.
template <class Q>
__global__ __launch_bounds__(256) void hist_quantile_kernel(const unsigned long long *hists, Q qs, int nq, double *out,
                                                            unsigned long long *counts)
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
    hipLaunchKernelGGL(hist_quantile_kernel<const double *>, dim3(nhist), dim3(256), 0, st, hists, qs, nq, out, counts);
    hip_check(hipGetLastError(), "histogram quantiles");
}

void launch_hist_quantiles_list(const unsigned long long *hists, int nhist, const QuantileList &qs, int nq, double *out,
                                unsigned long long *counts, hipStream_t st)
{
    if (nhist <= 0) return;
    hipLaunchKernelGGL(hist_quantile_kernel<QuantileList>, dim3(nhist), dim3(256), 0, st, hists, qs, nq, out, counts);
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
