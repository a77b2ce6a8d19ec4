// CAMBI (contrast-aware multiscale banding index) kernels for gfx950.
//
// The per-picture pipeline of Cambi::launch / Cambi::finish (src/cambi.cpp:88-184) as four kernel kinds:
//   cambi_base     scale 0: the spatial mask of the unfiltered picture (src/kernels.cu:9-40) and the 3x3 mode filter
//                  (src/kernels.cu:486-500) from one LDS-staged tile, packed as value | mask << 15;
//   cambi_scale    scales 1-4: decimation of the previous packed scale (src/kernels.cu:42-53) and the mode filter;
//   cambi_cvalues  c-values of every scale (src/kernels.cu:129-278): per 64-column tile one wave keeps the window's
//                  column counts of the tile's relevant values in LDS and turns them into window counts with a
//                  wave-wide prefix sum per row; exact sums of the positive c-values and a compact list of them;
//   topk_pass      the k-th largest positive c-value (src/kernels.cu:71-95, src/cambi.cpp:137-170) by a 3-pass
//                  9-bit-digit radix select done on the device, the last block of each pass walking the digits;
//   cambi_finish   the pooled scales and the score (src/cambi.cpp:172-184).
#include "device_util.hpp"

namespace es {

// ---------------------------------------------------------------------------------------------------------------
// Shared helpers
// ---------------------------------------------------------------------------------------------------------------

// The scale-0 value at (x, y): the sample, or with from8 cambi.c's 8-bit anti-dithering (src/kernels.cu:522-535).
__device__ __forceinline__ unsigned base_value(const PlaneIn &in, int w, int h, int x, int y)
{
    if (!in.from8) return raw_sample(in, x, y);
    if (x < w - 1 && y < h - 1)
        return (raw_sample(in, x, y) + raw_sample(in, x + 1, y) + raw_sample(in, x, y + 1) + raw_sample(in, x + 1, y + 1)) >> 2;
    if (y < h - 1) return (raw_sample(in, x, y) + raw_sample(in, x, y + 1)) >> 1;
    if (x < w - 1) return (raw_sample(in, x, y) + raw_sample(in, x + 1, y)) >> 1;
    return raw_sample(in, x, y);
}

// mode3 (src/kernels.cu:55-60)
__device__ __forceinline__ uint16_t mode3(uint16_t a, uint16_t b, uint16_t c)
{
    if (a == b || a == c) return a;
    if (b == c) return b;
    return min(a, min(b, c));
}

// ---------------------------------------------------------------------------------------------------------------
// Scale 0: spatial mask + mode filter, one 64 x 32 tile per block of 256 threads.
// ---------------------------------------------------------------------------------------------------------------
constexpr int kBaseW = 64, kBaseH = 32;

int base_tiles(int w, int h) { return ((w + kBaseW - 1) / kBaseW) * ((h + kBaseH - 1) / kBaseH); }

__global__ __launch_bounds__(256) void cambi_base_kernel(const CambiPic *pics)
{
    __shared__ uint16_t v[kBaseH + 8][kBaseW + 8];  // the picture around the tile, 4 pixels of margin
    __shared__ uint8_t flat[kBaseH + 6][kBaseW + 6];
    __shared__ uint8_t hs[kBaseH + 6][kBaseW];      // flat pixels among the 7 of each row around a column
    const CambiPic &pic = pics[blockIdx.y];
    const int w = pic.sw[0], h = pic.sh[0];
    const int across = (w + kBaseW - 1) / kBaseW;
    if ((int)blockIdx.x >= across * ((h + kBaseH - 1) / kBaseH)) return;
    const int x0 = (blockIdx.x % across) * kBaseW, y0 = (blockIdx.x / across) * kBaseH;
    const int tid = threadIdx.x;
    for (int i = tid; i < (kBaseH + 8) * (kBaseW + 8); i += 256) {
        const int r = i / (kBaseW + 8), c = i % (kBaseW + 8);
        const int x = x0 - 4 + c, y = y0 - 4 + r;
        v[r][c] = (x >= 0 && x < w && y >= 0 && y < h) ? (uint16_t)base_value(pic.in, w, h, x, y) : kNone;
    }
    __syncthreads();
    // flat: equal to the right and lower neighbours (the last column and row count as equal); never outside
    for (int i = tid; i < (kBaseH + 6) * (kBaseW + 6); i += 256) {
        const int r = i / (kBaseW + 6), c = i % (kBaseW + 6);
        const int x = x0 - 3 + c, y = y0 - 3 + r;
        uint8_t f = 0;
        if (x >= 0 && x < w && y >= 0 && y < h) {
            const uint16_t a = v[r + 1][c + 1];
            f = (uint8_t)((x == w - 1 || a == v[r + 1][c + 2]) && (y == h - 1 || a == v[r + 2][c + 1]));
        }
        flat[r][c] = f;
    }
    __syncthreads();
    for (int i = tid; i < (kBaseH + 6) * kBaseW; i += 256) {
        const int r = i / kBaseW, c = i % kBaseW;
        int s = 0;
#pragma unroll
        for (int d = 0; d < 7; d++) s += flat[r][c + d];
        hs[r][c] = (uint8_t)s;
    }
    __syncthreads();
    const int index = pic.consts->mask_index;
    const int c = tid & 63, g = tid >> 6, x = x0 + c;
    if (x >= w) return;
    uint16_t *out = pic.planes;
#pragma unroll 2
    for (int j = 0; j < kBaseH / 4; j++) {
        const int r = g * (kBaseH / 4) + j, y = y0 + r;
        if (y >= h) break;
        int count = 0;
#pragma unroll
        for (int d = 0; d < 7; d++) count += hs[r + d][c];
        const unsigned mask = count > index;
        const int vr = r + 4, vc = c + 4;
        uint16_t m;
        if (y == 0 || y == h - 1) {
            m = v[vr][vc];
        } else {
            uint16_t t[3];
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const int row = vr - 1 + q;
                t[q] = (x == 0 || x == w - 1) ? v[row][vc] : mode3(v[row][vc - 1], v[row][vc], v[row][vc + 1]);
            }
            m = mode3(t[0], t[1], t[2]);
        }
        out[(size_t)y * w + x] = (uint16_t)(m | (mask << 15));
    }
}

void launch_cambi_base(const CambiPic *pics, const int * /*tile_first*/, int npics, int max_tiles, hipStream_t st)
{
    if (npics <= 0) return;
    hipLaunchKernelGGL(cambi_base_kernel, dim3(max_tiles, npics), dim3(256), 0, st, pics);
    hip_check(hipGetLastError(), "cambi base");
}

// ---------------------------------------------------------------------------------------------------------------
// Scales 1-4: decimate + mode filter of the previous packed scale; the mask bit comes along from (2x, 2y).
// ---------------------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void cambi_scale_kernel(const CambiPic *pics, int s)
{
    const CambiPic &pic = pics[blockIdx.y];
    const int w = pic.sw[s], h = pic.sh[s];
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)w * h) return;
    const int x = (int)(i % w), y = (int)(i / w);
    const int pw = pic.sw[s - 1];
    const uint16_t *prev = pic.planes + pic.off[s - 1];
    auto d = [&](int xx, int yy) -> uint16_t { return prev[(size_t)(2 * yy) * pw + 2 * xx] & 0x7FFF; };
    const unsigned mask = prev[(size_t)(2 * y) * pw + 2 * x] >> 15;
    uint16_t m;
    if (y == 0 || y == h - 1) {
        m = d(x, y);
    } else {
        uint16_t t[3];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int yy = y - 1 + q;
            t[q] = (x == 0 || x == w - 1) ? d(x, yy) : mode3(d(x - 1, yy), d(x, yy), d(x + 1, yy));
        }
        m = mode3(t[0], t[1], t[2]);
    }
    pic.planes[pic.off[s] + i] = (uint16_t)(m | (mask << 15));
}

void launch_cambi_scale(const CambiPic *pics, const long long * /*px_first*/, int npics, long long max_px, int s,
                        hipStream_t st)
{
    if (npics <= 0) return;
    hipLaunchKernelGGL(cambi_scale_kernel, dim3((unsigned)((max_px + 255) / 256), npics), dim3(256), 0, st, pics, s);
    hip_check(hipGetLastError(), "cambi scale");
}

// ---------------------------------------------------------------------------------------------------------------
// c-values.  One wave per tile of 64 columns by tile_h rows of one scale of one picture.
//
// Setup: the values of the tile's masked in-band pixels ("served") and of the tile with its window margins
// ("occupied") are flagged; the relevant values, occupied and within 4 of a served one, are numbered in order (at
// most the band).  Values within 4 of each other get numbers within 4 of each other, so a pass over numbers
// b0 .. b0 + B - 1 serves every pixel whose own number is in [b0 + 4, b0 + B - 4) (from 0 in the first pass, to the
// end in the last).
//
// A pass keeps, for each of the 64 + 2 pad columns of the tile and its margins, the counts of the pass's B numbers in
// the vertical window around the current row (u8, 4 per word, in LDS); lane i owns columns i and 64 + i.  Per row, a
// wave-wide prefix sum over the columns of each pair of counts (16 bits each) gives every lane the window count
// P(x + span) - P(x) of its output column, with one ds_bpermute for the far end; the counts go to LDS for the lane's
// own lookups of p0 and the neighbours' counts.  No lane reads another lane's LDS words, so no barrier is needed in
// the row loop.
// ---------------------------------------------------------------------------------------------------------------
constexpr int kCvMaxB = 32;
constexpr int kCvFlags = 1040;  // >= kMaxBand + 2 * kDiffs, multiple of 16

struct CvShared {
    uint32_t cc[kCvMaxB / 4][128];         // column counts: bins 4q .. 4q + 3 of column c in word cc[q][c]
    uint32_t hh[kCvMaxB / 2][64];          // window counts: bins 2j, 2j + 1 of output column x in word hh[j][x]
    uint16_t rank[kMaxBand + 2 * kDiffs];  // number of value v at rank[v + 4], kNone if not relevant
};
static_assert(sizeof(uint32_t) * (kCvMaxB / 2) * 64 >= 2 * kCvFlags, "flags live in hh during the setup");

int cvalue_tiles(int w, int h, int tile_h) { return ((w + kCvTileW - 1) / kCvTileW) * ((h + tile_h - 1) / tile_h); }

struct CvTile {
    const uint16_t *img;
    int w, h, x0, y0, rows;
    int pad, band_base, band_size, vlt;
    int tvi[kDiffs], weight[kDiffs];
    const float *lut;
    float *plane;           // full c-value plane or nullptr
    float *list;
    unsigned int *count;
};

__device__ __forceinline__ unsigned cv_code(const CvTile &t, int x, int y)
{
    if (x < 0 || x >= t.w || y < 0 || y >= t.h) return kNone;
    const unsigned v = t.img[(size_t)y * t.w + x];
    const unsigned u = (v & 0x7FFFu) - (unsigned)t.band_base;
    return (v >> 15) && u < (unsigned)t.band_size ? u : kNone;
}

template <int B>
__device__ void cv_passes(CvShared &sh, const CvTile &t, int relevant, unsigned long long &sum_fx)
{
    constexpr int Q = B / 4;
    const int lane = threadIdx.x, pad = t.pad, span = 2 * pad + 1, cw = kCvTileW + 2 * pad;
    const int xl = t.x0 - pad;  // picture column of margin column 0
    for (int b0 = 0;;) {
        const int lo = b0 ? b0 + kDiffs : 0, hi = b0 + B >= relevant ? relevant : b0 + B - kDiffs;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            sh.cc[q][lane] = 0;
            sh.cc[q][lane + 64] = 0;
        }
        auto update = [&](int y, unsigned delta) {
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const int c = lane + 64 * k;
                if (c >= cw) continue;
                const unsigned u = cv_code(t, xl + c, y);
                if (u == kNone) continue;
                const unsigned b = (unsigned)sh.rank[u + kDiffs] - (unsigned)b0;
                if (b < (unsigned)B) atomicAdd(&sh.cc[b >> 2][c], delta << (8 * (b & 3)));
            }
        };
        const int ya = t.y0 - pad < 0 ? 0 : t.y0 - pad, yb = t.y0 + pad < t.h ? t.y0 + pad : t.h;
        for (int y = ya; y < yb; y++) update(y, 1u);
        for (int r = 0; r < t.rows; r++) {
            const int y = t.y0 + r;
            if (y + pad < t.h) update(y + pad, 1u);
            const unsigned u = cv_code(t, t.x0 + lane, y);
            const unsigned own = u == kNone ? (unsigned)kNone : sh.rank[u + kDiffs];
            const bool serve = own != kNone && (int)own >= lo && (int)own < hi;
            if (__any(serve)) {
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    const uint32_t qa = sh.cc[q][lane], qb = sh.cc[q][lane + 64];
#pragma unroll
                    for (int hf = 0; hf < 2; hf++) {
                        const uint32_t ra = qa >> (16 * hf), rb = qb >> (16 * hf);
                        const uint32_t a = (ra & 0xFFu) | ((ra & 0xFF00u) << 8), b = (rb & 0xFFu) | ((rb & 0xFF00u) << 8);
                        const uint32_t sa = wave_scan_incl(a), sb = wave_scan_incl(b);
                        const uint32_t ta = (uint32_t)__builtin_amdgcn_readlane((int)sa, 63);
                        const uint32_t tb = (uint32_t)__builtin_amdgcn_readlane((int)sb, 63);
                        const uint32_t pa = sa - a, pb = ta + sb - b;  // P(lane), P(64 + lane)
                        const uint32_t give = lane >= span ? pa : pb;  // what the lane span below this one needs
                        uint32_t far = (uint32_t)__builtin_amdgcn_ds_bpermute(((lane + span) & 63) << 2, (int)give);
                        if (lane + span == 128) far = ta + tb;
                        sh.hh[2 * q + hf][lane] = far - pa;
                    }
                }
                float c = 0.0f;
                if (serve) {
                    auto count = [&](unsigned rk) -> unsigned {
                        const unsigned b = rk - (unsigned)b0;
                        return b < (unsigned)B ? (sh.hh[b >> 1][lane] >> (16 * (b & 1))) & 0xFFFFu : 0u;
                    };
                    const int value = (int)u + t.band_base + kDiffs;
                    const unsigned p0 = count(own);
#pragma unroll
                    for (int d = 0; d < kDiffs; d++) {
                        if (value <= t.tvi[d] && value + d + 1 > t.vlt) {
                            const unsigned p1 = count(sh.rank[u + kDiffs + d + 1]), p2 = count(sh.rank[u + kDiffs - d - 1]);
                            const unsigned pm = p1 > p2 ? p1 : p2;
                            const float val = __fmul_rn((float)((unsigned)t.weight[d] * p0 * pm), t.lut[pm + p0]);
                            if (val > c) c = val;
                        }
                    }
                    if (t.plane) t.plane[(size_t)y * t.w + t.x0 + lane] = c;
                    if (c > 0.0f) sum_fx += fixed24(c);
                }
                const bool pos = serve && c > 0.0f;
                const unsigned long long m = __ballot(pos);
                if (m) {
                    unsigned base = 0;
                    if (lane == 0) base = atomicAdd(t.count, (unsigned)__popcll(m));
                    base = (unsigned)__builtin_amdgcn_readfirstlane((int)base);
                    if (pos) t.list[base + __popcll(m & ((1ull << lane) - 1))] = c;
                }
            }
            if (y - pad >= 0) update(y - pad, ~0u);
        }
        if (hi == relevant) break;
        b0 = hi - kDiffs;
    }
}

__device__ int find_slot(const int *first, int n, int t)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ void cv_tile(CvShared &sh, const CambiPic &pic, int s, int local, int tile_h)
{
    const CambiConsts &K = *pic.consts;
    CvTile t;
    t.w = pic.sw[s];
    t.h = pic.sh[s];
    t.img = pic.planes + pic.off[s];
    const int across = (t.w + kCvTileW - 1) / kCvTileW;
    t.x0 = (local % across) * kCvTileW;
    t.y0 = (local / across) * tile_h;
    t.rows = min(tile_h, t.h - t.y0);
    t.pad = K.pad;
    t.band_base = K.band_base;
    t.band_size = K.band_size;
    t.vlt = K.vlt;
#pragma unroll
    for (int d = 0; d < kDiffs; d++) {
        t.tvi[d] = K.tvi[d];
        t.weight[d] = K.weight[d];
    }
    t.lut = pic.lut;
    t.plane = pic.cvals ? pic.cvals + pic.off[s] : nullptr;
    t.list = pic.list + pic.off[s];
    t.count = pic.counts + s;
    const int lane = threadIdx.x, nv = t.band_size + 2 * kDiffs, cw = kCvTileW + 2 * t.pad;
    __syncthreads();  // the previous tile's reads of hh and rank are done
    uint8_t *served = (uint8_t *)&sh.hh[0][0], *occupied = served + kCvFlags;
    for (int i = lane; i < 2 * kCvFlags / 4; i += 64) ((uint32_t *)served)[i] = 0;
    __syncthreads();
    bool any = false;
    for (int r = 0; r < t.rows; r++) {
        const unsigned u = cv_code(t, t.x0 + lane, t.y0 + r);
        if (u != kNone) {
            served[u + kDiffs] = 1;
            any = true;
        }
    }
    if (!__any(any)) return;  // most tiles: nothing masked inside the band
    const int ya = max(t.y0 - t.pad, 0), yb = min(t.y0 + t.rows + t.pad, t.h);
    for (int y = ya; y < yb; y++)
        for (int c = lane; c < cw; c += 64) {
            const unsigned u = cv_code(t, t.x0 - t.pad + c, y);
            if (u != kNone) occupied[u + kDiffs] = 1;
        }
    __syncthreads();
    int relevant = 0;
    for (int i0 = 0; i0 < nv; i0 += 64) {
        const int i = i0 + lane;
        bool need = false;
        if (i < nv && occupied[i])
            for (int k = -kDiffs; k <= kDiffs; k++) need |= i + k >= 0 && i + k < nv && served[i + k];
        const unsigned long long m = __ballot(need);
        if (i < nv) sh.rank[i] = need ? (uint16_t)(relevant + __popcll(m & ((1ull << lane) - 1))) : kNone;
        relevant += __popcll(m);
    }
    __syncthreads();
    if (relevant == 0) return;
    unsigned long long sum_fx = 0;
    if (relevant <= 8) cv_passes<8>(sh, t, relevant, sum_fx);
    else if (relevant <= 16) cv_passes<16>(sh, t, relevant, sum_fx);
    else cv_passes<32>(sh, t, relevant, sum_fx);
    sum_fx = wave_sum(sum_fx);
    if (lane == 0 && sum_fx) atomicAdd(pic.sums + s, sum_fx);
}

__global__ __launch_bounds__(64) void cambi_cvalues_kernel(const CambiPic *pics, const int *tile_first, int nslots,
                                                           int tile_h, int persistent, unsigned int *counter)
{
    __shared__ CvShared sh;
    const int total = tile_first[nslots];
    for (int t = persistent ? 0 : (int)blockIdx.x;;) {
        if (persistent) {
            unsigned v = 0;
            if (threadIdx.x == 0) v = atomicAdd(counter, 1u);
            t = __builtin_amdgcn_readfirstlane((int)v);
        }
        if (t >= total) break;
        const int slot = find_slot(tile_first, nslots, t);
        const int p = slot / kScales, s = slot % kScales;
        cv_tile(sh, pics[p], s, t - tile_first[slot], tile_h);
        if (!persistent) break;
    }
}

void launch_cambi_cvalues(const CambiPic *pics, const int *tile_first, int nslots, int total_tiles, int tile_h,
                          int persistent, unsigned int *work_counter, int blocks, hipStream_t st)
{
    if (!persistent && total_tiles <= 0) return;
    const int grid = persistent ? blocks : total_tiles;
    hipLaunchKernelGGL(cambi_cvalues_kernel, dim3(grid), dim3(64), 0, st, pics, tile_first, nslots, tile_h, persistent,
                       work_counter);
    hip_check(hipGetLastError(), "cambi c-values");
}

// ---------------------------------------------------------------------------------------------------------------
// Top-k select.  Positive c-values are at least 0.5, so their bit patterns less 0x3F000000 are 27-bit keys in value
// order; three passes of 9-bit digits select the k-th largest.  Each pass counts (and sums exactly) the digit of the
// keys that match the digits chosen so far; the last block to finish walks the digits from the top, as the reference's
// host loop does (src/cambi.cpp:151-156), keeps the count and sum of the keys above the chosen digit, and clears the
// histogram for the next pass.  Jobs with no more than k positive values are skipped: their pooling needs no select.
// ---------------------------------------------------------------------------------------------------------------
constexpr unsigned kKeyBase = 0x3F000000u;

__global__ __launch_bounds__(256) void topk_pass_kernel(TopkJob *jobs, int pass)
{
    TopkJob &J = jobs[blockIdx.y];
    const unsigned n = *J.count;
    if ((unsigned long long)n <= J.k) return;
    __shared__ unsigned hist[512];
    __shared__ unsigned long long hs[512];
    __shared__ int last;
    const int tid = threadIdx.x;
    for (int i = tid; i < 512; i += 256) {
        hist[i] = 0;
        hs[i] = 0;
    }
    __syncthreads();
    const int shift = 18 - 9 * pass;
    const unsigned prefix = pass ? J.prefix : 0u;
    for (unsigned i = blockIdx.x * 256u + tid; i < n; i += gridDim.x * 256u) {
        const float c = J.list[i];
        const unsigned key = __float_as_uint(c) - kKeyBase;
        if ((key >> (shift + 9)) == (prefix >> (shift + 9))) {
            const unsigned d = (key >> shift) & 511u;
            atomicAdd(&hist[d], 1u);
            atomicAdd(&hs[d], fixed24(c));
        }
    }
    __syncthreads();
    for (int d = tid; d < 512; d += 256)
        if (hist[d]) {
            atomicAdd(&J.hist[d], hist[d]);
            atomicAdd(&J.hsum[d], hs[d]);
        }
    __threadfence();
    __syncthreads();
    if (tid == 0) last = atomicAdd(J.ticket, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    __threadfence();
    if (tid == 0) {
        const unsigned remaining = pass ? J.remaining : (unsigned)J.k;
        unsigned above = 0;
        unsigned long long above_fx = 0;
        int d = 511;
        for (; d > 0; d--) {
            const unsigned c = __hip_atomic_load(&J.hist[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (above + c >= remaining) break;
            above += c;
            above_fx += __hip_atomic_load(&J.hsum[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        J.remaining = remaining - above;
        J.prefix = prefix | ((unsigned)d << shift);
        J.above = (pass ? J.above : 0ull) + above;
        J.above_fx = (pass ? J.above_fx : 0ull) + above_fx;
    }
    __syncthreads();  // the walk has read the histogram before it is cleared
    for (int i = tid; i < 512; i += 256) {
        J.hist[i] = 0;
        J.hsum[i] = 0;
    }
    if (tid == 0) *J.ticket = 0;
}

void launch_topk_pass(TopkJob *jobs, int njobs, int pass, int blocks_per_job, hipStream_t st)
{
    if (njobs <= 0) return;
    hipLaunchKernelGGL(topk_pass_kernel, dim3(blocks_per_job, njobs), dim3(256), 0, st, jobs, pass);
    hip_check(hipGetLastError(), "top-k select");
}

// One thread per picture: pooled scales and the score (src/cambi.cpp:172-184), from the exact sums and the select.
__global__ void cambi_finish_kernel(const CambiPic *pics, TopkJob *jobs, int npics, double *scores)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npics) return;
    const CambiPic &pic = pics[p];
    const double wts[kScales] = {16, 8, 4, 2, 1};
    double total = 0;
    for (int s = 0; s < kScales; s++) {
        const TopkJob &J = jobs[p * kScales + s];
        const unsigned long long n = pic.counts[s], k = J.k;
        double pooled;
        if (n <= k) {
            pooled = (double)pic.sums[s] * (1.0 / 16777216.0) / (double)k;
        } else {
            const float thr = __uint_as_float(J.prefix + kKeyBase);
            pooled = __dadd_rn((double)J.above_fx * (1.0 / 16777216.0), __dmul_rn((double)(k - J.above), (double)thr)) / (double)k;
        }
        total = __dadd_rn(total, __dmul_rn(pooled, wts[s]));
    }
    const int odd = 2 * (pic.consts->window >> 1) + 1;
    const double score = total / (double)(odd * odd);
    scores[p] = score < 1000.0 ? score : 1000.0;
}

// Clears the per-batch counters: a kernel, so that a captured graph holds no memset node.
__global__ void clear_kernel(unsigned long long *a, long long na, unsigned int *b, long long nb)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < na) a[i] = 0;
    if (i < nb) b[i] = 0;
}

void launch_clear(unsigned long long *a, long long na, unsigned int *b, long long nb, hipStream_t st)
{
    const long long n = na > nb ? na : nb;
    if (n <= 0) return;
    hipLaunchKernelGGL(clear_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, na, b, nb);
    hip_check(hipGetLastError(), "clear");
}

void launch_cambi_finish(const CambiPic *pics, TopkJob *jobs, int npics, double *scores, hipStream_t st)
{
    if (npics <= 0) return;
    hipLaunchKernelGGL(cambi_finish_kernel, dim3((npics + 63) / 64), dim3(64), 0, st, pics, jobs, npics, scores);
    hip_check(hipGetLastError(), "cambi finish");
}

}  // namespace es
