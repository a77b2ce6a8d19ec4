// Synthetic luma for the benchmark configurations, integer-only so the host and the device produce the same frames.
//
// Random numbers are counter-based (SplitMix64 of seed, stream and coordinates).  Gaussian draws are Irwin-Hall:
// the sum of twelve 16-bit uniforms less its mean, standard deviation 65536.  Values are carried in 1/256 of a code.
//
// kind 0 (config 1): L0 + 30 x / (w - 1) + 30 r / r_max (r from a seeded centre) with grain sigma 2.0 everywhere; the
// encode (box3) is the noise-free ramp quantised down to 4-code steps, then a 3x3 box mean (edges repeated).  L0 is
// seeded in [300, 480], so the whole ramp lies in CAMBI's band at 10 bits (codes up to tvi[3] = 563) and the encode's
// steps are measured as banding.
// kind 1 (configs 2-5): 64x64 tiles of the source's picture, seeded per tile: 40% linear gradients (16 directions,
// slope 0.05-0.5 codes/px), 40% static band-limited texture (white noise of sigma 29.25 through a 5x5 binomial
// filter: sigma about 8), 20% dark flat tiles in [64, 120] with grain sigma 1.5 per frame.  An encode keeps a share
// of the grain, then is blurred, quantised and block-averaged as its description says.
#include "host_common.hpp"

#include <cmath>
#include <thread>

namespace es {

namespace {

__host__ __device__ inline uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t hash4(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b, uint64_t c)
{
    uint64_t h = splitmix64(seed ^ (stream << 56));
    h = splitmix64(h ^ (a & 0xFFFFFFFFull) ^ (b << 32));
    return splitmix64(h ^ c);
}

// Irwin-Hall: standard deviation 65536
__host__ __device__ inline long long gauss(uint64_t seed, uint64_t stream, long long frame, int x, int y)
{
    long long s = 0;
    for (int i = 0; i < 3; i++) {
        const uint64_t h = hash4(seed, stream, (uint64_t)frame, (uint64_t)(uint32_t)x | ((uint64_t)(uint32_t)y << 16), (uint64_t)i);
        for (int j = 0; j < 4; j++) s += (long long)((h >> (16 * j)) & 0xFFFFu);
    }
    return s - 393210;
}

__host__ __device__ inline long long floor_div(long long a, long long b)
{
    const long long q = a / b;
    return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

// round(sigma_q8 * g / 65536), half away from zero: a draw in 1/256 code
__host__ __device__ inline long long noise_q8(long long sigma_q8, long long g)
{
    const long long v = sigma_q8 * g;
    return v >= 0 ? (v + 32768) / 65536 : -((-v + 32768) / 65536);
}

__host__ __device__ inline unsigned long long isqrt64(unsigned long long n)
{
    unsigned long long r = 0, bit = 1ull << 62;
    while (bit > n) bit >>= 2;
    while (bit) {
        if (n >= r + bit) {
            n -= r + bit;
            r = (r >> 1) + bit;
        } else {
            r >>= 1;
        }
        bit >>= 2;
    }
    return r;
}

// cos and sin of i pi / 8 in 1/4096
__host__ __device__ inline int dir_cos(int i)
{
    const int t[16] = {4096, 3784, 2896, 1567, 0, -1567, -2896, -3784, -4096, -3784, -2896, -1567, 0, 1567, 2896, 3784};
    return t[i & 15];
}
__host__ __device__ inline int dir_sin(int i) { return dir_cos(i + 12); }

enum { kGrad = 0, kTex = 1, kDark = 2 };

struct Gen {
    int kind, full_w, full_h, crop_x, crop_y, w, h, depth;
    uint64_t seed;
    int salt, grain_q8, blur_q8, quant_step, quant_where, block, block_pct, box3;
    // kind 0
    long long l0, cx, cy, rmax_q8;
    // blur kernel
    int taps;
    int kw[16];
    long long ksum;
};

__host__ __device__ inline int tile_class(const Gen &g, int X, int Y, uint64_t *th)
{
    const uint64_t h = hash4(g.seed, 10, (uint64_t)(X >> 6), (uint64_t)(Y >> 6), 0);
    *th = h;
    const int u = (int)(h % 100);
    return u < 40 ? kGrad : u < 80 ? kTex : kDark;
}

// kind 0's noise-free ramp in 1/256 code
__host__ __device__ inline long long ramp_q8(const Gen &g, int X, int Y)
{
    const long long dx = X - g.cx, dy = Y - g.cy;
    const long long r = (long long)isqrt64((unsigned long long)(dx * dx + dy * dy) << 16);
    return g.l0 * 256 + (7680LL * X) / (g.full_w - 1) + (7680LL * r) / g.rmax_q8;
}

// The static content at source coordinates (X, Y) in 1/256 code, and its tile class.
__host__ __device__ inline long long content_q8(const Gen &g, int X, int Y, int *cls)
{
    if (g.kind == 0) {
        *cls = kTex;
        return ramp_q8(g, X, Y);
    }
    uint64_t th;
    const int c = tile_class(g, X, Y, &th);
    *cls = c;
    if (c == kGrad) {
        const long long level = 200 + (long long)((th >> 8) % 500);
        const int dir = (int)((th >> 20) % 16);
        const long long slope = 205 + (long long)((th >> 28) % 1844);  // 1/4096 code per px
        const long long dx = X - ((X >> 6) * 64 + 32), dy = Y - ((Y >> 6) * 64 + 32);
        const long long proj = dir_cos(dir) * dx + dir_sin(dir) * dy;  // 1/4096 px
        return level * 256 + floor_div(slope * proj, 65536);
    }
    if (c == kTex) {
        const long long level = 300 + (long long)((th >> 8) % 400);
        const int w5[5] = {1, 4, 6, 4, 1};
        long long s = 0;
        for (int j = -2; j <= 2; j++)
            for (int i = -2; i <= 2; i++) s += (long long)(w5[i + 2] * w5[j + 2]) * noise_q8(7488, gauss(g.seed, 11, 0, X + i, Y + j));
        return level * 256 + floor_div(s, 256);
    }
    return (64 + (long long)((th >> 8) % 57)) * 256;
}

// Stage 1: content plus the kept grain, 1/256 code, at picture coordinates (x, y).
__host__ __device__ inline long long stage1(const Gen &g, long long frame, int x, int y)
{
    const int X = x + g.crop_x, Y = y + g.crop_y;
    if (g.kind == 0 && g.box3) {
        // 4-code steps of the noise-free ramp, then the 3x3 mean (edges repeated)
        long long s = 0;
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int xx = min(max(x + dx, 0), g.w - 1) + g.crop_x, yy = min(max(y + dy, 0), g.h - 1) + g.crop_y;
                const long long code = floor_div(ramp_q8(g, xx, yy) + 128, 256);
                s += 4 * floor_div(code, 4);
            }
        return ((s + 4) / 9) * 256;
    }
    int cls;
    long long v = content_q8(g, X, Y, &cls);
    const bool grainy = g.kind == 0 || cls == kDark;
    if (grainy && g.grain_q8) {
        const long long sigma = g.kind == 0 ? 512 : 384;
        v += floor_div(noise_q8(sigma, gauss(g.seed, 12, frame, X, Y)) * g.grain_q8, 256);
    }
    return v;
}

__host__ __device__ inline int to_code(const Gen &g, long long v_q8, int x, int y)
{
    long long code = floor_div(v_q8 + 128, 256);
    if (g.quant_step > 1) {
        bool on = g.quant_where == 0;
        if (g.quant_where && g.kind == 1) {
            uint64_t th;
            const int cls = tile_class(g, x + g.crop_x, y + g.crop_y, &th);
            on = (g.quant_where == 1 && cls == kGrad) || (g.quant_where == 2 && cls == kDark);
        }
        if (on) code = g.quant_step * floor_div(code, g.quant_step);
    }
    return (int)code;
}

__host__ __device__ inline int finish_code(const Gen &g, long long code)
{
    if (g.depth == 8) {
        code = 16 + floor_div((code - 64) * 219 + 438, 876);
        return (int)min(max(code, 16LL), 235LL);
    }
    return (int)min(max(code, 64LL), 940LL);
}

__host__ __device__ inline bool block_on(const Gen &g, int bx, int by)
{
    return (int)(hash4(g.seed, 13, (uint64_t)bx, (uint64_t)by, (uint64_t)g.salt) % 100) < g.block_pct;
}

Gen make_gen(const es_synth_desc &d)
{
    require(d.kind == 0 || d.kind == 1, "synth kind must be 0 or 1");
    require(d.w >= 1 && d.h >= 1 && d.crop_x >= 0 && d.crop_y >= 0 && d.crop_x + d.w <= d.full_w && d.crop_y + d.h <= d.full_h,
            "synth crop outside the picture");
    require(d.depth == 8 || d.depth == 10, "synth depth must be 8 or 10");
    require(d.block == 0 || d.block == 8 || d.block == 16, "block size must be 8 or 16");
    Gen g{};
    g.kind = d.kind;
    g.full_w = d.full_w;
    g.full_h = d.full_h;
    g.crop_x = d.crop_x;
    g.crop_y = d.crop_y;
    g.w = d.w;
    g.h = d.h;
    g.depth = d.depth;
    g.seed = d.seed;
    g.salt = d.salt;
    g.grain_q8 = d.grain_q8;
    g.blur_q8 = d.blur_q8;
    g.quant_step = d.quant_step;
    g.quant_where = d.quant_where;
    g.block = d.block;
    g.block_pct = d.block_pct;
    g.box3 = d.box3;
    if (d.kind == 0) {
        g.l0 = 300 + (long long)(hash4(d.seed, 1, 0, 0, 0) % 181);
        g.cx = (long long)(hash4(d.seed, 2, 0, 0, 0) % (uint64_t)d.full_w);
        g.cy = (long long)(hash4(d.seed, 3, 0, 0, 0) % (uint64_t)d.full_h);
        long long rmax = 1;
        const long long cs[4][2] = {{0, 0}, {d.full_w - 1, 0}, {0, d.full_h - 1}, {d.full_w - 1, d.full_h - 1}};
        for (auto &c : cs) {
            const long long dx = c[0] - g.cx, dy = c[1] - g.cy;
            rmax = std::max(rmax, (long long)isqrt64((unsigned long long)(dx * dx + dy * dy) << 16));
        }
        g.rmax_q8 = rmax;
    }
    if (d.blur_q8 > 0) {
        const double sigma = d.blur_q8 / 256.0;
        g.taps = std::min(7, (int)std::ceil(3 * sigma));
        g.ksum = 0;
        for (int k = -g.taps; k <= g.taps; k++) {
            g.kw[k + g.taps] = (int)std::lround(65536.0 * std::exp(-(double)(k * k) / (2 * sigma * sigma)));
            g.ksum += g.kw[k + g.taps];
        }
    }
    return g;
}

// ---- device stages ----
__global__ void synth_stage1(Gen g, long long frame, long long *out)
{
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= g.w) return;
    out[(size_t)y * g.w + x] = stage1(g, frame, x, y);
}

__host__ __device__ inline long long blur_at(const Gen &g, const long long *in, int x, int y, int dir)
{
    long long s = 0;
    for (int k = -g.taps; k <= g.taps; k++) {
        const int xx = dir == 0 ? min(max(x + k, 0), g.w - 1) : x, yy = dir == 1 ? min(max(y + k, 0), g.h - 1) : y;
        s += (long long)g.kw[k + g.taps] * in[(size_t)yy * g.w + xx];
    }
    return floor_div(2 * s + g.ksum, 2 * g.ksum);
}

__global__ void synth_blur(Gen g, const long long *in, long long *out, int dir)
{
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= g.w) return;
    out[(size_t)y * g.w + x] = blur_at(g, in, x, y, dir);
}

__host__ __device__ inline int block_code(const Gen &g, const long long *v, int x, int y)
{
    const int bx = x / g.block, by = y / g.block;
    if (!block_on(g, bx, by)) return to_code(g, v[(size_t)y * g.w + x], x, y);
    long long s = 0, n = 0;
    for (int yy = by * g.block; yy < min((by + 1) * g.block, g.h); yy++)
        for (int xx = bx * g.block; xx < min((bx + 1) * g.block, g.w); xx++) {
            s += to_code(g, v[(size_t)yy * g.w + xx], xx, yy);
            n++;
        }
    return (int)floor_div(2 * s + n, 2 * n);
}

__global__ void synth_out(Gen g, const long long *v, void *out, int bytes)
{
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= g.w) return;
    const int code = finish_code(g, g.block ? block_code(g, v, x, y) : to_code(g, v[(size_t)y * g.w + x], x, y));
    if (bytes == 1) ((uint8_t *)out)[(size_t)y * g.w + x] = (uint8_t)code;
    else ((uint16_t *)out)[(size_t)y * g.w + x] = (uint16_t)code;
}

template <class F>
void host_rows(int h, int threads, F &&f)
{
    const int n = std::max(1, threads);
    std::vector<std::thread> ts;
    for (int t = 0; t < n; t++)
        ts.emplace_back([&, t] {
            for (int y = t; y < h; y += n) f(y);
        });
    for (auto &t : ts) t.join();
}

}  // namespace

}  // namespace es

using namespace es;

extern "C" {

int es_synth_frame(const es_synth_desc *d, int frame, void *dev_out, int bytes, es_stream s)
{
    return guard([&] {
        ensure_init();
        require(d && dev_out && (bytes == 1 || bytes == 2), "bad argument");
        const Gen g = make_gen(*d);
        const size_t n = (size_t)g.w * g.h;
        static thread_local DevBuf *tmp = nullptr;  // grows to the largest picture generated on this thread
        if (!tmp) tmp = new DevBuf();
        if (tmp->bytes < 2 * n * 8) {
            hip_check(hipStreamSynchronize(hs(s)), "synth");
            tmp->alloc(2 * n * 8);
        }
        long long *a = tmp->as<long long>(), *b = a + n;
        const dim3 grid((g.w + 255) / 256, g.h);
        hipLaunchKernelGGL(synth_stage1, grid, dim3(256), 0, hs(s), g, (long long)frame, a);
        if (g.blur_q8 > 0) {
            hipLaunchKernelGGL(synth_blur, grid, dim3(256), 0, hs(s), g, a, b, 0);
            hipLaunchKernelGGL(synth_blur, grid, dim3(256), 0, hs(s), g, b, a, 1);
        }
        hipLaunchKernelGGL(synth_out, grid, dim3(256), 0, hs(s), g, a, dev_out, bytes);
        hip_check(hipGetLastError(), "synth");
    });
}

int es_synth_frame_host(const es_synth_desc *d, int frame, void *host_out, int bytes, int threads)
{
    return guard([&] {
        require(d && host_out && (bytes == 1 || bytes == 2), "bad argument");
        const Gen g = make_gen(*d);
        const size_t n = (size_t)g.w * g.h;
        std::vector<long long> a(n), b;
        host_rows(g.h, threads, [&](int y) {
            for (int x = 0; x < g.w; x++) a[(size_t)y * g.w + x] = stage1(g, frame, x, y);
        });
        if (g.blur_q8 > 0) {
            b.resize(n);
            host_rows(g.h, threads, [&](int y) {
                for (int x = 0; x < g.w; x++) b[(size_t)y * g.w + x] = blur_at(g, a.data(), x, y, 0);
            });
            host_rows(g.h, threads, [&](int y) {
                for (int x = 0; x < g.w; x++) a[(size_t)y * g.w + x] = blur_at(g, b.data(), x, y, 1);
            });
        }
        host_rows(g.h, threads, [&](int y) {
            for (int x = 0; x < g.w; x++) {
                const int code = finish_code(g, g.block ? block_code(g, a.data(), x, y) : to_code(g, a[(size_t)y * g.w + x], x, y));
                if (bytes == 1) ((uint8_t *)host_out)[(size_t)y * g.w + x] = (uint8_t)code;
                else ((uint16_t *)host_out)[(size_t)y * g.w + x] = (uint16_t)code;
            }
        });
    });
}

}  // extern "C"
