// Frame ingest, PSNR and colour reductions for gfx950.
//
// frame_planes (src/kernels.cu:282-310): a decoded 4:2:0 frame, planar or semi-planar, 8 or 16 bits per sample,
// cropped and brought to the pipeline depth as packed luma and chroma planes.  One thread per 2x2 luma quad and its
// chroma sample, the quad's two luma rows written as one 32-bit word each.
//
// plane_sse / color_sums (src/kernels.cu:453-484): the squared error of two planes and the colour sums of a pair of
// chroma planes, in exact integers.  Each block adds its total to the output with one 64-bit integer atomic per
// figure: integer sums do not depend on the order of the adds, so the result is the same on every run.  The colour
// strength sqrtf(du^2 + dv^2) is rounded to single precision as in the reference, then added exactly as an integer in
// units of 2^-23 (every strength of a nonzero square is at least 1, so its float value is a multiple of 2^-23).
//
// frame_sums: the scorer's form, every (source, encode) pair of a round in one launch, reading the decoded frames in
// place at their crops and formats (no packed copies): luma and chroma squared errors and both colour sums.
#include "device_util.hpp"
#include "frame.hpp"

#include <type_traits>

namespace es {

// 2^23 sqrt(s) rounded to single precision (round to nearest), as an exact integer, for 0 <= s < 2^24.  The hardware
// square root r is within one unit in the last place; it is corrected by comparing s with the squares of the midpoints
// to its neighbours.  A midpoint has a 25-bit significand, so it and its square (50 bits) are exact in double, as is
// s.  In r's double the 29 significand bits below the float's 23 are zero: r + ulp / 2 is that pattern with bit 28
// set, and the pattern less 2^28 is the midpoint below, r - ulp / 2 (or, when r is a power of two, r - ulp / 4: the
// borrow crosses into the binade below, whose units are half as large).  No square root of an integer lies on a
// midpoint.  Branch-free: s = 0 gives 0 (the lower pattern is then a NaN, which compares false).
__device__ __forceinline__ unsigned long long strength_fixed(unsigned s)
{
    unsigned b = __float_as_uint(__builtin_amdgcn_sqrtf((float)s));
    const unsigned long long rb = (unsigned long long)__double_as_longlong((double)__uint_as_float(b));
    const double hi = __longlong_as_double((long long)(rb | (1ull << 28)));
    const double lo = __longlong_as_double((long long)(rb - (1ull << 28)));
    const double sd = (double)s;
    b = b + (sd > __dmul_rn(hi, hi) ? 1u : 0u) - (sd < __dmul_rn(lo, lo) ? 1u : 0u);
    const unsigned e = b >> 23, F = (b & 0x7FFFFFu) | 0x800000u;
    return s ? (unsigned long long)F << (e - 127) : 0ull;
}

// The colour figures of one chroma sample pair at the 10-bit scale.
__device__ __forceinline__ void color_add(unsigned u, unsigned v, unsigned long long &su, unsigned long long &sv,
                                          unsigned long long &ss)
{
    const int du = (int)u - 512, dv = (int)v - 512;
    su += u;
    sv += v;
    ss += strength_fixed((unsigned)(__mul24(du, du) + __mul24(dv, dv)));  // |du|, |dv| <= 512 at the 10-bit scale
}

// Adds N per-thread sums of a 256-thread block to out[0 .. N) with one atomic per nonzero block total.
template <int N>
__device__ __forceinline__ void block_add(unsigned long long (&v)[N], unsigned long long *out)
{
    __shared__ unsigned long long part[4][N];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < N; k++) part[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x < N) {
        const unsigned long long s = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
        if (s) atomicAdd(out + threadIdx.x, s);
    }
}

// ---------------------------------------------------------------------------------------------------------------
// frame_planes
// ---------------------------------------------------------------------------------------------------------------
constexpr int kQuadsPerBlock = 256;

template <int BYTES, int SEMI>
__global__ __launch_bounds__(256) void frame_planes_kernel(es_frame f, int x0, int y0, int lshift, uint16_t *y16,
                                                           uint16_t *uo, uint16_t *vo, int w, int h)
{
    using T = typename std::conditional<BYTES == 1, uint8_t, uint16_t>::type;
    const int w2 = w >> 1;
    const int qx = blockIdx.x * kQuadsPerBlock + threadIdx.x, qy = blockIdx.y;
    if (qx >= w2) return;
    const int x = 2 * qx, y = 2 * qy;
    const unsigned rs = (unsigned)f.shift;
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const ES_GLOBAL T *row = (const ES_GLOBAL T *)((const char *)f.y + (size_t)(y0 + y + r) * f.ypitch) + x0 + x;
        const unsigned a = ((unsigned)row[0] >> rs) << lshift, b = ((unsigned)row[1] >> rs) << lshift;
        *(ES_GLOBAL uint32_t *)(y16 + (size_t)(y + r) * w + x) = a | b << 16;
    }
    const int cx = qx + (x0 >> 1), cy = qy + (y0 >> 1);
    const ES_GLOBAL T *cu = (const ES_GLOBAL T *)((const char *)f.u + (size_t)cy * f.cpitch);
    unsigned u, v;
    if (SEMI) {
        u = cu[2 * cx];
        v = cu[2 * cx + 1];
    } else {
        const ES_GLOBAL T *cv = (const ES_GLOBAL T *)((const char *)f.v + (size_t)cy * f.cpitch);
        u = cu[cx];
        v = cv[cx];
    }
    const size_t ci = (size_t)qy * w2 + qx;
    ((ES_GLOBAL uint16_t *)uo)[ci] = (uint16_t)((u >> rs) << lshift);
    ((ES_GLOBAL uint16_t *)vo)[ci] = (uint16_t)((v >> rs) << lshift);
}

void launch_frame_planes(const es_frame &f, int x0, int y0, int lshift, uint16_t *y16, uint16_t *u, uint16_t *v, int w,
                         int h, hipStream_t st)
{
    const dim3 grid((w / 2 + kQuadsPerBlock - 1) / kQuadsPerBlock, h / 2);
    auto k = f.bytes == 1 ? (f.semi ? frame_planes_kernel<1, 1> : frame_planes_kernel<1, 0>)
                          : (f.semi ? frame_planes_kernel<2, 1> : frame_planes_kernel<2, 0>);
    hipLaunchKernelGGL(k, grid, dim3(kQuadsPerBlock), 0, st, f, x0, y0, lshift, y16, u, v, w, h);
    hip_check(hipGetLastError(), "frame planes");
}

// ---------------------------------------------------------------------------------------------------------------
// plane_sse and color_sums
// ---------------------------------------------------------------------------------------------------------------
constexpr int kSumThreads = 256;
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
constexpr int kSumMaxBlocks = 2048;

static int sum_blocks(size_t n)
{
    const size_t per_block = (size_t)kSumThreads * 16;  // about 16 samples per thread
    return (int)std::max<size_t>(1, std::min<size_t>((n + per_block - 1) / per_block, kSumMaxBlocks));
}

__device__ __forceinline__ unsigned sq_diff(unsigned a, unsigned b)
{
    const unsigned d = __usad(a, b, 0u);  // |a - b| < 2^16: an unsigned 24-bit product, below 2^32
    return __umul24(d, d);
}

// vec: a and b (or u and v) are 16-byte aligned; then eight samples per load, the remainder one at a time.
template <bool VEC>
__global__ __launch_bounds__(kSumThreads) void plane_sse_kernel(const uint16_t *a, const uint16_t *b, const float *fa,
                                                                const float *fb, size_t n, unsigned long long *sse)
{
    unsigned long long s[1] = {0};
    const size_t tid = (size_t)blockIdx.x * kSumThreads + threadIdx.x, stride = (size_t)gridDim.x * kSumThreads;
    if (a) {
        size_t i0 = 0;
        if (VEC) {
            const size_t n8 = n / 8;
            const ES_GLOBAL u32x4 *a8 = (const ES_GLOBAL u32x4 *)a, *b8 = (const ES_GLOBAL u32x4 *)b;
            for (size_t i = tid; i < n8; i += stride) {
                const u32x4 x = a8[i], y = b8[i];
                const unsigned xs[4] = {x.x, x.y, x.z, x.w}, ys[4] = {y.x, y.y, y.z, y.w};
#pragma unroll
                for (int k = 0; k < 4; k++)
                    s[0] += (unsigned long long)sq_diff(xs[k] & 0xFFFFu, ys[k] & 0xFFFFu) + sq_diff(xs[k] >> 16, ys[k] >> 16);
            }
            i0 = n8 * 8;
        }
        for (size_t i = i0 + tid; i < n; i += stride) s[0] += sq_diff(((const ES_GLOBAL uint16_t *)a)[i], ((const ES_GLOBAL uint16_t *)b)[i]);
    } else {
        for (size_t i = tid; i < n; i += stride) {
            const long long d = (long long)((const ES_GLOBAL float *)fa)[i] - (long long)((const ES_GLOBAL float *)fb)[i];
            s[0] += (unsigned long long)(d * d);
        }
    }
    block_add<1>(s, sse);
}

template <bool VEC>
__global__ __launch_bounds__(kSumThreads) void color_sums_kernel(const uint16_t *u, const uint16_t *v, size_t n, int up,
                                                                 unsigned long long *sums)
{
    unsigned long long s[3] = {0, 0, 0};
    const size_t tid = (size_t)blockIdx.x * kSumThreads + threadIdx.x, stride = (size_t)gridDim.x * kSumThreads;
    size_t i0 = 0;
    if (VEC) {
        const size_t n8 = n / 8;
        const ES_GLOBAL u32x4 *u8 = (const ES_GLOBAL u32x4 *)u, *v8 = (const ES_GLOBAL u32x4 *)v;
        for (size_t i = tid; i < n8; i += stride) {
            const u32x4 x = u8[i], y = v8[i];
            const unsigned xs[4] = {x.x, x.y, x.z, x.w}, ys[4] = {y.x, y.y, y.z, y.w};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                color_add((xs[k] & 0xFFFFu) << up, (ys[k] & 0xFFFFu) << up, s[0], s[1], s[2]);
                color_add((xs[k] >> 16) << up, (ys[k] >> 16) << up, s[0], s[1], s[2]);
            }
        }
        i0 = n8 * 8;
    }
    for (size_t i = i0 + tid; i < n; i += stride)
        color_add((unsigned)((const ES_GLOBAL uint16_t *)u)[i] << up, (unsigned)((const ES_GLOBAL uint16_t *)v)[i] << up,
                  s[0], s[1], s[2]);
    block_add<3>(s, sums);
}

static bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

void launch_plane_sse(const uint16_t *a, const uint16_t *b, const float *fa, const float *fb, size_t n,
                      unsigned long long *sse, hipStream_t st)
{
    if (n == 0) return;
    const bool vec = a && aligned16(a) && aligned16(b);
    hipLaunchKernelGGL(vec ? plane_sse_kernel<true> : plane_sse_kernel<false>, dim3(sum_blocks(n)), dim3(kSumThreads), 0,
                       st, a, b, fa, fb, n, sse);
    hip_check(hipGetLastError(), "plane sse");
}

void launch_color_sums(const uint16_t *u, const uint16_t *v, size_t n, int up, unsigned long long *sums, hipStream_t st)
{
    if (n == 0) return;
    const bool vec = aligned16(u) && aligned16(v);
    hipLaunchKernelGGL(vec ? color_sums_kernel<true> : color_sums_kernel<false>, dim3(sum_blocks(n)), dim3(kSumThreads),
                       0, st, u, v, n, up, sums);
    hip_check(hipGetLastError(), "color sums");
}

// The reference's color_sums output (src/kernels.cu:470-484) from the exact sums: adds sum u, sum v and the strength
// sum (units of 2^-23) to three doubles, as the reference's kernel adds into its caller's cleared sums.
__global__ void color_to_double_kernel(const unsigned long long *exact, double *sums)
{
    if (threadIdx.x == 0) {
        sums[0] += (double)exact[0];
        sums[1] += (double)exact[1];
        sums[2] += ldexp((double)exact[2], -23);
    }
}

void launch_color_to_double(const unsigned long long *exact, double *sums, hipStream_t st)
{
    hipLaunchKernelGGL(color_to_double_kernel, dim3(1), dim3(64), 0, st, exact, sums);
    hip_check(hipGetLastError(), "color sums");
}

// ---------------------------------------------------------------------------------------------------------------
// frame_sums: per block kFrameSumRows luma rows and half as many chroma rows of one job (luma outside the area the
// window-stats kernel sums).  A lane takes kSumUnroll
// columns 256 apart (clamped to the row, the surplus ones counted as zero) and walks down the block's rows two at a
// time, so that a wave keeps many loads in flight.  Planes are read as buffers: a column's byte offset is fixed for
// the walk and a row's offset is one scalar.  The sample sizes of the two pictures are template parameters, chosen
// once per block.
// ---------------------------------------------------------------------------------------------------------------
constexpr int kSumUnroll = 8;

// A plane of BYTES bytes per sample as a buffer over the `rows` rows the block may read: sample (column offset vo,
// row y) = (word >> rshift) << shift.
template <int BYTES>
struct SumPlane {
    __amdgpu_buffer_rsrc_t rs;
    unsigned pitch;  // bytes per row
    int rshift, shift;
    __device__ __forceinline__ SumPlane(const void *p, int pitch_samples, unsigned long long samples, int rsh, int sh)
        : rs(buffer_over(p, samples * BYTES)), pitch((unsigned)pitch_samples * BYTES), rshift(rsh), shift(sh) {}
    __device__ __forceinline__ unsigned at(unsigned vo, int y) const
    {
        const unsigned so = __builtin_amdgcn_readfirstlane((unsigned)y * pitch);
        const unsigned v = BYTES == 1 ? (unsigned)__builtin_amdgcn_raw_buffer_load_b8(rs, vo, so, 0)
                                      : (unsigned)__builtin_amdgcn_raw_buffer_load_b16(rs, vo, so, 0);
        return (v >> rshift) << shift;
    }
};

template <int BR, int BD>
__device__ __forceinline__ void frame_sums_block(const FrameSumJob &J, int r0)
{
    constexpr int U = kSumUnroll;
    const int w = J.w, r1 = min(r0 + kFrameSumRows, J.h);
    unsigned long long s[kFrameSums] = {};
    {
        // the planes' samples up to row r1 - 1, column w - 1
        const unsigned long long nr = (unsigned long long)(r1 - 1) * J.ry.pitch + w, nd = (unsigned long long)(r1 - 1) * J.dy.pitch + w;
        const SumPlane<BR> R(J.ry.ptr, J.ry.pitch, nr, J.ry.rshift, J.ry.shift);
        const SumPlane<BD> D(J.dy.ptr, J.dy.pitch, nd, J.dy.rshift, J.dy.shift);
        // the block's rows lie all inside or all below the window-stats area, whose columns [0, lw) are summed there
        const int lx0 = r0 < J.lh ? J.lw : 0;
        for (int xb = lx0 + (int)threadIdx.x; xb < w; xb += 256 * U) {
            unsigned cr[U], cd[U];
            bool in[U];
#pragma unroll
            for (int k = 0; k < U; k++) {
                const int x = xb + 256 * k;
                in[k] = x < w;
                cr[k] = (unsigned)min(x, w - 1) * BR;
                cd[k] = (unsigned)min(x, w - 1) * BD;
            }
            for (int y = r0; y < r1; y += 2) {
                const int y2 = min(y + 1, r1 - 1);  // a lone last row is read twice and counted once
                unsigned a[2][U], b[2][U];
#pragma unroll
                for (int k = 0; k < U; k++) {
                    a[0][k] = R.at(cr[k], y);
                    b[0][k] = D.at(cd[k], y);
                    a[1][k] = R.at(cr[k], y2);
                    b[1][k] = D.at(cd[k], y2);
                }
#pragma unroll
                for (int k = 0; k < U; k++) {
                    s[0] += in[k] ? sq_diff(a[0][k], b[0][k]) : 0u;
                    s[0] += in[k] && y + 1 < r1 ? sq_diff(a[1][k], b[1][k]) : 0u;
                }
            }
        }
    }
    const int w2 = w >> 1, c0 = r0 >> 1, c1 = r1 >> 1, up = J.up;
    if (c1 > c0) {
        const bool src_color = J.src_color;
        auto plane = [&](const ChromaIn &c, auto bytes) {
            using P = SumPlane<decltype(bytes)::value>;
            return P(c.ptr, c.pitch, (unsigned long long)(c1 - 1) * c.pitch + (unsigned long long)(w2 - 1) * c.step + 1, c.rshift,
                     c.shift);
        };
        const auto RU = plane(J.ru, std::integral_constant<int, BR>()), RV = plane(J.rv, std::integral_constant<int, BR>());
        const auto DU = plane(J.du, std::integral_constant<int, BD>()), DV = plane(J.dv, std::integral_constant<int, BD>());
        for (int xb = threadIdx.x; xb < w2; xb += 256 * U) {
            unsigned crr[U], cdd[U];
            bool in[U];
#pragma unroll
            for (int k = 0; k < U; k++) {
                const int x = xb + 256 * k;
                in[k] = x < w2;
                crr[k] = (unsigned)min(x, w2 - 1) * J.ru.step * BR;  // u and v of a picture share their step
                cdd[k] = (unsigned)min(x, w2 - 1) * J.du.step * BD;
            }
            for (int y = c0; y < c1; y++) {
                unsigned ur[U], vr[U], ud[U], vd[U];
#pragma unroll
                for (int k = 0; k < U; k++) {
                    ur[k] = RU.at(crr[k], y);
                    vr[k] = RV.at(crr[k], y);
                    ud[k] = DU.at(cdd[k], y);
                    vd[k] = DV.at(cdd[k], y);
                }
#pragma unroll
                for (int k = 0; k < U; k++) {
                    if (!in[k]) break;  // the lane's columns are in order: the rest are past the row too
                    s[1] += sq_diff(ur[k], ud[k]);
                    s[2] += sq_diff(vr[k], vd[k]);
                    color_add(ud[k] << up, vd[k] << up, s[3], s[4], s[5]);
                    if (src_color) color_add(ur[k] << up, vr[k] << up, s[6], s[7], s[8]);
                }
            }
        }
    }
    block_add<kFrameSums>(s, J.out);
}

__global__ __launch_bounds__(256) void frame_sums_kernel(const FrameSumJob *jobs)
{
    const FrameSumJob &J = jobs[blockIdx.y];
    const int r0 = blockIdx.x * kFrameSumRows;
    if (J.w <= 0 || r0 >= J.h) return;  // inactive job or past its rows: uniform over the block
    if (J.ry.bytes == 2) {
        if (J.dy.bytes == 2) frame_sums_block<2, 2>(J, r0);
        else frame_sums_block<2, 1>(J, r0);
    } else {
        if (J.dy.bytes == 2) frame_sums_block<1, 2>(J, r0);
        else frame_sums_block<1, 1>(J, r0);
    }
}

void launch_frame_sums(const FrameSumJob *jobs, int njobs, int max_h, hipStream_t st)
{
    if (njobs <= 0 || max_h <= 0) return;
    hipLaunchKernelGGL(frame_sums_kernel, dim3((max_h + kFrameSumRows - 1) / kFrameSumRows, njobs), dim3(256), 0, st, jobs);
    hip_check(hipGetLastError(), "frame sums");
}

}  // namespace es
