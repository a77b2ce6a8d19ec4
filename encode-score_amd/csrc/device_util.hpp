// Device helpers shared by the kernel files: sample reads, wave-wide scans and reductions, exact fixed point.
#pragma once
#include "es_internal.hpp"

// Global-memory pointers: loads and stores through them are global_* instructions, which (unlike flat ones) do not
// hold up the waits on LDS results.
#define ES_GLOBAL __attribute__((address_space(1)))

namespace es {

// The hardware hands a launch's workgroups to the 8 XCDs in turn (workgroup i, counting x fastest, to XCD i % 8), and
// each XCD has its own L2.  For tiles that share halo rows, halo columns or cache lines with their neighbours, this
// renumbers the workgroups so that each XCD takes one contiguous run of them (in x-fastest order): neighbours then
// meet in the same L2 instead of fetching the shared lines from HBM once per XCD.
__device__ __forceinline__ uint3 xcd_block3()
{
    const unsigned gx = gridDim.x, gy = gridDim.y, n = gx * gy * gridDim.z;
    const unsigned lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned xcd = lin % 8u, k = lin / 8u, q = n / 8u, r = n % 8u;
    const unsigned l = xcd * q + min(xcd, r) + k;  // XCD x takes q (+1 for the first r) consecutive numbers
    return make_uint3(l % gx, (l / gx) % gy, l / (gx * gy));
}

// The wave's index in its block as a wave-uniform value.  The compiler cannot tell that threadIdx.x >> 6 is the same
// on every lane of a 64-wide wave, and would otherwise keep what is derived from it (a strip's origin, its row
// counter, row addresses and row tests) in vector registers and compute it per lane.
__device__ __forceinline__ int wave_index() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// `bytes` bytes at p as a raw buffer (the address made wave-uniform).  Loads through it take a byte offset per lane
// and one in a scalar register, so a loop over rows adds one scalar per row instead of a 64-bit address per lane; a
// load past `bytes` returns 0.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_over(const void *p, unsigned long long bytes)
{
    const unsigned long long a = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a), hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    const unsigned n = __builtin_amdgcn_readfirstlane((unsigned)(bytes < 0xFFFFFFFFull ? bytes : 0xFFFFFFFFull));
    return __builtin_amdgcn_make_buffer_rsrc((void *)((unsigned long long)hi << 32 | lo), (short)0, (int)n, 0x00020000);
}

__device__ __forceinline__ unsigned raw_sample(const PlaneIn &in, int x, int y)
{
    const size_t i = (size_t)y * in.pitch + x;
    const unsigned v = in.bytes == 1 ? (unsigned)((const ES_GLOBAL uint8_t *)in.ptr)[i]
                                     : (unsigned)((const ES_GLOBAL uint16_t *)in.ptr)[i] >> in.rshift;
    return v << in.shift;
}

// Inclusive prefix sum over the 64 lanes of a wave with DPP: shifts of 1, 2, 4, 8 inside rows of 16 lanes, then the
// row sums broadcast from lane 15 to rows 1 and 3 and from lane 31 to rows 2 and 3.  Lanes without a source add 0.
__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);
    return v;
}

// One step of a DPP scan of a 64-bit value: both halves moved from the same source lane, then added.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ long long dpp_add64(long long v)
{
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)v, kCtrl, kRowMask, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)((unsigned long long)v >> 32), kCtrl, kRowMask, 0xF, false);
    return v + (long long)((unsigned long long)hi << 32 | lo);
}

// Inclusive prefix sum of a 64-bit value over the wave's lanes: the DPP steps of wave_scan_incl on both halves.
__device__ __forceinline__ long long wave_scan_incl64(long long v)
{
    v = dpp_add64<0x111, 0xF>(v);
    v = dpp_add64<0x112, 0xF>(v);
    v = dpp_add64<0x114, 0xF>(v);
    v = dpp_add64<0x118, 0xF>(v);
    v = dpp_add64<0x142, 0xA>(v);
    v = dpp_add64<0x143, 0xC>(v);
    return v;
}

// The sum of v over the wave's 64 lanes, on every lane: the inclusive DPP scan above on both halves of the 64-bit
// value (no LDS round trips), then lane 63's total.
__device__ __forceinline__ long long wave_total64(long long v)
{
    v = wave_scan_incl64(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)((unsigned long long)v >> 32), 63);
    return (long long)((unsigned long long)hi << 32 | lo);
}

// The wave totals of four 64-bit values at once; value k's total is left in lane 16 k + 15.  Lane swaps first pair
// each value's lanes with another value's (32 lanes apart: v0 with v2 and v1 with v3, then 16 apart), so that one
// add combines two values' partial sums and every lane keeps one value; a DPP scan inside each row of 16 lanes then
// sums the row.  About a third of the instructions of four separate wave_total64 calls.
__device__ __forceinline__ long long swap_add64_32(long long x, long long y)
{
    // lanes 0-31: x + x of lane + 32; lanes 32-63: y of lane - 32 + y
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)x, (unsigned)y, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)((unsigned long long)x >> 32),
                                                     (unsigned)((unsigned long long)y >> 32), false, false);
    return (long long)((unsigned long long)hi[0] << 32 | lo[0]) + (long long)((unsigned long long)hi[1] << 32 | lo[1]);
}
__device__ __forceinline__ long long swap_add64_16(long long x, long long y)
{
    // rows of 16 lanes: rows 0 and 2 take x + x of the next row, rows 1 and 3 y of the row before + y
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)x, (unsigned)y, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)((unsigned long long)x >> 32),
                                                     (unsigned)((unsigned long long)y >> 32), false, false);
    return (long long)((unsigned long long)hi[0] << 32 | lo[0]) + (long long)((unsigned long long)hi[1] << 32 | lo[1]);
}
__device__ __forceinline__ long long wave_total4x64(const long long v[4])
{
    // after the first swaps, lanes 0-31 hold v0 (of s) and v1 (of t) summed over lanes l and l + 32, lanes 32-63 v2
    // and v3; after the second, row k holds value k summed over 4 lanes
    long long u = swap_add64_16(swap_add64_32(v[0], v[2]), swap_add64_32(v[1], v[3]));
    u = dpp_add64<0x111, 0xF>(u);
    u = dpp_add64<0x112, 0xF>(u);
    u = dpp_add64<0x114, 0xF>(u);
    u = dpp_add64<0x118, 0xF>(u);
    return u;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    for (int o = 32; o; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// A positive c-value (at least 0.5) as an exact integer in units of 2^-24: with c = m 2^(e - 150), m the 24-bit
// significand and e >= 126 the biased exponent, c 2^24 = m << (e - 126).
__device__ __forceinline__ unsigned long long fixed24(float c)
{
    const unsigned b = __float_as_uint(c);
    return (unsigned long long)((b & 0x7FFFFFu) | 0x800000u) << ((b >> 23) - 126u);
}

// Histogram::add's bin (src/stats.hpp:16): 0.05-point bins over [-500, 500], rounded half away from zero.
__device__ __forceinline__ int hist_bin(double pct)
{
    const double c = pct < -500.0 ? -500.0 : pct > 500.0 ? 500.0 : pct;
    return (int)round((c - -500.0) / 0.05);
}

}  // namespace es
