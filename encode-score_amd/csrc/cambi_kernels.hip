// CAMBI (contrast-aware multiscale banding index) kernels for gfx950.
//
// The per-picture pipeline of Cambi::launch / Cambi::finish (src/cambi.cpp:88-184) as four kernel kinds:
//   cambi_base     scale 0: the spatial mask of the unfiltered picture (src/kernels.cu:9-40) and the 3x3 mode filter
//                  (src/kernels.cu:486-500) by wave strips, each pixel stored as its code (value | mask << 15 less
//                  a per-picture bias, so that a c-value tile indexes its tables with the code itself);
//   cambi_scale    scales 1-4: decimation of the previous packed scale (src/kernels.cu:42-53) and the mode filter;
//   cambi_cvalues  c-values of every scale (src/kernels.cu:129-278): per 64-column tile one wave keeps the window's
//                  column counts of the tile's relevant values in LDS and turns them into window counts with a
//                  wave-wide prefix sum per row; exact sums of the positive c-values and a compact list of them;
//   topk           the k-th largest positive c-value (src/kernels.cu:71-95, src/cambi.cpp:137-170) by a 3-pass
//                  9-bit-digit radix select done on the device, a walk kernel choosing each pass's digit;
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

// Neighbour lanes across the whole wave by DPP (lane i takes lane i + 1, or lane i - 1); the lane with no neighbour
// gets 0 and is replaced by its caller.
__device__ __forceinline__ int from_next_lane(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xF, 0xF, true); }
__device__ __forceinline__ int from_prev_lane(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xF, 0xF, true); }
// The same shifts with the lane that has no neighbour (63, or 0) taking `edge` instead.
__device__ __forceinline__ int from_next_lane_or(int v, int edge) { return __builtin_amdgcn_update_dpp(edge, v, 0x130, 0xF, 0xF, false); }
__device__ __forceinline__ int from_prev_lane_or(int v, int edge) { return __builtin_amdgcn_update_dpp(edge, v, 0x138, 0xF, 0xF, false); }

// mode3 (src/kernels.cu:55-60) of 16-bit values held in 32-bit registers
__device__ __forceinline__ unsigned mode3(unsigned a, unsigned b, unsigned c)
{
    if (a == b || a == c) return a;
    if (b == c) return b;
    return min(a, min(b, c));
}

// ---------------------------------------------------------------------------------------------------------------
// Scale 0: spatial mask + mode filter.  One wave walks down a strip of 64 columns and kBaseRows rows.  Per row each
// lane reads the pixels 32 left and 32 right of its output column (128 columns in all, the strip and its margins);
// two ballots give the flat bits of all of them, and the lane counts the 7 around its column with one popcount of a
// 128-bit window; a running sum over the last 7 rows gives the 7x7 count.  The horizontal 3-modes of the last three
// rows (neighbours by lane shuffles) give the vertical mode.  The mask lags the row read by 3 rows and the mode by 1,
// so the packed value is written 3 rows back.
// ---------------------------------------------------------------------------------------------------------------
constexpr int kBaseW = 64, kBaseRows = 64;

int base_tiles(int w, int h) { return ((w + kBaseW - 1) / kBaseW) * ((h + kBaseRows - 1) / kBaseRows); }

// Readers of the scale-0 samples.  col(x) and row(y) clamp a position into the picture and give its offsets (bytes, or
// the coordinates for the anti-dithering reader); load(col, row) reads the sample there.  `step` advances row() by
// one row.  The plain readers load through a buffer over the plane, so a row loop adds one scalar per row.
struct BaseRead16 {  // 16-bit words holding the samples as they are (10-bit planar)
    __amdgpu_buffer_rsrc_t rs;
    int w, h;
    unsigned step;
    __device__ __forceinline__ unsigned col(int x) const { return 2u * (unsigned)min(max(x, 0), w - 1); }
    __device__ __forceinline__ unsigned row(int y) const { return (unsigned)min(max(y, 0), h - 1) * step; }
    __device__ __forceinline__ int load(unsigned c, unsigned r) const { return (int)__builtin_amdgcn_raw_buffer_load_b16(rs, c, r, 0); }
};
struct BaseRead16S : BaseRead16 {  // 16-bit words shifted: (word >> rshift) << shift (P010's high bits)
    int rshift, shift;
    __device__ __forceinline__ int load(unsigned c, unsigned r) const
    {
        return (int)(((unsigned)__builtin_amdgcn_raw_buffer_load_b16(rs, c, r, 0) >> rshift) << shift);
    }
};
struct BaseRead8 {  // bytes, shifted to the pipeline depth
    __amdgpu_buffer_rsrc_t rs;
    int w, h;
    unsigned step;
    int shift;
    __device__ __forceinline__ unsigned col(int x) const { return (unsigned)min(max(x, 0), w - 1); }
    __device__ __forceinline__ unsigned row(int y) const { return (unsigned)min(max(y, 0), h - 1) * step; }
    __device__ __forceinline__ int load(unsigned c, unsigned r) const { return (int)((unsigned)__builtin_amdgcn_raw_buffer_load_b8(rs, c, r, 0) << shift); }
};
struct BaseReadFrom8 {  // cambi.c's 8-bit anti-dithering mean; offsets are the clamped coordinates
    PlaneIn in;
    int w, h;
    unsigned step;
    __device__ __forceinline__ unsigned col(int x) const { return (unsigned)min(max(x, 0), w - 1); }
    __device__ __forceinline__ unsigned row(int y) const { return (unsigned)min(max(y, 0), h - 1); }
    __device__ __forceinline__ int load(unsigned c, unsigned r) const { return (int)base_value(in, w, h, (int)c, (int)r); }
};

// One strip.  Positions outside the picture enter only as wave masks: margin columns outside it and rows outside it
// have no flat bits, and the mode filter reads no sample outside the picture for any pixel it writes.  An interior
// strip (kInterior: every row it reads lies inside the picture, rows 0, 1, h - 2 and h - 1 excepted) makes no row tests
// and reads its rows at a running offset.  With kLoadOwn the lane's own column is loaded as well (the shifting
// readers); otherwise it is taken from the margin columns' samples by a lane swap, one load fewer per row (the plain
// 10-bit reader and the anti-dithering one).
template <bool kLoadOwn, bool kInterior, typename Rd>
__device__ __forceinline__ void base_strip(const CambiPic &pic, const Rd &rd, int x0, int y0, int y1)
{
    const int w = pic.sw[0], h = pic.sh[0];
    const int lane = threadIdx.x & 63;
    const int xa = x0 - 32 + lane, xb = x0 + 32 + lane, x = x0 + lane;
    const int index = pic.consts->mask_index;
    // margin columns inside the picture, and the picture's last column (flat whatever its right neighbour)
    const unsigned long long in_a = __builtin_amdgcn_ballot_w64(xa >= 0 && xa < w), in_b = __builtin_amdgcn_ballot_w64(xb < w);
    const unsigned long long end_a = __builtin_amdgcn_ballot_w64(xa == w - 1), end_b = __builtin_amdgcn_ballot_w64(xb == w - 1);
    const bool edge = x == 0 || x >= w - 1;  // the picture's first and last columns keep their own value
    // columns x - 3 .. x + 3 are bits sft .. sft + 6 of the flat bits hi:lo of columns x0 - 32 .. x0 + 95
    const int sft = lane + 29;
    const bool low = sft < 64;
    const unsigned va = rd.col(xa), vb = rd.col(xb), vl = rd.col(x);
    // the packed plane as a buffer; a lane right of the picture stores past its end, where the store is dropped
    const __amdgpu_buffer_rsrc_t out = buffer_over(pic.planes, (unsigned long long)w * h * 2);
    const unsigned vo = x < w ? 2u * (unsigned)x : 0x80000000u;
    const unsigned bias = plane_bias(pic.consts->band_base);
    // samples of rows r, r + 1 and r + 2 (loading) in slots (r - y0 + 3) mod 3 of A (columns a), B (columns b) and L
    // (own column); the row loop is unrolled 3 times, so the slots rotate by renaming
    int A[3], B[3], L[3] = {0, 0, 0};
    {
        const unsigned r0 = rd.row(y0 - 3), r1 = rd.row(y0 - 2);
        A[0] = rd.load(va, r0);
        B[0] = rd.load(vb, r0);
        A[1] = rd.load(va, r1);
        B[1] = rd.load(vb, r1);
        if (kLoadOwn) {
            L[0] = rd.load(vl, r0);
            L[1] = rd.load(vl, r1);
        }
    }
    unsigned ring = 0;  // the flat counts of the last 7 rows, 4 bits each, the newest lowest
    int sum = 0;
    int H[3] = {0, 0, 0};  // horizontal modes of rows r - 2 .. r in their rows' slots
    int prev = 0;          // own value of row r - 1
    unsigned md = 0;       // modes of rows r - 3 (low half) and r - 2 (high half), written 3 rows back
    unsigned so = kInterior ? (unsigned)(y0 - 1) * rd.step : 0u;  // interior: offset of row r + 2
    // row r in slot i; stores the packed value of row r - 3 when kStore
    auto step = [&](const int r, const int i, const bool store) __attribute__((always_inline)) {
        const int j = (i + 1) % 3, k = (i + 2) % 3;  // slots of rows r + 1 and r + 2 (= r - 1)
        const int ca = A[i], cb = B[i], na = A[j], nb = B[j];
        const unsigned rr = kInterior ? so : rd.row(r + 2);
        A[k] = rd.load(va, rr);
        B[k] = rd.load(vb, rr);
        int cc = L[i];
        if (kLoadOwn) {
            L[k] = rd.load(vl, rr);
        } else {
            // the own column x0 + lane: column a of lane + 32 (lanes 0-31) or column b of lane - 32 (lanes 32-63)
            const auto sw = __builtin_amdgcn_permlane32_swap((unsigned)ca, (unsigned)cb, false, false);
            cc = (int)(lane < 32 ? sw[1] : sw[0]);
        }
        if (kInterior) so += rd.step;
        // flat bits of row r: equal to the right neighbour (or the last column) and to the one below (or the last
        // row); lane 63's right neighbour in a is lane 0 of b
        const int ra = from_next_lane_or(ca, __builtin_amdgcn_readlane(cb, 0));
        const int rb = from_next_lane(cb);
        const unsigned long long ea = __builtin_amdgcn_ballot_w64(ca == ra), eb = __builtin_amdgcn_ballot_w64(cb == rb);
        const unsigned long long da = __builtin_amdgcn_ballot_w64(ca == na), db = __builtin_amdgcn_ballot_w64(cb == nb);
        unsigned long long lo, hi;
        if (kInterior) {
            lo = in_a & (end_a | ea) & da;
            hi = in_b & (end_b | eb) & db;
        } else {
            lo = hi = 0;
            if (r >= 0 && r < h) {
                lo = in_a & (end_a | ea) & (r == h - 1 ? ~0ull : da);
                hi = in_b & (end_b | eb) & (r == h - 1 ? ~0ull : db);
            }
        }
        // the lane's 7 bits, without branching on the lane: both funnel forms, then a select
        const unsigned w0 = (unsigned)(lo >> (sft & 63)) | (unsigned)(hi << ((64 - sft) & 63));
        const unsigned w1 = (unsigned)(hi >> ((sft - 64) & 63));
        const unsigned count = (unsigned)__builtin_popcount((low ? w0 : w1) & 0x7Fu);
        sum += (int)count - (int)((ring >> 24) & 0xFu);
        ring = ring << 4 | count;
        // horizontal mode of row r at column x; every lane takes part in every shuffle, then the edge lanes take the
        // margin columns' values
        const int lft = from_prev_lane_or(cc, __builtin_amdgcn_readlane(ca, 31));
        const int rgt = from_next_lane_or(cc, __builtin_amdgcn_readlane(cb, 32));
        int hv = (int)mode3(lft, cc, rgt);
        asm("" : "+v"(hv));  // computed on every lane and selected, rather than branched around on the edge lanes
        const int h3 = edge ? cc : hv;
        // the mode of row r - 1 from rows r - 2 (slot j), r - 1 (slot k) and r; its first and last rows are copied
        const int ym = r - 1;
        const int m = !kInterior && (ym == 0 || ym == h - 1) ? prev : (int)mode3(H[j], H[k], h3);
        H[i] = h3;
        prev = cc;
        if (store) {  // the row whose 7x7 count is complete, r - 3; its mode is the low half of md
            const unsigned y = (unsigned)(r - 3);
            __builtin_amdgcn_raw_buffer_store_b16((uint16_t)((md | (sum > index ? 1u : 0u) << 15) - bias), out, vo,
                                                  __builtin_amdgcn_readfirstlane(2u * y * (unsigned)w), 0);
        }
        md = __builtin_amdgcn_alignbit((unsigned)m, md, 16);  // (m : md) >> 16: row r - 2's mode low, row r - 1's high
    };
    // rows y0 - 3 .. y0 + 2 fill the window; then rows y0 + 3 .. y1 + 2 write rows y0 .. y1 - 1, three at a time while
    // three remain; the slot of row r is (r - y0 + 3) mod 3 throughout
    step(y0 - 3, 0, false);
    step(y0 - 2, 1, false);
    step(y0 - 1, 2, false);
    step(y0, 0, false);
    step(y0 + 1, 1, false);
    step(y0 + 2, 2, false);
    int r = y0 + 3;
    for (; r + 2 < y1 + 3; r += 3) {
        step(r, 0, true);
        step(r + 1, 1, true);
        step(r + 2, 2, true);
    }
    if (r < y1 + 3) step(r, 0, true);
    if (r + 1 < y1 + 3) step(r + 1, 1, true);
}

// A strip of the reader's plane: interior strips take the row-test-free form.
template <bool kLoadOwn, typename Rd>
__device__ __forceinline__ void base_strip_any(const CambiPic &pic, const Rd &rd, unsigned bx)
{
    const int w = pic.sw[0], h = pic.sh[0];
    const int strip = (int)bx * 4 + wave_index();
    const int across = (w + kBaseW - 1) / kBaseW;
    if (strip >= across * ((h + kBaseRows - 1) / kBaseRows)) return;  // uniform over the wave
    const int x0 = (strip % across) * kBaseW, y0 = (strip / across) * kBaseRows;
    const int y1 = min(y0 + kBaseRows, h);
    // rows read: y0 - 3 .. y1 + 4, all inside the picture; rows tested for flatness: y0 - 3 .. y1 + 2, none the last;
    // rows filtered: y0 .. y1 - 1, neither the first nor the last
    if (y0 - 3 >= 2 && y1 + 4 <= h - 2) base_strip<kLoadOwn, true>(pic, rd, x0, y0, y1);
    else base_strip<kLoadOwn, false>(pic, rd, x0, y0, y1);
}

__global__ __launch_bounds__(256) void cambi_base_kernel(const CambiPic *pics)
{
    const uint3 b = xcd_block3();  // neighbouring strips share halo columns and rows
    const CambiPic &pic = pics[b.y];
    const int w = pic.sw[0], h = pic.sh[0];
    const PlaneIn in = pic.in;
    // one reader per plane format, chosen once per block
    if (in.from8) {
        base_strip_any<false>(pic, BaseReadFrom8{in, w, h, 1u}, b.x);
    } else if (in.bytes == 1) {
        base_strip_any<true>(pic, BaseRead8{buffer_over(in.ptr, (unsigned long long)in.pitch * h), w, h, (unsigned)in.pitch, in.shift}, b.x);
    } else {
        const BaseRead16 r16{buffer_over(in.ptr, (unsigned long long)in.pitch * h * 2), w, h, 2u * (unsigned)in.pitch};
        if (in.rshift == 0 && in.shift == 0) base_strip_any<false>(pic, r16, b.x);  // 10-bit samples, the common case
        else base_strip_any<true>(pic, BaseRead16S{r16, in.rshift, in.shift}, b.x);
    }
}

void launch_cambi_base(const CambiPic *pics, const int * /*tile_first*/, int npics, int max_tiles, hipStream_t st)
{
    if (npics <= 0) return;
    hipLaunchKernelGGL(cambi_base_kernel, dim3((max_tiles + 3) / 4, npics), dim3(256), 0, st, pics);
    hip_check(hipGetLastError(), "cambi base");
}

// ---------------------------------------------------------------------------------------------------------------
// Scales 1-4: decimate + mode filter of the previous packed scale; the mask bit comes along from (2x, 2y).  One wave
// per strip of 64 columns and kScaleRows rows: per row each lane gathers the decimated sample of its column (and the
// edge lanes the one beyond the strip), the horizontal 3-mode takes its neighbours by lane shuffles, and the vertical
// 3-mode of the row above is formed from the last three rows.  The next row's samples are loaded a row ahead.
// ---------------------------------------------------------------------------------------------------------------
constexpr int kScaleRows = 32;

int scale_strips(int w, int h) { return ((w + 63) / 64) * ((h + kScaleRows - 1) / kScaleRows); }

__global__ __launch_bounds__(256) void cambi_scale_kernel(const CambiPic *pics, int s)
{
    const uint3 b = xcd_block3();  // neighbouring strips share halo rows
    const CambiPic &pic = pics[b.y];
    const int w = pic.sw[s], h = pic.sh[s], pw = pic.sw[s - 1];
    const int lane = threadIdx.x & 63, strip = (int)b.x * 4 + wave_index();
    const int across = (w + 63) / 64;
    if (strip >= across * ((h + kScaleRows - 1) / kScaleRows)) return;  // uniform over the wave
    const int x0 = (strip % across) * 64, y0 = (strip / across) * kScaleRows, y1 = min(y0 + kScaleRows, h);
    const int x = x0 + lane;
    // the previous scale's packed plane and this scale's as buffers: rows at a scalar offset, and a lane right of the
    // picture stores past the end, where the store is dropped
    const __amdgpu_buffer_rsrc_t in = buffer_over(pic.planes + pic.off[s - 1], (unsigned long long)pw * pic.sh[s - 1] * 2);
    const __amdgpu_buffer_rsrc_t out = buffer_over(pic.planes + pic.off[s], (unsigned long long)w * h * 2);
    // this lane's column and its halo column (lane 0: the column left of the strip, lane 63: the one right of it), as
    // byte offsets of the decimated samples (2x, 2y) of the previous scale
    const unsigned vc = 4u * (unsigned)min(x, w - 1), vh = 4u * (unsigned)(lane == 0 ? max(x0 - 1, 0) : min(x0 + 64, w - 1));
    const unsigned vo = x < w ? 2u * (unsigned)x : 0x80000000u;
    // samples as value | mask << 15 (the codes with the bias added back)
    const unsigned bias = plane_bias(pic.consts->band_base);
    auto ld = [&](unsigned v, int r) {
        return ((unsigned)__builtin_amdgcn_raw_buffer_load_b16(in, v, (unsigned)(2 * min(max(r, 0), h - 1)) * (unsigned)pw * 2u, 0) + bias) & 0xFFFFu;
    };
    // the horizontal 3-mode of a row from its own and halo samples; the picture's first and last columns keep theirs
    auto hmode = [&](unsigned c, unsigned hv) -> unsigned {
        const unsigned v = c & 0x7FFFu, e = hv & 0x7FFFu;
        const unsigned left = (unsigned)from_prev_lane((int)v), right = (unsigned)from_next_lane((int)v);
        const unsigned l = lane == 0 ? e : left, rt = lane == 63 ? e : right;
        return (x == 0 || x >= w - 1) ? v : mode3(l, v, rt);
    };
    // rows y - 1 and y: t = horizontal modes, c = raw sample (value and mask bit); rows y + 1 .. y + 3 in flight in
    // slots (row - y0 - 1) mod 3, each loaded three rows before it is used (the loop is unrolled 3 times)
    unsigned C[3], E[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        C[k] = ld(vc, y0 + 1 + k);
        E[k] = ld(vh, y0 + 1 + k);
    }
    unsigned c1 = ld(vc, y0);
    unsigned t0 = hmode(ld(vc, y0 - 1), ld(vh, y0 - 1)), t1 = hmode(c1, ld(vh, y0));
    auto step = [&](const int y, const int i) __attribute__((always_inline)) {
        const unsigned c2 = C[i], e2 = E[i];
        C[i] = ld(vc, y + 4);
        E[i] = ld(vh, y + 4);
        const unsigned t2 = hmode(c2, e2);
        const unsigned m = (y == 0 || y == h - 1) ? (c1 & 0x7FFFu) : mode3(t0, t1, t2);
        __builtin_amdgcn_raw_buffer_store_b16((uint16_t)((m | (c1 & 0x8000u)) - bias), out, vo,
                                              __builtin_amdgcn_readfirstlane(2u * (unsigned)y * (unsigned)w), 0);
        t0 = t1;
        t1 = t2;
        c1 = c2;
    };
    int y = y0;
    for (; y + 2 < y1; y += 3) {
        step(y, 0);
        step(y + 1, 1);
        step(y + 2, 2);
    }
    if (y < y1) step(y, 0);
    if (y + 1 < y1) step(y + 1, 1);
}

void launch_cambi_scale(const CambiPic *pics, int npics, int max_strips, int s, hipStream_t st)
{
    if (npics <= 0) return;
    hipLaunchKernelGGL(cambi_scale_kernel, dim3((unsigned)((max_strips + 3) / 4), npics), dim3(256), 0, st, pics, s);
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
// the vertical window around the current row (u8, 4 per word, in LDS); lane i owns columns 2i and 2i + 1 and updates
// them as rows enter and leave the window.  Per row, every lane reads all its count words at once, then the wave
// runs one prefix sum per pair of numbers (16 bits each) over the column pairs, the scans of all pairs interleaved
// step by step; the prefixes P(c) of every column go to LDS and a served pixel at output column x reads the window
// count of each number it needs as P(x + span) - P(x).  The codes of the next row are loaded while the current row is
// processed, the setup's loads go out 32 at a time (a lane's samples of 32 rows, or of 16 rows of both its columns),
// and the c-value arithmetic is predicated rather than branched.  Positive c-values go to the tile's own segment of
// the list at a wave-uniform offset.
// ---------------------------------------------------------------------------------------------------------------
constexpr int kCvFlags = 592;   // >= kMaxBand + 2 * kDiffs, multiple of 16
constexpr unsigned kCvDummy = kCvFlags - 1;  // the flag an inactive sample sets: past every index a tile reads
constexpr int kCvCols = 130;    // prefix entries per number: columns 0 .. 128, padded to an even count

// LDS of one wave for passes of up to BMAX numbers: 14.5 KB for 32, 26.7 KB for 64 (fewer passes over tiles with
// many values, fewer waves per CU).
template <int BMAX>
struct CvShared {
    uint32_t cc[BMAX / 4][128];               // column counts: numbers 4q .. 4q + 3 of column c in cc[q][c], u8 each
    uint16_t pp[BMAX][kCvCols];               // prefix over columns of the column counts of number b: pp[b][c]
    uint16_t rank[kMaxBand + 2 * kDiffs];     // number of value v at rank[v + 4], kNone if not relevant
};
static_assert(sizeof(uint16_t) * 32 * kCvCols >= 2 * kCvFlags, "flags live in pp during the setup");
// Column counts are at most 2 pad + 1 <= 65 (u8), and a prefix over 128 columns at most 128 * 65 < 2^16 (u16).
static_assert(2 * kMaxPad + 1 < 256 && 128 * (2 * kMaxPad + 1) < 65536, "count widths");

int cvalue_tiles(int w, int h, int tile_h) { return ((w + kCvTileW - 1) / kCvTileW) * ((h + tile_h - 1) / tile_h); }

struct CvTile {
    __amdgpu_buffer_rsrc_t rs;  // the scale's packed plane as a buffer of w * h samples
    int w, h, x0, y0, rows;
    int pad, band_base, band_size, vlt;
    int tvi[kDiffs], weight[kDiffs];
    // the thresholds of difference d as one unsigned range test on a served sample's index i (its code, value less
    // band_base plus 5): value + 4 <= tvi[d] and value + 4 + d + 1 > vlt exactly when i - oklo[d] <= okspan[d]; an
    // empty range has
    // oklo = 2^31 and okspan = 0, which no index below 2^31 passes
    unsigned oklo[kDiffs], okspan[kDiffs];
    int keep;   // the full c-value plane is kept
    int pairs;  // an even width on a 4-byte aligned plane: every row's codes at even columns are 4-byte aligned
    // the full c-value plane (when kept; else an empty buffer) and this tile's segment of the list, as buffers
    __amdgpu_buffer_rsrc_t plane_rs, seg_rs;
};

// Samples are read as buffer loads at byte offset xo + yo: a load past the buffer's w * h samples returns 0, which is
// an inactive sample (mask bit clear), so positions outside the picture need no test and no lane mask.  A column
// outside the picture gets xo = 2^31, and a row outside it a yo of at least 2 w h (below) or of 2^32 - 2 w |y| with
// |y| <= 2 kMaxPad + 1 (above); every sum with such a term lies at or past 2 w h for pictures under 2^30 bytes.
constexpr unsigned kCvOut = 0x80000000u;
__device__ __forceinline__ unsigned cv_col(const CvTile &t, int x) { return (unsigned)x < (unsigned)t.w ? 2u * (unsigned)x : kCvOut; }
__device__ __forceinline__ unsigned cv_at(const CvTile &t, unsigned xo, int y)
{
    return __builtin_amdgcn_raw_buffer_load_b16(t.rs, xo + (unsigned)(y * t.w * 2), 0, 0);
}

// A sample's code less kDiffs + 1: the value less band_base when the sample is masked and not below the band; an
// unmasked sample, a masked one below the band or the 0 read outside the picture wraps far above any band size.  So the
// sample is active exactly when the result is below band_size.
__device__ __forceinline__ unsigned cv_off(const CvTile &, unsigned v) { return v - (unsigned)(kDiffs + 1); }

// Index into the rank table of a packed sample: value - band_base + kDiffs when it is active, else 0 (whose entry is
// kNone: no value below the band is relevant).
__device__ __forceinline__ unsigned cv_index(const CvTile &t, unsigned v)
{
    const unsigned u = cv_off(t, v);
    return u < (unsigned)t.band_size ? u + kDiffs : 0u;
}

// RN(1/n) for 1 <= n <= 8192, the entry of the reference's reciprocal table: the hardware reciprocal (within 1 ulp)
// refined by one Newton step with fused multiply-adds.  The residual e = 1 - n r0 is exact, and r0 + e r0 differs from
// 1/n by r0 e^2 / (1 - e), below 2^-44 relative, while 1/n lies at least 2^-37 relative from any rounding boundary
// (n < 2^13), so the rounded step is 1/n correctly rounded.
__device__ __forceinline__ float recip_rn(unsigned n)
{
    const float x = (float)n;
    const float r0 = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, r0, 1.0f);
    return __builtin_fmaf(e, r0, r0);
}

// The value less band_base of an active (masked, in-band) packed sample, else kNone.
__device__ __forceinline__ unsigned cv_decode(const CvTile &t, unsigned v)
{
    const unsigned u = cv_off(t, v);
    return u < (unsigned)t.band_size ? u : kNone;
}

template <int B, typename Sh>
__device__ void cv_passes(Sh &sh, const CvTile &t, int relevant, double &sum_c, unsigned &tcount)
{
    constexpr int Q = B / 4, NC = B / 8;  // count words per column; chunks of 8 numbers (2 words, 4 pairs)
    const int lane = threadIdx.x, pad = t.pad, span = 2 * pad + 1;
    const int c0 = 2 * lane, xl = t.x0 - pad + c0;  // this lane's margin columns c0, c0 + 1 at picture column xl
    const bool col0 = c0 < kCvTileW + 2 * pad, col1 = c0 + 1 < kCvTileW + 2 * pad;
    const int x = t.x0 + lane;  // this lane's output column
    // byte offsets of this lane's margin columns and output column (kCvOut: none, or outside the picture)
    const unsigned xo0 = col0 ? cv_col(t, xl) : kCvOut, xo1 = col1 ? cv_col(t, xl + 1) : kCvOut, xoo = cv_col(t, x);
    for (int b0 = 0;;) {
        const int lo = b0 ? b0 + kDiffs : 0, hi = b0 + B >= relevant ? relevant : b0 + B - kDiffs;
#pragma unroll
        for (int q = 0; q < Q; q++) *(uint2 *)&sh.cc[q][c0] = make_uint2(0, 0);
        // A sample of rank rk entering (delta 1) or leaving (delta ~0) column c's count: an increment of zero when
        // it is inactive or outside the pass, so that every lane makes the same LDS operation.
        auto bump = [&](unsigned rk, int c, unsigned delta) {
            const unsigned b = rk - (unsigned)b0;
            const unsigned inc = b < (unsigned)B ? delta << ((b & 3u) << 3) : 0u;
            atomicAdd(&sh.cc[(b >> 2) & (Q - 1)][c], inc);
        };
        // the window rows above the first row, eight at a time
        const int ya = max(t.y0 - pad, 0), yb = min(t.y0 + pad, t.h);
        for (int y = ya; y < yb; y += 8) {
            unsigned v[16];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int yy = y + j < yb ? y + j : -1;
                v[2 * j] = cv_at(t, xo0, yy);
                v[2 * j + 1] = cv_at(t, xo1, yy);
            }
            unsigned rk[16];
#pragma unroll
            for (int j = 0; j < 16; j++) rk[j] = sh.rank[cv_index(t, v[j])];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                bump(rk[2 * j], c0, 1u);
                bump(rk[2 * j + 1], c0 + 1, 1u);
            }
        }
        // codes of the row: entering (y + pad), own (y), leaving (y - pad); loaded a row ahead.  The leaving samples
        // of row y are removed at the start of row y + 1, with that row's entering ones.
        auto load = [&](int y, unsigned *v) {
            v[0] = cv_at(t, xo0, y + pad);
            v[1] = cv_at(t, xo1, y + pad);
            v[2] = cv_at(t, xoo, y);
            v[3] = cv_at(t, xo0, y - pad);
            v[4] = cv_at(t, xo1, y - pad);
        };
        unsigned cur[5], nxt[5] = {0, 0, 0, 0, 0}, gone0 = 0, gone1 = 0;
        load(t.y0, cur);
        for (int r = 0; r < t.rows; r++) {
            const int y = t.y0 + r;
            if (r + 1 < t.rows) load(y + 1, nxt);
            // the ranks of the samples entering and leaving and of the own sample, read at once; those of the values
            // within kDiffs of it only for a row the pass serves
            const unsigned io = cv_index(t, cur[2]), base = io ? io - kDiffs : 0u;
            const unsigned re0 = sh.rank[cv_index(t, cur[0])], re1 = sh.rank[cv_index(t, cur[1])];
            const unsigned rl0 = sh.rank[cv_index(t, gone0)], rl1 = sh.rank[cv_index(t, gone1)];
            const unsigned rown = sh.rank[base + kDiffs];
            bump(rl0, c0, ~0u);
            bump(rl1, c0 + 1, ~0u);
            bump(re0, c0, 1u);
            bump(re1, c0 + 1, 1u);
            const unsigned own = io ? rown : (unsigned)kNone;
            const bool serve = own != kNone && (int)own >= lo && (int)own < hi;
            if (__builtin_amdgcn_ballot_w64(serve)) {
                unsigned rk[2 * kDiffs + 1];
#pragma unroll
                for (int k = 0; k < 2 * kDiffs + 1; k++) rk[k] = k == kDiffs ? rown : sh.rank[base + k];
                // the chunks of 8 numbers some served pixel reads: its own number and those within 4 of it
                const int rel = (int)own - b0;
                unsigned need = 1;
                if (NC > 1) {
                    need = 0;
#pragma unroll
                    for (int k = 0; k < NC; k++)
                        need |= __builtin_amdgcn_ballot_w64(serve && rel + kDiffs >= 8 * k && rel - kDiffs < 8 * k + 8) ? 1u << k : 0u;
                }
                uint2 cw[Q];
#pragma unroll
                for (int q = 0; q < Q; q++) cw[q] = *(const uint2 *)&sh.cc[q][c0];
#pragma unroll
                for (int k = 0; k < NC; k++) {
                    if (!(need >> k & 1u)) continue;
                    // column counts of both columns, numbers 8k .. 8k + 7 widened to 16 bits per number: pair j
                    // holds numbers 8k + 2j (low half) and 8k + 2j + 1 (high half)
                    const uint2 ca = cw[2 * k], cb = cw[2 * k + 1];
                    uint32_t xs[4], e[4], s[4];
                    e[0] = __builtin_amdgcn_perm(0u, ca.x, 0x0C010C00u);
                    e[1] = __builtin_amdgcn_perm(0u, ca.x, 0x0C030C02u);
                    e[2] = __builtin_amdgcn_perm(0u, cb.x, 0x0C010C00u);
                    e[3] = __builtin_amdgcn_perm(0u, cb.x, 0x0C030C02u);
                    xs[0] = e[0] + __builtin_amdgcn_perm(0u, ca.y, 0x0C010C00u);
                    xs[1] = e[1] + __builtin_amdgcn_perm(0u, ca.y, 0x0C030C02u);
                    xs[2] = e[2] + __builtin_amdgcn_perm(0u, cb.y, 0x0C010C00u);
                    xs[3] = e[3] + __builtin_amdgcn_perm(0u, cb.y, 0x0C030C02u);
#pragma unroll
                    for (int j = 0; j < 4; j++) s[j] = xs[j];
#define ES_SCAN_STEP(ctrl, rmask)                                                                        \
    _Pragma("unroll") for (int j = 0; j < 4; j++) s[j] += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)s[j], ctrl, rmask, 0xF, false);
                    ES_SCAN_STEP(0x111, 0xF)
                    ES_SCAN_STEP(0x112, 0xF)
                    ES_SCAN_STEP(0x114, 0xF)
                    ES_SCAN_STEP(0x118, 0xF)
                    ES_SCAN_STEP(0x142, 0xA)
                    ES_SCAN_STEP(0x143, 0xC)
#undef ES_SCAN_STEP
                    // P(c0) = s - x, P(c0 + 1) = P(c0) + count(c0), P(c0 + 2) = s (the next lane's P(c0) as well,
                    // written by both with the same value), split into one u16 row per number
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const uint32_t p = s[j] - xs[j], q = p + e[j];
                        const int n = 8 * k + 2 * j;
                        *(uint32_t *)&sh.pp[n][c0] = __builtin_amdgcn_perm(q, p, 0x05040100u);
                        *(uint32_t *)&sh.pp[n + 1][c0] = __builtin_amdgcn_perm(q, p, 0x07060302u);
                        sh.pp[n][c0 + 2] = (uint16_t)s[j];
                        sh.pp[n + 1][c0 + 2] = (uint16_t)(s[j] >> 16);
                    }
                }
                __syncthreads();  // the prefixes of all columns are in before any lane reads them
                // window counts of p0 and of the neighbours at +-(d + 1): P(x + span) - P(x) in margin columns; a
                // number outside the pass reads a valid entry and counts 0
                const uint16_t *p1 = &sh.pp[0][0] + lane, *p2 = p1 + span;
                unsigned cnt[2 * kDiffs + 1];
#pragma unroll
                for (int k = 0; k < 2 * kDiffs + 1; k++) {
                    const unsigned b = rk[k] - (unsigned)b0, bb = b & (B - 1);
                    const unsigned d = (unsigned)p2[bb * kCvCols] - (unsigned)p1[bb * kCvCols];
                    cnt[k] = b < (unsigned)B ? d : 0u;
                }
                // every count is needed below on some lanes: have all of them read here, together, rather than each
                // under its own condition
#pragma unroll
                for (int k = 0; k < 2 * kDiffs + 1; k++) asm volatile("" ::"v"(cnt[k]));
                const int value = (int)(io ? io - kDiffs : 0u) + t.band_base + kDiffs;
                const unsigned p0 = cnt[kDiffs];
                float c = 0.0f;
#pragma unroll
                for (int d = 0; d < kDiffs; d++) {
                    const unsigned pa = cnt[kDiffs + d + 1], pb = cnt[kDiffs - d - 1];
                    const unsigned pm = pa > pb ? pa : pb;
                    // w * p0 * pm * RN(1 / (p0 + pm)), the reciprocal as the reference's table holds it; computed on
                    // every lane and selected, the tests combined without short-circuiting.  On a served pixel
                    // p0 + pm is at least 1 (its own count) and pm = 0 gives 0, which never exceeds c, so the
                    // reference's pm > 0 test needs no instruction; elsewhere a NaN (0 / 0) exceeds nothing either
                    const float rcp = recip_rn(p0 + pm);
                    float val = __fmul_rn((float)__umul24(__umul24((unsigned)t.weight[d], p0), pm), rcp);
                    asm("" : "+v"(val));
                    const bool take = (value <= t.tvi[d]) & (value + d + 1 > t.vlt) & (val > c);
                    c = take ? val : c;
                }
                c = serve ? c : 0.0f;
                __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(c), t.plane_rs, serve ? 4u * (unsigned)x : 0x80000000u,
                                                      __builtin_amdgcn_readfirstlane(4u * (unsigned)y * (unsigned)t.w), 0);
                sum_c += (double)c;  // exact: see cv_tile
                const unsigned long long m = __builtin_amdgcn_ballot_w64(c > 0.0f);
                const unsigned below = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(c), t.seg_rs, c > 0.0f ? 4u * below : 0x80000000u,
                                                      __builtin_amdgcn_readfirstlane(4u * tcount), 0);
                tcount += (unsigned)__popcll(m);
                __syncthreads();  // every lane has read the prefixes before the next row writes them
            }
            gone0 = cur[3];
            gone1 = cur[4];
#pragma unroll
            for (int k = 0; k < 5; k++) cur[k] = nxt[k];
        }
        if (hi == relevant) break;
        b0 = hi - kDiffs;
    }
}

// ---------------------------------------------------------------------------------------------------------------
// c-values on the matrix cores.  The same tiles, numbering and passes; per row the window count of every number of
// the pass at every output column comes from one banded matrix product instead of prefix scans:
//   W[n][x] = sum over margin columns c of C[n][c] A[c][x],   A[c][x] = [x <= c <= x + 2 pad],
// with C the column counts (u8, at most 2 pad + 1) and W at most (2 pad + 1)^2, exact in i32.  The 64 output columns
// are two blocks of 32 and the margin columns (64 + 2 pad <= 128) four blocks of 32; a block product is one
// v_mfma_i32_32x32x32_i8 for 32 numbers, and only the blocks the band touches are multiplied (at most 3 per output
// block).  The counts are the A operand, read 16 bytes at a time from a [number][column] byte table; a lane holds the
// band bytes of its output column over 16 margin columns of each block (the B operand), built once per tile.  The
// lane layout of the K dimension is the same for both operands (lane / 32 picks the 16-column half of the block), so
// the sum over K does not depend on the hardware's order of it.  The product leaves each lane the counts of 16
// numbers at one output column; they go to a [column][number] u16 table, one row per output column, from which the
// served pixel of that column reads the counts of its numbers.
// ---------------------------------------------------------------------------------------------------------------
typedef int cv_v4i __attribute__((ext_vector_type(4)));
typedef int cv_v16i __attribute__((ext_vector_type(16)));
constexpr int kMCnt = 144;  // bytes per number in the count table: 128 columns, rows 16-byte aligned on spread banks
constexpr int kMChunks = (kMaxBand + 2 * kDiffs + 63) / 64;  // 64-value chunks of the relevance masks

// u16 entries per output column in the window table: the B numbers, a zero at B and one of spread, so that a row is
// an odd number of 4-byte words (17 for B = 32): the 32 columns a half-wave writes or reads at one number fall in 32
// distinct banks.
template <int NB> constexpr int kMWinT = 32 * NB + 2;

template <int NB>
struct CvSharedM {
    // column counts of number n at margin column c; row B takes the updates of samples outside the pass
    alignas(16) uint8_t cnt[32 * NB + 1][kMCnt];
    uint16_t win[kCvTileW][kMWinT<NB>];  // window counts at output column x of number n; entry B: zero
    // twice the pass's row of each value (twice its number less b0, or 2 B outside the pass or not relevant) at
    // row[4 + i], i the value's index (its code clamped to nv + 1: value less band_base plus 5 for a masked in-band
    // value), with 2 B at 0 .. 4 and from 5 + nv on, so that every index from 0 to nv + 1 and the 4 either side of it
    // are entries, and the three 4-byte words from row[i & ~3] hold row[i .. i + 8]
    alignas(4) uint8_t row[kMaxBand + 6 * kDiffs + 4];
    unsigned long long relevant[kMChunks];  // bit i of chunk j: value 64 j + i is relevant (numbered in value order)
    __device__ uint8_t *flags() { return (uint8_t *)&win[0][0]; }
};
static_assert(sizeof(uint16_t) * kCvTileW * kMWinT<1> >= 2 * kCvFlags, "flags live in the window table during the setup");
static_assert(sizeof(CvSharedM<1>) <= 10240, "16 waves of one CU share its 160 KB of LDS");
static_assert(2 * 32 * 2 < 256, "twice a row fits a byte");

// The 16 band bytes [xr <= c <= xr + span] over margin columns c = cb .. cb + 15, one byte per column.
__device__ __forceinline__ cv_v4i band_bytes(int xr, int cb, int span)
{
    const int lo = min(max(xr - cb, 0), 16), hi = min(max(xr + span - cb + 1, 0), 16);
    const unsigned m = ((1u << hi) - 1u) & ~((1u << lo) - 1u);  // bit t: column cb + t in the band
    cv_v4i v;
#pragma unroll
    for (int d = 0; d < 4; d++) v[d] = (int)((((m >> (4 * d)) & 0xFu) * 0x00204081u) & 0x01010101u);  // bit i -> byte i
    return v;
}

// Passes of NBU blocks of 32 numbers (NBU <= NB, the blocks the LDS layout holds); the band of every output block
// touches NJ margin blocks.  k16: one pass of 16 numbers, for a tile with at most 16 relevant values: the products are
// v_mfma_i32_16x16x64_i8 over output blocks of 16 columns and margin blocks of 64 (half the operand reads and product
// stores of a 32-number pass), and the window table is laid out as rows of 18 entries per output column.  Per pass each value's row in the tables is set once (row[], doubled): its number less
// b0, or B when it is outside the pass or not relevant.  Window counts of such values read the zero entry B of the
// pixel's column, and their count updates go to row B of cnt, which nothing reads.  A sample's index is its code
// clamped to nv + 1 (one instruction): value - band_base + 5 for a masked sample in the band, and every index outside
// 5 .. nv - 4 has row B (0 for a sample outside the picture).
template <int NB, int NBU, int NJ, bool kPairs, bool k16>
__device__ void cv_passes_m(CvSharedM<NB> &sh, const CvTile &t, int relevant, double &sum_c, unsigned &tcount)
{
    static_assert(NBU >= 1 && NBU <= NB && NJ >= 1 && NJ <= 3 && (!k16 || NBU == 1), "blocks of numbers, band blocks");
    constexpr int B = k16 ? 16 : 32 * NBU;
    // k16: u16 entries per output column in the window table (16 numbers, a zero at 16 and one of spread: 9 words, an
    // odd stride) and the words between output blocks of 16 columns
    constexpr int kW16 = 18, kOb16 = 16 * kW16 / 2;
    const int lane = threadIdx.x, pad = t.pad, span = 2 * pad, cwid = kCvTileW + 2 * pad;
    // this lane's two margin columns c0 and c0 + 1: 2 lane and 2 lane + 1, or with kPairs the 4-byte aligned pair of
    // picture columns, 2 lane - sh and 2 lane + 1 - sh (sh = 1 when the window starts at an odd column), read as one
    // load.  A column outside the window (-1, or cwid and up) is read but never counted: its increment is 0.
    const int sh_ = kPairs ? ((t.x0 - pad) & 1) : 0;
    const int c0 = 2 * lane - sh_, xl = t.x0 - pad + c0;
    const bool col0 = c0 >= 0 && c0 < cwid, col1 = c0 + 1 < cwid;
    const int x = t.x0 + lane;
    const unsigned xo0 = col0 ? cv_col(t, xl) : kCvOut, xo1 = col1 ? cv_col(t, xl + 1) : kCvOut, xoo = cv_col(t, x);
    // kPairs: the pair is inside or outside the picture as a whole (picture columns 2k, 2k + 1)
    const unsigned xop = xl >= 0 && xl < t.w ? 2u * (unsigned)xl : kCvOut;
    // band operands of output block mb and margin block mb + j: this lane's output column 32 mb + lane % 32 and its
    // half of the block's margin columns; k16: of output block ob and margin block kb, output column 16 ob + lane % 16
    // and its quarter of the block's margin columns 64 kb ..
    const int g = lane >> 5, r32 = lane & 31, g4 = lane >> 4, r16 = lane & 15;
    cv_v4i band[k16 ? 4 : 2][k16 ? 2 : NJ];
    if constexpr (k16) {
#pragma unroll
        for (int ob = 0; ob < 4; ob++)
#pragma unroll
            for (int kb = 0; kb < 2; kb++) band[ob][kb] = band_bytes(16 * ob + r16, 64 * kb + 16 * g4, span);
    } else {
#pragma unroll
        for (int mb = 0; mb < 2; mb++)
#pragma unroll
            for (int j = 0; j < NJ; j++) band[mb][j] = band_bytes(32 * mb + r32, 32 * (mb + j) + 16 * g, span);
    }
    // the count words of the lane's two margin columns in every number's row (the same word unless sh = 1): their
    // byte offsets in the row, and the increments of the columns' bytes (a decrement is the negated increment; 0 for
    // a column outside the window)
    const unsigned cword0 = (unsigned)(max(c0, 0) >> 2) * 4u, cword1 = (unsigned)((c0 + 1) >> 2) * 4u;
    const unsigned inc0 = col0 ? 1u << ((c0 & 3) * 8) : 0u, inc1 = col1 ? 1u << (((c0 + 1) & 3) * 8) : 0u;
    uint8_t *const cnt = &sh.cnt[0][0];
    // a doubled row r2 = 2 row: the row's bytes start at r2 * kMCnt / 2
    auto bump0 = [&](unsigned r2, unsigned inc) { atomicAdd((uint32_t *)(cnt + r2 * (kMCnt / 2) + cword0), inc); };
    auto bump1 = [&](unsigned r2, unsigned inc) { atomicAdd((uint32_t *)(cnt + r2 * (kMCnt / 2) + cword1), inc); };
    // the codes of the lane's two margin columns in row y
    auto ld2 = [&](int y, unsigned &a, unsigned &b) __attribute__((always_inline)) {
        if (kPairs) {
            const unsigned p = __builtin_amdgcn_raw_buffer_load_b32(t.rs, xop + (unsigned)(y * t.w * 2), 0, 0);
            a = p & 0xFFFFu;
            b = p >> 16;
        } else {
            a = cv_at(t, xo0, y);
            b = cv_at(t, xo1, y);
        }
    };
    // this lane's start in the count table (A operand: number r32, its 16 margin columns of a block; k16: number r16,
    // its 16 margin columns of a block of 64)
    const uint8_t *const crow = k16 ? &sh.cnt[r16][16 * g4] : &sh.cnt[r32][16 * g];
    // the window table's words of the products this lane writes (output column 32 mb + r32, numbers 4 g ..; k16:
    // output columns 16 ob + r16 for ob 0, 1 and 2, 3, numbers 4 g4 ..), as one LDS address per output block (k16: per
    // pair of them) held in a register, so that every store is that address plus an immediate offset
    typedef __attribute__((address_space(3))) uint32_t lds_word;
    uint16_t *const w16 = &sh.win[0][0];  // k16: the table as rows of kW16 entries
    lds_word *wcol[2] = {(lds_word *)(k16 ? w16 + r16 * kW16 + 4 * g4 : &sh.win[r32][4 * g]),
                         (lds_word *)(k16 ? w16 + (32 + r16) * kW16 + 4 * g4 : &sh.win[32 + r32][4 * g])};
    asm volatile("" : "+v"(wcol[0]), "+v"(wcol[1]));
    // the row of this lane's own output column, read by its served pixel at byte offset 2 n for number n
    typedef __attribute__((address_space(3))) const uint8_t lds_byte;
    lds_byte *const wmine = (lds_byte *)(k16 ? w16 + lane * kW16 : &sh.win[lane][0]);
    const unsigned nv = (unsigned)(t.band_size + 2 * kDiffs);
    auto index = [&](unsigned v) { return min(v, nv + 1); };
    const uint8_t *const rows = &sh.row[kDiffs];  // rows[i]: the doubled row of index i (i from -4 to nv + 5)
    for (int b0 = 0;;) {
        const int lo = b0 ? b0 + kDiffs : 0, hi = b0 + B >= relevant ? relevant : b0 + B - kDiffs;
        for (int i = lane; i < B * kMCnt / 16; i += 64) ((cv_v4i *)cnt)[i] = cv_v4i{0, 0, 0, 0};
        // the zero entry B of every column (the setup's flags lived here)
        ((uint32_t *)(k16 ? w16 + lane * kW16 + B : &sh.win[lane][B]))[0] = 0u;
        // the rows of the pass: a relevant value's number is the count of relevant values below it
        for (int j = 0, below = 0; j * 64 < (int)nv; j++) {
            const unsigned long long m = sh.relevant[j];
            const unsigned n = (unsigned)below + __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
            const int i = j * 64 + lane;
            if (i < (int)nv) sh.row[kDiffs + 1 + i] = (uint8_t)(2u * ((m >> lane & 1ull) ? min(n - (unsigned)b0, (unsigned)B) : (unsigned)B));
            below += __popcll(m);
        }
        if (lane <= kDiffs) sh.row[lane] = (uint8_t)(2 * B);
        if (lane < 2 * kDiffs + 1) sh.row[kDiffs + 1 + nv + lane] = (uint8_t)(2 * B);
        __syncthreads();  // one wave: orders the table writes before the reads below
        const int ya = max(t.y0 - pad, 0), yb = min(t.y0 + pad, t.h);
        for (int y = ya; y < yb; y += 8) {
            unsigned v[16];
#pragma unroll
            for (int j = 0; j < 8; j++) ld2(y + j < yb ? y + j : -1, v[2 * j], v[2 * j + 1]);
            unsigned rw[16];
#pragma unroll
            for (int j = 0; j < 16; j++) rw[j] = rows[index(v[j])];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                bump0(rw[2 * j], inc0);
                bump1(rw[2 * j + 1], inc1);
            }
        }
        // codes of a row: entering (y + pad), own (y), leaving (y - pad); loaded a row ahead into the other of two
        // arrays (the row loop is unrolled twice, so they swap roles by renaming).  The leaving samples of row y are
        // removed at the start of row y + 1, with that row's entering ones.
        auto load = [&](int y, unsigned *v) {
            ld2(y + pad, v[0], v[1]);
            v[2] = cv_at(t, xoo, y);
            ld2(y - pad, v[3], v[4]);
        };
        const unsigned lo2 = 2u * (unsigned)(lo - b0), hi2 = 2u * (unsigned)(hi - b0);
        auto step = [&](const int r, const unsigned *cur, unsigned *nxt, unsigned gone0, unsigned gone1) __attribute__((always_inline)) {
            const int y = t.y0 + r;
            const unsigned io = index(cur[2]);
            const unsigned re0 = rows[index(cur[0])], re1 = rows[index(cur[1])];
            const unsigned rl0 = rows[index(gone0)], rl1 = rows[index(gone1)];
            const unsigned own = rows[io];
            load(y + 1, nxt);  // past the tile's last row it reads unused samples (past the picture, zeros)
            bump0(rl0, 0u - inc0);
            bump1(rl1, 0u - inc1);
            bump0(re0, inc0);
            bump1(re1, inc1);
            const bool serve = own >= lo2 && own < hi2;
            const unsigned long long sm = __builtin_amdgcn_ballot_w64(serve);
            if (sm) {
                // the LDS operations of one wave complete in order: the row's updates are in before these reads
                asm volatile("" ::: "memory");
                // the doubled rows of the own value and the 4 values either side, rows[io - 4 .. io + 4] =
                // row[io .. io + 8], as three aligned words, read only for a row the pass serves, after the products
                // are issued (they are not needed until after the products are stored)
                unsigned d0, d1, d2;
                // the products of the output blocks some pixel of the row is served in (U0: columns 0 .. 31, U1:
                // 32 .. 63), the two blocks' chains interleaved so that neither waits on its own previous product
                auto products = [&](auto U0, auto U1) __attribute__((always_inline)) {
                    constexpr bool use[2] = {decltype(U0)::value, decltype(U1)::value};
                    if constexpr (k16) {
                        // output blocks 0, 1 are the left half, 2, 3 the right; each the sum over two margin blocks
                        cv_v4i aop[2], acc[4];
#pragma unroll
                        for (int kb = 0; kb < 2; kb++) aop[kb] = *(const cv_v4i *)(crow + 64 * kb);
#pragma unroll
                        for (int ob = 0; ob < 4; ob++)
                            if (use[ob / 2]) acc[ob] = __builtin_amdgcn_mfma_i32_16x16x64_i8(aop[0], band[ob][0], cv_v4i{}, 0, 0, 0);
#pragma unroll
                        for (int ob = 0; ob < 4; ob++)
                            if (use[ob / 2]) acc[ob] = __builtin_amdgcn_mfma_i32_16x16x64_i8(aop[1], band[ob][1], acc[ob], 0, 0, 0);
                        const uint32_t *const rw = (const uint32_t *)&sh.row[io & ~3u];
                        d0 = rw[0];
                        d1 = rw[1];
                        d2 = rw[2];
                        // D's register i of this lane: number 4 (lane / 16) + i, output column 16 ob + lane % 16; two
                        // numbers per 4-byte store
#pragma unroll
                        for (int ob = 0; ob < 4; ob++)
#pragma unroll
                            for (int e = 0; e < 2; e++)
                                if (use[ob / 2])
                                    wcol[ob / 2][(ob & 1) * kOb16 + e] = (unsigned)acc[ob][2 * e] | (unsigned)acc[ob][2 * e + 1] << 16;
                        return;
                    }
                    constexpr int kb0 = use[0] ? 0 : 1, kb1 = use[1] ? NJ : NJ - 1;
                    cv_v4i aop[NBU][NJ + 1];
#pragma unroll
                    for (int nb = 0; nb < NBU; nb++)
#pragma unroll
                        for (int kb = kb0; kb <= kb1; kb++) aop[nb][kb] = *(const cv_v4i *)(crow + (32 * nb * kMCnt + 32 * kb));
#pragma unroll
                    for (int nb = 0; nb < NBU; nb++) {
                        cv_v16i acc[2];
#pragma unroll
                        for (int mb = 0; mb < 2; mb++)
                            if (use[mb]) acc[mb] = __builtin_amdgcn_mfma_i32_32x32x32_i8(aop[nb][mb], band[mb][0], cv_v16i{}, 0, 0, 0);
#pragma unroll
                        for (int j = 1; j < NJ; j++)
#pragma unroll
                            for (int mb = 0; mb < 2; mb++)
                                if (use[mb]) acc[mb] = __builtin_amdgcn_mfma_i32_32x32x32_i8(aop[nb][mb + j], band[mb][j], acc[mb], 0, 0, 0);
                        if (nb == NBU - 1) {
                            const uint32_t *const rw = (const uint32_t *)&sh.row[io & ~3u];
                            d0 = rw[0];
                            d1 = rw[1];
                            d2 = rw[2];
                        }
                        // D's register 4 q + i of this lane: number 32 nb + 8 q + 4 (lane / 32) + i, output column
                        // 32 mb + lane % 32; two numbers per 4-byte store
#pragma unroll
                        for (int mb = 0; mb < 2; mb++)
#pragma unroll
                            for (int q = 0; q < 4; q++)
#pragma unroll
                                for (int e = 0; e < 2; e++)
                                    if (use[mb])
                                        wcol[mb][16 * nb + 4 * q + e] = (unsigned)acc[mb][4 * q + 2 * e] | (unsigned)acc[mb][4 * q + 2 * e + 1] << 16;
                    }
                };
                typedef std::integral_constant<bool, true> On;
                typedef std::integral_constant<bool, false> Off;
                if ((unsigned)sm == 0u) products(Off(), On());
                else if ((sm >> 32) == 0ull) products(On(), Off());
                else products(On(), On());
                // byte k of the nine: the doubled row of value index io - 4 + k
                const unsigned w0 = __builtin_amdgcn_alignbyte(d1, d0, io), w1 = __builtin_amdgcn_alignbyte(d2, d1, io),
                               w2 = __builtin_amdgcn_alignbyte(d2, d2, io);
                asm volatile("" ::: "memory");
                unsigned cnt9[2 * kDiffs + 1];
#pragma unroll
                for (int k = 0; k < 2 * kDiffs + 1; k++) {
                    const unsigned r2 = ((k < 4 ? w0 : k < 8 ? w1 : w2) >> (8 * (k & 3))) & 0xFFu;
                    cnt9[k] = *(const __attribute__((address_space(3))) uint16_t *)(wmine + r2);
                }
                // the index of a served pixel, and past every threshold range on the other lanes, so that one range
                // test per difference takes both the thresholds and the serving
                const unsigned ios = serve ? io : 0x40000000u;
                const unsigned p0 = cnt9[kDiffs];
                float c = 0.0f;
#pragma unroll
                for (int d = 0; d < kDiffs; d++) {
                    // a difference no served pixel of the row passes is skipped by the whole wave (bright rows pass
                    // only the larger differences)
                    const bool ok = ios - t.oklo[d] <= t.okspan[d];
                    if (!__builtin_amdgcn_ballot_w64(ok)) continue;
                    const unsigned pa = cnt9[kDiffs + d + 1], pb = cnt9[kDiffs - d - 1];
                    const unsigned pm = pa > pb ? pa : pb;
                    // w * p0 * pm * RN(1 / (p0 + pm)), the reciprocal as the reference's table holds it; computed on
                    // every lane and selected.  On a served pixel p0 + pm is at least 1 (its own count) and pm = 0
                    // gives 0, which never exceeds c, so the reference's pm > 0 test needs no instruction; elsewhere a
                    // NaN (0 / 0) exceeds nothing either
                    const float rcp = recip_rn(p0 + pm);
                    float val = __fmul_rn((float)__umul24(__umul24((unsigned)t.weight[d], p0), pm), rcp);
                    asm("" : "+v"(val));
                    const bool take = ok & (val > c);
                    c = take ? val : c;
                }
                // c stays 0 on a lane that is not served.  The full plane (when kept) and the tile's list, as buffer
                // stores: a lane with nothing to store stores past the end, where the store is dropped
                if (t.keep)
                    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(c), t.plane_rs, serve ? 4u * (unsigned)x : 0x80000000u,
                                                          __builtin_amdgcn_readfirstlane(4u * (unsigned)y * (unsigned)t.w), 0);
                sum_c += (double)c;  // exact: see cv_tile
                const unsigned long long m = __builtin_amdgcn_ballot_w64(c > 0.0f);
                const unsigned below = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(c), t.seg_rs, c > 0.0f ? 4u * below : 0x80000000u,
                                                      __builtin_amdgcn_readfirstlane(4u * tcount), 0);
                tcount += (unsigned)__popcll(m);
                asm volatile("" ::: "memory");  // every read of the table before the next row writes it
            }
        };
        unsigned ca[5], cb[5] = {0, 0, 0, 0, 0};
        load(t.y0, ca);
        // the first row removes nothing: index(0) is 0, whose row is B
        step(0, ca, cb, 0u, 0u);
        for (int r = 1; r < t.rows; r += 2) {
            step(r, cb, ca, ca[3], ca[4]);
            if (r + 1 >= t.rows) break;
            step(r + 1, ca, cb, cb[3], cb[4]);
        }
        if (hi == relevant) break;
        b0 = hi - kDiffs;
        __syncthreads();  // the pass's reads of the tables are done before the next pass rewrites them
    }
}

template <int BMAX>
__device__ __forceinline__ uint8_t *cv_flags(CvShared<BMAX> &sh) { return (uint8_t *)&sh.pp[0][0]; }
template <int NB>
__device__ __forceinline__ uint8_t *cv_flags(CvSharedM<NB> &sh) { return sh.flags(); }

template <int BMAX>
__device__ __forceinline__ void cv_run(CvShared<BMAX> &sh, const CvTile &t, int relevant, double &sum_c,
                                       unsigned &tcount)
{
    if (relevant <= 8) cv_passes<8>(sh, t, relevant, sum_c, tcount);
    else if (relevant <= 16) cv_passes<16>(sh, t, relevant, sum_c, tcount);
    else if (BMAX == 32 || relevant <= 32) cv_passes<32>(sh, t, relevant, sum_c, tcount);
    else cv_passes<(BMAX > 32 ? BMAX : 32)>(sh, t, relevant, sum_c, tcount);
}
template <int NB, int NJ>
__device__ __forceinline__ void cv_run_m(CvSharedM<NB> &sh, const CvTile &t, int relevant, double &sum_c,
                                         unsigned &tcount)
{
    if (t.pairs) {
        if (NB == 1 || relevant <= 32) cv_passes_m<NB, 1, NJ, true, false>(sh, t, relevant, sum_c, tcount);
        else cv_passes_m<NB, (NB > 1 ? 2 : 1), NJ, true, false>(sh, t, relevant, sum_c, tcount);
    } else {
        if (NB == 1 || relevant <= 32) cv_passes_m<NB, 1, NJ, false, false>(sh, t, relevant, sum_c, tcount);
        else cv_passes_m<NB, (NB > 1 ? 2 : 1), NJ, false, false>(sh, t, relevant, sum_c, tcount);
    }
}
template <int NB>
__device__ __forceinline__ void cv_run(CvSharedM<NB> &sh, const CvTile &t, int relevant, double &sum_c,
                                       unsigned &tcount)
{
    // a tile with at most 16 relevant values takes one pass of 16 numbers
    if (relevant <= 16) {
        if (t.pairs) cv_passes_m<NB, 1, 1, true, true>(sh, t, relevant, sum_c, tcount);
        else cv_passes_m<NB, 1, 1, false, true>(sh, t, relevant, sum_c, tcount);
        return;
    }
    // the margin blocks the band of an output block touches: 2 pad + 1 columns from the block's first
    const int nj = min((31 + 2 * t.pad) / 32 + 1, 3);
    if (nj == 3) cv_run_m<NB, 3>(sh, t, relevant, sum_c, tcount);
    else if (nj == 2) cv_run_m<NB, 2>(sh, t, relevant, sum_c, tcount);
    else cv_run_m<NB, 1>(sh, t, relevant, sum_c, tcount);
}

// The numbering of the relevant values: lane c holds chunk c's relevance mask (values 64 c .. 64 c + 63) and the
// number of relevant values below the chunk.  The prefix scans keep a rank per value; the matrix-core form keeps the
// masks, from which each pass numbers.
template <int BMAX>
__device__ __forceinline__ void cv_number(CvShared<BMAX> &sh, unsigned long long need, unsigned before, int nchunks, int nv)
{
    const int lane = threadIdx.x;
    for (int c = 0; c < nchunks; c++) {
        const unsigned long long m = (unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(unsigned)need, c) |
                                     (unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(unsigned)(need >> 32), c) << 32;
        const unsigned b = (unsigned)__builtin_amdgcn_readlane((int)before, c);
        const int i = 64 * c + lane;
        if (i < nv) sh.rank[i] = (m >> lane & 1ull) ? (uint16_t)(b + __popcll(m & ((1ull << lane) - 1))) : kNone;
    }
}
template <int NB>
__device__ __forceinline__ void cv_number(CvSharedM<NB> &sh, unsigned long long need, unsigned, int nchunks, int)
{
    if ((int)threadIdx.x < nchunks) sh.relevant[threadIdx.x] = need;
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

// The matrix-core layouts (a pass of 32 numbers costs the same however many of them are used).
template <typename Sh> constexpr bool kCvMatrix = false;
template <int NB> constexpr bool kCvMatrix<CvSharedM<NB>> = true;

template <typename Sh>
__device__ void cv_tile(Sh &sh, const CambiPic &pic, int s, int local)
{
    const CambiConsts &K = *pic.consts;
    CvTile t;
    t.w = pic.sw[s];
    t.h = pic.sh[s];
    const int tile_h = pic.th[s];
    t.rs = buffer_over(pic.planes + pic.off[s], (unsigned long long)t.w * t.h * 2);  // one tile per wave
    const int across = (t.w + kCvTileW - 1) / kCvTileW;
    t.x0 = (local % across) * kCvTileW;
    t.y0 = (local / across) * tile_h;
    t.rows = min(tile_h, t.h - t.y0);
    // the picture's constants, the same on every lane
    t.pad = __builtin_amdgcn_readfirstlane(K.pad);
    t.band_base = __builtin_amdgcn_readfirstlane(K.band_base);
    t.band_size = __builtin_amdgcn_readfirstlane(K.band_size);
    t.vlt = __builtin_amdgcn_readfirstlane(K.vlt);
#pragma unroll
    for (int d = 0; d < kDiffs; d++) {
        t.tvi[d] = __builtin_amdgcn_readfirstlane(K.tvi[d]);
        t.weight[d] = __builtin_amdgcn_readfirstlane(K.weight[d]) & 0xFF;  // 1 .. 4: 24-bit products stay exact
        // i - 1 + band_base <= tvi[d] and i - 1 + band_base + d + 1 > vlt, with i >= 1
        const int lo = max(t.vlt - t.band_base - d, 0) + 1, hi = t.tvi[d] - t.band_base + 1;
        t.oklo[d] = hi >= lo ? (unsigned)lo : 0x80000000u;
        t.okspan[d] = hi >= lo ? (unsigned)(hi - lo) : 0u;
    }
    t.keep = pic.cvals != nullptr;
    t.pairs = (t.w & 1) == 0 && ((uintptr_t)(pic.planes + pic.off[s]) & 3u) == 0;
    t.plane_rs = buffer_over(pic.cvals ? pic.cvals + pic.off[s] : nullptr, pic.cvals ? (unsigned long long)t.w * t.h * 4 : 0);
    t.seg_rs = buffer_over(pic.list + pic.loff[s] + (long long)local * kCvTileW * tile_h, (unsigned long long)kCvTileW * tile_h * 4);
    unsigned *seg_count = pic.seg_count + pic.toff[s] + local;
    const int lane = threadIdx.x, nv = t.band_size + 2 * kDiffs;
    const int c0 = 2 * lane, xl = t.x0 - t.pad + c0, cw = kCvTileW + 2 * t.pad;
    __syncthreads();  // the previous tile's reads of the tables are done
    // the served values (of the tile's masked in-band pixels) and the occupied ones (of the tile and its margins) as
    // byte flags over value indices (value - band_base + 4); an inactive sample sets the unread flag kCvDummy
    uint8_t *const served = cv_flags(sh), *const occupied = served + kCvFlags;
    for (int i = lane; i < 2 * kCvFlags / 4; i += 64) ((uint32_t *)served)[i] = 0u;
    __syncthreads();
    auto flag = [&](uint8_t *flags, unsigned v) {
        const unsigned u = cv_off(t, v);
        const bool active = u < (unsigned)t.band_size;
        flags[active ? u + kDiffs : kCvDummy] = 1;
        return active;
    };
    bool any = false;
    // the setup waits on memory once per batch: 32 loads per lane in flight (the row loops below keep fewer)
    if (t.pairs) {
        // aligned pairs: lane l reads picture columns x0 + 2 (l % 32) and the next of row r + l / 32, so that one load
        // covers two rows of the tile
        const int xp = t.x0 + 2 * (lane & 31), rl = lane >> 5;
        const unsigned xo = xp < t.w ? 2u * (unsigned)xp : kCvOut;
        for (int r = 0; r < t.rows; r += 64) {
            unsigned v[32];
#pragma unroll
            for (int j = 0; j < 32; j++) {
                const int rr = r + 2 * j + rl;
                v[j] = __builtin_amdgcn_raw_buffer_load_b32(t.rs, xo + (unsigned)((rr < t.rows ? t.y0 + rr : -1) * t.w * 2), 0, 0);
            }
#pragma unroll
            for (int j = 0; j < 32; j++) {
                any |= flag(served, v[j] & 0xFFFFu);
                any |= flag(served, v[j] >> 16);
            }
        }
    } else {
        for (int r = 0; r < t.rows; r += 32) {
            unsigned v[32];
#pragma unroll
            for (int j = 0; j < 32; j++) v[j] = cv_at(t, cv_col(t, t.x0 + lane), r + j < t.rows ? t.y0 + r + j : -1);
#pragma unroll
            for (int j = 0; j < 32; j++) any |= flag(served, v[j]);
        }
    }
    if (!__any(any)) {  // most tiles of most pictures: nothing masked inside the band
        if (lane == 0) *seg_count = 0;
        return;
    }
    __syncthreads();  // the served flags are in before they are read
    // the served set, 64 values per chunk, lane c holding chunk c, and the values within 4 of a served one (the set
    // widened by 4 either way, across chunk edges): all below nv
    const int nchunks = (nv + 63) / 64;
    unsigned long long sc = 0;  // this lane's chunk of the served set
    for (int c = 0; c < nchunks; c++) {
        const unsigned long long sm = __builtin_amdgcn_ballot_w64(64 * c + lane < nv && served[64 * c + lane]);
        if (lane == c) sc = sm;
    }
    const unsigned long long sp = __shfl_up(sc, 1), sn = __shfl_down(sc, 1);
    unsigned long long near = sc;
#pragma unroll
    for (int m = 1; m <= kDiffs; m++) near |= sc >> m | (lane + 1 < nchunks ? sn : 0ull) << (64 - m) | sc << m | (lane ? sp : 0ull) >> (64 - m);
    unsigned long long need = lane < nchunks ? near : 0ull;
    // The relevant values: those within 4 of a served one that occur in the tile with its margins (numbered in value
    // order).  When the matrix-core form takes all the values within 4 of a served one in one pass of 32, the scan of
    // the margins is left out: a value that does not occur adds a number whose counts stay 0, and the tile takes one
    // pass either way.
    if (!kCvMatrix<Sh> || __builtin_amdgcn_readlane(wave_scan_incl((unsigned)__popcll(need)), 63) > 32) {
        const int ya = max(t.y0 - t.pad, 0), yb = min(t.y0 + t.rows + t.pad, t.h);
        if (t.pairs) {
            // An even width on a 4-byte aligned plane keeps every row's samples at even picture columns 4-byte
            // aligned: a lane reads the two codes of one aligned pair in one load, window columns 2 lane - sh and
            // 2 lane + 1 - sh (sh = 1 when the window starts at an odd column, and its column -1 is then left out).  A
            // pair is inside or outside the picture as a whole (picture columns 2k, 2k + 1).
            const int sh_ = (t.x0 - t.pad) & 1, cl = 2 * lane - sh_, xlo = t.x0 - t.pad + cl;
            const unsigned xo = xlo >= 0 && xlo < t.w ? 2u * (unsigned)xlo : kCvOut;
            const bool lo_ok = cl >= 0 && cl < cw, hi_ok = cl + 1 < cw;
            for (int y = ya; y < yb; y += 32) {
                unsigned v[32];
#pragma unroll
                for (int j = 0; j < 32; j++) {
                    const int yy = y + j < yb ? y + j : -1;
                    v[j] = __builtin_amdgcn_raw_buffer_load_b32(t.rs, xo + (unsigned)(yy * t.w * 2), 0, 0);
                }
#pragma unroll
                for (int j = 0; j < 32; j++) {
                    flag(occupied, lo_ok ? v[j] & 0xFFFFu : 0u);
                    flag(occupied, hi_ok ? v[j] >> 16 : 0u);
                }
            }
        } else {
            for (int y = ya; y < yb; y += 16) {
                unsigned v[32];
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    v[2 * j] = cv_at(t, c0 < cw ? cv_col(t, xl) : kCvOut, y + j < yb ? y + j : -1);
                    v[2 * j + 1] = cv_at(t, c0 + 1 < cw ? cv_col(t, xl + 1) : kCvOut, y + j < yb ? y + j : -1);
                }
#pragma unroll
                for (int j = 0; j < 32; j++) flag(occupied, v[j]);
            }
        }
        __syncthreads();
        unsigned long long oc = 0;  // this lane's chunk of the occupied set
        for (int c = 0; c < nchunks; c++) {
            const unsigned long long om = __builtin_amdgcn_ballot_w64(64 * c + lane < nv && occupied[64 * c + lane]);
            if (lane == c) oc = om;
        }
        need &= oc;
    }
    const unsigned count = (unsigned)__popcll(need), after = wave_scan_incl(count);
    const int relevant = (int)__builtin_amdgcn_readlane(after, 63);
    cv_number(sh, need, after - count, nchunks, nv);
    __syncthreads();
    // A lane's positive c-values are at least 0.5, so multiples of 2^-24, each at most 4 * 65^2 / 2 < 2^14, and it serves
    // at most one pixel per row of the tile (at most kMaxTileH = 2^12 rows): its double sum stays below 2^26 in whole
    // units of 2^-24, so it is exact, and so is the product by 2^24 below (an integer under 2^50), the c-values' exact
    // sum in units of 2^-24.
    double sum_c = 0.0;
    unsigned tcount = 0;
    if (relevant > 0) cv_run(sh, t, relevant, sum_c, tcount);
    const unsigned long long sum_fx = wave_sum((unsigned long long)(sum_c * 16777216.0));
    if (lane == 0) {
        *seg_count = tcount;
        if (tcount) atomicAdd(pic.counts + s, tcount);
        if (sum_fx) atomicAdd(pic.sums + s, sum_fx);
    }
}

// Waves per SIMD each layout is built for, as its LDS allows: four for the matrix-core form over 32 numbers (10 KB:
// 16 waves per CU), two for the 64-number layouts, three for the prefix-scan form over 32 numbers.
template <typename Sh> constexpr int kCvWaves = 3;
template <> constexpr int kCvWaves<CvSharedM<1>> = 4;
template <> constexpr int kCvWaves<CvSharedM<2>> = 2;
template <> constexpr int kCvWaves<CvShared<64>> = 2;

// The tile of claim v on XCD queue k, or -1 when the queue is empty: queue k holds the k-th eighth of every scale's
// tiles (tile_first lists the slots by scale, coarsest first, np pictures each), coarsest scale first.
__device__ int queue_tile(const int *tile_first, int np, unsigned k, unsigned v)
{
    for (int j = 0; j < kScales; j++) {
        const unsigned a = (unsigned)tile_first[j * np], n = (unsigned)tile_first[(j + 1) * np] - a;
        const unsigned lo = (unsigned)((unsigned long long)n * k / kXcds), hi = (unsigned)((unsigned long long)n * (k + 1) / kXcds);
        if (v < hi - lo) return (int)(a + lo + v);
        v -= hi - lo;
    }
    return -1;
}

// Sh: CvShared<BMAX> (prefix scans on the vector units) or CvSharedM<NB> (window counts on the matrix cores).
// Persistent waves take tiles from queues.  With xcd_queues there is one queue per XCD (counter[k]): the blocks
// launched to XCD k (blockIdx.x % 8) take the k-th eighth of every scale's tiles, coarsest scale first, so that
// neighbouring tiles, which share halo rows and columns, run together on one XCD and meet in its L2, and every XCD
// has the same share of every scale's work.  A wave whose queue is empty goes on to the next queues and stops when all
// eight are empty; each queue's claims rise, so an empty queue stays empty, and every tile is in exactly one queue.
// Without xcd_queues every wave takes tiles in order from counter[0].
template <typename Sh>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kCvWaves<Sh>))) void cambi_cvalues_kernel(
    const CambiPic *pics, const int *tile_first, int nslots, int persistent, unsigned int *counter, int xcd_queues)
{
    __shared__ Sh sh;
    const int total = tile_first[nslots], np = nslots / kScales;
    const unsigned home = blockIdx.x % kXcds;
    unsigned q = 0;  // queues left behind as empty
    for (int t = persistent ? 0 : (int)blockIdx.x;;) {
        if (persistent && xcd_queues) {
            t = -1;
            for (; q < kXcds && t < 0; ) {
                const unsigned k = (home + q) % kXcds;
                unsigned v = 0;
                if (threadIdx.x == 0) v = atomicAdd(counter + k, 1u);
                t = queue_tile(tile_first, np, k, __builtin_amdgcn_readfirstlane(v));
                if (t < 0) q++;
            }
            if (t < 0) break;
        } else if (persistent) {
            unsigned v = 0;
            if (threadIdx.x == 0) v = atomicAdd(counter, 1u);
            t = __builtin_amdgcn_readfirstlane((int)v);
        }
        if (t >= total) break;
        const int slot = find_slot(tile_first, nslots, t);  // slots by scale, coarsest first, then by picture
        const int np = nslots / kScales, p = slot % np, s = kScales - 1 - slot / np;
        cv_tile(sh, pics[p], s, t - tile_first[slot]);
        if (!persistent) break;
    }
}

void launch_cambi_cvalues(const CambiPic *pics, const int *tile_first, int nslots, int total_tiles, int persistent,
                          unsigned int *work_counter, int blocks, int bmax, hipStream_t st, int mfma,
                          int xcd_queues)
{
    if (!persistent && total_tiles <= 0) return;
    const int grid = persistent ? blocks : total_tiles;
    auto k = mfma ? (bmax == 64 ? cambi_cvalues_kernel<CvSharedM<2>> : cambi_cvalues_kernel<CvSharedM<1>>)
                  : (bmax == 64 ? cambi_cvalues_kernel<CvShared<64>> : cambi_cvalues_kernel<CvShared<32>>);
    hipLaunchKernelGGL(k, dim3(grid), dim3(64), 0, st, pics, tile_first, nslots, persistent, work_counter, xcd_queues);
    hip_check(hipGetLastError(), "cambi c-values");
}

// ---------------------------------------------------------------------------------------------------------------
// Top-k select.  Positive c-values are at least 0.5, so their bit patterns less 0x3F000000 are 27-bit keys in value
// order; three passes of 9-bit digits select the k-th largest.  Each pass counts (and sums exactly) the digit of the
// keys that match the digits chosen so far, several blocks per job each into its own slot; then one block per job
// adds the slots and walks the digits from the top, as the reference's host loop does (src/cambi.cpp:151-156), keeping
// the count and sum of the keys above the chosen digit.  Jobs with no more than k positive values are skipped: their
// pooling needs no select.
// ---------------------------------------------------------------------------------------------------------------
constexpr unsigned kKeyBase = 0x3F000000u;

__global__ __launch_bounds__(256) void topk_count_kernel(TopkJob *jobs, int pass)
{
    const TopkJob &J = jobs[blockIdx.y];
    if ((unsigned long long)*J.count <= J.k) return;
    __shared__ unsigned hist[512];
    __shared__ unsigned long long hs[512];
    const int tid = threadIdx.x;
    for (int i = tid; i < 512; i += 256) {
        hist[i] = 0;
        hs[i] = 0;
    }
    __syncthreads();
    const int shift = 18 - 9 * pass;
    const unsigned prefix = pass ? J.prefix : 0u;
    for (int sg = blockIdx.x; sg < J.nseg; sg += gridDim.x) {
        const unsigned cnt = J.seg_count[sg];
        const float *seg = J.list + (size_t)sg * J.seg_cap;
        for (unsigned i = tid; i < cnt; i += 256) {
            const float c = seg[i];
            const unsigned key = __float_as_uint(c) - kKeyBase;
            if ((key >> (shift + 9)) == (prefix >> (shift + 9))) {
                const unsigned d = (key >> shift) & 511u;
                atomicAdd(&hist[d], 1u);
                atomicAdd(&hs[d], fixed24(c));
            }
        }
    }
    __syncthreads();
    // this block's counts and sums, in its own slot
    unsigned *ho = J.hist + (size_t)blockIdx.x * 512;
    unsigned long long *so = J.hsum + (size_t)blockIdx.x * 512;
    for (int d = tid; d < 512; d += 256) {
        ho[d] = hist[d];
        so[d] = hs[d];
    }
}

// One block per job: the digit counts and sums of the counting blocks added up, then the walk from the top digit,
// in parallel.  Thread t holds digit g = 511 - t, so an inclusive scan over the threads (DPP within each wave, the
// waves' totals through LDS) gives each digit the count of keys at or above it (incl) and strictly above it (excl).
// The host loop (src/cambi.cpp:151-156) stops at the highest digit g >= 1 with excl + count >= remaining, or at 0:
// since incl only grows with t, that is the one thread with (t == 511 or incl >= remaining) and (t == 0 or
// excl < remaining), which records the digit and the count and exact sum of the keys above it.
__global__ __launch_bounds__(512) void topk_walk_kernel(TopkJob *jobs, int pass, int blocks)
{
    TopkJob &J = jobs[blockIdx.x];
    if ((unsigned long long)*J.count <= J.k) return;
    __shared__ unsigned wc[8];
    __shared__ unsigned long long wf[8];
    const int t = threadIdx.x, g = 511 - t, lane = t & 63, wave = t >> 6;
    // the job's state, read by every thread before the barrier below, after which one thread rewrites it
    const unsigned remaining = pass ? J.remaining : (unsigned)J.k;
    const unsigned prefix = pass ? J.prefix : 0u;
    const unsigned long long above0 = pass ? J.above : 0ull, fx0 = pass ? J.above_fx : 0ull;
    unsigned c = 0;
    unsigned long long sm = 0;
#pragma unroll 4
    for (int b = 0; b < blocks; b++) {
        c += J.hist[(size_t)b * 512 + g];
        sm += J.hsum[(size_t)b * 512 + g];
    }
    const unsigned ci = wave_scan_incl(c);
    const unsigned long long fi = (unsigned long long)wave_scan_incl64((long long)sm);
    if (lane == 63) {
        wc[wave] = ci;
        wf[wave] = fi;
    }
    __syncthreads();
    unsigned off = 0;
    unsigned long long foff = 0;
    for (int v = 0; v < wave; v++) {
        off += wc[v];
        foff += wf[v];
    }
    const unsigned incl = off + ci, excl = incl - c;
    if ((t == 511 || incl >= remaining) && (t == 0 || excl < remaining)) {
        J.remaining = remaining - excl;
        J.prefix = prefix | ((unsigned)g << (18 - 9 * pass));
        J.above = above0 + excl;
        J.above_fx = fx0 + (foff + fi - sm);
    }
}

void launch_topk_pass(TopkJob *jobs, int njobs, int pass, int blocks, hipStream_t st)
{
    if (njobs <= 0) return;
    hipLaunchKernelGGL(topk_count_kernel, dim3(blocks, njobs), dim3(256), 0, st, jobs, pass);
    hipLaunchKernelGGL(topk_walk_kernel, dim3(njobs), dim3(512), 0, st, jobs, pass, blocks);
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
