// Internal declarations shared by the HIP translation units of libencscore_hip: the device-side descriptors of
// pictures and jobs, and the host-side launch wrappers each kernel file exports.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace es {

constexpr int kScales = 5;
constexpr int kDiffs = 4;
constexpr uint16_t kNone = 0xFFFF;
constexpr int kCvTileW = 64;     // c-value tiles are one wave wide
constexpr int kMaxPad = 32;      // one wave covers the tile and both margins as 2 columns per lane: window <= 65
// band values: the display model fixes the band (band_base 0 to the threshold of difference 4 plus 5: 568 values of
// 10-bit codes); the parameters' set-up checks it, and the c-value kernel's LDS tables are sized by it
constexpr int kMaxBand = 576;
constexpr int kXcds = 8;          // XCDs of an MI355X: a launch's workgroups go to them in turn
constexpr int kHistBins = 20001; // per-window histograms: 0.05-point bins over [-500, 500]
constexpr int kRing = 13, kRadius = 6;
constexpr int kMaxTileH = 4096;  // c-value tile heights: a lane's exact double sum of a tile's c-values needs < 2^20

// Thrown by every host function on a HIP error or a bad argument; the C ABI turns it into a status code.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t r, const char *what)
{
    if (r != hipSuccess) throw Error(2, std::string(what) + ": " + hipGetErrorString(r));
}

// CAMBI constants of one picture size (the reference's CambiParams less the per-scale size).
struct CambiConsts {
    int pad, band_base, band_size, vlt, mask_index, window;
    int tvi[kDiffs], weight[kDiffs];
};

// A luma plane as the kernels read it: sample(x, y) = (plane[y * pitch + x] >> rshift) << shift, `bytes` 1 or 2 per
// sample (rshift only with 2: 6 for P010's high-bit samples), and with from8 the 8-bit anti-dithering mean of cambi.c
// applied after the shifts.
struct PlaneIn {
    const void *ptr;
    int pitch, bytes, shift, from8;
    int rshift;
};

// The code of a packed CAMBI sample: (value | mask << 15) - plane_bias(band_base) modulo 2^16.  It is 1 + (value -
// band_base + 4) for a masked value from band_base - 4 on, 0 for band_base - 5 and for what a buffer load past a
// plane returns, and far above the band for an unmasked value or a masked one below band_base - 5; adding the bias
// back gives value | mask << 15.
__host__ __device__ constexpr unsigned plane_bias(int band_base) { return 0x8000u + (unsigned)band_base - (kDiffs + 1); }

// One picture for the CAMBI kernels.  The packed scale planes hold each pixel's code (above), scale s at
// planes + off[s] (and the optional full c-value planes at cvals + off[s]).  The positive c-values of a scale go to
// per-tile segments: tile j of scale s (64 columns by th[s] rows) writes its values from list + loff[s] + j * 64 *
// th[s] and their number to seg_count[toff[s] + j], so no tile waits on another.
struct CambiPic {
    PlaneIn in;
    int sw[kScales], sh[kScales];
    long long off[kScales], loff[kScales];
    int toff[kScales];
    int th[kScales];              // c-value tile height of each scale
    uint16_t *planes;
    float *cvals;                 // full c-value planes or nullptr
    float *list;                  // per-tile segments of positive c-values
    unsigned int *seg_count;      // positive c-values per tile
    unsigned long long *sums;     // [kScales] exact sums, units of 2^-24
    unsigned int *counts;         // [kScales] positive c-values of the scale
    const CambiConsts *consts;    // device pointer
};

// Radix-select state of one (picture, scale) top-k: the k-th largest positive c-value when more than k are positive.
struct TopkJob {
    const float *list;             // the scale's segments, seg_cap values apart
    const unsigned int *seg_count; // values per segment
    int nseg, seg_cap;
    const unsigned int *count;
    unsigned long long k;
    unsigned int *hist;            // [blocks][512] each counting block's counts of the current digit
    unsigned long long *hsum;      // [blocks][512] and its exact sums per digit
    unsigned int prefix, remaining;
    unsigned long long above, above_fx;
};

// The c-value tile height of each scale: tile_h0 rows at full resolution (0: tile_h), tile_h rows at scale 1 and at
// most kCoarseTileH rows at scales 2-4.  Full-resolution tiles mostly take one pass, so taller ones spread the
// window's first rows and the tile's set-up over more rows; a coarse scale's tile spans many more distinct values and
// takes many passes (10-13 on a 4K mosaic at scale 4), so its tiles stay short: fewer values each, and enough tiles
// to spread over the GPU.
constexpr int kCoarseTileH = 32;
inline void cv_tile_heights(int tile_h, int tile_h0, int *th)
{
    for (int s = 0; s < kScales; s++)
        th[s] = s == 0 ? (tile_h0 > 0 ? tile_h0 : tile_h) : s == 1 ? tile_h : std::min(tile_h, kCoarseTileH);
}

// The per-tile segment layout of a picture's c-value lists for tiles of th[s] rows at scale s: list offsets and tile
// offsets per scale; returns the capacity in values and the number of tiles.
inline long long cambi_list_layout(const int *sw, const int *sh, const int *th, long long *loff, int *toff, int *tiles)
{
    long long cap = 0;
    int t = 0;
    for (int s = 0; s < kScales; s++) {
        const int n = ((sw[s] + kCvTileW - 1) / kCvTileW) * ((sh[s] + th[s] - 1) / th[s]);
        loff[s] = cap;
        toff[s] = t;
        cap += (long long)n * kCvTileW * th[s];
        t += n;
    }
    if (tiles) *tiles = t;
    return cap;
}

// Host-side launch wrappers (cambi_kernels.hip).
void launch_cambi_base(const CambiPic *pics, const int *tile_first, int npics, int total_tiles, hipStream_t st);
int scale_strips(int w, int h);
void launch_cambi_scale(const CambiPic *pics, int npics, int max_strips, int s, hipStream_t st);
// tile_first[k] (device) is the first tile of slot k, which holds scale kScales - 1 - k / npics of picture k % npics
// (nslots = npics * kScales: coarsest scale first); tile_first[nslots] holds the total; total_tiles sizes the grid
// when not persistent.
// mfma: window counts on the matrix cores instead of prefix scans (same results).
void launch_cambi_cvalues(const CambiPic *pics, const int *tile_first, int nslots, int total_tiles, int persistent,
                          unsigned int *work_counter, int blocks, int bmax, hipStream_t st, int mfma,
                          int xcd_queues = 0);
void launch_topk_pass(TopkJob *jobs, int njobs, int pass, int blocks, hipStream_t st);
void launch_cambi_finish(const CambiPic *pics, TopkJob *jobs, int npics, double *scores, hipStream_t st);
void launch_clear(unsigned long long *a, long long na, unsigned int *b, long long nb, hipStream_t st);
int base_tiles(int w, int h);
int cvalue_tiles(int w, int h, int tile_h);

}  // namespace es
