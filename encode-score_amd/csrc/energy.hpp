// Job descriptors and launch wrappers of the window-energy, lasting-texture and collect kernels.
#pragma once
#include "es_internal.hpp"

namespace es {

// One (source, encode) pair for window_stats: both planes cropped to w x h (the source's pointer already offset to
// its crop), luma brought to the 10-bit scale by `up`.  mid_r / mid_d receive M = 81 mid on the area whole windows
// cover (either may be null); sums receives 12 exact sums per window.  sse (may be null) += the luma squared error of
// the pair over that same area at the pipeline depth (the samples before `up`), the share of the frame-sums job whose
// lw x lh area it is.
struct WsPair {
    PlaneIn r, d;
    int w, h, up;
    int32_t *mid_r, *mid_d;
    long long *sums;
    unsigned long long *sse;
    // FP64 form only (any may be null): mid = M / 81 as float planes, and the reference's 8 figures per window as
    // floats (src/kernels.cu:340-347)
    float *midf_r, *midf_d;
    float *vals;
};

// One lasting-texture job: ring[j] is the source mid plane of offset j - 6 from the centre frame.
//   plain:  mid_d = the encode's plane of the centre, out = 5 sums per window
//   source: st receives ST = SA + SB, out = [M_r(c) ST, SA SB, SA^2, SB^2] per window
//   encode: mid_d and st are read, out = [M_d(c) ST] per window
struct PersistJob {
    const int32_t *ring[kRing];
    const int32_t *mid_d;
    int32_t *st;
    int w, h;
    long long *out;
};

// Lasting texture of the batched form, once per crop and round: the centres centre0 .. centre0 + ncentres - 1 in
// order, with the source's SA and SB carried from one centre to the next as exact running sums
// (SA(c) = SB(c - 1) - m(c) + m(c + 6), SB(c) = SA(c - 1) - m(c - 7) + m(c - 1); computed in full at the first centre
// of the run), and each of the crop's encodes' projections on ST.  Centre c is round slot c + 6 - f0.
constexpr int kMaxCropEncodes = 16;
struct LastingJob {
    const int32_t *ring;              // the crop's source mid planes, `slots` of them, slot of frame f at f % slots
    const int32_t *emid[kMaxCropEncodes];  // the encodes' mid planes, same slots
    long long *penc[kMaxCropEncodes];     // per encode: [slot][window] M_d(c) ST
    int last_centre[kMaxCropEncodes];     // per encode: its last centre with lasting texture
    int nenc, slots, w, h;
    int32_t *sa, *sb;                 // running sums of the previous centre, per pixel
    long long *psrc;                  // [slot][window][4]: M_r(c) ST, SA SB, SA^2, SB^2
    int centre0, ncentres, slot0;
};

// The reference-shaped lasting texture on float mid planes (16 x 16 windows): ring[j] is the source's plane of offset
// j - 6 from the centre frame; out receives 5 floats per window.
struct PersistF32Job {
    const float *ring[kRing];
    const float *mid_d;
    int w, h;
    float *out;
};

void launch_window_stats(const WsPair *pairs, int npairs, int max_w, int max_h, int win, bool fp64, hipStream_t st);
void launch_persist_f32(const PersistF32Job *jobs, int njobs, int max_w, int max_h, hipStream_t st);
void launch_lasting(const LastingJob *jobs, int njobs, int max_w, int max_h, int win, hipStream_t st);
void launch_persist(const PersistJob *jobs, int njobs, int max_w, int max_h, int win, int kind, hipStream_t st);

// Device collect of one (encode, frame): structure threshold, mask, histogram entries and frame totals; and of the
// lasting texture of a centre frame.  See collect_kernels.hip.
struct CollectJob {
    const long long *sums;        // windows x 12
    int windows;
    uint8_t *mask;                // out: this frame's structure mask (windows)
    unsigned long long *hist;     // the segment's 6 histograms: amount[3], matching[2], lasting (6 x 20001)
    long long *tot_int;           // out: 12 exact totals: all[6], structured[6]
    double *tot_orient;           // out: 4 totals: all_r, all_d, st_r, st_d of the oriented energy
    double *quantile;             // out: the structure threshold (diagnostic)
    // lasting texture of the centre frame (unused when psrc is null)
    const long long *psrc;        // windows x 4 (batched) or windows x 5 (plain sums), see lasting_plain
    const long long *penc;        // windows x 1 (batched)
    const uint8_t *cmask;         // the centre frame's structure mask
    unsigned long long *lhist;    // the centre segment's histograms (lasting at index 5)
    long long *tot_lasting;       // out: [own, mine, count] exact
    int lasting_plain;
};

// phase 0: the frame part of every job; phase 1: the lasting texture (after every job's phase 0 wrote its mask).
// Working memory of the collect: the keys of the structure threshold, the per-chunk partial results (chunks of 256
// windows) and the windows' histogram bins.
struct CollectWork {
    unsigned long long *keys;  // njobs x max_windows
    long long *cint;           // njobs x max_chunks x 12
    double *cdbl;              // njobs x max_chunks x 4
    long long *clast;          // njobs x max_chunks x 3
    uint16_t *bins;            // njobs x max_windows x 6: each window's bin in the 6 histograms, or 0xFFFF
    unsigned *fhist;           // njobs x 6 x 20001: each frame's counts
    int max_windows, max_chunks;
};
// jobs[f * nenc + e] is frame f of the round for encode e.
void launch_collect(const CollectJob *jobs, int nframes, int nenc, const CollectWork &w, hipStream_t st);
void launch_hist_quantiles(const unsigned long long *hists, int nhist, const double *qs, int nq, double *out,
                           unsigned long long *counts, hipStream_t st);
// The quantiles of a launch carried in its arguments.
constexpr int kMaxListQuantiles = 16;
struct QuantileList {
    double q[kMaxListQuantiles];
    __host__ __device__ double operator[](int j) const { return q[j]; }
};
void launch_hist_quantiles_list(const unsigned long long *hists, int nhist, const QuantileList &qs, int nq, double *out,
                                unsigned long long *counts, hipStream_t st);
void launch_select_u64(const unsigned long long *keys, long long n, long long rank, unsigned long long *out,
                       hipStream_t st);
void launch_copy(void *dst, const void *src, size_t bytes, hipStream_t st);
void launch_select_u32(const unsigned int *keys, long long n, long long rank, unsigned int *out, hipStream_t st);

}  // namespace es
