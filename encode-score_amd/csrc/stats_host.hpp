// Host accumulators: the Histogram and Accum of src/stats.hpp:8-82 (measure fields), the per-frame fold of the
// canonical frame totals, and the host collect of the plain form (src/main.cpp:576-643).
#pragma once
#include "encscore.h"
#include "es_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace es {

// Histogram::add's bin (src/stats.hpp:16)
inline int hist_bin_host(double pct)
{
    const double c = pct < -500.0 ? -500.0 : pct > 500.0 ? 500.0 : pct;
    return (int)std::lround((c - -500.0) / 0.05);
}

// Histogram::quantile (src/stats.hpp:21-30) of bins[0 .. kHistBins) with n entries.
inline double hist_quantile_host(const uint64_t *bins, uint64_t n, double q)
{
    if (!n) return NAN;
    const double target = q * (double)(n - 1);
    uint64_t seen = 0;
    for (int b = 0; b < kHistBins; b++) {
        seen += bins[b];
        if ((double)seen > target) return -500.0 + (double)b * 0.05;
    }
    return 500.0;
}

struct Accum {
    uint64_t frames = 0;
    double cambi = 0, cambi_source = 0;
    double all_r[3] = {}, all_d[3] = {}, st_r[3] = {}, st_d[3] = {};
    double sh_all_r[2] = {}, sh_all_x[2] = {}, sh_st_r[2] = {}, sh_st_x[2] = {};
    double own = 0, mine = 0;
    std::vector<uint64_t> hist = std::vector<uint64_t>((size_t)6 * kHistBins, 0);  // amount x3, matching x2, lasting
    // beside the measures (src/stats.hpp:37-39): PSNR and mean colour sums of the frames with chroma, and VMAF
    double psnr[3] = {}, color_r[3] = {}, color_d[3] = {};
    uint64_t cframes = 0;
    double vmaf = 0, vmaf_min = 1e9;
    uint64_t vframes = 0;

    void add(int k, double pct) { hist[(size_t)k * kHistBins + hist_bin_host(pct)]++; }
    uint64_t count(int k) const
    {
        uint64_t n = 0;
        for (int b = 0; b < kHistBins; b++) n += hist[(size_t)k * kHistBins + b];
        return n;
    }
    // Accum::merge (src/stats.hpp:50-64)
    void merge(const Accum &o)
    {
        frames += o.frames;
        cambi += o.cambi;
        cambi_source += o.cambi_source;
        own += o.own;
        mine += o.mine;
        for (int i = 0; i < 3; i++) {
            all_r[i] += o.all_r[i];
            all_d[i] += o.all_d[i];
            st_r[i] += o.st_r[i];
            st_d[i] += o.st_d[i];
        }
        for (int i = 0; i < 2; i++) {
            sh_all_r[i] += o.sh_all_r[i];
            sh_all_x[i] += o.sh_all_x[i];
            sh_st_r[i] += o.sh_st_r[i];
            sh_st_x[i] += o.sh_st_x[i];
        }
        for (size_t b = 0; b < hist.size(); b++) hist[b] += o.hist[b];
        for (int i = 0; i < 3; i++) {
            psnr[i] += o.psnr[i];
            color_r[i] += o.color_r[i];
            color_d[i] += o.color_d[i];
        }
        cframes += o.cframes;
        vmaf += o.vmaf;
        vmaf_min = std::min(vmaf_min, o.vmaf_min);
        vframes += o.vframes;
    }
    // A frame's PSNR (y, cb, cr) and the mean colours of the source and the encode (src/main.cpp:594, :601-605).
    void add_chroma(const double *p, const double *cr, const double *cd)
    {
        for (int i = 0; i < 3; i++) {
            psnr[i] += p[i];
            color_r[i] += cr[i];
            color_d[i] += cd[i];
        }
        cframes++;
    }
    void add_vmaf(double v)
    {
        vmaf += v;
        vmaf_min = std::min(vmaf_min, v);
        vframes++;
    }
    // A frame's exact totals ([all 6, structured 6] window sums, oriented totals all_r, all_d, st_r, st_d).
    void add_frame(double score, double score_src, const long long *ti, const double *to)
    {
        frames++;
        cambi += score;
        cambi_source += score_src;
        const double div[2] = {81.0, 6561.0};
        for (int key = 0; key < 2; key++) {
            all_r[key] += (double)ti[key] / div[key];
            all_d[key] += (double)ti[2 + key] / div[key];
            st_r[key] += (double)ti[6 + key] / div[key];
            st_d[key] += (double)ti[8 + key] / div[key];
            sh_all_r[key] += (double)ti[key] / div[key];
            sh_all_x[key] += (double)ti[4 + key] / div[key];
            sh_st_r[key] += (double)ti[6 + key] / div[key];
            sh_st_x[key] += (double)ti[10 + key] / div[key];
        }
        all_r[2] += to[0];
        all_d[2] += to[1];
        st_r[2] += to[2];
        st_d[2] += to[3];
    }
    void add_lasting(const long long *tl)
    {
        own += (double)tl[0] / 78732.0;
        mine += (double)tl[1] / 78732.0;
    }
    void fill(es_segment *o) const;
};

inline double clamp100(double v) { return std::min(std::max(v, 0.0), 100.0); }

inline void Accum::fill(es_segment *o) const
{
    o->frames = frames;
    o->cambi = cambi;
    o->cambi_source = cambi_source;
    for (int i = 0; i < 3; i++) {
        o->all_r[i] = all_r[i];
        o->all_d[i] = all_d[i];
        o->st_r[i] = st_r[i];
        o->st_d[i] = st_d[i];
    }
    for (int i = 0; i < 2; i++) {
        o->sh_all_r[i] = sh_all_r[i];
        o->sh_all_x[i] = sh_all_x[i];
        o->sh_st_r[i] = sh_st_r[i];
        o->sh_st_x[i] = sh_st_x[i];
    }
    o->own = own;
    o->mine = mine;
    uint64_t n[6];
    double med[6], ten[6];
    for (int k = 0; k < 6; k++) {
        n[k] = count(k);
        med[k] = hist_quantile_host(hist.data() + (size_t)k * kHistBins, n[k], 0.5);
        ten[k] = hist_quantile_host(hist.data() + (size_t)k * kHistBins, n[k], 0.1);
    }
    for (int k = 0; k < 3; k++) {
        o->n_amount[k] = n[k];
        o->median_amount[k] = med[k];
        o->tenth_amount[k] = ten[k];
    }
    for (int k = 0; k < 2; k++) {
        o->n_matching[k] = n[3 + k];
        o->median_matching[k] = med[3 + k];
        o->tenth_matching[k] = ten[3 + k];
    }
    o->n_lasting = n[5];
    o->median_lasting = med[5];
    o->tenth_lasting = ten[5];
    for (int i = 0; i < 3; i++) {
        o->psnr[i] = psnr[i];
        o->color_source[i] = color_r[i];
        o->color_encode[i] = color_d[i];
    }
    o->chroma_frames = cframes;
    o->vmaf = vmaf;
    o->vmaf_min = vmaf_min;
    o->vmaf_frames = vframes;
    // Accum::score (src/stats.hpp:70-82); fidelity only when every frame has its VMAF
    o->fidelity = frames && vframes == frames ? clamp100(vmaf / (double)frames) : NAN;
    o->texture = n[5] >= 100 ? clamp100(med[5]) / 2 + clamp100(ten[5]) / 2 : NAN;
    o->banding = frames ? clamp100(100.0 * (1 - (cambi - cambi_source) / (double)frames / 5.0)) : NAN;
    const double kept = all_r[0] > 0 ? 100.0 * std::sqrt(all_d[0] / all_r[0]) : NAN;
    o->grain = clamp100(100.0 - std::fabs(kept - 100.0));
    if (std::isnan(kept)) o->grain = NAN;
    const double v[4] = {o->fidelity, o->texture, o->banding, o->grain}, w[4] = {0.30, 0.30, 0.25, 0.15};
    double total = 0, weight = 0;
    for (int i = 0; i < 4; i++)
        if (!std::isnan(v[i])) {
            total += v[i] * w[i];
            weight += w[i];
        }
    o->score = weight > 0 ? total / weight : NAN;
}

// The window figures (src/kernels.cu:421-425) from the exact sums.
inline void window_values_host(const long long *z, double *v)
{
    v[0] = (double)z[0] / 81.0;
    v[1] = (double)z[1] / 6561.0;
    v[2] = (double)z[2] / 81.0;
    v[3] = (double)z[3] / 6561.0;
    v[4] = (double)z[4] / 81.0;
    v[5] = (double)z[5] / 6561.0;
    for (int p = 0; p < 2; p++) {
        const double d = (double)(z[6 + 3 * p] - z[7 + 3 * p]), x = (double)z[8 + 3 * p];
        v[6 + p] = std::sqrt(d * d + 4.0 * (x * x)) / 81.0;
    }
}

// The frame part of collect on the host (src/main.cpp:606-629): numpy's linear 0.8 quantile of the oriented
// energies as the reference evaluates it (src/main.cpp:279-289: float keys, the difference of the two order
// statistics in float, the interpolation in double rounded to float), the mask (each window's oriented energy
// rounded once to float against it), histogram entries and the canonical frame totals.
inline double collect_frame_host(Accum &acc, const long long *sums, int windows, uint8_t *mask, long long *ti, double *to)
{
    std::vector<double> v((size_t)windows * 8);
    std::vector<float> orient(windows), keys;
    for (int i = 0; i < windows; i++) {
        window_values_host(sums + (size_t)i * 12, &v[(size_t)i * 8]);
        orient[i] = (float)v[(size_t)i * 8 + 6];
    }
    keys = orient;
    const double pos = 0.8 * (double)(windows - 1);
    const size_t lo = (size_t)pos;
    float q = 0.0f;
    if (windows > 0) {
        std::nth_element(keys.begin(), keys.begin() + lo, keys.end());
        q = keys[lo];
        if (lo + 1 < keys.size()) {
            const float a = q, b = *std::min_element(keys.begin() + lo + 1, keys.end());
            const float d = b - a;
            q = (float)((double)a + (double)d * (pos - (double)lo));
        }
    }
    for (int k = 0; k < 12; k++) ti[k] = 0;
    double tree[4][256];
    for (int k = 0; k < 4; k++) to[k] = 0.0;
    for (int b = 0; b < windows; b += 256) {
        for (int t = 0; t < 256; t++) {
            const int i = b + t;
            double o[4] = {0, 0, 0, 0};
            if (i < windows) {
                const double *s = &v[(size_t)i * 8];
                const long long *z = sums + (size_t)i * 12;
                const bool m = orient[i] >= q && orient[i] > 0.0f;
                mask[i] = m;
                for (int k = 0; k < 6; k++) {
                    ti[k] += z[k];
                    if (m) ti[6 + k] += z[k];
                }
                o[0] = s[6];
                o[1] = s[7];
                o[2] = m ? s[6] : 0.0;
                o[3] = m ? s[7] : 0.0;
                if (m) {
                    const double er[3] = {s[0], s[1], s[6]}, ed[3] = {s[2], s[3], s[7]};
                    for (int k = 0; k < 3; k++) acc.add(k, 100.0 * std::sqrt(ed[k] / std::max(er[k], 1e-9)));
                    for (int k = 0; k < 2; k++) acc.add(3 + k, 100.0 * s[4 + k] / std::max(er[k], 1e-9));
                }
            }
            for (int k = 0; k < 4; k++) tree[k][t] = o[k];
        }
        for (int s = 128; s > 0; s >>= 1)
            for (int t = 0; t < s; t++)
                for (int k = 0; k < 4; k++) tree[k][t] += tree[k][t + s];
        for (int k = 0; k < 4; k++) to[k] += tree[k][0];
    }
    return q;
}

// The lasting-texture part of collect (src/main.cpp:630-641) from the plain form's 5 sums per window.
inline void collect_lasting_host(Accum &acc, const long long *psums, const uint8_t *cmask, int windows, long long *tl)
{
    tl[0] = tl[1] = tl[2] = 0;
    for (int i = 0; i < windows; i++) {
        const long long *p = psums + (size_t)i * 5;
        const double v0 = (double)p[0] / 78732.0, v1 = (double)p[1] / 78732.0;
        const double v2 = (double)p[2] / 236196.0, v3 = (double)p[3] / 236196.0, v4 = (double)p[4] / 236196.0;
        const double agree = v2 / std::sqrt(std::max(v3 * v4, 1e-9));
        if (cmask[i] && agree > 0.5 && v0 > 0) {
            tl[0] += p[0];
            tl[1] += p[1];
            tl[2]++;
            acc.add(5, 100.0 * v1 / v0);
        }
    }
}

}  // namespace es
