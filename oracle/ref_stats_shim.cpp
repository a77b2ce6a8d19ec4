// A C shim over the reference's own src/stats.hpp (compiled from the reference tree, never copied), so tests can pin
// the oracle's Histogram / Accum::score restatement against the reference's code.  Test infrastructure only.
#include "stats.hpp"

#include <cstdint>
#include <cstring>

extern "C" {

// Histogram::add of n percentages, then its bins (20001), count and quantile(q) for each of nq quantiles.
void ref_hist(const double *pcts, long n, uint64_t *bins, uint64_t *count, const double *qs, int nq, double *out)
{
    Histogram h;
    for (long i = 0; i < n; i++) h.add(pcts[i]);
    if (bins) std::memcpy(bins, h.bins.data(), h.bins.size() * sizeof(uint64_t));
    if (count) *count = h.n;
    for (int i = 0; i < nq; i++) out[i] = h.quantile(qs[i]);
}

// Accum::score of an accumulator with the given frames, totals and lasting-texture percentages.
void ref_score(uint64_t frames, double vmaf, double cambi, double cambi_source, double all_d0, double all_r0,
               const double *lasting, long nlasting, double *parts)
{
    Accum a;
    a.frames = frames;
    a.vmaf = vmaf;
    a.cambi = cambi;
    a.cambi_source = cambi_source;
    a.all_d[0] = all_d0;
    a.all_r[0] = all_r0;
    for (long i = 0; i < nlasting; i++) a.lasting.add(lasting[i]);
    Accum::Parts p = a.score();
    parts[0] = p.fidelity;
    parts[1] = p.texture;
    parts[2] = p.banding;
    parts[3] = p.grain;
    parts[4] = p.score;
}
}
