// es_cambi_*: the per-picture CAMBI object (replaces class Cambi, src/cambi.hpp:9-35, src/cambi.cpp:31-184).
#include "cambi_host.hpp"

#include <algorithm>

namespace es {

static constexpr int kDefaultTileH = 32;
static constexpr int kTopkBlocks = 64;

CambiUnit::CambiUnit(int w, int h, int depth, int window_base, bool keep_cvalues, int tile_h, int tile_h0)
    : info_(cambi_info(w, h, depth, window_base)), keep_(keep_cvalues)
{
    if (tile_h > kMaxTileH || tile_h0 > kMaxTileH) throw Error(ES_EINVAL, "tile_h must be at most 4096");
    cv_tile_heights(tile_h > 0 ? tile_h : kDefaultTileH, tile_h0, th_);
    ensure_init();
    const size_t total = (size_t)info_.off[kScales];
    CambiConsts k{};
    k.pad = info_.pad;
    k.band_base = info_.band_base;
    k.band_size = info_.band_size;
    k.vlt = info_.vlt;
    k.mask_index = info_.mask_index;
    k.window = info_.window;
    for (int d = 0; d < kDiffs; d++) {
        k.tvi[d] = info_.tvi[d];
        k.weight[d] = info_.weight[d];
    }
    consts_.alloc(sizeof k);
    hip_check(hipMemcpy(consts_.p, &k, sizeof k, hipMemcpyHostToDevice), "cambi constants");
    planes_.alloc(total * 2);
    if (keep_) cvals_.alloc(total * 4);
    int ntiles = 0;
    const long long cap = cambi_list_layout(info_.sw, info_.sh, th_, loff_, toff_, &ntiles);
    list_.alloc((size_t)cap * 4);
    segs_.alloc((size_t)ntiles * 4);
    sums_.alloc(kScales * 8);
    counts_.alloc(kScales * 4);
    hist_.alloc((size_t)kScales * kTopkBlocks * 512 * 4);
    hsum_.alloc((size_t)kScales * kTopkBlocks * 512 * 8);
    TopkJob jobs[kScales] = {};
    for (int s = 0; s < kScales; s++) {
        jobs[s].list = list_.as<float>() + loff_[s];
        jobs[s].seg_count = segs_.as<unsigned>() + toff_[s];
        jobs[s].nseg = cvalue_tiles(info_.sw[s], info_.sh[s], th_[s]);
        jobs[s].seg_cap = kCvTileW * th_[s];
        jobs[s].count = counts_.as<unsigned>() + s;
        jobs[s].k = (unsigned long long)info_.k[s];
        jobs[s].hist = hist_.as<unsigned>() + (size_t)s * kTopkBlocks * 512;
        jobs[s].hsum = hsum_.as<unsigned long long>() + (size_t)s * kTopkBlocks * 512;
    }
    jobs_.alloc(sizeof jobs);
    hip_check(hipMemcpy(jobs_.p, jobs, sizeof jobs, hipMemcpyHostToDevice), "top-k jobs");
    // the c-value tiles, coarsest scale first (slot k holds scale kScales - 1 - k)
    int first[kScales + 1];
    first[0] = 0;
    for (int k = 0; k < kScales; k++)
        first[k + 1] = first[k] + cvalue_tiles(info_.sw[kScales - 1 - k], info_.sh[kScales - 1 - k], th_[kScales - 1 - k]);
    tiles_ = first[kScales];
    base_tiles_ = base_tiles(w, h);
    first_.alloc(sizeof first);
    hip_check(hipMemcpy(first_.p, first, sizeof first, hipMemcpyHostToDevice), "tile table");
    pic_.alloc(sizeof(CambiPic));
    score_.alloc(8);
    stage_.alloc(sizeof(CambiPic));
    results_.alloc(sizeof(Results));
    hip_check(hipStreamCreateWithFlags(&own_, hipStreamNonBlocking), "creating a stream");
    hip_check(hipEventCreateWithFlags(&copied_, hipEventDisableTiming), "creating an event");
}

CambiUnit::~CambiUnit()
{
    if (own_) {
        (void)hipStreamSynchronize(own_);
        (void)hipStreamDestroy(own_);
    }
    if (copied_) (void)hipEventDestroy(copied_);
}

void CambiUnit::launch(const void *img, int pitch, int bytes, int shift, hipStream_t caller, int rshift)
{
    require(img != nullptr, "null picture");
    require(bytes == 1 || bytes == 2, "samples must be 1 or 2 bytes");
    require(pitch >= info_.w, "pitch below the width");
    require(shift >= 0 && shift <= 8, "bad shift");
    require(rshift >= 0 && rshift <= 8 && (bytes == 2 || rshift == 0), "bad right shift");
    require(!pending_, "launch while the previous picture is not finished");
    CambiPic *p = stage_.as<CambiPic>();
    std::memset(p, 0, sizeof *p);
    p->in.ptr = img;
    p->in.pitch = pitch;
    p->in.bytes = bytes;
    p->in.shift = info_.depth == 8 ? 2 + shift : shift;
    p->in.from8 = info_.depth == 8;
    p->in.rshift = rshift;
    for (int s = 0; s < kScales; s++) {
        p->sw[s] = info_.sw[s];
        p->sh[s] = info_.sh[s];
        p->off[s] = info_.off[s];
        p->loff[s] = loff_[s];
        p->toff[s] = toff_[s];
        p->th[s] = th_[s];
    }
    p->planes = planes_.as<uint16_t>();
    p->cvals = keep_ ? cvals_.as<float>() : nullptr;
    p->list = list_.as<float>();
    p->seg_count = segs_.as<unsigned>();
    p->sums = sums_.as<unsigned long long>();
    p->counts = counts_.as<unsigned>();
    p->consts = consts_.as<CambiConsts>();
    const CambiPic *dpic = pic_.as<CambiPic>();
    hip_check(hipMemcpyAsync(pic_.p, p, sizeof *p, hipMemcpyHostToDevice, caller), "cambi descriptor");
    launch_cambi_base(dpic, nullptr, 1, base_tiles_, caller);
    hip_check(hipEventRecord(copied_, caller), "cambi");
    hip_check(hipStreamWaitEvent(own_, copied_, 0), "cambi");
    hip_check(hipMemsetAsync(sums_.p, 0, sums_.bytes, own_), "clearing sums");
    hip_check(hipMemsetAsync(counts_.p, 0, counts_.bytes, own_), "clearing counts");
    if (keep_) hip_check(hipMemsetAsync(cvals_.p, 0, cvals_.bytes, own_), "clearing c-values");
    for (int s = 1; s < kScales; s++) launch_cambi_scale(dpic, 1, scale_strips(info_.sw[s], info_.sh[s]), s, own_);
    // one block per tile (window counts on the matrix cores unless set otherwise)
    launch_cambi_cvalues(dpic, first_.as<int>(), kScales, tiles_, 0, nullptr, 0, bmax_, own_, mfma_ ? 1 : 0);
    for (int pass = 0; pass < 3; pass++) launch_topk_pass(jobs_.as<TopkJob>(), kScales, pass, kTopkBlocks, own_);
    launch_cambi_finish(dpic, jobs_.as<TopkJob>(), 1, score_.as<double>(), own_);
    Results *r = results_.as<Results>();
    hip_check(hipMemcpyAsync(&r->score, score_.p, 8, hipMemcpyDeviceToHost, own_), "reading the score");
    hip_check(hipMemcpyAsync(r->sums, sums_.p, sums_.bytes, hipMemcpyDeviceToHost, own_), "reading sums");
    hip_check(hipMemcpyAsync(r->counts, counts_.p, counts_.bytes, hipMemcpyDeviceToHost, own_), "reading counts");
    hip_check(hipMemcpyAsync(r->jobs, jobs_.p, jobs_.bytes, hipMemcpyDeviceToHost, own_), "reading the select");
    pending_ = true;
}

void CambiUnit::set_method(bool mfma, int bmax)
{
    require(bmax == 32 || bmax == 64, "the c-value layout holds 32 or 64 numbers per pass");
    require(!pending_, "method change while a picture is not finished");
    mfma_ = mfma;
    bmax_ = bmax;
}

double CambiUnit::finish()
{
    require(pending_, "finish without launch");
    hip_check(hipStreamSynchronize(own_), "cambi");
    pending_ = false;
    return results_.as<Results>()->score;
}

void CambiUnit::scales(es_cambi_scale out[kScales]) const
{
    const Results *r = results_.as<Results>();
    for (int s = 0; s < kScales; s++) {
        es_cambi_scale &o = out[s];
        const unsigned long long k = (unsigned long long)info_.k[s];
        o.sum_fx = r->sums[s];
        o.count = r->counts[s];
        o.k = k;
        if (o.count <= k) {
            o.thr_bits = 0;
            o.above = 0;
            o.above_fx = 0;
            o.pooled = (double)o.sum_fx * (1.0 / 16777216.0) / (double)k;
        } else {
            o.thr_bits = r->jobs[s].prefix + 0x3F000000u;
            o.above = r->jobs[s].above;
            o.above_fx = r->jobs[s].above_fx;
            float thr;
            std::memcpy(&thr, &o.thr_bits, 4);
            o.pooled = ((double)o.above_fx * (1.0 / 16777216.0) + (double)(k - o.above) * (double)thr) / (double)k;
        }
    }
}

void CambiUnit::plane(int s, uint16_t *host) const
{
    require(s >= 0 && s < kScales && host, "bad scale or output");
    hip_check(hipStreamSynchronize(own_), "cambi");
    hip_check(hipMemcpy(host, planes_.as<uint16_t>() + info_.off[s], (size_t)info_.n[s] * 2, hipMemcpyDeviceToHost), "reading a plane");
    // the device holds codes; the caller gets value | mask << 15
    const unsigned bias = plane_bias(info_.band_base);
    for (long long i = 0; i < info_.n[s]; i++) host[i] = (uint16_t)(host[i] + bias);
}

void CambiUnit::cvalues(int s, float *host) const
{
    require(s >= 0 && s < kScales && host, "bad scale or output");
    require(keep_, "this instance does not keep its c-values");
    hip_check(hipStreamSynchronize(own_), "cambi");
    hip_check(hipMemcpy(host, cvals_.as<float>() + info_.off[s], (size_t)info_.n[s] * 4, hipMemcpyDeviceToHost), "reading c-values");
}

}  // namespace es

using namespace es;

struct es_cambi_s {
    CambiUnit unit;
    es_cambi_s(int w, int h, int depth, int base, bool keep) : unit(w, h, depth, base, keep) {}
};

extern "C" {

int es_cambi_create_ex(int w, int h, int depth, int window_base, int keep_cvalues, es_cambi **out)
{
    return guard([&] {
        require(out != nullptr, "null argument");
        *out = new es_cambi_s(w, h, depth, window_base, keep_cvalues != 0);
    });
}

int es_cambi_create(int w, int h, int depth, es_cambi **out) { return es_cambi_create_ex(w, h, depth, 65, 0, out); }

int es_cambi_set_method(es_cambi *c, int cv_mfma, int cv_bmax)
{
    return guard([&] {
        require(c != nullptr, "null instance");
        c->unit.set_method(cv_mfma != 0, cv_bmax);
    });
}

int es_cambi_destroy(es_cambi *c)
{
    return guard([&] { delete c; });
}

int es_cambi_launch_ex(es_cambi *c, const void *img, int pitch, int bytes, int shift, es_stream stream)
{
    return guard([&] {
        require(c != nullptr, "null instance");
        c->unit.launch(img, pitch, bytes, shift, hs(stream));
    });
}

int es_cambi_launch(es_cambi *c, const void *img, int pitch, es_stream stream) { return es_cambi_launch_ex(c, img, pitch, 2, 0, stream); }

int es_cambi_finish(es_cambi *c, double *score)
{
    return guard([&] {
        require(c && score, "null argument");
        *score = c->unit.finish();
    });
}

int es_cambi_scales(es_cambi *c, es_cambi_scale out[5])
{
    return guard([&] {
        require(c && out, "null argument");
        c->unit.scales(out);
    });
}

int es_cambi_plane(es_cambi *c, int scale, uint16_t *host_out)
{
    return guard([&] {
        require(c != nullptr, "null instance");
        c->unit.plane(scale, host_out);
    });
}

int es_cambi_cvalues(es_cambi *c, int scale, float *host_out)
{
    return guard([&] {
        require(c != nullptr, "null instance");
        c->unit.cvalues(scale, host_out);
    });
}

}  // extern "C"
