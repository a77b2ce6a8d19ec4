// One CAMBI instance of a picture size: the per-picture form of Cambi (src/cambi.hpp, src/cambi.cpp) over the gfx950
// kernels, with its buffers, its own stream and the select on the device.
#pragma once
#include "host_common.hpp"

namespace es {

class CambiUnit {
public:
    // c-value tiles of tile_h rows (0: the default), tile_h0 at full resolution (0: tile_h)
    CambiUnit(int w, int h, int depth, int window_base, bool keep_cvalues, int tile_h = 0, int tile_h0 = 0);
    ~CambiUnit();
    CambiUnit(const CambiUnit &) = delete;
    CambiUnit &operator=(const CambiUnit &) = delete;
    // Reads the plane in stream order on `caller`; everything after the first kernel runs on the own stream.
    // the picture is read in place: sample = (word >> rshift) << shift, at `bytes` per sample
    void launch(const void *img, int pitch, int bytes, int shift, hipStream_t caller, int rshift = 0);
    double finish();
    void scales(es_cambi_scale out[kScales]) const;
    void plane(int s, uint16_t *host) const;
    void cvalues(int s, float *host) const;
    const CambiInfo &info() const { return info_; }
    // the c-value kernel of the next launches: window counts on the matrix cores (mfma) or by prefix scans, with an
    // LDS layout of `bmax` (32 or 64) numbers per pass; the results are the same
    void set_method(bool mfma, int bmax);

private:
    struct Results {
        double score;
        unsigned long long sums[kScales];
        unsigned int counts[kScales];
        TopkJob jobs[kScales];
    };
    CambiInfo info_;
    bool keep_;
    int th_[kScales], tiles_ = 0, base_tiles_ = 0;
    bool mfma_ = true;
    int bmax_ = 32;
    DevBuf consts_, planes_, cvals_, list_, segs_, sums_, counts_, jobs_, hist_, hsum_, pic_, first_, score_;
    long long loff_[kScales] = {};
    int toff_[kScales] = {};
    HostBuf stage_, results_;
    hipStream_t own_ = nullptr;
    hipEvent_t copied_ = nullptr;
    bool pending_ = false;
};

}  // namespace es
