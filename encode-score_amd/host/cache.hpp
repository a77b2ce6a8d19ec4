// Drop-in for the reference's src/cache.hpp: the source's CAMBI per frame, kept on disk between runs
// (es_cambi_cache_*).  The same constructor, find and add; the files have the reference's format and naming with
// this library's own version in the key.
#pragma once
#include "gpu.hpp"

#include <cstdint>
#include <optional>
#include <string>

class CambiCache {
public:
    // The scores of `path` cropped to `w` by `h` at (`x0`, `y0`) and scored at `depth` bits (src/cache.hpp:13-15).
    CambiCache(const std::string &path, int x0, int y0, int w, int h, int depth)
    {
        es_check(es_cambi_cache_open(path.c_str(), x0, y0, w, h, depth, nullptr, &c_), "source CAMBI cache");
    }
    ~CambiCache()
    {
        if (c_) es_cambi_cache_close(c_);
    }
    CambiCache(const CambiCache &) = delete;
    CambiCache &operator=(const CambiCache &) = delete;
    // The score of the frame with presentation timestamp `pts`, if an earlier run computed it.
    std::optional<double> find(int64_t pts) const
    {
        double v = 0;
        int found = 0;
        es_check(es_cambi_cache_find(c_, pts, &v, &found), "source CAMBI cache");
        if (!found) return std::nullopt;
        return v;
    }
    void add(int64_t pts, double score) { es_check(es_cambi_cache_add(c_, pts, score), "source CAMBI cache"); }
    std::string file() const { return es_cambi_cache_file(c_); }

private:
    es_cambi_cache *c_ = nullptr;
};
