// The source's per-frame CAMBI kept on disk between runs (class CambiCache, src/cache.hpp, src/cache.cpp): it
// depends on the source alone, so a later run against the same source reads it instead of computing it again.
#pragma once
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

namespace es {

class CambiCacheFile {
public:
    // Version of this library's CAMBI scores in the cache key: different from the reference's (1), so files written
    // by another CAMBI implementation are never read as this one's.
    static constexpr int kVersion = 355;
    static constexpr size_t kSample = 1 << 20;  // the first and last mebibyte of the source are hashed
    static constexpr size_t kFlush = 240;       // records written per append

    // The cache of `path` cropped to w x h at (x0, y0), scored at `depth` bits, in `dir` (empty: $XDG_CACHE_HOME or
    // $HOME/.cache, then /encode-score).  Inert (finds and keeps nothing) when the source or the directory cannot be
    // used, as the reference's is.
    CambiCacheFile(const std::string &path, int x0, int y0, int w, int h, int depth, const std::string &dir);
    ~CambiCacheFile();
    CambiCacheFile(const CambiCacheFile &) = delete;
    CambiCacheFile &operator=(const CambiCacheFile &) = delete;

    bool find(int64_t pts, double *score) const;
    void add(int64_t pts, double score);
    void flush();
    const std::string &file() const { return file_; }

    static uint64_t fnv1a(uint64_t h, const void *data, size_t n);

private:
    struct Record {
        int64_t pts;
        double score;
    };
    std::string file_;
    std::unordered_map<int64_t, double> scores_;
    std::vector<Record> pending_;
    int fd_ = -1;
};

}  // namespace es
