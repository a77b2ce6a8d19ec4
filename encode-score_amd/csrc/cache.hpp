// The source-CAMBI store: a source's per-frame CAMBI kept on disk between runs, so that a later run against the same
// source (and crop and depth) reads the scores instead of computing them.  Replaces class CambiCache (src/cache.hpp,
// src/cache.cpp); the file format, the naming and the location are the reference's, the implementation is this
// library's own.
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace es {

// Incremental 64-bit FNV-1a (Fowler/Noll/Vo: offset basis 0xcbf29ce484222325, prime 0x100000001b3).
class Fnv1a64 {
public:
    void update(const void *data, size_t n);
    uint64_t value() const { return h_; }

private:
    uint64_t h_ = 0xcbf29ce484222325ull;
};

class SourceCambiStore {
public:
    // The CAMBI version in the store's key; the reference writes 1.  Scores of another CAMBI implementation are
    // never read as this one's, since libvmaf, the reference's pin, is absent here.
    static constexpr int kVersion = 355;
    static constexpr size_t kSampleBytes = size_t(1) << 20;  // the source's first and last MiB enter the name
    static constexpr size_t kBatch = 240;                    // records appended to the file at a time
    static constexpr size_t kRecordBytes = 16;               // {int64 pts, double score}, host byte order

    // The store of source `path` cropped to w x h at (x0, y0) and scored at `depth` bits, in `dir` (empty:
    // $XDG_CACHE_HOME/encode-score, else $HOME/.cache/encode-score).  When the source cannot be read or the directory
    // cannot be used, the store is inert: it finds nothing and keeps nothing.
    SourceCambiStore(const std::string &path, int x0, int y0, int w, int h, int depth, const std::string &dir);
    ~SourceCambiStore();
    SourceCambiStore(const SourceCambiStore &) = delete;
    SourceCambiStore &operator=(const SourceCambiStore &) = delete;

    bool lookup(int64_t pts, double *score) const;
    void keep(int64_t pts, double score);
    void write_pending();
    const std::string &file() const { return file_; }

    // The file name ("<16 hex digits>.cambi") of a source's store, or "" when the source cannot be read.
    static std::string name_of(const std::string &path, int x0, int y0, int w, int h, int depth);

private:
    void load();
    std::string file_;            // empty when inert
    std::map<int64_t, double> known_;
    std::vector<unsigned char> unwritten_;
};

}  // namespace es
