// CambiCacheFile: the on-disk source-CAMBI cache (src/cache.cpp:14-88).  The file format is the reference's: an
// append-only sequence of 16-byte records {int64 pts, double score} in host byte order, named by a 64-bit FNV-1a
// hash (16 hex digits, ".cambi") of the key "cambi <version> <source size> <x0> <y0> <w> <h> <depth>" followed by the
// source's first and last mebibyte.  A record cut short by an interrupted run is dropped when the file is opened.
#include "cache.hpp"
#include "host_common.hpp"

#include <cstdio>
#include <cstdlib>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

namespace es {

uint64_t CambiCacheFile::fnv1a(uint64_t h, const void *data, size_t n)
{
    const unsigned char *p = (const unsigned char *)data;
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 0x100000001b3ull;  // the 64-bit FNV prime
    }
    return h;
}

CambiCacheFile::CambiCacheFile(const std::string &path, int x0, int y0, int w, int h, int depth, const std::string &dir_in)
{
    std::string dir = dir_in;
    if (dir.empty()) {
        const char *xdg = std::getenv("XDG_CACHE_HOME"), *home = std::getenv("HOME");
        if (xdg && *xdg) dir = xdg;
        else if (home) dir = std::string(home) + "/.cache";
        else return;
        mkdir(dir.c_str(), 0755);
        dir += "/encode-score";
    }
    mkdir(dir.c_str(), 0755);
    FILE *src = std::fopen(path.c_str(), "rb");
    if (!src) return;
    std::fseek(src, 0, SEEK_END);
    const long size = std::ftell(src);
    char key[112];
    const int len = std::snprintf(key, sizeof key, "cambi %d %ld %d %d %d %d %d", kVersion, size, x0, y0, w, h, depth);
    uint64_t hash = fnv1a(0xcbf29ce484222325ull, key, (size_t)len);  // the 64-bit FNV offset basis
    std::vector<unsigned char> buf(kSample);
    const long last = size > (long)kSample ? size - (long)kSample : 0;
    for (long at : {0L, last}) {
        std::fseek(src, at, SEEK_SET);
        const size_t got = std::fread(buf.data(), 1, kSample, src);
        hash = fnv1a(hash, buf.data(), got);
    }
    std::fclose(src);
    char name[40];
    std::snprintf(name, sizeof name, "/%016llx.cambi", (unsigned long long)hash);
    file_ = dir + name;
    fd_ = open(file_.c_str(), O_RDWR | O_CREAT | O_APPEND, 0644);
    if (fd_ < 0) return;
    struct stat st;
    if (fstat(fd_, &st) != 0) {
        close(fd_);
        fd_ = -1;
        return;
    }
    const off_t whole = st.st_size - st.st_size % (off_t)sizeof(Record);
    if (whole != st.st_size && ftruncate(fd_, whole) != 0) {
        close(fd_);
        fd_ = -1;
        return;
    }
    std::vector<Record> recs((size_t)whole / sizeof(Record));
    if (whole > 0 && pread(fd_, recs.data(), (size_t)whole, 0) == (ssize_t)whole)
        for (const Record &r : recs) scores_[r.pts] = r.score;
}

CambiCacheFile::~CambiCacheFile()
{
    flush();
    if (fd_ >= 0) close(fd_);
}

bool CambiCacheFile::find(int64_t pts, double *score) const
{
    if (pts == INT64_MIN) return false;  // no timestamp (AV_NOPTS_VALUE)
    const auto it = scores_.find(pts);
    if (it == scores_.end()) return false;
    *score = it->second;
    return true;
}

void CambiCacheFile::add(int64_t pts, double score)
{
    if (fd_ < 0 || pts == INT64_MIN) return;
    scores_[pts] = score;
    pending_.push_back(Record{pts, score});
    if (pending_.size() >= kFlush) flush();
}

void CambiCacheFile::flush()
{
    if (fd_ >= 0 && !pending_.empty() && write(fd_, pending_.data(), pending_.size() * sizeof(Record)) < 0) {
        close(fd_);
        fd_ = -1;
    }
    pending_.clear();
}

}  // namespace es

using namespace es;

struct es_cambi_cache_s {
    CambiCacheFile f;
    es_cambi_cache_s(const char *path, int x0, int y0, int w, int h, int depth, const char *dir)
        : f(path, x0, y0, w, h, depth, dir ? dir : "") {}
};

extern "C" {

int es_cambi_cache_open(const char *path, int x0, int y0, int w, int h, int depth, const char *dir, es_cambi_cache **out)
{
    return guard([&] {
        require(path && out, "null argument");
        *out = new es_cambi_cache_s(path, x0, y0, w, h, depth, dir);
    });
}

int es_cambi_cache_find(const es_cambi_cache *c, int64_t pts, double *score, int *found)
{
    return guard([&] {
        require(c && score && found, "null argument");
        *found = c->f.find(pts, score) ? 1 : 0;
    });
}

int es_cambi_cache_add(es_cambi_cache *c, int64_t pts, double score)
{
    return guard([&] {
        require(c != nullptr, "null cache");
        c->f.add(pts, score);
    });
}

const char *es_cambi_cache_file(const es_cambi_cache *c) { return c ? c->f.file().c_str() : ""; }

int es_cambi_cache_close(es_cambi_cache *c)
{
    return guard([&] { delete c; });
}

}  // extern "C"
