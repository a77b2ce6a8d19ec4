// SourceCambiStore (see cache.hpp).  On disk (the reference's format, src/cache.cpp:14-88): an append-only sequence of
// 16-byte records {int64 pts, double score} in host byte order, in a file named by the 64-bit FNV-1a hash of the key
// "cambi <version> <source size> <x0> <y0> <w> <h> <depth>" continued over the source's first and then last MiB (the
// whole file twice when it is shorter), as 16 hex digits and ".cambi".  A record cut short by an interrupted run is
// cut off when the store is opened; later records of the same timestamp replace earlier ones.
#include "cache.hpp"
#include "host_common.hpp"

#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>

namespace fs = std::filesystem;

namespace es {

void Fnv1a64::update(const void *data, size_t n)
{
    const unsigned char *p = static_cast<const unsigned char *>(data);
    for (const unsigned char *end = p + n; p != end; ++p) h_ = (h_ ^ *p) * 0x100000001b3ull;
}

namespace {

// The store's directory: `dir` as given, or the user's cache directory with "encode-score" under it; created when
// missing.  Empty when no directory can be named or made.
std::string store_dir(const std::string &dir)
{
    fs::path d;
    if (!dir.empty()) {
        d = dir;
    } else {
        const char *xdg = std::getenv("XDG_CACHE_HOME"), *home = std::getenv("HOME");
        if (xdg && *xdg) d = fs::path(xdg) / "encode-score";
        else if (home && *home) d = fs::path(home) / ".cache" / "encode-score";
        else return "";
    }
    std::error_code ec;
    fs::create_directories(d, ec);
    return fs::is_directory(d, ec) ? d.string() : "";
}

// Hashes `n` bytes of `in` from `offset` (fewer when the file ends first).
void hash_span(std::ifstream &in, uint64_t offset, size_t n, Fnv1a64 &h)
{
    std::vector<char> buf(n);
    in.clear();
    in.seekg((std::streamoff)offset);
    in.read(buf.data(), (std::streamsize)n);
    h.update(buf.data(), (size_t)in.gcount());
}

}  // namespace

std::string SourceCambiStore::name_of(const std::string &path, int x0, int y0, int w, int h, int depth)
{
    std::error_code ec;
    const uintmax_t size = fs::file_size(path, ec);
    std::ifstream in(path, std::ios::binary);
    if (ec || !in) return "";
    const std::string key = "cambi " + std::to_string(kVersion) + " " + std::to_string(size) + " " + std::to_string(x0) +
                            " " + std::to_string(y0) + " " + std::to_string(w) + " " + std::to_string(h) + " " +
                            std::to_string(depth);
    Fnv1a64 hash;
    hash.update(key.data(), key.size());
    hash_span(in, 0, kSampleBytes, hash);
    hash_span(in, size > kSampleBytes ? size - kSampleBytes : 0, kSampleBytes, hash);
    static const char digits[] = "0123456789abcdef";
    std::string name(16, '0');
    for (int i = 0; i < 16; i++) name[i] = digits[(hash.value() >> (60 - 4 * i)) & 15];
    return name + ".cambi";
}

SourceCambiStore::SourceCambiStore(const std::string &path, int x0, int y0, int w, int h, int depth, const std::string &dir)
{
    const std::string where = store_dir(dir);
    const std::string name = where.empty() ? "" : name_of(path, x0, y0, w, h, depth);
    if (name.empty()) return;  // inert
    file_ = (fs::path(where) / name).string();
    load();
}

// Reads the whole records of the file (cutting a partial last one off it).
void SourceCambiStore::load()
{
    std::error_code ec;
    if (!fs::exists(file_, ec)) return;
    const uintmax_t size = fs::file_size(file_, ec);
    if (ec) {
        file_.clear();
        return;
    }
    const uintmax_t whole = size - size % kRecordBytes;
    if (whole != size) {
        fs::resize_file(file_, whole, ec);
        if (ec) {
            file_.clear();
            return;
        }
    }
    std::ifstream in(file_, std::ios::binary);
    std::vector<unsigned char> raw((size_t)whole);
    if (!in.read(reinterpret_cast<char *>(raw.data()), (std::streamsize)whole)) {
        file_.clear();
        return;
    }
    for (size_t at = 0; at < raw.size(); at += kRecordBytes) {
        int64_t pts;
        double score;
        std::memcpy(&pts, &raw[at], 8);
        std::memcpy(&score, &raw[at + 8], 8);
        known_[pts] = score;
    }
}

SourceCambiStore::~SourceCambiStore() { write_pending(); }

bool SourceCambiStore::lookup(int64_t pts, double *score) const
{
    if (pts == INT64_MIN) return false;  // a frame without a timestamp (AV_NOPTS_VALUE)
    const auto it = known_.find(pts);
    if (it == known_.end()) return false;
    *score = it->second;
    return true;
}

void SourceCambiStore::keep(int64_t pts, double score)
{
    if (file_.empty() || pts == INT64_MIN) return;
    known_[pts] = score;
    unsigned char rec[kRecordBytes];
    std::memcpy(rec, &pts, 8);
    std::memcpy(rec + 8, &score, 8);
    unwritten_.insert(unwritten_.end(), rec, rec + kRecordBytes);
    if (unwritten_.size() >= kBatch * kRecordBytes) write_pending();
}

void SourceCambiStore::write_pending()
{
    if (file_.empty() || unwritten_.empty()) return;
    std::ofstream out(file_, std::ios::binary | std::ios::app);
    out.write(reinterpret_cast<const char *>(unwritten_.data()), (std::streamsize)unwritten_.size());
    out.flush();
    if (!out) file_.clear();  // the file cannot be written: keep nothing more
    unwritten_.clear();
}

}  // namespace es

using namespace es;

struct es_cambi_cache_s {
    SourceCambiStore store;
    es_cambi_cache_s(const char *path, int x0, int y0, int w, int h, int depth, const char *dir)
        : store(path, x0, y0, w, h, depth, dir ? dir : "") {}
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
        *found = c->store.lookup(pts, score) ? 1 : 0;
    });
}

int es_cambi_cache_add(es_cambi_cache *c, int64_t pts, double score)
{
    return guard([&] {
        require(c != nullptr, "null cache");
        c->store.keep(pts, score);
    });
}

const char *es_cambi_cache_file(const es_cambi_cache *c) { return c ? c->store.file().c_str() : ""; }

int es_cambi_cache_close(es_cambi_cache *c)
{
    return guard([&] { delete c; });
}

}  // extern "C"
