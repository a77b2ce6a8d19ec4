// The run's JSON report in the reference's format (src/main.cpp:693-777): the same field names in the same order and
// the same digits, so that readers of the reference's files read these unchanged.  A figure that could not be measured
// (no lasting texture in a short run, or VMAF, PSNR or colour not supplied) is written as null.
#include "report.hpp"

#include <cmath>
#include <string>

namespace es {

namespace {

// A streaming JSON emitter: it writes the separators itself (", " between members and elements, or "," in compact
// arrays), so the report below only names what it writes.
class JsonText {
public:
    explicit JsonText(FILE *out) : out_(out) {}

    JsonText &object() { return open('{', false); }
    JsonText &array(bool compact = false) { return open('[', compact); }
    JsonText &end()
    {
        std::fputc(levels_.back().closer, out_);
        levels_.pop_back();
        return *this;
    }
    // A member's name: the value written next belongs to it.
    JsonText &key(const char *name)
    {
        element();
        put_string(name);
        std::fputs(": ", out_);
        keyed_ = true;
        return *this;
    }
    // %.{digits}f, as the reference prints its figures; null when not finite.
    JsonText &fixed(double v, int digits)
    {
        element();
        if (std::isfinite(v)) std::fprintf(out_, "%.*f", digits, v);
        else std::fputs("null", out_);
        return *this;
    }
    JsonText &integer(long long v)
    {
        element();
        std::fprintf(out_, "%lld", v);
        return *this;
    }
    JsonText &shortest(double v)  // %g
    {
        element();
        std::fprintf(out_, "%g", v);
        return *this;
    }
    JsonText &text(const char *s)
    {
        element();
        put_string(s ? s : "");
        return *this;
    }
    JsonText &fixed_list(std::initializer_list<double> vs, int digits)
    {
        array();
        for (double v : vs) fixed(v, digits);
        return end();
    }

private:
    struct Level {
        char closer;
        bool compact, empty;
    };
    JsonText &open(char c, bool compact)
    {
        element();
        std::fputc(c, out_);
        levels_.push_back(Level{c == '{' ? '}' : ']', compact, true});
        return *this;
    }
    // Before a member or an element: the separator from the one before it (a member's value follows its name).
    void element()
    {
        if (keyed_) {
            keyed_ = false;
            return;
        }
        if (levels_.empty()) return;
        Level &l = levels_.back();
        if (!l.empty) std::fputs(l.compact ? "," : ", ", out_);
        l.empty = false;
    }
    // A JSON string: quotes and backslashes escaped, control characters as \u00XX (the reference writes paths as they
    // are, which is the same for ordinary paths).
    void put_string(const char *s)
    {
        std::fputc('"', out_);
        for (const unsigned char *p = (const unsigned char *)s; *p; p++) {
            if (*p == '"' || *p == '\\') std::fprintf(out_, "\\%c", *p);
            else if (*p < 0x20) std::fprintf(out_, "\\u%04x", *p);
            else std::fputc(*p, out_);
        }
        std::fputc('"', out_);
    }

    FILE *out_;
    std::vector<Level> levels_;
    bool keyed_ = false;
};

// Accum::kept (src/stats.hpp:67): the share of the source's energy the encode keeps, in per cent.
double kept_share(double d, double r) { return r > 0 ? 100.0 * std::sqrt(d / r) : NAN; }

// A per-frame figure's mean over a segment, or NaN unless every frame of the segment has it.
double frame_mean(double sum, uint64_t have, uint64_t frames) { return frames && have == frames ? sum / (double)frames : NAN; }

// A segment's figures (the reference's parts object, src/main.cpp:737-759).
void segment_figures(JsonText &j, const es_segment &a)
{
    const uint64_t n = a.frames;
    auto chroma = [&](double sum) { return frame_mean(sum, a.chroma_frames, n); };
    j.object();
    j.key("frames").integer((long long)n);
    j.key("score").fixed(a.score, 4);
    j.key("fidelity").fixed(a.fidelity, 4);
    j.key("texture").fixed(a.texture, 4);
    j.key("banding").fixed(a.banding, 4);
    j.key("grain").fixed(a.grain, 4);
    j.key("vmaf").fixed(frame_mean(a.vmaf, a.vmaf_frames, n), 6);
    j.key("vmaf_min").fixed(n && a.vmaf_frames == n ? a.vmaf_min : NAN, 6);
    j.key("psnr_y").fixed(chroma(a.psnr[0]), 4);
    j.key("psnr_cb").fixed(chroma(a.psnr[1]), 4);
    j.key("psnr_cr").fixed(chroma(a.psnr[2]), 4);
    j.key("cambi").fixed(a.cambi / (double)n, 6);
    j.key("cambi_source").fixed(a.cambi_source / (double)n, 6);
    j.key("lasting").object();
    j.key("windows").integer((long long)a.n_lasting);
    j.key("median").fixed(a.median_lasting, 3);
    j.key("worst_tenth").fixed(a.tenth_lasting, 3);
    j.key("total").fixed(a.own > 0 ? 100.0 * a.mine / a.own : NAN, 3);
    j.end();
    j.key("amount").object();
    j.key("all").fixed_list({kept_share(a.all_d[0], a.all_r[0]), kept_share(a.all_d[1], a.all_r[1]),
                             kept_share(a.all_d[2], a.all_r[2])}, 3);
    j.key("structure").fixed_list({kept_share(a.st_d[0], a.st_r[0]), kept_share(a.st_d[1], a.st_r[1]),
                                   kept_share(a.st_d[2], a.st_r[2])}, 3);
    j.key("median").fixed_list({a.median_amount[0], a.median_amount[1], a.median_amount[2]}, 3);
    j.key("worst_tenth").fixed_list({a.tenth_amount[0], a.tenth_amount[1], a.tenth_amount[2]}, 3);
    j.end();
    j.key("matching").object();
    j.key("all").fixed_list({100.0 * a.sh_all_x[0] / a.sh_all_r[0], 100.0 * a.sh_all_x[1] / a.sh_all_r[1]}, 3);
    j.key("structure").fixed_list({100.0 * a.sh_st_x[0] / a.sh_st_r[0], 100.0 * a.sh_st_x[1] / a.sh_st_r[1]}, 3);
    j.key("median").fixed_list({a.median_matching[0], a.median_matching[1]}, 3);
    j.key("worst_tenth").fixed_list({a.tenth_matching[0], a.tenth_matching[1]}, 3);
    j.end();
    j.key("color_source").fixed_list({chroma(a.color_source[0]), chroma(a.color_source[1]), chroma(a.color_source[2])}, 4);
    j.key("color_encode").fixed_list({chroma(a.color_encode[0]), chroma(a.color_encode[1]), chroma(a.color_encode[2])}, 4);
    j.end();
}

// A per-frame list (compact, as the reference writes them): n entries, null past the figures held or where not finite.
void per_frame(JsonText &j, const char *name, const std::vector<double> &v, size_t n, int digits)
{
    j.key(name).array(true);
    for (size_t i = 0; i < n; i++) j.fixed(i < v.size() ? v[i] : NAN, digits);
    j.end();
}

}  // namespace

void write_json(FILE *out, const es_run_info &info, long frames, const std::vector<ReportEncode> &encs)
{
    JsonText j(out);
    j.object();
    j.key("source").text(info.source);
    j.key("start").fixed(info.start, 6);
    j.key("offset").fixed(info.offset, 6);
    j.key("frames").integer(frames);
    j.key("fps").fixed(info.fps, 6);
    j.key("width").integer(info.width);
    j.key("height").integer(info.height);
    j.key("model").text(info.model);
    j.key("segment_seconds").shortest(info.segment_seconds);
    j.key("encodes").array();
    for (size_t e = 0; e < encs.size(); e++) {
        const ReportEncode &c = encs[e];
        j.object();
        j.key("path").text(info.paths ? info.paths[e] : "");
        j.key("whole");
        segment_figures(j, c.whole);
        j.key("segments").array();
        for (const es_segment &s : c.segments) segment_figures(j, s);
        j.end();
        const size_t n = c.cambi.size();
        j.key("per_frame").object();
        per_frame(j, "vmaf", c.vmaf, n, 6);
        per_frame(j, "cambi", c.cambi, n, 6);
        per_frame(j, "cambi_source", c.cambi_source, n, 6);
        per_frame(j, "psnr_y", c.psnr_y, n, 3);
        j.end();
        j.end();
    }
    j.end();
    j.end();
    std::fputc('\n', out);
}

}  // namespace es
