// The JSON report and the text summary of a run (src/main.cpp:693-777), field for field: the same names, order and
// digits, so that readers of the reference's files read these unchanged.  A figure that could not be measured (no
// lasting texture in a short run, or VMAF, PSNR or colour not supplied) is written as null.
#include "report.hpp"

#include <algorithm>
#include <cmath>
#include <string>

namespace es {

namespace {

std::string num(double v, int digits)
{
    if (!std::isfinite(v)) return "null";
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.*f", digits, v);
    return buf;
}

std::string list(std::initializer_list<double> vs, int digits)
{
    std::string out = "[";
    for (double v : vs) out += (out.size() > 1 ? ", " : "") + num(v, digits);
    return out + "]";
}

// A path as a JSON string; the reference writes paths as they are, which is the same for paths without quotes,
// backslashes or control characters.
std::string quoted(const char *s)
{
    std::string out = "\"";
    for (const char *p = s ? s : ""; *p; p++) {
        const unsigned char c = (unsigned char)*p;
        if (c == '"' || c == '\\') {
            out += '\\';
            out += (char)c;
        } else if (c < 0x20) {
            char buf[8];
            std::snprintf(buf, sizeof buf, "\\u%04x", c);
            out += buf;
        } else {
            out += (char)c;
        }
    }
    return out + "\"";
}

double kept(double d, double r) { return r > 0 ? 100.0 * std::sqrt(d / r) : NAN; }  // Accum::kept, src/stats.hpp:67

// The per-frame means of a segment's sums, or NaN when not every frame has them.
double mean_of(double sum, uint64_t have, uint64_t frames) { return have == frames && frames ? sum / (double)frames : NAN; }

// parts (src/main.cpp:737-759)
std::string parts(const es_segment &a)
{
    const uint64_t f = a.frames;
    const double vmaf_min = a.vmaf_frames == f && f ? a.vmaf_min : NAN;
    auto c = [&](double v) { return mean_of(v, a.chroma_frames, f); };
    return "{\"frames\": " + std::to_string(a.frames) + ", \"score\": " + num(a.score, 4) + ", \"fidelity\": " +
           num(a.fidelity, 4) + ", \"texture\": " + num(a.texture, 4) + ", \"banding\": " + num(a.banding, 4) +
           ", \"grain\": " + num(a.grain, 4) + ", \"vmaf\": " + num(mean_of(a.vmaf, a.vmaf_frames, f), 6) +
           ", \"vmaf_min\": " + num(vmaf_min, 6) + ", \"psnr_y\": " + num(c(a.psnr[0]), 4) + ", \"psnr_cb\": " +
           num(c(a.psnr[1]), 4) + ", \"psnr_cr\": " + num(c(a.psnr[2]), 4) + ", \"cambi\": " +
           num(a.cambi / (double)f, 6) + ", \"cambi_source\": " + num(a.cambi_source / (double)f, 6) +
           ", \"lasting\": {\"windows\": " + std::to_string(a.n_lasting) + ", \"median\": " + num(a.median_lasting, 3) +
           ", \"worst_tenth\": " + num(a.tenth_lasting, 3) + ", \"total\": " +
           num(a.own > 0 ? 100.0 * a.mine / a.own : NAN, 3) + "}" + ", \"amount\": {\"all\": " +
           list({kept(a.all_d[0], a.all_r[0]), kept(a.all_d[1], a.all_r[1]), kept(a.all_d[2], a.all_r[2])}, 3) +
           ", \"structure\": " + list({kept(a.st_d[0], a.st_r[0]), kept(a.st_d[1], a.st_r[1]), kept(a.st_d[2], a.st_r[2])}, 3) +
           ", \"median\": " + list({a.median_amount[0], a.median_amount[1], a.median_amount[2]}, 3) +
           ", \"worst_tenth\": " + list({a.tenth_amount[0], a.tenth_amount[1], a.tenth_amount[2]}, 3) + "}" +
           ", \"matching\": {\"all\": " +
           list({100.0 * a.sh_all_x[0] / a.sh_all_r[0], 100.0 * a.sh_all_x[1] / a.sh_all_r[1]}, 3) +
           ", \"structure\": " + list({100.0 * a.sh_st_x[0] / a.sh_st_r[0], 100.0 * a.sh_st_x[1] / a.sh_st_r[1]}, 3) +
           ", \"median\": " + list({a.median_matching[0], a.median_matching[1]}, 3) +
           ", \"worst_tenth\": " + list({a.tenth_matching[0], a.tenth_matching[1]}, 3) + "}" + ", \"color_source\": " +
           list({c(a.color_source[0]), c(a.color_source[1]), c(a.color_source[2])}, 4) + ", \"color_encode\": " +
           list({c(a.color_encode[0]), c(a.color_encode[1]), c(a.color_encode[2])}, 4) + "}";
}

void frame_list(FILE *out, const std::vector<double> &v, size_t n, const char *fmt)
{
    for (size_t i = 0; i < n; i++) {
        if (i) std::fputc(',', out);
        if (i < v.size() && std::isfinite(v[i])) std::fprintf(out, fmt, v[i]);
        else std::fputs("null", out);
    }
}

// line (src/main.cpp:705-710)
void line(FILE *out, const char *label, const es_segment &a)
{
    const double f = (double)a.frames;
    const double vmaf = mean_of(a.vmaf, a.vmaf_frames, a.frames), vmaf_min = a.vmaf_frames == a.frames ? a.vmaf_min : NAN;
    std::fprintf(out,
                 "  %-14s score %5.1f  fidelity %5.1f  texture %5.1f  banding %5.1f  grain %5.1f | VMAF %6.2f (worst frame %5.1f)  "
                 "CAMBI %.2f vs %.2f  lasting texture median %3.0f%% worst tenth %4.0f%%  fine detail %3.0f%%\n",
                 label, a.score, a.fidelity, a.texture, a.banding, a.grain, vmaf, vmaf_min, a.cambi / f, a.cambi_source / f,
                 a.median_lasting, a.tenth_lasting, kept(a.all_d[0], a.all_r[0]));
}

}  // namespace

void write_json(FILE *out, const es_run_info &info, long frames, const std::vector<ReportEncode> &encs)
{
    std::fprintf(out,
                 "{\"source\": %s, \"start\": %.6f, \"offset\": %.6f, \"frames\": %ld, \"fps\": %.6f, \"width\": %d, "
                 "\"height\": %d, \"model\": %s, \"segment_seconds\": %g, \"encodes\": [",
                 quoted(info.source).c_str(), info.start, info.offset, frames, info.fps, info.width, info.height,
                 quoted(info.model).c_str(), info.segment_seconds);
    for (size_t ci = 0; ci < encs.size(); ci++) {
        const ReportEncode &c = encs[ci];
        std::fprintf(out, "%s{\"path\": %s, \"whole\": ", ci ? ", " : "", quoted(info.paths ? info.paths[ci] : "").c_str());
        std::fputs(parts(c.whole).c_str(), out);
        std::fputs(", \"segments\": [", out);
        for (size_t i = 0; i < c.segments.size(); i++) {
            if (i) std::fputs(", ", out);
            std::fputs(parts(c.segments[i]).c_str(), out);
        }
        const size_t n = c.cambi.size();
        std::fputs("], \"per_frame\": {\"vmaf\": [", out);
        frame_list(out, c.vmaf, n, "%.6f");
        std::fputs("], \"cambi\": [", out);
        frame_list(out, c.cambi, n, "%.6f");
        std::fputs("], \"cambi_source\": [", out);
        frame_list(out, c.cambi_source, n, "%.6f");
        std::fputs("], \"psnr_y\": [", out);
        frame_list(out, c.psnr_y, n, "%.3f");
        std::fputs("]}}", out);
    }
    std::fputs("]}\n", out);
}

void write_summary(FILE *out, const es_run_info &info, double seconds, long frames, const std::vector<ReportEncode> &encs)
{
    std::fprintf(out, "%ld frames (%.1f min) of %dx%d in %.0f s (%.0f fps), VMAF model %s\n", frames,
                 frames / info.fps / 60, info.width, info.height, seconds, frames / seconds, info.model ? info.model : "");
    for (size_t ci = 0; ci < encs.size(); ci++) {
        const ReportEncode &c = encs[ci];
        std::fprintf(out, "\n%s\n", info.paths ? info.paths[ci] : "");
        line(out, "whole run", c.whole);
        std::vector<size_t> order(c.segments.size());
        for (size_t i = 0; i < order.size(); i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t a, size_t b) { return c.segments[a].score < c.segments[b].score; });
        std::fprintf(out, "  worst segments:\n");
        for (size_t i = 0; i < std::min<size_t>(5, order.size()); i++) {
            char label[32];
            const double t = info.start + (double)order[i] * info.segment_seconds;
            std::snprintf(label, sizeof label, "%02d:%02d:%02d", (int)(t / 3600), (int)(t / 60) % 60, (int)t % 60);
            line(out, label, c.segments[order[i]]);
        }
    }
}

}  // namespace es
