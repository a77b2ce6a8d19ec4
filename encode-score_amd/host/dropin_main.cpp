// A caller shaped like the reference's src/main.cpp (the queue/collect loop of :513-643) driving this library through
// the drop-in headers only: Kernels::get / Kernels::launch with the reference kernels' argument arrays, class Cambi,
// two instances per file alternating by frame % 2, the source's mid planes in a ring of 13 and the encode's in a ring
// of 7, the figures copied to host memory while the next frame runs, and the host collect with the reference's
// arithmetic on the reference's float figures.  Decoding is replaced by synthetic frames (es_synth_frame), VMAF is not
// computed.  It prints the per-frame figures and each segment's accumulator as JSON on standard output.
//
// usage: dropin_main W H w h frames per_segment seed
//   the source is W x H; the encode is its centred w x h crop (offsets (W - w) / 2 & ~1, as src/main.cpp:466-467),
//   with banding, blur and no grain; every frame carries 4:2:0 chroma planes.
#include "cambi.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

namespace {

constexpr int kRing = 13, kRadius = 6, kStats = 8, kPersist = 5;

// The caller's distribution of per-window percentages (src/stats.hpp:10-32): 0.05-point bins over [-500, 500].
struct Percentages {
    std::vector<uint64_t> bins = std::vector<uint64_t>(20001, 0);
    uint64_t n = 0;
    void add(double pct)
    {
        const double c = std::min(std::max(pct, -500.0), 500.0);
        bins[(size_t)std::lround((c + 500.0) / 0.05)]++;
        n++;
    }
    double at(double q) const
    {
        if (!n) return NAN;
        uint64_t seen = 0;
        for (size_t b = 0; b < bins.size(); b++)
            if ((double)(seen += bins[b]) > q * (double)(n - 1)) return -500.0 + 0.05 * (double)b;
        return 500.0;
    }
};

// The caller's per-segment sums (the measure fields of src/stats.hpp:36-48).
struct Segment {
    uint64_t frames = 0;
    double cambi = 0, cambi_source = 0, psnr[3] = {}, color_r[3] = {}, color_d[3] = {};
    double all_r[3] = {}, all_d[3] = {}, st_r[3] = {}, st_d[3] = {};
    double sh_all_r[2] = {}, sh_all_x[2] = {}, sh_st_r[2] = {}, sh_st_x[2] = {};
    double own = 0, mine = 0;
    Percentages amount[3], matching[2], lasting;
};

double psnr_of(unsigned long long sse, size_t n, int depth)  // src/main.cpp:273-277
{
    const double mse = (double)sse / (double)n, peak = (double)((1 << depth) - 1);
    return std::min(10.0 * std::log10(peak * peak / std::max(mse, 1e-16)), 6.0 * depth + 12);
}

// numpy's linear quantile as the reference's caller evaluates it on float figures (src/main.cpp:279-289)
float quantile_of(std::vector<float> v, double q)
{
    const double pos = q * (double)(v.size() - 1);
    const size_t lo = (size_t)pos;
    std::nth_element(v.begin(), v.begin() + (long)lo, v.end());
    const float a = v[lo];
    if (lo + 1 >= v.size()) return a;
    const float b = *std::min_element(v.begin() + (long)lo + 1, v.end());
    return (float)(a + (b - a) * (pos - (double)lo));
}

struct Device {
    void *p = nullptr;
    explicit Device(size_t bytes) { es_check(es_malloc(&p, bytes), "allocating"); }
    ~Device() { es_free(p); }
    template <class T>
    T *as() const { return (T *)p; }
};

struct Host {
    void *p = nullptr;
    explicit Host(size_t bytes) { es_check(es_host_alloc(&p, bytes), "allocating"); }
    ~Host() { es_host_free(p); }
    template <class T>
    T *as() const { return (T *)p; }
};

// Packed 4:2:0 planes of one picture at the pipeline depth (the reference's Planes).
struct Planes {
    Device y16, u, v;
    explicit Planes(size_t n) : y16(n * 2), u(n / 2), v(n / 2) {}
};

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 8) {
        std::fprintf(stderr, "usage: %s W H w h frames per_segment seed\n", argv[0]);
        return 2;
    }
    const int W = std::atoi(argv[1]), H = std::atoi(argv[2]), w = std::atoi(argv[3]), h = std::atoi(argv[4]);
    const long frames = std::atol(argv[5]), per_segment = std::atol(argv[6]);
    const unsigned long long seed = std::strtoull(argv[7], nullptr, 0);
    try {
        Kernels kernels;
        kernels.load(0);
        const es_kernel frame_planes = kernels.get("frame_planes"), window_stats = kernels.get("window_stats"),
                        persist_stats = kernels.get("persist_stats"), plane_sse = kernels.get("plane_sse"),
                        color_sums = kernels.get("color_sums");
        es_stream stream = Kernels::stream();
        const int pd = 10, up10 = 0;
        const size_t n = (size_t)w * h, plane = n;
        const int ww = w / 16, wh = h / 16, windows = ww * wh;
        const int sx = (W - w) / 2 & ~1, sy = (H - h) / 2 & ~1;

        // the decoded frames: the source's full picture and the encode's, luma and chroma (16-bit, the value in the
        // low 10 bits)
        es_synth_desc sd{}, ed{};
        sd.kind = ed.kind = 1;
        sd.seed = ed.seed = seed;
        sd.full_w = ed.full_w = W;
        sd.full_h = ed.full_h = H;
        sd.w = W, sd.h = H, sd.depth = 10, sd.grain_q8 = 256;
        ed.salt = 1, ed.crop_x = sx, ed.crop_y = sy, ed.w = w, ed.h = h, ed.depth = 10;
        ed.quant_step = 2, ed.quant_where = 2, ed.block = 8, ed.block_pct = 50, ed.blur_q8 = 128;
        auto half = [](es_synth_desc d, int salt) {
            d.w /= 2, d.h /= 2, d.full_w /= 2, d.full_h /= 2, d.crop_x /= 2, d.crop_y /= 2, d.salt += salt;
            return d;
        };
        Device sy_(W * (size_t)H * 2), su_(W * (size_t)H / 2), sv_(W * (size_t)H / 2);
        Device ey_(n * 2), eu_(n / 2), ev_(n / 2);

        Planes src(n), enc(n);
        Device ring(n * 4 * kRing);  // the source's mid planes by frame % 13
        std::vector<std::unique_ptr<Device>> mid;  // the encode's mid planes by frame % 7
        for (int i = 0; i <= kRadius; i++) mid.push_back(std::make_unique<Device>(n * 4));
        Device stats((size_t)windows * kStats * 4), persist((size_t)windows * kPersist * 4), sse(3 * 8),
            scolor(3 * 8), ecolor(3 * 8);
        std::unique_ptr<Cambi> source_cambi[2], encode_cambi[2];
        for (int i = 0; i < 2; i++) {
            source_cambi[i] = std::make_unique<Cambi>(kernels, w, h, pd);
            encode_cambi[i] = std::make_unique<Cambi>(kernels, w, h, pd);
        }
        // per frame % 2: the figures in host memory and an event for when they are there
        struct Results {
            Host mem;
            unsigned long long *sse;
            double *scolor, *ecolor;
            float *stats, *persist;
            explicit Results(int windows)
                : mem(72 + (size_t)windows * (kStats + kPersist) * 4), sse(mem.as<unsigned long long>()),
                  scolor(mem.as<double>() + 3), ecolor(mem.as<double>() + 6), stats((float *)(mem.as<double>() + 9)),
                  persist(stats + (size_t)windows * kStats) {}
        };
        std::unique_ptr<Results> results[2] = {std::make_unique<Results>(windows), std::make_unique<Results>(windows)};
        es_event done[2];
        for (es_event &e : done) es_check(es_event_create(&e), "creating an event");
        std::vector<Segment> segments;
        std::vector<uint8_t> mask_ring[kRadius + 1];
        for (auto &m : mask_ring) m.assign(windows, 0);
        std::vector<double> cambi_of, cambi_source_of, psnr_of_frame;

        auto copy_back = [&](void *dst, const Device &from, size_t bytes) {
            es_check(es_memcpy_dtoh(dst, from.p, bytes, stream), "reading figures");
        };
        auto clear = [&](const Device &d, size_t bytes) { es_check(es_memset(d.p, 0, bytes, stream), "clearing"); };
        auto to_planes = [&](const es_frame &f, int x0, int y0, Planes &p) {
            int xx = x0, yy = y0, depth = pd, pw = w, ph = h;
            void *a[] = {(void *)&f, &xx, &yy, &depth, &p.y16.p, &p.u.p, &p.v.p, &pw, &ph};
            Kernels::launch(frame_planes, w, h, a, stream);
        };

        // queue (src/main.cpp:519-576): the GPU work of frame f
        auto queue = [&](long f) {
            const int at = (int)(f % 2), slot = (int)(f % kRing), cslot = (int)(f % (kRadius + 1));
            es_synth_desc s1 = half(sd, 100), s2 = half(sd, 200), e1 = half(ed, 100), e2 = half(ed, 200);
            es_check(es_synth_frame(&sd, (int)f, sy_.p, 2, stream), "decoding");
            es_check(es_synth_frame(&s1, (int)f, su_.p, 2, stream), "decoding");
            es_check(es_synth_frame(&s2, (int)f, sv_.p, 2, stream), "decoding");
            es_check(es_synth_frame(&ed, (int)f, ey_.p, 2, stream), "decoding");
            es_check(es_synth_frame(&e1, (int)f, eu_.p, 2, stream), "decoding");
            es_check(es_synth_frame(&e2, (int)f, ev_.p, 2, stream), "decoding");
            es_frame sf{}, ef{};
            es_check(es_frame_view(ES_FMT_YUV420P10, sy_.p, su_.p, sv_.p, W * 2, W, &sf), "frame");
            es_check(es_frame_view(ES_FMT_YUV420P10, ey_.p, eu_.p, ev_.p, w * 2, w, &ef), "frame");
            to_planes(sf, sx, sy, src);
            source_cambi[at]->launch(src.y16.p, w, stream);
            Results &r = *results[at];
            clear(scolor, 24);
            size_t cn = n / 4;
            int up = up10;
            { void *a[] = {&src.u.p, &src.v.p, &cn, &up, &scolor.p}; Kernels::launch(color_sums, w, h, a, stream); }
            copy_back(r.scolor, scolor, 24);
            to_planes(ef, 0, 0, enc);
            encode_cambi[at]->launch(enc.y16.p, w, stream);
            int pw = w, ph = h;
            void *ring_slot = ring.as<char>() + (size_t)slot * plane * 4, *mid_d = mid[cslot]->p;
            { void *a[] = {&src.y16.p, &enc.y16.p, &up, &pw, &ph, &ring_slot, &mid_d, &stats.p};
              Kernels::launch(window_stats, w, h, a, stream); }
            clear(sse, 24);
            clear(ecolor, 24);
            void *null = nullptr, *s0 = sse.p, *s1p = sse.as<char>() + 8, *s2p = sse.as<char>() + 16;
            size_t nn = n;
            { void *a[] = {&src.y16.p, &enc.y16.p, &null, &null, &nn, &s0}; Kernels::launch(plane_sse, w, h, a, stream); }
            { void *a[] = {&src.u.p, &enc.u.p, &null, &null, &cn, &s1p}; Kernels::launch(plane_sse, w, h, a, stream); }
            { void *a[] = {&src.v.p, &enc.v.p, &null, &null, &cn, &s2p}; Kernels::launch(plane_sse, w, h, a, stream); }
            { void *a[] = {&enc.u.p, &enc.v.p, &cn, &up, &ecolor.p}; Kernels::launch(color_sums, w, h, a, stream); }
            copy_back(r.sse, sse, 24);
            copy_back(r.ecolor, ecolor, 24);
            copy_back(r.stats, stats, (size_t)windows * kStats * 4);
            if (f >= 2 * kRadius) {  // the lasting texture of the frame kRadius back
                const long centre = f - kRadius;
                int cring = (int)(centre % kRing);
                void *mid_c = mid[centre % (kRadius + 1)]->p;
                size_t pl = plane;
                void *a[] = {&ring.p, &pl, &cring, &mid_c, &pw, &persist.p};
                Kernels::launch(persist_stats, w, h, a, stream);
                copy_back(r.persist, persist, (size_t)windows * kPersist * 4);
            }
            es_check(es_event_record(done[at], stream), "event");
        };

        // collect (src/main.cpp:576-643): the figures of frame f, once in host memory
        auto collect = [&](long f) {
            const int at = (int)(f % 2), cslot = (int)(f % (kRadius + 1));
            es_check(es_event_synchronize(done[at]), "waiting for a frame");
            const Results &r = *results[at];
            const double banding_source = source_cambi[at]->finish(), banding = encode_cambi[at]->finish();
            cambi_of.push_back(banding);
            cambi_source_of.push_back(banding_source);
            const double p[3] = {psnr_of(r.sse[0], n, pd), psnr_of(r.sse[1], n / 4, pd), psnr_of(r.sse[2], n / 4, pd)};
            psnr_of_frame.insert(psnr_of_frame.end(), p, p + 3);
            const size_t seg = (size_t)(f / per_segment);
            if (segments.size() <= seg) segments.resize(seg + 1);
            Segment &acc = segments[seg];
            acc.frames++;
            acc.cambi += banding;
            acc.cambi_source += banding_source;
            for (int i = 0; i < 3; i++) {
                acc.psnr[i] += p[i];
                acc.color_r[i] += r.scolor[i] / (double)(n / 4);
                acc.color_d[i] += r.ecolor[i] / (double)(n / 4);
            }
            std::vector<float> oriented(windows);
            for (int i = 0; i < windows; i++) oriented[i] = r.stats[(size_t)i * kStats + 6];
            const float q = quantile_of(oriented, 0.8);
            std::vector<uint8_t> &mask = mask_ring[cslot];
            for (int i = 0; i < windows; i++) {
                const float *s = &r.stats[(size_t)i * kStats];
                const bool m = s[6] >= q && s[6] > 0;
                mask[i] = m;
                const double er[3] = {s[0], s[1], s[6]}, ed_[3] = {s[2], s[3], s[7]};
                for (int k = 0; k < 3; k++) {
                    acc.all_r[k] += er[k];
                    acc.all_d[k] += ed_[k];
                    if (m) {
                        acc.st_r[k] += er[k];
                        acc.st_d[k] += ed_[k];
                        acc.amount[k].add(100.0 * std::sqrt(ed_[k] / std::max(er[k], 1e-9)));
                    }
                }
                for (int k = 0; k < 2; k++) {
                    acc.sh_all_r[k] += er[k];
                    acc.sh_all_x[k] += s[4 + k];
                    if (m) {
                        acc.sh_st_r[k] += er[k];
                        acc.sh_st_x[k] += s[4 + k];
                        acc.matching[k].add(100.0 * s[4 + k] / std::max(er[k], 1e-9));
                    }
                }
            }
            if (f >= 2 * kRadius) {
                const long centre = f - kRadius;
                Segment &cacc = segments[(size_t)(centre / per_segment)];
                const std::vector<uint8_t> &cmask = mask_ring[centre % (kRadius + 1)];
                for (int i = 0; i < windows; i++) {
                    const float *s = &r.persist[(size_t)i * kPersist];
                    const double agree = s[2] / std::sqrt(std::max((double)s[3] * s[4], 1e-9));
                    if (cmask[i] && agree > 0.5 && s[0] > 0) {
                        cacc.own += s[0];
                        cacc.mine += s[1];
                        cacc.lasting.add(100.0 * s[1] / s[0]);
                    }
                }
            }
        };

        for (long f = 0; f < frames; f++) {
            queue(f);
            if (f) collect(f - 1);
        }
        if (frames) collect(frames - 1);

        auto num = [](double v) {  // a JSON number, or null when not finite
            if (std::isfinite(v)) std::printf("%.17g", v);
            else std::printf("null");
        };
        auto list = [&](const double *v, int k) {
            std::printf("[");
            for (int i = 0; i < k; i++) {
                if (i) std::printf(", ");
                num(v[i]);
            }
            std::printf("]");
        };
        std::printf("{\"cambi\": ");
        list(cambi_of.data(), (int)cambi_of.size());
        std::printf(", \"cambi_source\": ");
        list(cambi_source_of.data(), (int)cambi_source_of.size());
        std::printf(", \"psnr\": ");
        list(psnr_of_frame.data(), (int)psnr_of_frame.size());
        std::printf(", \"segments\": [");
        for (size_t i = 0; i < segments.size(); i++) {
            const Segment &a = segments[i];
            std::printf("%s{\"frames\": %llu, \"cambi\": %.17g, \"cambi_source\": %.17g, \"psnr\": ", i ? ", " : "",
                        (unsigned long long)a.frames, a.cambi, a.cambi_source);
            list(a.psnr, 3);
            std::printf(", \"color_source\": ");
            list(a.color_r, 3);
            std::printf(", \"color_encode\": ");
            list(a.color_d, 3);
            const double *groups[] = {a.all_r, a.all_d, a.st_r, a.st_d};
            const char *names[] = {"all_r", "all_d", "st_r", "st_d"};
            for (int g = 0; g < 4; g++) {
                std::printf(", \"%s\": ", names[g]);
                list(groups[g], 3);
            }
            const double *pairs[] = {a.sh_all_r, a.sh_all_x, a.sh_st_r, a.sh_st_x};
            const char *pnames[] = {"sh_all_r", "sh_all_x", "sh_st_r", "sh_st_x"};
            for (int g = 0; g < 4; g++) {
                std::printf(", \"%s\": ", pnames[g]);
                list(pairs[g], 2);
            }
            std::printf(", \"own\": %.17g, \"mine\": %.17g", a.own, a.mine);
            std::printf(", \"n_amount\": [%llu, %llu, %llu], \"n_matching\": [%llu, %llu], \"n_lasting\": %llu",
                        (unsigned long long)a.amount[0].n, (unsigned long long)a.amount[1].n,
                        (unsigned long long)a.amount[2].n, (unsigned long long)a.matching[0].n,
                        (unsigned long long)a.matching[1].n, (unsigned long long)a.lasting.n);
            const double med_amount[3] = {a.amount[0].at(0.5), a.amount[1].at(0.5), a.amount[2].at(0.5)};
            const double ten_amount[3] = {a.amount[0].at(0.1), a.amount[1].at(0.1), a.amount[2].at(0.1)};
            const double med_matching[2] = {a.matching[0].at(0.5), a.matching[1].at(0.5)};
            const double ten_matching[2] = {a.matching[0].at(0.1), a.matching[1].at(0.1)};
            std::printf(", \"median_amount\": ");
            list(med_amount, 3);
            std::printf(", \"tenth_amount\": ");
            list(ten_amount, 3);
            std::printf(", \"median_matching\": ");
            list(med_matching, 2);
            std::printf(", \"tenth_matching\": ");
            list(ten_matching, 2);
            std::printf(", \"median_lasting\": ");
            num(a.lasting.at(0.5));
            std::printf(", \"tenth_lasting\": ");
            num(a.lasting.at(0.1));
            std::printf("}");
        }
        std::printf("]}\n");
        for (es_event e : done) es_event_destroy(e);
        es_stream_destroy(stream);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "encode-score: %s\n", e.what());  // as src/main.cpp:780-783
        return 1;
    }
    return 0;
}
