// es_scorer: the measure path of one source against its encodes, frame step by frame step (src/main.cpp:516-643).
//
// Plain form: the reference's structure on the GPU.  Per frame, one CAMBI instance per picture (two per file,
// alternating by frame parity), one window-stats and one lasting-texture launch per encode in double precision, and
// the collect on the host, overlapped with the GPU's work on the next frame (src/main.cpp:666-687).
//
// Batched form: `batch` frame steps per round.  Every kernel covers all pictures or pairs of the round in one launch
// (CAMBI's c-values as a persistent kernel pulling tiles from a device counter), the energies in exact integers, the
// source side of the lasting texture once per crop, and the collect and selects on the device; the round is captured
// once as a HIP graph per staging buffer and replayed.  Per-pair crops and uneven encode lengths are handled by
// descriptors: an encode that has ended leaves its pictures and jobs inactive.
//
// Frames are read in place in their decoded formats (es_frame: P010, NV12 and planar 4:2:0 at 8 or 16 bits per
// sample); with chroma planes each frame step also measures PSNR and the mean colours (src/main.cpp:543-556).
#include "cache.hpp"
#include "cambi_host.hpp"
#include "energy.hpp"
#include "frame.hpp"
#include "host_common.hpp"
#include "report.hpp"
#include "stats_host.hpp"

#include <algorithm>
#include <chrono>
#include <memory>

namespace es {

static constexpr int kTopkBlocks = 16;  // counting blocks per top-k job of a round

struct Crop {
    int w, h, cx, cy, windows, ww, wh;
    long len;  // frames while any of its encodes runs
    CambiInfo ci;
};

struct EncodeState {
    es_encode_desc d;
    int crop, windows;
    std::vector<Accum> seg;
    std::vector<double> cambi, cambi_src, quantile;
    std::vector<double> psnr;  // 3 per frame
    std::vector<double> vmaf;  // attached by the caller
};

class ScorerBase {
public:
    explicit ScorerBase(const es_scorer_config &c) : cfg(c)
    {
        require(c.n_encodes >= 1 && c.encodes, "at least one encode");
        require(c.src_bytes == 1 || c.src_bytes == 2, "samples must be 1 or 2 bytes");
        require(c.depth == 8 || c.depth == 10, "depth must be 8 or 10");
        require(c.win == 16 || c.win == 32 || c.win == 64, "window size must be 16, 32 or 64");
        require(c.per_segment >= 1, "per_segment must be positive");
        require(c.src_pitch >= c.src_w, "source pitch below its width");
        window_base = c.window_base > 0 ? c.window_base : 65;
        for (int e = 0; e < c.n_encodes; e++) {
            const es_encode_desc &d = c.encodes[e];
            require(d.w >= c.win && d.h >= c.win && d.w <= c.src_w && d.h <= c.src_h, "encode size must fit the source");
            require(d.length >= 0 && d.pitch >= d.w && (d.bytes == 1 || d.bytes == 2), "bad encode description");
            EncodeState s;
            s.d = d;
            s.crop = -1;
            for (size_t k = 0; k < crops.size(); k++)
                if (crops[k].w == d.w && crops[k].h == d.h) s.crop = (int)k;
            if (s.crop < 0) {
                Crop cr{};
                cr.w = d.w;
                cr.h = d.h;
                cr.cx = ((c.src_w - d.w) / 2) & ~1;  // centred, even (src/main.cpp:466-467)
                cr.cy = ((c.src_h - d.h) / 2) & ~1;
                cr.ww = d.w / c.win;
                cr.wh = d.h / c.win;
                cr.windows = cr.ww * cr.wh;
                cr.len = 0;
                cr.ci = cambi_info(d.w, d.h, c.depth, window_base);
                crops.push_back(cr);
                s.crop = (int)crops.size() - 1;
            }
            Crop &cr = crops[s.crop];
            cr.len = std::max(cr.len, (long)d.length);
            s.windows = cr.windows;
            s.seg.resize(d.length ? (size_t)((d.length + c.per_segment - 1) / c.per_segment) : 0);
            enc.push_back(std::move(s));
        }
        encodes_copy.assign(c.encodes, c.encodes + c.n_encodes);
        cfg.encodes = encodes_copy.data();
        up = 10 - c.depth;
    }
    virtual ~ScorerBase() = default;
    // pts: the source frames' timestamps for the source-CAMBI caches, or null
    virtual void submit(int n, const es_frame *src, const es_frame *encp, const int64_t *pts, hipStream_t st) = 0;
    virtual void sync() = 0;
    virtual void segment_hist(int e, int seg, Accum &a) { (void)e, (void)seg, (void)a; }

    // Segment k of encode e with its histograms and the attached VMAF of its frames.
    Accum seg_accum(int e, int k)
    {
        Accum a = enc[e].seg[k];
        segment_hist(e, k, a);
        const std::vector<double> &v = enc[e].vmaf;
        const long f0 = (long)k * cfg.per_segment, f1 = std::min<long>(f0 + cfg.per_segment, (long)v.size());
        for (long f = f0; f < f1; f++) a.add_vmaf(v[f]);
        return a;
    }
    // The segments of encode e that have scored frames.
    int scored_segments(int e) const
    {
        const long n = (long)enc[e].cambi.size();
        return (int)((n + cfg.per_segment - 1) / cfg.per_segment);
    }

    es_scorer_config cfg;
    std::vector<es_encode_desc> encodes_copy;
    std::vector<Crop> crops;
    std::vector<EncodeState> enc;
    int window_base, up;
    long frames = 0;  // frame steps submitted
    int chroma = -1;  // whether the run's frames carry chroma planes (-1 before the first frame)
    std::vector<std::unique_ptr<SourceCambiStore>> caches;  // per crop, when the run uses source-CAMBI caches
    long cache_hits = 0;

    void use_cache(const std::string &path, const std::string &dir)
    {
        caches.clear();
        for (const Crop &cr : crops)
            caches.push_back(std::make_unique<SourceCambiStore>(path, cr.cx, cr.cy, cr.w, cr.h, cfg.depth, dir));
    }
    // The cached CAMBI of crop c's source picture in the frame of timestamp pts.
    bool cached_source(int c, int64_t pts, double *score) const
    {
        return !caches.empty() && caches[c]->lookup(pts, score);
    }
    void keep_source(int c, int64_t pts, double score)
    {
        if (!caches.empty()) caches[c]->keep(pts, score);
    }

    // The frames of a run as plain luma planes of the config's formats (es_scorer_submit).
    void luma_frames(int n, const void *const *src, const void *const *encp, std::vector<es_frame> &fs,
                     std::vector<es_frame> &fe) const
    {
        const int E = (int)enc.size();
        fs.assign(n, es_frame{});
        fe.assign((size_t)n * E, es_frame{});
        for (int i = 0; i < n; i++) {
            es_frame &f = fs[i];
            f.y = src[i];
            f.ypitch = cfg.src_pitch * cfg.src_bytes;
            f.bytes = cfg.src_bytes;
            f.depth = cfg.depth - cfg.src_shift;
            for (int e = 0; e < E; e++) {
                const es_encode_desc &d = enc[e].d;
                es_frame &g = fe[(size_t)i * E + e];
                g.y = encp[(size_t)i * E + e];
                g.ypitch = d.pitch * d.bytes;
                g.bytes = d.bytes;
                g.depth = cfg.depth - d.shift;
            }
        }
    }

    // Checks the frames of steps frames .. frames + n - 1 and fixes whether the run carries chroma.
    void check_frames(int n, const es_frame *src, const es_frame *encp)
    {
        const int E = (int)enc.size();
        for (int i = 0; i < n; i++) {
            const long f = frames + i;
            const bool c = src[i].u != nullptr;
            if (chroma < 0) {
                chroma = c;
                if (c)
                    for (const Crop &cr : crops) require(cr.w % 2 == 0 && cr.h % 2 == 0, "picture sizes must be even with chroma");
            }
            require(c == (chroma == 1), "every frame of a run must carry chroma planes, or none");
            check_frame(src[i], cfg.depth, c);
            require(src[i].ypitch / src[i].bytes >= cfg.src_w, "source pitch below its width");
            for (int e = 0; e < E; e++) {
                if (!active(e, f)) continue;
                const es_frame &g = encp[(size_t)i * E + e];
                require((g.u != nullptr) == c, "every frame of a run must carry chroma planes, or none");
                check_frame(g, cfg.depth, c);
                require(g.ypitch / g.bytes >= enc[e].d.w, "encode pitch below its width");
            }
        }
    }

protected:
    bool active(int e, long f) const { return f < enc[e].d.length; }
    bool crop_active(int c, long f) const
    {
        for (size_t e = 0; e < enc.size(); e++)
            if (enc[e].crop == c && active((int)e, f)) return true;
        return false;
    }
    PlaneIn src_plane(const es_frame &f, int c) const { return frame_luma(f, crops[c].cx, crops[c].cy, cfg.depth); }
    PlaneIn enc_plane(const es_frame &f) const { return frame_luma(f, 0, 0, cfg.depth); }
    // The PSNR and colour job of encode e's frame pair; src_color: this job also sums the source's colour.
    FrameSumJob frame_job(const es_frame &src, const es_frame &g, int e, bool src_color, unsigned long long *out) const
    {
        const Crop &cr = crops[enc[e].crop];
        FrameSumJob j{};
        j.ry = src_plane(src, enc[e].crop);
        j.dy = enc_plane(g);
        j.ru = frame_chroma(src, 0, cr.cx, cr.cy, cfg.depth);
        j.rv = frame_chroma(src, 1, cr.cx, cr.cy, cfg.depth);
        j.du = frame_chroma(g, 0, 0, 0, cfg.depth);
        j.dv = frame_chroma(g, 1, 0, 0, cfg.depth);
        j.w = cr.w;
        j.h = cr.h;
        j.up = up;
        j.src_color = src_color;
        j.out = out;
        return j;
    }
    // The encode whose job sums the source's colour of crop c in frame f: the first active encode of the crop.
    int color_encode(int c, long f) const
    {
        for (size_t e = 0; e < enc.size(); e++)
            if (enc[e].crop == c && active((int)e, f)) return (int)e;
        return -1;
    }
    // Folds a frame's PSNR and colour sums (kFrameSums of them; the source's colour from its crop's first encode).
    void fold_chroma(int e, long f, const unsigned long long *z, const unsigned long long *zs)
    {
        const Crop &cr = crops[enc[e].crop];
        const size_t n = (size_t)cr.w * cr.h, nc = n / 4;
        double p[3], col_r[3], col_d[3];
        for (int k = 0; k < 3; k++) {
            const double mse = (double)z[k] / (double)(k ? nc : n), peak = (double)((1 << cfg.depth) - 1);
            p[k] = std::min(10.0 * std::log10(peak * peak / std::max(mse, 1e-16)), 6.0 * cfg.depth + 12);  // src/main.cpp:273-277
            const double sd = k < 2 ? (double)z[3 + k] : std::ldexp((double)z[5], -23);
            const double sr = k < 2 ? (double)zs[6 + k] : std::ldexp((double)zs[8], -23);
            col_d[k] = sd / (double)nc;
            col_r[k] = sr / (double)nc;
        }
        segment(e, f).add_chroma(p, col_r, col_d);
        enc[e].psnr.insert(enc[e].psnr.end(), p, p + 3);
    }
    void no_chroma(int e) { enc[e].psnr.insert(enc[e].psnr.end(), 3, NAN); }
    Accum &segment(int e, long f) { return enc[e].seg[(size_t)(f / cfg.per_segment)]; }
};

// ---------------------------------------------------------------------------------------------------------------
// Plain form
// ---------------------------------------------------------------------------------------------------------------
class PlainScorer final : public ScorerBase {
public:
    explicit PlainScorer(const es_scorer_config &c) : ScorerBase(c)
    {
        for (const Crop &cr : crops) {
            CropBufs b;
            for (auto &u : b.cambi) u = std::make_unique<CambiUnit>(cr.w, cr.h, c.depth, window_base, false, c.tile_h, c.tile_h0);
            b.ring.alloc((size_t)kRing * cr.w * cr.h * 4);
            hip_check(hipMemset(b.ring.p, 0, b.ring.bytes), "clearing");
            cb.push_back(std::move(b));
        }
        for (size_t e = 0; e < enc.size(); e++) {
            const Crop &cr = crops[enc[e].crop];
            EncBufs b;
            for (auto &u : b.cambi) u = std::make_unique<CambiUnit>(cr.w, cr.h, c.depth, window_base, false, c.tile_h, c.tile_h0);
            b.mid.alloc((size_t)(kRadius + 1) * cr.w * cr.h * 4);
            hip_check(hipMemset(b.mid.p, 0, b.mid.bytes), "clearing");
            for (int k = 0; k < 2; k++) {
                b.sums[k].alloc((size_t)cr.windows * 12 * 8);
                b.psums[k].alloc((size_t)cr.windows * 5 * 8);
                b.hsums[k].alloc((size_t)cr.windows * 12 * 8);
                b.hpsums[k].alloc((size_t)cr.windows * 5 * 8);
                b.fs[k].alloc(kFrameSums * 8);
                b.hfs[k].alloc(kFrameSums * 8);
            }
            b.masks.assign(kRadius + 1, std::vector<uint8_t>(cr.windows, 0));
            eb.push_back(std::move(b));
        }
        for (auto &e : done) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "creating an event");
    }
    ~PlainScorer() override
    {
        for (auto &e : done)
            if (e) (void)hipEventDestroy(e);
    }

    void submit(int n, const es_frame *src, const es_frame *encp, const int64_t *pts, hipStream_t st) override
    {
        check_frames(n, src, encp);
        for (int i = 0; i < n; i++) {
            queue(frames, src[i], encp + (size_t)i * enc.size(), pts ? pts[i] : INT64_MIN, st);
            if (frames > 0) collect(frames - 1);
            frames++;
        }
        pending = frames > 0;
    }

    void sync() override
    {
        if (pending && frames > 0) collect(frames - 1);
        pending = false;
    }

private:
    struct CropBufs {
        std::unique_ptr<CambiUnit> cambi[2];
        DevBuf ring;  // 13 mid planes by frame % 13
        bool hit[2] = {};  // the source CAMBI of the frame in slot at came from the cache
        double cached[2] = {};
    };
    int64_t slot_pts[2] = {INT64_MIN, INT64_MIN};
    struct EncBufs {
        std::unique_ptr<CambiUnit> cambi[2];
        DevBuf mid;  // 7 mid planes by frame % 7
        DevBuf sums[2], psums[2], fs[2];
        HostBuf hsums[2], hpsums[2], hfs[2];
        std::vector<std::vector<uint8_t>> masks;
    };
    std::vector<CropBufs> cb;
    std::vector<EncBufs> eb;
    hipEvent_t done[2] = {};
    bool pending = false;
    long collected = -1;

    int32_t *ring_plane(int c, long f) { return cb[c].ring.as<int32_t>() + (size_t)(f % kRing) * crops[c].w * crops[c].h; }
    int32_t *mid_plane(int e, long f)
    {
        const Crop &cr = crops[enc[e].crop];
        return eb[e].mid.as<int32_t>() + (size_t)(f % (kRadius + 1)) * cr.w * cr.h;
    }

    // queue (src/main.cpp:516-574): the GPU work of frame f
    void queue(long f, const es_frame &src, const es_frame *encp, int64_t pts, hipStream_t st)
    {
        const int at = (int)(f % 2);
        slot_pts[at] = pts;
        for (size_t c = 0; c < crops.size(); c++) {
            cb[c].hit[at] = cached_source((int)c, pts, &cb[c].cached[at]);
            if (!crop_active((int)c, f) || cb[c].hit[at]) continue;
            const PlaneIn p = src_plane(src, (int)c);
            cb[c].cambi[at]->launch(p.ptr, p.pitch, p.bytes, p.shift, st, p.rshift);
        }
        std::vector<bool> wrote(crops.size(), false);
        for (size_t e = 0; e < enc.size(); e++) {
            if (!active((int)e, f)) continue;
            const int c = enc[e].crop;
            const Crop &cr = crops[c];
            const PlaneIn pe = enc_plane(encp[e]);
            eb[e].cambi[at]->launch(pe.ptr, pe.pitch, pe.bytes, pe.shift, st, pe.rshift);
            WsPair pr{};
            pr.r = src_plane(src, c);
            pr.d = pe;
            pr.w = cr.w;
            pr.h = cr.h;
            pr.up = up;
            pr.mid_r = wrote[c] ? nullptr : ring_plane(c, f);
            wrote[c] = true;
            pr.mid_d = mid_plane((int)e, f);
            pr.sums = eb[e].sums[at].as<long long>();
            const auto *dp = (const WsPair *)arena().stage(&pr, sizeof pr, st);
            launch_window_stats(dp, 1, cr.w, cr.h, cfg.win, true, st);
            arena().done(st);
            hip_check(hipMemcpyAsync(eb[e].hsums[at].p, eb[e].sums[at].p, eb[e].sums[at].bytes, hipMemcpyDeviceToHost, st), "reading figures");
            if (chroma == 1) {
                // plane_sse x 3 and color_sums (src/main.cpp:543-554) as one job
                unsigned long long *out = eb[e].fs[at].as<unsigned long long>();
                launch_clear(out, kFrameSums, nullptr, 0, st);
                const FrameSumJob j = frame_job(src, encp[e], (int)e, color_encode(c, f) == (int)e, out);
                const auto *dj = (const FrameSumJob *)arena().stage(&j, sizeof j, st);
                launch_frame_sums(dj, 1, cr.h, st);
                arena().done(st);
                hip_check(hipMemcpyAsync(eb[e].hfs[at].p, out, kFrameSums * 8, hipMemcpyDeviceToHost, st), "reading figures");
            }
            if (f >= 2 * kRadius) {
                const long centre = f - kRadius;
                PersistJob j{};
                for (int o = -kRadius; o <= kRadius; o++) j.ring[o + kRadius] = ring_plane(c, centre + o);
                j.mid_d = mid_plane((int)e, centre);
                j.w = cr.w;
                j.h = cr.h;
                j.out = eb[e].psums[at].as<long long>();
                const auto *dj = (const PersistJob *)arena().stage(&j, sizeof j, st);
                launch_persist(dj, 1, cr.w, cr.h, cfg.win, 0, st);
                arena().done(st);
                hip_check(hipMemcpyAsync(eb[e].hpsums[at].p, eb[e].psums[at].p, eb[e].psums[at].bytes, hipMemcpyDeviceToHost, st), "reading figures");
            }
        }
        hip_check(hipEventRecord(done[at], st), "event");
    }

    // collect (src/main.cpp:576-643): the figures of frame f into the segments
    void collect(long f)
    {
        if (f <= collected) return;
        collected = f;
        const int at = (int)(f % 2);
        hip_check(hipEventSynchronize(done[at]), "waiting for a frame");
        std::vector<double> src_score(crops.size(), 0.0);
        for (size_t c = 0; c < crops.size(); c++) {
            if (!crop_active((int)c, f)) continue;
            if (cb[c].hit[at]) {
                src_score[c] = cb[c].cached[at];
                cache_hits++;
            } else {
                src_score[c] = cb[c].cambi[at]->finish();
                keep_source((int)c, slot_pts[at], src_score[c]);
            }
        }
        for (size_t e = 0; e < enc.size(); e++) {
            if (!active((int)e, f)) continue;
            EncodeState &s = enc[e];
            const double score = eb[e].cambi[at]->finish();
            s.cambi.push_back(score);
            s.cambi_src.push_back(src_score[s.crop]);
            long long ti[12];
            double to[4];
            Accum &acc = segment((int)e, f);
            const double q = collect_frame_host(acc, eb[e].hsums[at].as<long long>(), s.windows,
                                                eb[e].masks[f % (kRadius + 1)].data(), ti, to);
            s.quantile.push_back(q);
            acc.add_frame(score, src_score[s.crop], ti, to);
            if (chroma == 1) {
                const int ce = color_encode(s.crop, f);
                fold_chroma((int)e, f, eb[e].hfs[at].as<unsigned long long>(), eb[ce].hfs[at].as<unsigned long long>());
            } else {
                no_chroma((int)e);
            }
            if (f >= 2 * kRadius) {
                const long centre = f - kRadius;
                long long tl[3];
                Accum &cacc = segment((int)e, centre);
                collect_lasting_host(cacc, eb[e].hpsums[at].as<long long>(), eb[e].masks[centre % (kRadius + 1)].data(),
                                     s.windows, tl);
                cacc.add_lasting(tl);
            }
        }
    }
};

// ---------------------------------------------------------------------------------------------------------------
// Batched form
// ---------------------------------------------------------------------------------------------------------------
class BatchedScorer final : public ScorerBase {
public:
    explicit BatchedScorer(const es_scorer_config &c) : ScorerBase(c)
    {
        B = c.batch > 0 ? c.batch : 16;
        require(B <= 128, "batch must be at most 128 frames");
        R = B + 16;
        require(c.tile_h <= kMaxTileH && c.tile_h0 <= kMaxTileH, "tile_h must be at most 4096");
        cv_tile_heights(c.tile_h > 0 ? c.tile_h : 32, c.tile_h0, th);
        int cus = 0;
        hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0), "device attributes");
        cv_blocks = std::max(1, cus) * (c.cv_waves > 0 ? c.cv_waves : 16);
        require(c.cv_bmax == 0 || c.cv_bmax == 32 || c.cv_bmax == 64, "cv_bmax must be 32 or 64");
        cv_bmax = c.cv_bmax ? c.cv_bmax : 32;
        cv_tile_blocks = c.cv_tile_blocks != 0;
        cv_mfma = c.cv_mfma != 0;
        cv_queues = c.cv_queues > 0;
        const int C = (int)crops.size(), E = (int)enc.size();
        roles = C + E;
        P = B * roles;
        // CAMBI slots: one per (frame of the round, role); roles 0..C-1 are the source's crops, then the encodes
        consts.alloc((size_t)C * sizeof(CambiConsts));
        std::vector<CambiConsts> hk(C);
        for (int k = 0; k < C; k++) {
            const CambiInfo &ci = crops[k].ci;
            CambiConsts &K = hk[k];
            K.pad = ci.pad;
            K.band_base = ci.band_base;
            K.band_size = ci.band_size;
            K.vlt = ci.vlt;
            K.mask_index = ci.mask_index;
            K.window = ci.window;
            for (int d = 0; d < kDiffs; d++) {
                K.tvi[d] = ci.tvi[d];
                K.weight[d] = ci.weight[d];
            }
        }
        hip_check(hipMemcpy(consts.p, hk.data(), consts.bytes, hipMemcpyHostToDevice), "cambi constants");
        slot_off.resize(P + 1);
        list_off.resize(P + 1);
        seg_off.resize(P + 1);
        slot_off[0] = list_off[0] = seg_off[0] = 0;
        lay.resize(C);
        for (int k = 0; k < C; k++) {
            const CambiInfo &ci = crops[k].ci;
            lay[k].cap = cambi_list_layout(ci.sw, ci.sh, th, lay[k].loff, lay[k].toff, &lay[k].tiles);
        }
        for (int p = 0; p < P; p++) {
            const int k = crop_of_role(p % roles);
            slot_off[p + 1] = slot_off[p] + (size_t)crops[k].ci.off[kScales];
            list_off[p + 1] = list_off[p] + (size_t)lay[k].cap;
            seg_off[p + 1] = seg_off[p] + (size_t)lay[k].tiles;
        }
        planes.alloc(slot_off[P] * 2);
        list.alloc(list_off[P] * 4);
        segs.alloc(seg_off[P] * 4);
        sums.alloc((size_t)P * kScales * 8);
        counts.alloc((size_t)P * kScales * 4 + 64);  // + the work counter
        thist.alloc((size_t)P * kScales * kTopkBlocks * 512 * 4);
        thsum.alloc((size_t)P * kScales * kTopkBlocks * 512 * 8);
        std::vector<TopkJob> jobs((size_t)P * kScales);
        for (int p = 0; p < P; p++) {
            const int kc = crop_of_role(p % roles);
            const CambiInfo &ci = crops[kc].ci;
            for (int s = 0; s < kScales; s++) {
                TopkJob &j = jobs[(size_t)p * kScales + s];
                j = TopkJob{};
                j.list = list.as<float>() + list_off[p] + lay[kc].loff[s];
                j.seg_count = segs.as<unsigned>() + seg_off[p] + lay[kc].toff[s];
                j.nseg = cvalue_tiles(ci.sw[s], ci.sh[s], th[s]);
                j.seg_cap = kCvTileW * th[s];
                j.count = counts.as<unsigned>() + (size_t)p * kScales + s;
                j.k = (unsigned long long)ci.k[s];
                j.hist = thist.as<unsigned>() + ((size_t)p * kScales + s) * kTopkBlocks * 512;
                j.hsum = thsum.as<unsigned long long>() + ((size_t)p * kScales + s) * kTopkBlocks * 512;
            }
        }
        topk.alloc(jobs.size() * sizeof(TopkJob));
        hip_check(hipMemcpy(topk.p, jobs.data(), topk.bytes, hipMemcpyHostToDevice), "top-k jobs");
        scores.alloc((size_t)P * 8);
        max_base_tiles = 0;
        max_px.assign(kScales, 0);
        max_strips.assign(kScales, 0);
        for (const Crop &cr : crops) {
            max_base_tiles = std::max(max_base_tiles, base_tiles(cr.w, cr.h));
            for (int s = 0; s < kScales; s++) {
                max_px[s] = std::max(max_px[s], cr.ci.n[s]);
                max_strips[s] = std::max(max_strips[s], scale_strips(cr.ci.sw[s], cr.ci.sh[s]));
            }
            max_w = std::max(max_w, cr.w);
            max_h = std::max(max_h, cr.h);
            max_windows = std::max(max_windows, cr.windows);
        }
        // energies, lasting texture, collect
        for (int k = 0; k < C; k++) {
            CropDev d;
            const size_t n = (size_t)crops[k].w * crops[k].h;
            d.ring.alloc((size_t)R * n * 4);
            hip_check(hipMemset(d.ring.p, 0, d.ring.bytes), "clearing");
            d.sa.alloc(n * 4);
            d.sb.alloc(n * 4);
            d.psrc = lout_words;
            lout_words += (size_t)B * crops[k].windows * 4;
            cd.push_back(std::move(d));
        }
        for (int e = 0; e < E; e++) {
            EncDev d;
            const Crop &cr = crops[enc[e].crop];
            const size_t n = (size_t)cr.w * cr.h;
            d.mid.alloc((size_t)R * n * 4);
            hip_check(hipMemset(d.mid.p, 0, d.mid.bytes), "clearing");
            d.masks.alloc((size_t)R * cr.windows);
            hip_check(hipMemset(d.masks.p, 0, d.masks.bytes), "clearing");
            d.sums.alloc((size_t)B * cr.windows * 12 * 8);
            d.penc = lout_words;
            lout_words += (size_t)B * cr.windows;
            d.hist.alloc(std::max<size_t>(1, enc[e].seg.size()) * 6 * kHistBins * 8);
            hip_check(hipMemset(d.hist.p, 0, d.hist.bytes), "clearing");
            ed.push_back(std::move(d));
        }
        for (int e = 0; e < E; e++) require(std::count_if(enc.begin(), enc.end(), [&](const EncodeState &x) { return x.crop == enc[e].crop; }) <= kMaxCropEncodes,
                                            "too many encodes of one crop");
        lout.alloc(lout_words * 8);
        tot_int.alloc((size_t)B * E * 12 * 8);
        tot_or.alloc((size_t)B * E * 4 * 8);
        tot_last.alloc((size_t)B * E * 3 * 8);
        quant.alloc((size_t)B * E * 8);
        fsums.alloc((size_t)B * E * kFrameSums * 8);
        scratch.alloc((size_t)B * E * max_windows * 8);
        {
            const int chunks = (max_windows + 255) / 256;
            cwork_int.alloc((size_t)B * E * chunks * 12 * 8);
            cwork_dbl.alloc((size_t)B * E * chunks * 4 * 8);
            cwork_last.alloc((size_t)B * E * chunks * 3 * 8);
            cwork_bins.alloc((size_t)B * E * max_windows * 6 * 2);
            cwork_fhist.alloc((size_t)B * E * 6 * kHistBins * 4);
            cwork = CollectWork{scratch.as<unsigned long long>(), cwork_int.as<long long>(), cwork_dbl.as<double>(),
                                cwork_last.as<long long>(), cwork_bins.as<uint16_t>(), cwork_fhist.as<unsigned>(),
                                max_windows, chunks};
        }
        // descriptors: device arrays, staged from pinned host memory per round (two staging buffers)
        desc_bytes = align(P * sizeof(CambiPic)) + align(((size_t)P * kScales + 1) * 4) + align((size_t)B * E * sizeof(WsPair)) +
                     align((size_t)C * sizeof(LastingJob)) + align((size_t)B * E * sizeof(CollectJob)) +
                     align((size_t)B * E * sizeof(FrameSumJob));
        for (auto &d : desc) d.alloc(desc_bytes);
        res_bytes = align((size_t)P * 8) + align((size_t)B * E * 12 * 8) + align((size_t)B * E * 4 * 8) +
                    align((size_t)B * E * 3 * 8) + align((size_t)B * E * 8) + align((size_t)B * E * kFrameSums * 8);
        for (int k = 0; k < 2; k++) {
            stage_buf[k].alloc(desc_bytes);
            res[k].alloc(res_bytes);
            hip_check(hipEventCreateWithFlags(&done[k], hipEventDisableTiming), "creating an event");
            hip_check(hipEventCreateWithFlags(&done_e[k], hipEventDisableTiming), "creating an event");
            hip_check(hipEventCreateWithFlags(&fork_ev[k], hipEventDisableTiming), "creating an event");
        }
        hip_check(hipStreamCreateWithFlags(&work, hipStreamNonBlocking), "creating a stream");
        hip_check(hipStreamCreateWithFlags(&side, hipStreamNonBlocking), "creating a stream");
        hip_check(hipEventCreateWithFlags(&joined, hipEventDisableTiming), "creating an event");
        hip_check(hipEventCreateWithFlags(&side_joined, hipEventDisableTiming), "creating an event");
        use_graph = c.use_graph != 0;
        profile = c.profile != 0;
        serial = c.serial_stages != 0;
        early = c.early_energies != 0;
        if (profile)
            for (auto &row : ev)
                for (auto &pair : row)
                    for (auto &e : pair) hip_check(hipEventCreate(&e), "creating an event");
    }

    ~BatchedScorer() override
    {
        for (int k = 0; k < 2; k++) {
            for (hipEvent_t e : {done[k], done_e[k], fork_ev[k]})
                if (e) {
                    (void)hipEventSynchronize(e);
                    (void)hipEventDestroy(e);
                }
            for (int j = 0; j < 3; j++) {
                if (pexec[k][j]) (void)hipGraphExecDestroy(pexec[k][j]);
                if (pgraph[k][j]) (void)hipGraphDestroy(pgraph[k][j]);
            }
            for (int j = 0; j < ES_STAGES; j++) {
                if (sexec[k][j]) (void)hipGraphExecDestroy(sexec[k][j]);
                if (sgraph[k][j]) (void)hipGraphDestroy(sgraph[k][j]);
            }
        }
        for (hipStream_t *q : {&work, &side})
            if (*q) {
                (void)hipStreamSynchronize(*q);
                (void)hipStreamDestroy(*q);
            }
        for (hipEvent_t e : {joined, side_joined})
            if (e) (void)hipEventDestroy(e);
        for (auto &row : ev)
            for (auto &pair : row)
                for (auto &e : pair)
                    if (e) (void)hipEventDestroy(e);
    }

    double stage_ms[ES_STAGES] = {};
    long profiled_rounds = 0;

    // The rounds run on the scorer's own stream (a graph cannot be captured on the null stream), ordered after the
    // caller's earlier work on `st`, and the caller's later work on `st` waits for them.
    void submit(int n, const es_frame *src, const es_frame *encp, const int64_t *pts, hipStream_t st) override
    {
        check_frames(n, src, encp);
        hip_check(hipEventRecord(joined, st), "event");
        hip_check(hipStreamWaitEvent(work, joined, 0), "event");
        for (int i = 0; i < n; i += B) {
            const int m = std::min(B, n - i);
            round(m, src + i, encp + (size_t)i * enc.size(), pts ? pts + i : nullptr, work);
        }
        hip_check(hipEventRecord(joined, work), "event");
        hip_check(hipEventRecord(side_joined, side), "event");
        hip_check(hipStreamWaitEvent(st, joined, 0), "event");
        hip_check(hipStreamWaitEvent(st, side_joined, 0), "event");
    }

    void sync() override
    {
        // fold the outstanding rounds in order
        for (int k = 0; k < 2; k++) {
            const int p = (int)((rounds + k) % 2);
            if (pend[p].n) fold(p);
        }
    }

    void segment_hist(int e, int seg, Accum &a) override
    {
        hip_check(hipMemcpy(a.hist.data(), ed[e].hist.as<unsigned long long>() + (size_t)seg * 6 * kHistBins,
                            (size_t)6 * kHistBins * 8, hipMemcpyDeviceToHost),
                  "reading histograms");
    }

private:
    static size_t align(size_t n) { return (n + 255) & ~(size_t)255; }
    int crop_of_role(int r) const { return r < (int)crops.size() ? r : enc[r - crops.size()].crop; }

    struct CropDev {
        DevBuf ring, sa, sb;
        size_t psrc = 0;  // offset in lout: [slot][window][4]
    };
    struct EncDev {
        DevBuf mid, masks, sums, hist;
        size_t penc = 0;  // offset in lout: [slot][window]
    };
    struct Pending {
        long f0 = 0;
        int n = 0;
        std::vector<int64_t> pts;      // per frame of the round
        std::vector<char> hit;         // per (frame, crop): the source CAMBI came from the cache
        std::vector<double> cached;
    };

    int B, R, P, roles, th[kScales], cv_bmax = 32, cv_blocks = 2048, max_base_tiles = 0, max_w = 0, max_h = 0, max_windows = 0;
    bool cv_tile_blocks = false, cv_mfma = false;
    bool cv_queues = false;  // persistent c-value waves take tiles from one queue per XCD (else one shared queue)
    std::vector<long long> max_px;
    std::vector<int> max_strips;
    struct Layout {
        long long loff[kScales], cap;
        int toff[kScales], tiles;
    };
    std::vector<Layout> lay;
    std::vector<size_t> slot_off, list_off, seg_off;
    DevBuf consts, planes, list, segs, sums, counts, thist, thsum, topk, scores;
    std::vector<CropDev> cd;
    std::vector<EncDev> ed;
    DevBuf tot_int, tot_or, tot_last, quant, fsums, scratch, lout;
    DevBuf desc[2];  // a round's descriptors, one buffer per round parity: the energy chain of a round can still
                     // be reading its own while the next round's are uploaded
    DevBuf cwork_int, cwork_dbl, cwork_last, cwork_bins, cwork_fhist;
    CollectWork cwork{};
    size_t lout_words = 0;
    size_t desc_bytes, res_bytes;
    HostBuf stage_buf[2], res[2];
    // done: a round's CAMBI chain and its scores copy (on the work stream); done_e: its energy chain and copies (on the
    // side stream, unless serial); fork_ev: the point of the work stream where the round's energy chain may start
    hipEvent_t done[2] = {}, done_e[2] = {}, fork_ev[2] = {}, joined = nullptr, side_joined = nullptr;
    hipEvent_t ev[2][ES_STAGES][2] = {};  // profiled: each stage's start and end on the stream it runs on
    bool profile = false;
    hipStream_t work = nullptr, side = nullptr;  // the round's main stream and the energy chain's
    bool serial = false, early = false;
    hipGraph_t sgraph[2][ES_STAGES] = {};
    hipGraphExec_t sexec[2][ES_STAGES] = {};
    hipGraph_t pgraph[2][3] = {};  // graph form, chains overlapped: the work stream's two parts and the side's part
    hipGraphExec_t pexec[2][3] = {};
    Pending pend[2];
    long rounds = 0;
    bool use_graph = false;

    template <class T>
    T *dptr(int p, size_t off) const { return (T *)(desc[p].as<char>() + off); }

    // offsets of the descriptor arrays
    size_t o_pics() const { return 0; }
    size_t o_first() const { return align((size_t)P * sizeof(CambiPic)); }
    size_t o_ws() const { return o_first() + align(((size_t)P * kScales + 1) * 4); }
    size_t o_last() const { return o_ws() + align((size_t)B * enc.size() * sizeof(WsPair)); }
    size_t o_col() const { return o_last() + align((size_t)crops.size() * sizeof(LastingJob)); }
    size_t o_fs() const { return o_col() + align((size_t)B * enc.size() * sizeof(CollectJob)); }
    size_t r_scores() const { return 0; }
    size_t r_ti() const { return align((size_t)P * 8); }
    size_t r_to() const { return r_ti() + align((size_t)B * enc.size() * 12 * 8); }
    size_t r_tl() const { return r_to() + align((size_t)B * enc.size() * 4 * 8); }
    size_t r_q() const { return r_tl() + align((size_t)B * enc.size() * 3 * 8); }
    size_t r_fs() const { return r_q() + align((size_t)B * enc.size() * 8); }

    void fill(int p, long f0, int n, const es_frame *src, const es_frame *encp)
    {
        char *h = stage_buf[p].as<char>();
        std::memset(h, 0, desc_bytes);
        const int C = (int)crops.size(), E = (int)enc.size();
        CambiPic *pics = (CambiPic *)(h + o_pics());
        int *first = (int *)(h + o_first());
        WsPair *ws = (WsPair *)(h + o_ws());
        LastingJob *lj = (LastingJob *)(h + o_last());
        CollectJob *cj = (CollectJob *)(h + o_col());
        FrameSumJob *fj = (FrameSumJob *)(h + o_fs());
        int t = 0;
        for (int i = 0; i < B; i++) {
            const long f = f0 + i;
            for (int r = 0; r < roles; r++) {
                const int pidx = i * roles + r;
                CambiPic &pc = pics[pidx];
                const int c = crop_of_role(r);
                const CambiInfo &ci = crops[c].ci;
                const bool on = i < n && (r < C ? crop_active(r, f) && !pend[p].hit[(size_t)i * C + r] : active(r - C, f));
                for (int s = 0; s < kScales; s++) {
                    pc.sw[s] = on ? ci.sw[s] : 0;
                    pc.sh[s] = on ? ci.sh[s] : 0;
                    pc.off[s] = ci.off[s];
                    pc.loff[s] = lay[c].loff[s];
                    pc.toff[s] = lay[c].toff[s];
                    pc.th[s] = th[s];
                }
                if (on) {
                    pc.in = r < C ? src_plane(src[i], r) : enc_plane(encp[(size_t)i * E + (r - C)]);
                    if (cfg.depth == 8) {
                        pc.in.shift += 2;
                        pc.in.from8 = 1;
                    }
                }
                pc.planes = planes.as<uint16_t>() + slot_off[pidx];
                pc.cvals = nullptr;
                pc.list = list.as<float>() + list_off[pidx];
                pc.seg_count = segs.as<unsigned>() + seg_off[pidx];
                pc.sums = sums.as<unsigned long long>() + (size_t)pidx * kScales;
                pc.counts = counts.as<unsigned>() + (size_t)pidx * kScales;
                pc.consts = consts.as<CambiConsts>() + c;
            }
        }
        // the c-value tiles by scale, coarsest first, then by picture: a coarse tile holds many more relevant values
        // (passes) than a fine one, so the long tiles are handed out first and the launch ends on short ones
        for (int k = 0; k < kScales; k++)
            for (int pidx = 0; pidx < P; pidx++) {
                const int s = kScales - 1 - k;
                first[k * P + pidx] = t;
                t += cvalue_tiles(pics[pidx].sw[s], pics[pidx].sh[s], th[s]);
            }
        first[P * kScales] = t;
        for (int i = 0; i < B; i++) {
            const long f = f0 + i;
            std::vector<bool> wrote(C, false);
            for (int e = 0; e < E; e++) {
                const int c = enc[e].crop;
                const Crop &cr = crops[c];
                const size_t np = (size_t)cr.w * cr.h;
                const bool on = i < n && active(e, f);
                WsPair &w = ws[(size_t)i * E + e];
                CollectJob &j = cj[(size_t)i * E + e];
                if (!on) continue;
                w.r = src_plane(src[i], c);
                w.d = enc_plane(encp[(size_t)i * E + e]);
                if (chroma == 1) {
                    // the window-stats kernel adds the luma squared error of the area whole windows cover
                    FrameSumJob &fo = fj[(size_t)i * E + e];
                    fo = frame_job(src[i], encp[(size_t)i * E + e], e, color_encode(c, f) == e,
                                   fsums.as<unsigned long long>() + ((size_t)i * E + e) * kFrameSums);
                    fo.lw = cr.w / cfg.win * cfg.win;
                    fo.lh = cr.h / cfg.win * cfg.win;
                    w.sse = fo.out;
                }
                w.w = cr.w;
                w.h = cr.h;
                w.up = up;
                w.mid_r = wrote[c] ? nullptr : cd[c].ring.as<int32_t>() + (size_t)(f % R) * np;
                wrote[c] = true;
                w.mid_d = ed[e].mid.as<int32_t>() + (size_t)(f % R) * np;
                w.sums = ed[e].sums.as<long long>() + (size_t)i * cr.windows * 12;
                j.sums = w.sums;
                j.windows = cr.windows;
                j.mask = ed[e].masks.as<uint8_t>() + (size_t)(f % R) * cr.windows;
                j.hist = ed[e].hist.as<unsigned long long>() + (size_t)(f / cfg.per_segment) * 6 * kHistBins;
                j.tot_int = tot_int.as<long long>() + ((size_t)i * E + e) * 12;
                j.tot_orient = tot_or.as<double>() + ((size_t)i * E + e) * 4;
                j.quantile = quant.as<double>() + (size_t)i * E + e;
                j.tot_lasting = tot_last.as<long long>() + ((size_t)i * E + e) * 3;
                if (f >= 2 * kRadius) {
                    const long centre = f - kRadius;
                    j.psrc = lout.as<long long>() + cd[c].psrc + (size_t)i * cr.windows * 4;
                    j.penc = lout.as<long long>() + ed[e].penc + (size_t)i * cr.windows;
                    j.cmask = ed[e].masks.as<uint8_t>() + (size_t)(centre % R) * cr.windows;
                    j.lhist = ed[e].hist.as<unsigned long long>() + (size_t)(centre / cfg.per_segment) * 6 * kHistBins;
                    j.lasting_plain = 0;
                }
            }
        }
        // lasting texture per crop: the centres of the round's frames from frame 12 on while the crop is active
        for (int c = 0; c < C; c++) {
            LastingJob &L = lj[c];
            const Crop &cr = crops[c];
            const size_t np = (size_t)cr.w * cr.h;
            L.ring = cd[c].ring.as<int32_t>();
            L.slots = R;
            L.w = cr.w;
            L.h = cr.h;
            L.sa = cd[c].sa.as<int32_t>();
            L.sb = cd[c].sb.as<int32_t>();
            L.psrc = lout.as<long long>() + cd[c].psrc;
            L.ncentres = 0;
            for (int i = 0; i < n; i++) {
                const long f = f0 + i;
                if (f < 2 * kRadius || !crop_active(c, f)) continue;
                if (!L.ncentres) {
                    L.centre0 = (int)(f - kRadius);
                    L.slot0 = i;
                }
                L.ncentres++;
            }
            L.nenc = 0;
            for (int e = 0; e < E; e++) {
                if (enc[e].crop != c) continue;
                L.emid[L.nenc] = ed[e].mid.as<int32_t>();
                L.penc[L.nenc] = lout.as<long long>() + ed[e].penc;
                L.last_centre[L.nenc] = enc[e].d.length - 1 - kRadius;
                L.nenc++;
            }
            (void)np;
        }
    }

    // The round's pipeline on `st`: the descriptor copy, the kernels and the copy of the results.
    // Stage k of the round's pipeline on `st` (ES_STAGE_NAMES): the descriptor copy and clear, the CAMBI kernels,
    // the energies, the lasting texture, the collect and the copy of the results.
    void stage(int k, int p, hipStream_t st)
    {
        const int C = (int)crops.size(), E = (int)enc.size();
        const CambiPic *pics = dptr<CambiPic>(p, o_pics());
        switch (k) {
        case 0:
            hip_check(hipMemcpyAsync(desc[p].p, stage_buf[p].p, desc_bytes, hipMemcpyHostToDevice, st), "descriptors");
            launch_clear(sums.as<unsigned long long>(), (long long)P * kScales, counts.as<unsigned>(),
                         (long long)P * kScales + kXcds, st);
            break;
        case 1: launch_cambi_base(pics, nullptr, P, max_base_tiles, st); break;
        case 2:
            for (int s = 1; s < kScales; s++) launch_cambi_scale(pics, P, max_strips[s], s, st);
            break;
        case 3:
            // persistent waves taking tiles from a counter, or one block per tile of the round's largest layout (the
            // blocks past the round's tiles exit at once)
            launch_cambi_cvalues(pics, dptr<int>(p, o_first()), P * kScales, (int)seg_off[P], cv_tile_blocks ? 0 : 1,
                                 counts.as<unsigned>() + (size_t)P * kScales, cv_blocks, cv_bmax, st, cv_mfma ? 1 : 0,
                                 cv_queues ? 1 : 0);
            break;
        case 4:
            for (int pass = 0; pass < 3; pass++) launch_topk_pass(topk.as<TopkJob>(), P * kScales, pass, kTopkBlocks, st);
            break;
        case 5:
            launch_cambi_finish(pics, topk.as<TopkJob>(), P, scores.as<double>(), st);
            hip_check(hipMemcpyAsync(res[p].as<char>() + r_scores(), scores.p, (size_t)P * 8, hipMemcpyDeviceToHost, st),
                      "results");
            break;
        case 6:
            // the energy chain's own clears (on its stream: the previous round's energy copies are done with them)
            if (cfg.win > 16) launch_clear(lout.as<unsigned long long>(), (long long)lout_words, nullptr, 0, st);
            if (chroma == 1) launch_clear(fsums.as<unsigned long long>(), (long long)B * E * kFrameSums, nullptr, 0, st);
            launch_window_stats(dptr<WsPair>(p, o_ws()), B * E, max_w, max_h, cfg.win, false, st);
            break;
        case 7:
            if (chroma == 1) launch_frame_sums(dptr<FrameSumJob>(p, o_fs()), B * E, max_h, st);
            break;
        case 8: launch_lasting(dptr<LastingJob>(p, o_last()), C, max_w, max_h, cfg.win, st); break;
        case 9:
            launch_collect(dptr<CollectJob>(p, o_col()), B, E, cwork, st);
            break;
        case 10: {
            char *r = res[p].as<char>();
            hip_check(hipMemcpyAsync(r + r_ti(), tot_int.p, tot_int.bytes, hipMemcpyDeviceToHost, st), "results");
            hip_check(hipMemcpyAsync(r + r_to(), tot_or.p, tot_or.bytes, hipMemcpyDeviceToHost, st), "results");
            hip_check(hipMemcpyAsync(r + r_tl(), tot_last.p, tot_last.bytes, hipMemcpyDeviceToHost, st), "results");
            hip_check(hipMemcpyAsync(r + r_q(), quant.p, quant.bytes, hipMemcpyDeviceToHost, st), "results");
            if (chroma == 1) hip_check(hipMemcpyAsync(r + r_fs(), fsums.p, fsums.bytes, hipMemcpyDeviceToHost, st), "results");
            break;
        }
        }
    }

    hipGraphExec_t capture(int k0, int k1, int p, hipStream_t st, hipGraph_t *g)
    {
        hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "graph capture");
        for (int k = k0; k < k1; k++) stage(k, p, st);
        hip_check(hipStreamEndCapture(st, g), "graph capture");
        hipGraphExec_t e;
        hip_check(hipGraphInstantiate(&e, *g, nullptr, nullptr, 0), "graph instantiate");
        return e;
    }

    // The order of the round's stages.  Serial: all on `st` in order.  Otherwise the CAMBI chain (stages 0-5, the
    // scores copy with the finish) runs on `st` and the energy chain (6-10: its clears, window stats, frame sums,
    // lasting texture, collect and its copies) on the side stream; they need nothing from each other.  The energy
    // chain is released at a point of the CAMBI chain (after the pyramid, or after the upload when early) and only
    // the host and the caller wait for it, so the next round's CAMBI chain (upload, pyramid) runs beside this round's
    // energies instead of after them.  The buffers the two chains write are their own (the descriptors one set per
    // round parity, which the host reuses only after both chains of that round are done).  part(i, k0, k1, stream)
    // runs stages k0 .. k1 - 1 as part i.
    template <class F>
    void schedule(int p, hipStream_t st, F &&part)
    {
        if (serial) {
            part(0, 0, ES_STAGES, st);
            hip_check(hipEventRecord(done_e[p], st), "event");
            return;
        }
        const int fork_after = early ? 0 : 2;
        part(0, 0, fork_after + 1, st);
        hip_check(hipEventRecord(fork_ev[p], st), "event");
        hip_check(hipStreamWaitEvent(side, fork_ev[p], 0), "event");
        part(1, fork_after + 1, 6, st);
        part(2, 6, ES_STAGES, side);
        hip_check(hipEventRecord(done_e[p], side), "event");
    }

    // The round: eager launches, one graph per part, or (profiled) one graph per stage with events around each on its
    // stream.
    void pipeline(int p, hipStream_t st)
    {
        auto timed = [&](int k, hipStream_t s, auto &&launch) {
            if (profile) hip_check(hipEventRecord(ev[p][k][0], s), "stage event");
            launch();
            if (profile) hip_check(hipEventRecord(ev[p][k][1], s), "stage event");
        };
        if (!use_graph) {
            schedule(p, st, [&](int, int k0, int k1, hipStream_t s) {
                for (int k = k0; k < k1; k++) timed(k, s, [&] { stage(k, p, s); });
            });
            return;
        }
        if (!profile) {
            schedule(p, st, [&](int i, int k0, int k1, hipStream_t s) {
                if (!pexec[p][i]) pexec[p][i] = capture(k0, k1, p, s, &pgraph[p][i]);
                hip_check(hipGraphLaunch(pexec[p][i], s), "graph launch");
            });
            return;
        }
        schedule(p, st, [&](int, int k0, int k1, hipStream_t s) {
            for (int k = k0; k < k1; k++)
                timed(k, s, [&] {
                    if (k == 7 && chroma != 1) return;  // no work in the stage: nothing to capture
                    if (!sexec[p][k]) sexec[p][k] = capture(k, k + 1, p, s, &sgraph[p][k]);
                    hip_check(hipGraphLaunch(sexec[p][k], s), "graph launch");
                });
        });
    }

    void round(int n, const es_frame *src, const es_frame *encp, const int64_t *pts, hipStream_t st)
    {
        const int p = (int)(rounds % 2);
        if (pend[p].n) fold(p);
        const int C = (int)crops.size();
        Pending &pd = pend[p];
        pd.pts.assign(B, INT64_MIN);
        pd.hit.assign((size_t)B * C, 0);
        pd.cached.assign((size_t)B * C, 0.0);
        for (int i = 0; i < n; i++) {
            pd.pts[i] = pts ? pts[i] : INT64_MIN;
            for (int c = 0; c < C; c++) {
                double v;
                if (crop_active(c, frames + i) && cached_source(c, pd.pts[i], &v)) {
                    pd.hit[(size_t)i * C + c] = 1;
                    pd.cached[(size_t)i * C + c] = v;
                }
            }
        }
        fill(p, frames, n, src, encp);
        pipeline(p, st);
        hip_check(hipEventRecord(done[p], st), "event");
        pend[p].f0 = frames;
        pend[p].n = n;
        frames += n;
        rounds++;
    }

    void fold(int p)
    {
        hip_check(hipEventSynchronize(done[p]), "waiting for a round");
        hip_check(hipEventSynchronize(done_e[p]), "waiting for a round");
        if (profile) {
            for (int k = 0; k < ES_STAGES; k++) {
                float ms = 0;
                hip_check(hipEventElapsedTime(&ms, ev[p][k][0], ev[p][k][1]), "stage time");
                stage_ms[k] += ms;
            }
            profiled_rounds++;
        }
        const char *r = res[p].as<char>();
        const double *sc = (const double *)(r + r_scores());
        const long long *ti = (const long long *)(r + r_ti());
        const double *to = (const double *)(r + r_to());
        const long long *tl = (const long long *)(r + r_tl());
        const double *q = (const double *)(r + r_q());
        const unsigned long long *fs = (const unsigned long long *)(r + r_fs());
        const int C = (int)crops.size(), E = (int)enc.size();
        for (int i = 0; i < pend[p].n; i++) {
            const long f = pend[p].f0 + i;
            for (int c = 0; c < C; c++) {
                if (!crop_active(c, f)) continue;
                if (pend[p].hit[(size_t)i * C + c]) cache_hits++;
                else keep_source(c, pend[p].pts[i], sc[i * roles + c]);
            }
            for (int e = 0; e < E; e++) {
                if (!active(e, f)) continue;
                EncodeState &s = enc[e];
                const size_t k = (size_t)i * E + e;
                const size_t kc = (size_t)i * C + s.crop;
                const double score = sc[i * roles + C + e];
                const double src = pend[p].hit[kc] ? pend[p].cached[kc] : sc[i * roles + s.crop];
                s.cambi.push_back(score);
                s.cambi_src.push_back(src);
                s.quantile.push_back(q[k]);
                segment(e, f).add_frame(score, src, ti + k * 12, to + k * 4);
                if (chroma == 1)
                    fold_chroma(e, f, fs + k * kFrameSums, fs + ((size_t)i * E + color_encode(s.crop, f)) * kFrameSums);
                else
                    no_chroma(e);
                if (f >= 2 * kRadius) segment(e, f - kRadius).add_lasting(tl + k * 3);
            }
        }
        pend[p].n = 0;
    }
};

}  // namespace es

using namespace es;

struct es_scorer_s {
    std::unique_ptr<ScorerBase> impl;
};

extern "C" {

int es_scorer_create(const es_scorer_config *cfg, es_scorer **out)
{
    return guard([&] {
        ensure_init();
        require(cfg && out, "null argument");
        auto s = std::make_unique<es_scorer_s>();
        if (cfg->mode == ES_PLAIN) s->impl = std::make_unique<PlainScorer>(*cfg);
        else if (cfg->mode == ES_BATCHED) s->impl = std::make_unique<BatchedScorer>(*cfg);
        else throw Error(ES_EINVAL, "mode must be ES_PLAIN or ES_BATCHED");
        *out = s.release();
    });
}

int es_scorer_destroy(es_scorer *s)
{
    return guard([&] { delete s; });
}

int es_scorer_submit(es_scorer *s, int nframes, const void *const *src, const void *const *enc, es_stream stream)
{
    return guard([&] {
        require(s && src && enc && nframes >= 0, "bad argument");
        ensure_init();
        std::vector<es_frame> fs, fe;
        s->impl->luma_frames(nframes, src, enc, fs, fe);
        s->impl->submit(nframes, fs.data(), fe.data(), nullptr, hs(stream));
    });
}

int es_scorer_submit_frames(es_scorer *s, int nframes, const es_frame *src, const es_frame *enc, es_stream stream)
{
    return guard([&] {
        require(s && src && enc && nframes >= 0, "bad argument");
        ensure_init();
        s->impl->submit(nframes, src, enc, nullptr, hs(stream));
    });
}

int es_scorer_submit_pts(es_scorer *s, int nframes, const es_frame *src, const es_frame *enc, const int64_t *pts,
                         es_stream stream)
{
    return guard([&] {
        require(s && src && enc && pts && nframes >= 0, "bad argument");
        ensure_init();
        s->impl->submit(nframes, src, enc, pts, hs(stream));
    });
}

int es_scorer_use_cache(es_scorer *s, const char *source_path, const char *dir)
{
    return guard([&] {
        require(s && source_path, "null argument");
        require(s->impl->frames == 0, "the cache must be chosen before the first frame");
        s->impl->use_cache(source_path, dir ? dir : "");
    });
}

int es_scorer_cache_hits(es_scorer *s, long *hits)
{
    return guard([&] {
        require(s && hits, "null argument");
        *hits = s->impl->cache_hits;
    });
}

int es_scorer_cache_file(es_scorer *s, int crop, const char **file)
{
    return guard([&] {
        require(s && file, "null argument");
        ScorerBase &b = *s->impl;
        require(crop >= 0 && crop < (int)b.crops.size(), "no such crop");
        *file = b.caches.empty() ? "" : b.caches[crop]->file().c_str();
    });
}

int es_scorer_sync(es_scorer *s)
{
    return guard([&] {
        require(s != nullptr, "null scorer");
        s->impl->sync();
    });
}

int es_scorer_frames(es_scorer *s, long *frames)
{
    return guard([&] {
        require(s && frames, "null argument");
        *frames = s->impl->frames;
    });
}

int es_scorer_segments(es_scorer *s, int encode, int *count)
{
    return guard([&] {
        require(s && count, "null argument");
        require(encode >= 0 && encode < (int)s->impl->enc.size(), "no such encode");
        *count = (int)s->impl->enc[encode].seg.size();
    });
}

int es_scorer_segment(es_scorer *s, int encode, int seg, es_segment *out, uint64_t *hist)
{
    return guard([&] {
        require(s && out, "null argument");
        ScorerBase &b = *s->impl;
        require(encode >= 0 && encode < (int)b.enc.size(), "no such encode");
        auto &segs = b.enc[encode].seg;
        require(seg < (int)segs.size(), "no such segment");
        Accum a;
        if (seg >= 0) {
            a = b.seg_accum(encode, seg);
        } else {
            for (size_t k = 0; k < segs.size(); k++) a.merge(b.seg_accum(encode, (int)k));
        }
        a.fill(out);
        if (hist) std::memcpy(hist, a.hist.data(), a.hist.size() * 8);
    });
}

int es_scorer_cambi(es_scorer *s, int encode, long n, double *cambi, double *cambi_source)
{
    return guard([&] {
        require(s != nullptr, "null scorer");
        ScorerBase &b = *s->impl;
        require(encode >= 0 && encode < (int)b.enc.size(), "no such encode");
        const EncodeState &e = b.enc[encode];
        require(n >= 0 && n <= (long)e.cambi.size(), "more frames than scored");
        if (cambi) std::copy(e.cambi.begin(), e.cambi.begin() + n, cambi);
        if (cambi_source) std::copy(e.cambi_src.begin(), e.cambi_src.begin() + n, cambi_source);
    });
}

int es_scorer_psnr(es_scorer *s, int encode, long n, double *psnr)
{
    return guard([&] {
        require(s && psnr, "null argument");
        ScorerBase &b = *s->impl;
        require(encode >= 0 && encode < (int)b.enc.size(), "no such encode");
        const EncodeState &e = b.enc[encode];
        require(n >= 0 && 3 * (size_t)n <= e.psnr.size(), "more frames than scored");
        std::copy(e.psnr.begin(), e.psnr.begin() + 3 * n, psnr);
    });
}

int es_scorer_set_vmaf(es_scorer *s, int encode, long n, const double *vmaf)
{
    return guard([&] {
        require(s && (vmaf || n == 0), "null argument");
        ScorerBase &b = *s->impl;
        require(encode >= 0 && encode < (int)b.enc.size(), "no such encode");
        EncodeState &e = b.enc[encode];
        require(n >= 0 && n <= (long)e.cambi.size(), "more frames than scored");
        e.vmaf.assign(vmaf, vmaf + n);
    });
}

}  // extern "C"

namespace {

// The report's figures of every encode: the segments that have frames, their merge and the per-frame lists.
std::vector<ReportEncode> report_of(ScorerBase &b)
{
    std::vector<ReportEncode> out(b.enc.size());
    for (size_t e = 0; e < b.enc.size(); e++) {
        ReportEncode &r = out[e];
        Accum whole;
        for (int k = 0; k < b.scored_segments((int)e); k++) {
            const Accum a = b.seg_accum((int)e, k);
            whole.merge(a);
            es_segment sg;
            a.fill(&sg);
            r.segments.push_back(sg);
        }
        whole.fill(&r.whole);
        const EncodeState &st = b.enc[e];
        r.vmaf = st.vmaf;
        r.cambi = st.cambi;
        r.cambi_source = st.cambi_src;
        for (size_t i = 0; i < st.psnr.size(); i += 3) r.psnr_y.push_back(st.psnr[i]);
    }
    return out;
}

void check_info(const es_run_info *info)
{
    require(info != nullptr, "null run description");
    require(info->fps > 0, "frame rate must be positive");
}

}  // namespace

extern "C" {

int es_scorer_write_json(es_scorer *s, const es_run_info *info, const char *path)
{
    return guard([&] {
        require(s && path, "null argument");
        check_info(info);
        const std::vector<ReportEncode> r = report_of(*s->impl);
        FILE *f = std::fopen(path, "w");
        require(f != nullptr, std::string("cannot write ") + path);
        write_json(f, *info, s->impl->frames, r);
        require(std::fclose(f) == 0, std::string("cannot write ") + path);
    });
}

int es_scorer_stage_ms(es_scorer *s, double *ms, long *rounds)
{
    return guard([&] {
        require(s && ms, "null argument");
        auto *b = dynamic_cast<BatchedScorer *>(s->impl.get());
        require(b != nullptr, "stage times are kept by the batched form");
        for (int k = 0; k < ES_STAGES; k++) ms[k] = b->stage_ms[k];
        if (rounds) *rounds = b->profiled_rounds;
    });
}

int es_scorer_thresholds(es_scorer *s, int encode, long n, double *q)
{
    return guard([&] {
        require(s && q, "null argument");
        ScorerBase &b = *s->impl;
        require(encode >= 0 && encode < (int)b.enc.size(), "no such encode");
        const EncodeState &e = b.enc[encode];
        require(n >= 0 && n <= (long)e.quantile.size(), "more frames than scored");
        std::copy(e.quantile.begin(), e.quantile.begin() + n, q);
    });
}

// The autotuner: times the batched form on the caller's frames, one setting at a time from the defaults: the c-value
// method, the tile height and LDS layout together, then the c-value waves per CU (or one block per tile), when the
// energy chain starts, and frames per batch.  Each trial builds a scorer, runs one batch to warm up (graph capture
// included), then times two batches three times and keeps the best.
int es_scorer_autotune(const es_scorer_config *cfg, int nframes, const void *const *src, const void *const *enc,
                       es_tune_result *out)
{
    return guard([&] {
        ensure_init();
        require(cfg && src && enc && out, "null argument");
        es_scorer_config base = *cfg;
        base.mode = ES_BATCHED;
        base.use_graph = 1;
        base.profile = 0;
        es_tune_result best{base.batch ? base.batch : 16, base.tile_h ? base.tile_h : 32, base.cv_bmax ? base.cv_bmax : 32,
                            base.cv_waves ? base.cv_waves : 16, 0.0, 0, base.cv_tile_blocks, base.early_energies,
                            base.cv_mfma, base.cv_queues, base.tile_h0};
        const int ne = base.n_encodes;
        auto trial = [&](const es_tune_result &t) -> double {
            es_scorer_config c = base;
            c.batch = t.batch;
            c.tile_h = t.tile_h;
            c.tile_h0 = t.tile_h0;
            c.cv_bmax = t.cv_bmax;
            c.cv_waves = t.cv_waves;
            c.cv_tile_blocks = t.cv_tile_blocks;
            c.early_energies = t.early_energies;
            c.cv_mfma = t.cv_mfma;
            c.cv_queues = t.cv_queues;
            const int need = 3 * t.batch;
            if (nframes < need) return -1.0;  // too few frames for this batch
            // one warm-up batch (graph capture included), then the same two batches timed kRepeats times, the best
            // kept: short rounds are noisy.  Every encode runs through all of them.
            constexpr int kRepeats = 3;
            std::vector<es_encode_desc> descs(base.encodes, base.encodes + ne);
            for (auto &d : descs) d.length = t.batch * (1 + 2 * kRepeats);
            c.encodes = descs.data();
            BatchedScorer sc(c);
            hipStream_t st = nullptr;
            hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "creating a stream");
            std::vector<es_frame> fs, fe;
            sc.luma_frames(need, src, enc, fs, fe);
            sc.submit(t.batch, fs.data(), fe.data(), nullptr, st);
            sc.sync();
            hip_check(hipStreamSynchronize(st), "tuning");
            double best_ms = 0;
            for (int rep = 0; rep < kRepeats; rep++) {
                const auto t0 = std::chrono::steady_clock::now();
                sc.submit(2 * t.batch, fs.data() + t.batch, fe.data() + (size_t)t.batch * ne, nullptr, st);
                sc.sync();
                hip_check(hipStreamSynchronize(st), "tuning");
                const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                best_ms = rep ? std::min(best_ms, ms) : ms;
            }
            hip_check(hipStreamDestroy(st), "tuning");
            out->trials++;
            return best_ms / (2 * t.batch);
        };
        out->trials = 0;
        best.ms_per_step = trial(best);
        require(best.ms_per_step > 0, "too few frames to tune: give at least three batches");
        auto try_set = [&](auto set, std::initializer_list<int> values) {
            es_tune_result keep = best;
            for (int v : values) {
                es_tune_result t = keep;
                set(t, v);
                if (t.batch == keep.batch && t.tile_h == keep.tile_h && t.tile_h0 == keep.tile_h0 && t.cv_bmax == keep.cv_bmax &&
                    t.cv_waves == keep.cv_waves && t.cv_tile_blocks == keep.cv_tile_blocks &&
                    t.early_energies == keep.early_energies && t.cv_mfma == keep.cv_mfma && t.cv_queues == keep.cv_queues)
                    continue;
                const double ms = trial(t);
                if (ms > 0 && ms < best.ms_per_step) {
                    best = t;
                    best.ms_per_step = ms;
                }
            }
        };
        // window counts on the matrix cores or by prefix scans
        try_set([](es_tune_result &t, int v) { t.cv_mfma = v; }, {0, 1});
        // tile height and LDS layout together: taller tiles see more values, which favours the wider passes
        try_set([](es_tune_result &t, int v) { t.tile_h = v % 1000, t.cv_bmax = v / 1000; },
                {32016, 32024, 32032, 32048, 32064, 32096, 32128, 64016, 64024, 64032, 64048, 64064, 64096, 64128});
        // full-resolution tiles, mostly one pass each, taller than the others (0: the same height)
        try_set([](es_tune_result &t, int v) { t.tile_h0 = v; }, {0, 64, 96, 128, 192});
        // with the energy chain overlapping, fewer resident c-value waves can leave it room
        // (0: one block per tile, which the hardware interleaves with the overlapping energy chain)
        try_set([](es_tune_result &t, int v) { t.cv_waves = v, t.cv_tile_blocks = v == 0; }, {0, 6, 8, 12, 16, 24});
        // the persistent waves' queues: one shared queue (0) or one per XCD (1)
        if (!best.cv_tile_blocks) try_set([](es_tune_result &t, int v) { t.cv_queues = v; }, {0, 1});
        if (!base.serial_stages) try_set([](es_tune_result &t, int v) { t.early_energies = v; }, {0, 1});
        // larger batches spread each round's launches and its c-value kernel's tail over more frames; a batch needs three
        // batches of the caller's frames to be tried.  96 and 128 are left out: at 4K their trials measured
        // faster than 64 and whole runs slower
        try_set([](es_tune_result &t, int v) { t.batch = v; }, {8, 16, 32, 64});
        const int trials = out->trials;
        *out = best;
        out->trials = trials;
    });
}

}  // extern "C"
