/*
 * libencscore_hip: the C ABI of encode-score's per-frame measure path on AMD Instinct MI355X (gfx950).
 *
 * Every function returns 0 on success or a nonzero status (ES_EINVAL, ES_EHIP, ES_ENOMEM, ES_ESTATE) and leaves a
 * message for es_last_error().  Plain pointers and sizes only; device pointers are memory of the device es_init made
 * current (es_malloc, or any HIP allocation of the same process).  Streams and events are opaque handles.
 *
 * Each declaration names the reference interface it replaces (file:line in wakamex/encode-score v0.1.0).
 */
#ifndef ENCSCORE_H
#define ENCSCORE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ES_OK 0
#define ES_EINVAL 1
#define ES_EHIP 2
#define ES_ENOMEM 3
#define ES_ESTATE 4

typedef struct es_stream_s *es_stream;
typedef struct es_event_s *es_event;
typedef struct es_cambi_s es_cambi;
typedef struct es_scorer_s es_scorer;

/* ---- runtime (replaces cu_check and Kernels, src/gpu.hpp:7-23, src/gpu.cpp:4-19) ---- */

/* The message of the last failed call on this thread ("" when none). Replaces cu_check's message (src/gpu.hpp:7-13). */
const char *es_last_error(void);
/* Makes `device` current and loads the kernels. Replaces Kernels::load (src/gpu.cpp:4-6) and the context set-up of
 * src/main.cpp:359-362. */
int es_init(int device);
int es_device_count(int *count);
int es_device_synchronize(void);
/* Replaces cuStreamCreate(CU_STREAM_NON_BLOCKING) / cuStreamDestroy (src/cambi.cpp:37, :85, src/main.cpp:368). */
int es_stream_create(es_stream *out);
int es_stream_destroy(es_stream s);
int es_stream_synchronize(es_stream s);
/* Replaces cuEventCreate / cuEventRecord / cuStreamWaitEvent / cuEventSynchronize (src/cambi.cpp:38, :103-104,
 * src/main.cpp:573, :581); events here also time (es_event_elapsed_ms). */
int es_event_create(es_event *out);
int es_event_destroy(es_event e);
int es_event_record(es_event e, es_stream s);
int es_stream_wait_event(es_stream s, es_event e);
int es_event_synchronize(es_event e);
int es_event_elapsed_ms(es_event start, es_event end, float *ms);
/* Replaces cuMemAlloc / cuMemFree / cuMemAllocHost / cuMemcpy*Async / cuMemsetD8Async (src/cambi.cpp:59-86,
 * src/main.cpp:513-514). A NULL stream means the null stream and a synchronous copy. */
int es_malloc(void **ptr, size_t bytes);
int es_free(void *ptr);
int es_host_alloc(void **ptr, size_t bytes);
int es_host_free(void *ptr);
int es_memcpy_htod(void *dst, const void *src, size_t bytes, es_stream s);
int es_memcpy_dtoh(void *dst, const void *src, size_t bytes, es_stream s);
int es_memcpy_dtod(void *dst, const void *src, size_t bytes, es_stream s);
int es_memset(void *dst, int value, size_t bytes, es_stream s);

/* The by-name kernels (replaces Kernels::get / Kernels::launch, src/gpu.hpp:15-23, src/gpu.cpp:8-19, and the launches
 * by name of src/main.cpp:229-234, :513-566).  es_kernel_get knows the measure path's kernels: "frame_planes",
 * "window_stats", "persist_stats", "plane_sse" and "color_sums" (ES_EINVAL for any other name).  es_kernel_launch
 * queues one on `s` with `args` the reference kernel's argument array (one pointer per argument, in the order of
 * src/kernels.cu's signatures; frame_planes takes an es_frame where the reference's takes its struct Frame, which has
 * the same fields); w x h is the picture the launch covers, whose height persist_stats takes from it.  The outputs
 * are the reference's: window_stats and persist_stats as es_window_stats_f32 and es_persist_stats_f32, plane_sse adds
 * to *sse, color_sums adds sum u, sum v and the strength sum to its three doubles. */
typedef struct es_kernel_s *es_kernel;
int es_kernel_get(const char *name, es_kernel *out);
int es_kernel_launch(es_kernel k, unsigned w, unsigned h, void **args, es_stream s);

/* ---- CAMBI (replaces class Cambi, src/cambi.hpp:9-35, src/cambi.cpp:31-184) ---- */

/* The parameters Cambi::Cambi derives from the picture size (src/cambi.cpp:39-58); window_base is 65 in the
 * reference (src/cambi.cpp:39). */
typedef struct {
    int w, h, depth, window, pad, mask_index;
    int vlt, band_base, band_size;
    int tvi[4], weight[4];
    int sw[5], sh[5];
    long long n[5], k[5];
} es_cambi_info;
int es_cambi_params(int w, int h, int depth, int window_base, es_cambi_info *out);

/* Per-scale pooling figures of the last finished picture (Cambi::finish / topk_exact, src/cambi.cpp:137-181). */
typedef struct {
    unsigned long long sum_fx;   /* exact sum of the positive c-values, units of 2^-24 */
    unsigned long long count;    /* positive c-values */
    unsigned long long k;        /* clip((int)(0.6 n), 1, n) */
    unsigned int thr_bits;       /* the k-th largest c-value's bits when count > k, else 0 */
    unsigned long long above;    /* c-values strictly above it */
    unsigned long long above_fx; /* their exact sum */
    double pooled;
} es_cambi_scale;

/* Cambi(const Kernels&, int w, int h, int depth) (src/cambi.hpp:12). */
int es_cambi_create(int w, int h, int depth, es_cambi **out);
/* As es_cambi_create with the window's base constant (65 in the reference) and full c-value planes kept for
 * es_cambi_cvalues when keep_cvalues is nonzero. */
int es_cambi_create_ex(int w, int h, int depth, int window_base, int keep_cvalues, es_cambi **out);
int es_cambi_destroy(es_cambi *c);
/* The c-value kernel of the instance's next launches (no counterpart in the reference, whose cambi_cvalues has one
 * form, src/kernels.cu:129-278): window counts on the matrix cores (cv_mfma != 0, the default) or by prefix scans,
 * with an LDS layout of cv_bmax (32, the default, or 64) band values per pass.  Every form gives the same c-values. */
int es_cambi_set_method(es_cambi *c, int cv_mfma, int cv_bmax);
/* Cambi::launch(CUdeviceptr img, int pitch, CUstream) (src/cambi.hpp:15-18): img is a device plane of uint16
 * samples with `pitch` in pixels; it is read in stream order on `stream` and the rest runs on the instance's own
 * stream, so the caller may overwrite the plane with work queued on `stream` afterwards. */
int es_cambi_launch(es_cambi *c, const void *img, int pitch, es_stream stream);
/* As es_cambi_launch for planes of `bytes` (1 or 2) per sample, shifted left by `shift` on reading. */
int es_cambi_launch_ex(es_cambi *c, const void *img, int pitch, int bytes, int shift, es_stream stream);
/* double Cambi::finish() (src/cambi.hpp:19-20): blocks and returns the score, clamped to 1000. */
int es_cambi_finish(es_cambi *c, double *score);
int es_cambi_scales(es_cambi *c, es_cambi_scale out[5]);
/* Diagnostics of the last finished picture: the packed plane of a scale (value | mask << 15, sw x sh) and, when the
 * instance keeps them, its c-values (sw x sh floats), into host memory. */
int es_cambi_plane(es_cambi *c, int scale, uint16_t *host_out);
int es_cambi_cvalues(es_cambi *c, int scale, float *host_out);

/* ---- window energies and lasting texture (replace the launches of window_stats and persist_stats,
 *      src/main.cpp:539-542, :558-566, kernels src/kernels.cu:346-347, :435) ---- */

/* window_stats(yr, yd, up, w, h, mid_r, mid_d, out): for win x win windows (16 in the reference) of the pair of
 * uint16 planes (pitch w), 12 exact sums per window into `sums` (int64, windows = (w / win) * (h / win)):
 * [F_r^2, M_r^2, F_d^2, M_d^2, F_r F_d, M_r M_d, GX_r^2, GY_r^2, GX_r GY_r, GX_d^2, GY_d^2, GX_d GY_d] with F = 9 fine,
 * M = 81 mid, GX/GY = 9 times box3's central differences; mid_r / mid_d (either may be NULL) receive M as int32 on
 * the area the windows cover.  fp64 != 0 runs the double-precision form of the same arithmetic. */
int es_window_stats(const uint16_t *yr, const uint16_t *yd, int up, int w, int h, int win, int32_t *mid_r,
                    int32_t *mid_d, int64_t *sums, int fp64, es_stream s);
/* window_stats(yr, yd, up, w, h, mid_r, mid_d, out) with the reference's outputs (src/kernels.cu:331-347, launched at
 * src/main.cpp:539-542): 16 x 16 windows, mid_r / mid_d the float mid planes (box3 - box9 of each picture on the area
 * whole windows cover; either may be NULL), out 8 floats per window [fine_r^2, mid_r^2, fine_d^2, mid_d^2,
 * fine_r fine_d, mid_r mid_d, oriented_r, oriented_d] (sums over the window).  Each figure is the exact value rounded
 * once to float (the reference sums in float, so its low bits differ by that rounding). */
int es_window_stats_f32(const uint16_t *yr, const uint16_t *yd, int up, int w, int h, float *mid_r, float *mid_d,
                        float *out, es_stream s);
/* es_window_values: the reference's 8 figures per window (kernels.cu:340-341) from the 12 sums, on the host. */
int es_window_values(const int64_t *sums, int windows, double *values);
/* persist_stats(ring, plane, centre, mid_d, w, out): `ring` holds 13 int32 mid planes `plane` elements apart, the
 * centre frame's at slot `centre` (offset o at slot (centre + o + 13) % 13); 5 exact sums per window into `sums`:
 * [M_r(c) ST, M_d(c) ST, SA SB, SA^2, SB^2] with SA/SB the sums at even/odd offsets and ST = SA + SB. */
int es_persist_stats(const int32_t *ring, size_t plane, int centre, const int32_t *mid_d, int w, int h, int win,
                     int64_t *sums, es_stream s);
/* persist_stats(ring, plane, centre, mid_d, w, out) with the reference's inputs and outputs (src/kernels.cu:430-451,
 * launched at src/main.cpp:558-566 over (w / 16) x (h / 16) windows): `ring` holds 13 float mid planes `plane`
 * elements apart, the centre frame's at slot `centre`; per 16 x 16 window out receives 5 floats [sum m(c) T,
 * sum mid_d T, sum A B, sum A^2, sum B^2] with A, B the means at offsets +-2, 4, 6 and +-1, 3, 5 and T = (A + B) / 2,
 * evaluated in double and rounded to float. */
int es_persist_stats_f32(const float *ring, size_t plane, int centre, const float *mid_d, int w, int h, float *out,
                         es_stream s);

/* ---- selects (replace the radix select of src/kernels.cu:71-84 + src/cambi.cpp:144-157, quantile of
 *      src/main.cpp:279-289 and Histogram::quantile of src/stats.hpp:21-30) ---- */

/* The key of 0-based ascending rank `rank` among n device keys, into host *out. */
int es_select_rank_u32(const uint32_t *keys, size_t n, size_t rank, uint32_t *out, es_stream s);
int es_select_rank_u64(const uint64_t *keys, size_t n, size_t rank, uint64_t *out, es_stream s);
/* Histogram::quantile of nhist device histograms of 20001 uint64 bins each, for nq quantiles: host out[h * nq + j]
 * (NaN for an empty histogram) and host counts[h] (may be NULL). */
int es_hist_quantiles(const uint64_t *hists, int nhist, const double *qs, int nq, double *out, uint64_t *counts);
/* es_hist_quantiles without allocating: no counterpart in the reference; for callers that cannot allocate device or
 * pinned memory (code under graph capture, tight serving loops, harnesses that forbid it).  The same results, queued
 * on `s` and waited for; the results are staged in a static device array of the library (one per GPU, about 544 KiB,
 * held until the process ends), so at most ES_NOALLOC_MAX_HISTS histograms and ES_NOALLOC_MAX_QUANTILES quantiles per
 * call (ES_EINVAL beyond), and calls that overlap on one GPU race on it. */
#define ES_NOALLOC_MAX_HISTS 4096
#define ES_NOALLOC_MAX_QUANTILES 16
int es_hist_quantiles_noalloc(const uint64_t *hists, int nhist, const double *qs, int nq, double *out, uint64_t *counts,
                              es_stream s);

/* ---- frame ingest and the PSNR and colour reductions (src/kernels.cu:280-310, :453-484; src/main.cpp:128-153,
 *      :273-277, :496-505, :543-556) ---- */

/* Decoded 4:2:0 pixel formats at 8 or 10 bits (the table of view_of, src/main.cpp:135-147). */
#define ES_FMT_YUV420P 0   /* planar, 1 byte per sample */
#define ES_FMT_YUV420P10 1 /* planar, 2 bytes per sample, the value in the low 10 bits */
#define ES_FMT_NV12 2      /* a luma plane and one plane of interleaved u, v; 1 byte per sample */
#define ES_FMT_P010 3      /* as NV12 with 2 bytes per sample, the value in the high 10 bits */

/* A decoded frame as the kernels read it (struct Frame, src/kernels.cu:284-287): planar or semi-planar (u and v
 * interleaved in plane u; v unused) 4:2:0; a sample's value is the stored word shifted right by `shift` and has
 * `depth` bits.  Pitches are in bytes. */
typedef struct {
    const void *y, *u, *v;
    int ypitch, cpitch, semi, bytes, shift, depth;
} es_frame;
/* The description of a frame of `format` (view_of, src/main.cpp:128-153): ES_EINVAL for an unsupported format. */
int es_frame_view(int format, const void *y, const void *u, const void *v, int ypitch, int cpitch, es_frame *out);
/* frame_planes (src/kernels.cu:297-310): the frame cropped at (x0, y0) to w x h (both even) as packed luma y16
 * (w * h, 4-byte aligned) and chroma u, v ((w / 2) * (h / 2)), brought to the pipeline depth pd (8 or 10; an 8-bit
 * frame in a 10-bit pipeline is shifted up by 2, as libvmaf does).  Queued on s. */
int es_frame_planes(const es_frame *f, int x0, int y0, int pd, uint16_t *y16, uint16_t *u, uint16_t *v, int w, int h,
                    es_stream s);
/* plane_sse (src/kernels.cu:454-465): adds to the device *sse the sum of squared differences of n samples of a and b,
 * or, when a is NULL, of the floats fa and fb truncated to integers.  Integer adds: the same total on every run. */
int es_plane_sse(const uint16_t *a, const uint16_t *b, const float *fa, const float *fb, size_t n, uint64_t *sse,
                 es_stream s);
/* color_sums (src/kernels.cu:470-484) in exact integers: adds to the device sums[0 .. 2] the sums over n samples of
 * u << up, v << up, and 2^23 times the colour strength sqrtf(du^2 + dv^2) (du = (u << up) - 512, likewise dv) rounded
 * to single precision as the reference's is.  Samples must be at most 10-bit after the shift. */
int es_color_sums(const uint16_t *u, const uint16_t *v, size_t n, int up, uint64_t *sums, es_stream s);
/* The mean colour of n samples from es_color_sums's host copy: out[0 .. 2] = mean u, mean v, mean strength (each
 * frame's contribution to Accum's color, src/main.cpp:603-604). */
int es_color_values(const uint64_t *sums, size_t n, double *out);
/* psnr (src/main.cpp:273-277): libvmaf's PSNR of a squared error over n samples at bit depth `depth`, peak
 * 2^depth - 1, capped at 6 depth + 12 dB. */
int es_psnr(uint64_t sse, size_t n, int depth, double *out);

/* ---- the source-CAMBI cache (replaces class CambiCache, src/cache.hpp:9-29, src/cache.cpp:14-88) ---- */

#define ES_NOPTS INT64_MIN /* a frame without a timestamp (AV_NOPTS_VALUE): never found, never kept */
typedef struct es_cambi_cache_s es_cambi_cache;
/* CambiCache(path, x0, y0, w, h, depth): the per-frame CAMBI of the source file `path` cropped to w x h at (x0, y0)
 * and scored at `depth` bits, kept in `dir` (NULL: $XDG_CACHE_HOME/encode-score, else $HOME/.cache/encode-score).
 * The file format and naming are the reference's (16-byte {int64 pts, double score} records; FNV-1a of the key and
 * the source's first and last MiB) with this library's own version in the key, so the scores of a different CAMBI
 * implementation are never read as this one's.  When the source or the directory cannot be used the cache is inert:
 * it finds nothing and keeps nothing. */
int es_cambi_cache_open(const char *path, int x0, int y0, int w, int h, int depth, const char *dir,
                        es_cambi_cache **out);
/* CambiCache::find: *found = 1 and *score when the frame of timestamp pts has a cached score. */
int es_cambi_cache_find(const es_cambi_cache *c, int64_t pts, double *score, int *found);
/* CambiCache::add: keeps the score (written to the file every 240 records and on close). */
int es_cambi_cache_add(es_cambi_cache *c, int64_t pts, double score);
/* The cache file's path ("" when inert), valid until es_cambi_cache_close. */
const char *es_cambi_cache_file(const es_cambi_cache *c);
int es_cambi_cache_close(es_cambi_cache *c);

/* ---- the whole measure path of one source against its encodes (src/main.cpp:516-643 queue/collect,
 *      src/stats.hpp Accum) ---- */

#define ES_PLAIN 0   /* one launch per picture and per kernel, FP64 energies, collect on the host */
#define ES_BATCHED 1 /* many frames per launch, integer energies, collect and selects on the device, graph replay */

typedef struct {
    int w, h;       /* the encode's picture; the source is centre-cropped to it (offsets (W - w) / 2 & ~1) */
    int length;     /* frames of the encode; it leaves the run after them */
    int pitch;      /* samples per row of its planes */
    int bytes;      /* 1 or 2 bytes per sample */
    int shift;      /* left shift of its samples to the pipeline depth */
} es_encode_desc;

typedef struct {
    int mode;         /* ES_PLAIN or ES_BATCHED */
    int src_w, src_h, src_pitch, src_bytes, src_shift;
    int depth;        /* pipeline depth: 10, or 8 when every file is 8-bit (src/main.cpp:382, src/cambi.cpp:93) */
    int win;          /* window of the energies: 16 (reference), 32 or 64 */
    int per_segment;  /* frames per segment (src/main.cpp:434) */
    int window_base;  /* CAMBI window base constant, 65 (src/cambi.cpp:39) */
    int batch;        /* frames per batch in the batched form (1 to 128; 0 = 16) */
    int tile_h;       /* c-value tile height, 0 = default */
    int use_graph;    /* batched form: capture the per-batch pipeline once and replay it */
    int n_encodes;
    const es_encode_desc *encodes;
    int profile;      /* batched form: time each stage of every round with events (es_scorer_stage_ms) */
    int cv_bmax;      /* c-value LDS layout: numbers per pass, 32 or 64 (0 = 32) */
    int cv_waves;     /* batched form: persistent c-value waves per CU (0 = 16) */
    int serial_stages; /* batched form: run the CAMBI and energy chains of a round one after the other (0: overlap them
                        * on two streams) */
    int cv_tile_blocks; /* batched form: launch the c-value kernel as one block per tile instead of persistent waves,
                         * which lets the hardware interleave it with the overlapping energy chain */
    int early_energies; /* batched form, overlapped: start the energy chain with the CAMBI pyramid (1) instead of after
                         * it, with the c-value kernel (0) */
    int cv_mfma;        /* batched form: c-value window counts on the matrix cores (1) instead of prefix scans (0);
                         * the results are the same */
    int cv_queues;      /* batched form, persistent c-value waves: one tile queue shared by every XCD (0), or one
                         * queue per XCD, each with an eighth of every scale's tiles, so that tiles sharing halo rows
                         * and columns meet in one L2 (1) */
    int tile_h0;        /* c-value tile height at full resolution (scale 0), 0 = tile_h; at most 4096 */
} es_scorer_config;

/* One segment of one encode (Accum, src/stats.hpp:36-82), with its histograms' quantiles and the score parts.
 * Fidelity is VMAF's, which this path does not compute: with VMAF attached to every frame of the segment
 * (es_scorer_set_vmaf) it is clamp(vmaf / frames) as in the reference, otherwise NaN and left out of the weighted
 * mean as a part that could not be measured is.  The PSNR and colour fields are sums over the frames, as Accum keeps
 * them; they stay 0 for frames submitted without chroma planes (chroma_frames counts those with). */
typedef struct {
    uint64_t frames;
    double cambi, cambi_source;
    double all_r[3], all_d[3], st_r[3], st_d[3];
    double sh_all_r[2], sh_all_x[2], sh_st_r[2], sh_st_x[2];
    double own, mine;
    uint64_t n_amount[3], n_matching[2], n_lasting;
    double median_amount[3], tenth_amount[3], median_matching[2], tenth_matching[2], median_lasting, tenth_lasting;
    double texture, banding, grain, score;
    double psnr[3];                           /* y, cb, cr */
    double color_source[3], color_encode[3];  /* mean u, mean v, mean strength, at the 10-bit scale */
    uint64_t chroma_frames;
    double vmaf, vmaf_min;                    /* sum and minimum (1e9 when none) of the attached VMAF */
    uint64_t vmaf_frames;
    double fidelity;
} es_segment;

int es_scorer_create(const es_scorer_config *cfg, es_scorer **out);
/* Settings the autotuner chose and the time per frame step it measured with them. */
typedef struct {
    int batch, tile_h, cv_bmax, cv_waves;
    double ms_per_step;
    int trials;
    int cv_tile_blocks, early_energies, cv_mfma, cv_queues;
    int tile_h0;
} es_tune_result;
/* Autotuner (no counterpart in the reference): times the batched form of cfg's workload on the caller's frames
 * (nframes steps as for es_scorer_submit, at least three batches of the largest batch tried), trying the c-value
 * method (matrix cores or prefix scans), the tile height with the c-value LDS layout (together), then the tile height
 * at full resolution, then the c-value waves per CU or one block per tile, then the persistent waves' queues, then
 * when the energy chain starts, then the batch, and returns the fastest.  Each trial allocates a scorer of its
 * own. */
int es_scorer_autotune(const es_scorer_config *cfg, int nframes, const void *const *src, const void *const *enc,
                       es_tune_result *out);
int es_scorer_destroy(es_scorer *s);
/* Queues `nframes` frame steps: src[i] is the source's device plane of step i (the full picture), enc[i * n_encodes
 * + e] encode e's (ignored once the encode has ended). Device planes must stay unchanged until es_scorer_sync. */
int es_scorer_submit(es_scorer *s, int nframes, const void *const *src, const void *const *enc, es_stream stream);
/* Queues `nframes` frame steps of decoded frames, as es_scorer_submit: src[i] the source's frame of step i (the full
 * picture), enc[i * n_encodes + e] encode e's.  Each frame carries its own format and pitches (the config's pitch,
 * bytes and shift are not used) and is read in place.  With chroma planes (u not NULL) the frame step also measures
 * the PSNR of luma and both chroma planes and the mean colour of both pictures (src/main.cpp:525, :543-556,
 * :594-605); every frame of a run must carry chroma, or none.  Crops and picture sizes must be even. */
int es_scorer_submit_frames(es_scorer *s, int nframes, const es_frame *src, const es_frame *enc, es_stream stream);
/* Uses source-CAMBI caches of the source file `source_path` (one per crop, keyed as CambiCache keys them,
 * src/main.cpp:454-455; dir as for es_cambi_cache_open): frames submitted with timestamps whose source CAMBI is
 * cached skip the source's CAMBI, and the scores computed for the others are added (src/main.cpp:520-521, :585-587). */
int es_scorer_use_cache(es_scorer *s, const char *source_path, const char *dir);
/* As es_scorer_submit_frames, with the source frames' timestamps pts[i] (ES_NOPTS for none). */
int es_scorer_submit_pts(es_scorer *s, int nframes, const es_frame *src, const es_frame *enc, const int64_t *pts,
                         es_stream stream);
/* Source pictures whose CAMBI came from the cache so far, and the cache file of each crop (as for es_cambi_cache_file;
 * crop k is the k-th distinct encode size). */
int es_scorer_cache_hits(es_scorer *s, long *hits);
int es_scorer_cache_file(es_scorer *s, int crop, const char **file);
/* Waits for the queued work and folds its figures into the segments. */
int es_scorer_sync(es_scorer *s);
int es_scorer_frames(es_scorer *s, long *frames);
int es_scorer_segments(es_scorer *s, int encode, int *count);
/* Segment `seg` of encode `encode`, or the whole run merged when seg < 0; hist (may be NULL) receives its 6
 * histograms (amount x3, matching x2, lasting), 20001 bins each. */
int es_scorer_segment(es_scorer *s, int encode, int seg, es_segment *out, uint64_t *hist);
/* Per-frame CAMBI of the encode and of the source cropped to it, for frames [0, n). */
int es_scorer_cambi(es_scorer *s, int encode, long n, double *cambi, double *cambi_source);
/* Per-frame PSNR of the encode for frames [0, n): psnr[3 i + k] for y, cb, cr (NaN for frames without chroma). */
int es_scorer_psnr(es_scorer *s, int encode, long n, double *psnr);
/* Attaches VMAF scores of the encode's frames [0, n) computed elsewhere (libvmaf in the reference, src/main.cpp:698-704):
 * they fill the segments' vmaf figures and the fidelity part of the score. */
int es_scorer_set_vmaf(es_scorer *s, int encode, long n, const double *vmaf);

/* What the run's report names besides the figures (src/main.cpp:693-695). */
typedef struct {
    const char *source;            /* the source's path */
    const char *const *paths;      /* each encode's path, n_encodes of them */
    const char *model;             /* the VMAF model's name */
    double start, offset, fps, segment_seconds;
    int width, height;             /* the scored picture */
} es_run_info;
/* The run's JSON report in the reference's format (src/main.cpp:693-777: the header, per encode the whole run, the
 * segments and the per-frame lists; the same field names and digits, null for a figure that could not be measured)
 * written to `path`. */
int es_scorer_write_json(es_scorer *s, const es_run_info *info, const char *path);
/* Batched form with profile set: the device milliseconds of each stage summed over the synced rounds, for
 * ES_STAGES stages in the order ES_STAGE_NAMES lists, and the number of rounds.  Each stage is timed on its own
 * stream, so with overlapping chains the stages' times add up to more than the round's. */
#define ES_STAGES 11
#define ES_STAGE_NAMES "upload,cambi_base,cambi_scale,cambi_cvalues,cambi_topk,cambi_finish,window_stats,frame_sums,persist,collect,results"
int es_scorer_stage_ms(es_scorer *s, double *ms, long *rounds);
/* Per-frame diagnostics of the last synced frames: the 12 window sums of (encode, frame) when still held, the
 * structure threshold of each frame (NaN when not kept). */
int es_scorer_thresholds(es_scorer *s, int encode, long n, double *q);

/* A plain copy kernel of `bytes` (a multiple of 16) between device buffers, for calibrating HBM byte counters; no
 * counterpart in the reference. */
int es_copy_kernel(void *dst, const void *src, size_t bytes, es_stream s);

/* ---- synthetic inputs (the generator of the benchmark configurations; no counterpart in the reference, which
 *      decodes files) ---- */

typedef struct {
    int kind;        /* 0: gradient + radial ramp with grain (config 1); 1: 64x64 tile mosaic (configs 2-5) */
    uint64_t seed;   /* content and grain seed, shared by a source and its encodes */
    int salt;        /* per-file seed of the block choice */
    int full_w, full_h;   /* the source's picture, in whose coordinates content is defined */
    int crop_x, crop_y, w, h;
    int depth;       /* 10: codes in [64, 940]; 8: codes in [16, 235] */
    /* distortions (0 = off) */
    int grain_q8;        /* grain retention in 1/256 (256 = the source's grain) */
    int blur_q8;         /* Gaussian blur sigma in 1/256 px */
    int quant_step;      /* quantise to multiples of this many codes */
    int quant_where;     /* 0 everywhere, 1 gradient tiles, 2 dark tiles */
    int block;           /* block-mean size (8 or 16) */
    int block_pct;       /* percentage of blocks averaged */
    int box3;            /* config 1's encode: 3x3 box mean of the 4-code-quantised ramp, no grain */
} es_synth_desc;
/* Frame `frame` of the described file into a device plane (pitch w samples, `bytes` 1 or 2 per sample). */
int es_synth_frame(const es_synth_desc *d, int frame, void *dev_out, int bytes, es_stream s);
/* The same frame computed on the host (bit-identical), into host memory. */
int es_synth_frame_host(const es_synth_desc *d, int frame, void *host_out, int bytes, int threads);

#ifdef __cplusplus
}
#endif
#endif
