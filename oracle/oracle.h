/*
 * CPU oracle for encode-score's per-frame measure path: CAMBI banding, the detail / structure-tensor window energies,
 * the lasting-texture window sums, the order-statistic selects and the per-segment accumulation.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library calls, links or loads this code: it is the checker that
 * the tests, the smoke check and the CPU-baseline leg of the bench compare the HIP path against.
 *
 * Every function is a plain-C restatement of the reference's algorithm; each cites the reference file:line it follows
 * (paths relative to the reference tree).  Where the reference computes in FP32 in an order that depends on its launch
 * shape, this oracle computes the mathematically exact value instead (integer sums of integer-valued quantities) and
 * rounds once in FP64; those canonical forms are documented at each function.
 *
 * How each function is pinned (checked against something other than itself) is recorded in the tests: brute force on
 * tiny pictures, closed forms, the reference's own known-answer tests, and the reference's stats.hpp compiled from
 * its own source under oracle/_ref.
 */
#ifndef ENCSCORE_ORACLE_H
#define ENCSCORE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_SCALES 5
#define ORC_DIFFS 4
#define ORC_NONE 0xFFFF
#define ORC_HIST_BINS 20001

/* CAMBI parameters of one picture size (src/cambi.cpp:31-70). */
typedef struct {
    int w, h, depth, window, pad, mask_index;
    int vlt, band_base, band_size;
    int tvi[ORC_DIFFS], weight[ORC_DIFFS];
    int sw[ORC_SCALES], sh[ORC_SCALES];
    long long n[ORC_SCALES], k[ORC_SCALES];
    int lut_size; /* (2 pad + 1)^2 + 1 entries of the reciprocal table */
} orc_cambi_params;

/* Per-scale pooling results and the score of one picture (src/cambi.cpp:137-184). */
typedef struct {
    unsigned long long sum_fx[ORC_SCALES];   /* exact sum of the positive c-values, in units of 2^-24 */
    unsigned long long count[ORC_SCALES];    /* number of positive c-values */
    unsigned long long k[ORC_SCALES];
    unsigned int thr_bits[ORC_SCALES];       /* bits of the k-th largest c-value when count > k, else 0 */
    unsigned long long above[ORC_SCALES];    /* c-values strictly above it */
    unsigned long long above_fx[ORC_SCALES]; /* their exact sum, units of 2^-24 */
    double pooled[ORC_SCALES];
    double score;
} orc_cambi_result;

double orc_bt1886_luminance(int sample);
int orc_tvi_for_diff(int diff, double tvi_threshold);
int orc_vlt_luma(double vlt_threshold);
int orc_cambi_params_init(orc_cambi_params *p, int w, int h, int depth, int window_base);
void orc_recip_lut(float *lut, int n);

void orc_from8(const uint16_t *img, int pitch, int w, int h, uint16_t *out);
void orc_spatial_mask(const uint16_t *img, int w, int h, int index, uint8_t *mask);
void orc_decimate16(const uint16_t *src, int sw, int dw, int dh, uint16_t *dst);
void orc_decimate8(const uint8_t *src, int sw, int dw, int dh, uint8_t *dst);
void orc_mode_filter(const uint16_t *img, int w, int h, uint16_t *out);
void orc_pyramid(const orc_cambi_params *p, const uint16_t *img, int pitch, uint16_t *images, uint8_t *masks);
void orc_cvalues(const orc_cambi_params *p, const uint16_t *img, const uint8_t *mask, int w, int h, const float *lut,
                 float *out, int brute, int threads);
unsigned long long orc_fixed24(float c);
void orc_pool(const float *c, long long n, long long k, orc_cambi_result *r, int s);
double orc_cambi_score_of(const orc_cambi_params *p, orc_cambi_result *r);
int orc_cambi(const orc_cambi_params *p, const uint16_t *img, int pitch, float *cvalues, orc_cambi_result *r, int brute,
              int threads);

/* Window energies (src/kernels.cu:331-428, reference/encode_score/compare.py:77-108). */
void orc_window_stats(const uint16_t *yr, int pitch_r, const uint16_t *yd, int pitch_d, int up, int w, int h, int win,
                      int32_t *mid_r, int32_t *mid_d, long long *sums, double *stats);
void orc_window_values(const long long *sums, double *stats);
void orc_persist(const int32_t *const *ring, const int32_t *mid_d, int w, int h, int win, long long *sums);
void orc_persist_values(const long long *sums, double *vals);

/* Selects and the histogram (src/main.cpp:279-289, src/stats.hpp:8-32). */
double orc_quantile_linear(const double *v, long long n, double q);
float orc_quantile_f32(const float *v, long long n, double q);
unsigned long long orc_select_rank_u64(const unsigned long long *keys, long long n, long long rank);
double orc_tree_sum256(const double *v, long long n);

typedef struct {
    uint64_t bins[ORC_HIST_BINS];
    uint64_t n;
} orc_hist;
int orc_hist_bin(double pct);
void orc_hist_add(orc_hist *h, double pct);
double orc_hist_quantile(const orc_hist *h, double q);
void orc_hist_merge(orc_hist *h, const orc_hist *o);

/* The per-segment accumulator (src/stats.hpp:36-82), measure parts only. */
typedef struct {
    uint64_t frames;
    double cambi, cambi_source;
    double all_r[3], all_d[3], st_r[3], st_d[3];
    double sh_all_r[2], sh_all_x[2], sh_st_r[2], sh_st_x[2];
    double own, mine;
    orc_hist amount[3], matching[2], lasting;
    /* beside the measures: PSNR sums, mean colour sums (source, encode) and VMAF (src/stats.hpp:37-39) */
    double psnr[3], color_r[3], color_d[3];
    double vmaf, vmaf_min;
} orc_accum;

typedef struct {
    double fidelity, texture, banding, grain, score;
} orc_parts;

void orc_accum_clear(orc_accum *a);
void orc_accum_merge(orc_accum *a, const orc_accum *o);
void orc_accum_score(const orc_accum *a, double vmaf_total, orc_parts *p);
void orc_collect_frame(orc_accum *acc, const long long *sums, int windows, uint8_t *mask_out);
void orc_collect_lasting(orc_accum *acc, const long long *psums, const uint8_t *cmask, int windows);

/* Frame ingest, PSNR and colour (src/kernels.cu:282-310, :453-484; src/main.cpp:273-277, :594-605).  orc_frame has
 * the layout of struct Frame (src/kernels.cu:284-287); pitches in bytes. */
typedef struct {
    const void *y, *u, *v;
    int ypitch, cpitch, semi, bytes, shift, depth;
} orc_frame;
void orc_frame_planes(const orc_frame *f, int x0, int y0, int pd, uint16_t *y16, uint16_t *u, uint16_t *v, int w, int h);
unsigned long long orc_plane_sse(const uint16_t *a, const uint16_t *b, long long n);
unsigned long long orc_plane_sse_f(const float *a, const float *b, long long n);
unsigned long long orc_strength_fixed(unsigned s);
void orc_color_sums(const uint16_t *u, const uint16_t *v, long long n, int up, unsigned long long *sums);
double orc_psnr(unsigned long long sse, long long n, int depth);

/* Whole-run driver for one (source, encode) pair, following src/main.cpp:516-643. */
typedef struct {
    int w, h, depth, up, win, per_segment, window_base, threads;
} orc_run_config;
typedef struct orc_run orc_run;
orc_run *orc_run_create(const orc_run_config *c);
void orc_run_destroy(orc_run *r);
int orc_run_frame(orc_run *r, const uint16_t *src, int src_pitch, const uint16_t *enc, int enc_pitch, double *cambi,
                  double *cambi_source);
/* The same frame with its chroma: y planes w x h (pitch w), u and v (w / 2) x (h / 2), all at the pipeline depth;
 * adds the frame's PSNR and both mean colours to its segment (src/main.cpp:594-605). */
int orc_run_frame_yuv(orc_run *r, const uint16_t *const *src, const uint16_t *const *enc, double *cambi,
                      double *cambi_source, double *psnr);
int orc_run_segments(const orc_run *r);
const orc_accum *orc_run_segment(const orc_run *r, int s);

#ifdef __cplusplus
}
#endif
#endif
