/*
 * CPU oracle for the per-frame measure path.  TEST INFRASTRUCTURE ONLY: the product library never calls, links or
 * loads this file; only the tests, the smoke check and the bench's CPU-baseline leg do, as the checker.
 *
 * Reference citations are file:line in the reference tree (wakamex/encode-score v0.1.0).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------------------------------------
 * Luminance tables.  The reference calls libvmaf's vmaf_luminance_init_luma_range(10, limited),
 * vmaf_luminance_init_eotf("bt1886"), get_tvi_for_diff and get_vlt_luma (src/cambi.cpp:8-15, :41-53).  libvmaf is not
 * in the reference tree; these restate its published algorithm: the ITU-R BT.1886 EOTF L = a * max(V + b, 0)^2.4 with
 * white 300 cd/m^2 and black 0.01 cd/m^2 on the limited 10-bit range [64, 940], V = (clip(s) - 64) / 876.
 * ------------------------------------------------------------------------------------------------------------------ */
static const double kGamma = 2.4, kWhite = 300.0, kBlack = 0.01;
static const int kFoot = 64, kHead = 940;

double orc_bt1886_luminance(int sample)
{
    const int s = sample < kFoot ? kFoot : sample > kHead ? kHead : sample;
    const double v = (double)(s - kFoot) / (double)(kHead - kFoot);
    const double lw = pow(kWhite, 1.0 / kGamma), lb = pow(kBlack, 1.0 / kGamma);
    const double a = pow(lw - lb, kGamma), b = lb / (lw - lb);
    const double x = v + b > 0 ? v + b : 0;
    return a * pow(x, kGamma);
}

/* The step of `diff` codes above `sample` is visible when it raises the luminance by more than tvi_threshold of it. */
static int tvi_condition(int sample, int diff, double thr)
{
    const double mean = orc_bt1886_luminance(sample), up = orc_bt1886_luminance(sample + diff);
    return up - mean > thr * mean;
}

/* get_tvi_for_diff: the largest code of the range whose `diff` step is still visible.  The condition holds below a
 * point and fails above it (the relative step of a power law falls with the level), so libvmaf's bisection for the
 * code where it holds and fails one code up returns the last code where it holds; a scan gives the same code. */
int orc_tvi_for_diff(int diff, double thr)
{
    int last = kFoot;
    for (int s = kFoot; s <= kHead; s++) {
        if (!tvi_condition(s, diff, thr)) break;
        last = s;
    }
    return last;
}

/* get_vlt_luma: the first code, scanning up from 0, whose luminance reaches the visibility threshold.  With the
 * default threshold 0.0 (src/cambi.cpp:50) that is code 0, which leaves every code visible. */
int orc_vlt_luma(double thr)
{
    for (int s = 0; s < 1024; s++)
        if (orc_bt1886_luminance(s) >= thr) return s;
    return 1023;
}

/* ceil_log2 (src/cambi.cpp:21-26) */
static int ceil_log2(uint32_t n)
{
    int s = 0;
    for (uint32_t t = n ? n - 1 : 0; t; t >>= 1) s++;
    return s;
}

/* Cambi::Cambi parameter derivation (src/cambi.cpp:39-58) and topk_count (src/cambi.cpp:29). */
int orc_cambi_params_init(orc_cambi_params *p, int w, int h, int depth, int window_base)
{
    if (w < 1 || h < 1 || (depth != 8 && depth != 10)) return -1;
    memset(p, 0, sizeof *p);
    p->w = w;
    p->h = h;
    p->depth = depth;
    p->window = ((window_base * (w + h) / 375) >> 4) | 1;
    p->pad = p->window >> 1;
    p->mask_index = (49 + 3 * (ceil_log2((uint32_t)((w >> 6) * (h >> 6))) - 11) - 1) >> 1;
    for (int d = 0; d < ORC_DIFFS; d++) {
        p->tvi[d] = orc_tvi_for_diff(d + 1, 0.019) + ORC_DIFFS;
        p->weight[d] = d + 1;
    }
    p->vlt = orc_vlt_luma(0.0);
    p->band_base = p->vlt - 3 * ORC_DIFFS + 1 > 0 ? p->vlt - 3 * ORC_DIFFS + 1 : 0;
    p->band_size = p->tvi[ORC_DIFFS - 1] + 1 - p->band_base;
    p->sw[0] = w;
    p->sh[0] = h;
    for (int s = 1; s < ORC_SCALES; s++) {
        p->sw[s] = (p->sw[s - 1] + 1) >> 1;
        p->sh[s] = (p->sh[s - 1] + 1) >> 1;
    }
    for (int s = 0; s < ORC_SCALES; s++) {
        const long long n = (long long)p->sw[s] * p->sh[s];
        long long k = (long long)(int)(0.6 * (double)n);
        if (k < 1) k = 1;
        if (k > n) k = n;
        p->n[s] = n;
        p->k[s] = k;
    }
    const int span = 2 * p->pad + 1;
    p->lut_size = span * span + 1;
    return 0;
}

/* libvmaf's reciprocal table (used at src/kernels.cu:200): entry i is 1/i rounded to single precision. */
void orc_recip_lut(float *lut, int n)
{
    if (n > 0) lut[0] = 0.0f;
    for (int i = 1; i < n; i++) lut[i] = 1.0f / (float)i;
}

/* cambi_from8 (src/kernels.cu:522-535): 8-bit values to 10 bits, then the 2x2 anti-dithering mean. */
void orc_from8(const uint16_t *img, int pitch, int w, int h, uint16_t *out)
{
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
#define AT(xx, yy) ((unsigned)img[(size_t)(yy) * pitch + (xx)] << 2)
            unsigned v;
            if (x < w - 1 && y < h - 1) v = (AT(x, y) + AT(x + 1, y) + AT(x, y + 1) + AT(x + 1, y + 1)) >> 2;
            else if (y < h - 1) v = (AT(x, y) + AT(x, y + 1)) >> 1;
            else if (x < w - 1) v = (AT(x, y) + AT(x + 1, y)) >> 1;
            else v = AT(x, y);
#undef AT
            out[(size_t)y * w + x] = (uint16_t)v;
        }
}

/* cambi_mask (src/kernels.cu:9-40): 1 where more than `index` pixels of the 7x7 square around the pixel are flat. A
 * pixel is flat if it equals its right and lower neighbours; the last column and row count as equal; pixels outside
 * the picture are never flat.  Brute force over the 49 pixels. */
void orc_spatial_mask(const uint16_t *img, int w, int h, int index, uint8_t *mask)
{
    uint8_t *flat = (uint8_t *)malloc((size_t)w * h);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const uint16_t v = img[(size_t)y * w + x];
            const int right = x == w - 1 || v == img[(size_t)y * w + x + 1];
            const int down = y == h - 1 || v == img[(size_t)(y + 1) * w + x];
            flat[(size_t)y * w + x] = (uint8_t)(right && down);
        }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int count = 0;
            for (int dy = -3; dy <= 3; dy++) {
                const int yy = y + dy;
                if (yy < 0 || yy >= h) continue;
                for (int dx = -3; dx <= 3; dx++) {
                    const int xx = x + dx;
                    if (xx >= 0 && xx < w) count += flat[(size_t)yy * w + xx];
                }
            }
            mask[(size_t)y * w + x] = (uint8_t)(count > index);
        }
    free(flat);
}

/* cambi_decimate16/8 (src/kernels.cu:42-53): every second pixel of every second row. */
void orc_decimate16(const uint16_t *src, int sw, int dw, int dh, uint16_t *dst)
{
    for (int y = 0; y < dh; y++)
        for (int x = 0; x < dw; x++) dst[(size_t)y * dw + x] = src[(size_t)(2 * y) * sw + 2 * x];
}

void orc_decimate8(const uint8_t *src, int sw, int dw, int dh, uint8_t *dst)
{
    for (int y = 0; y < dh; y++)
        for (int x = 0; x < dw; x++) dst[(size_t)y * dw + x] = src[(size_t)(2 * y) * sw + 2 * x];
}

/* mode3 (src/kernels.cu:55-60) */
static uint16_t mode3(uint16_t a, uint16_t b, uint16_t c)
{
    if (a == b || a == c) return a;
    if (b == c) return b;
    uint16_t m = a < b ? a : b;
    return m < c ? m : c;
}

/* cambi_mode (src/kernels.cu:486-500): the vertical mode of three horizontal modes; first and last rows copied, the
 * first and last columns keep their own value in the horizontal step. */
void orc_mode_filter(const uint16_t *img, int w, int h, uint16_t *out)
{
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const size_t i = (size_t)y * w + x;
            if (y == 0 || y == h - 1) {
                out[i] = img[i];
                continue;
            }
            uint16_t m[3];
            for (int k = 0; k < 3; k++) {
                const uint16_t *row = img + (size_t)(y - 1 + k) * w;
                m[k] = (x == 0 || x == w - 1) ? row[x] : mode3(row[x - 1], row[x], row[x + 1]);
            }
            out[i] = mode3(m[0], m[1], m[2]);
        }
}

static long long scale_offset(const orc_cambi_params *p, int s)
{
    long long off = 0;
    for (int i = 0; i < s; i++) off += p->n[i];
    return off;
}

/* The pictures and masks of the 5 scales (Cambi::launch, src/cambi.cpp:88-120): scale 0 is the (8-bit: promoted and
 * anti-dithered) picture through the mode filter, with the mask taken on the unfiltered picture; scale s is the mode
 * filter of scale s-1 decimated, with the mask of scale s-1 decimated.  `images` and `masks` hold the scales one
 * after another. */
void orc_pyramid(const orc_cambi_params *p, const uint16_t *img, int pitch, uint16_t *images, uint8_t *masks)
{
    const int w = p->w, h = p->h;
    uint16_t *scratch = (uint16_t *)malloc((size_t)w * h * 2);
    if (p->depth == 8) orc_from8(img, pitch, w, h, scratch);
    else
        for (int y = 0; y < h; y++) memcpy(scratch + (size_t)y * w, img + (size_t)y * pitch, (size_t)w * 2);
    orc_spatial_mask(scratch, w, h, p->mask_index, masks);
    orc_mode_filter(scratch, w, h, images);
    for (int s = 1; s < ORC_SCALES; s++) {
        const long long po = scale_offset(p, s - 1), so = scale_offset(p, s);
        orc_decimate16(images + po, p->sw[s - 1], p->sw[s], p->sh[s], scratch);
        orc_decimate8(masks + po, p->sw[s - 1], p->sw[s], p->sh[s], masks + so);
        orc_mode_filter(scratch, p->sw[s], p->sh[s], images + so);
    }
    free(scratch);
}

/* cv_value (src/kernels.cu:113-119): the value less band_base of a masked pixel inside the band, else NONE. */
static inline int cv_value(const orc_cambi_params *p, const uint16_t *img, const uint8_t *mask, int w, int x, int y)
{
    const size_t k = (size_t)y * w + x;
    const unsigned u = (unsigned)img[k] - (unsigned)p->band_base;
    return mask[k] && u < (unsigned)p->band_size ? (int)u : ORC_NONE;
}

/* The c-value of one pixel of value u (less band_base) from the window counts H[4 + j] of value u + j, j in -4..4,
 * zero outside the band: cambi.c's c_value_pixel as src/kernels.cu:188-206 computes it. */
static inline float cval(const orc_cambi_params *p, int u, const unsigned *H, const float *lut)
{
    const int value = u + p->band_base + ORC_DIFFS;
    const unsigned p0 = H[ORC_DIFFS];
    float c = 0.0f;
    for (int d = 0; d < ORC_DIFFS; d++) {
        if (value <= p->tvi[d] && value + d + 1 > p->vlt) {
            const unsigned p1 = H[ORC_DIFFS + d + 1], p2 = H[ORC_DIFFS - d - 1];
            const unsigned pm = p1 > p2 ? p1 : p2;
            const float val = (float)((unsigned)p->weight[d] * p0 * pm) * lut[pm + p0];
            if (val > c) c = val;
        }
    }
    return c;
}

/* The c-values of one scale: for every masked pixel inside the band, the counts of masked in-band pixels of the
 * (2 pad + 1)^2 window (clipped to the picture) with each value within 4 of its own, then cval; 0 elsewhere.
 * brute != 0 counts each window directly (O(window^2) per pixel); otherwise a histogram of the whole band slides
 * along each row. */
void orc_cvalues(const orc_cambi_params *p, const uint16_t *img, const uint8_t *mask, int w, int h, const float *lut,
                 float *out, int brute, int threads)
{
    const int pad = p->pad;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel
    {
        unsigned *hist = (unsigned *)calloc((size_t)p->band_size + 2 * ORC_DIFFS, sizeof(unsigned));
#pragma omp for schedule(dynamic, 4)
        for (int y = 0; y < h; y++) {
            float *row = out + (size_t)y * w;
            int any = 0;
            for (int x = 0; x < w; x++) any |= cv_value(p, img, mask, w, x, y) != ORC_NONE;
            if (!any) {
                memset(row, 0, (size_t)w * sizeof(float));
                continue;
            }
            const int y0 = y - pad < 0 ? 0 : y - pad, y1 = y + pad >= h ? h - 1 : y + pad;
            if (brute) {
                for (int x = 0; x < w; x++) {
                    const int u = cv_value(p, img, mask, w, x, y);
                    if (u == ORC_NONE) {
                        row[x] = 0.0f;
                        continue;
                    }
                    unsigned H[2 * ORC_DIFFS + 1] = {0};
                    const int x0 = x - pad < 0 ? 0 : x - pad, x1 = x + pad >= w ? w - 1 : x + pad;
                    for (int yy = y0; yy <= y1; yy++)
                        for (int xx = x0; xx <= x1; xx++) {
                            const int v = cv_value(p, img, mask, w, xx, yy);
                            if (v != ORC_NONE && v >= u - ORC_DIFFS && v <= u + ORC_DIFFS) H[v - u + ORC_DIFFS]++;
                        }
                    row[x] = cval(p, u, H, lut);
                }
                continue;
            }
            /* hist[v + ORC_DIFFS] counts value v; the margins of ORC_DIFFS on either side stay 0 */
            memset(hist, 0, ((size_t)p->band_size + 2 * ORC_DIFFS) * sizeof(unsigned));
            for (int c = 0; c <= pad && c < w; c++)
                for (int yy = y0; yy <= y1; yy++) {
                    const int v = cv_value(p, img, mask, w, c, yy);
                    if (v != ORC_NONE) hist[v + ORC_DIFFS]++;
                }
            for (int x = 0; x < w; x++) {
                if (x > 0) {
                    const int leave = x - pad - 1, enter = x + pad;
                    if (leave >= 0)
                        for (int yy = y0; yy <= y1; yy++) {
                            const int v = cv_value(p, img, mask, w, leave, yy);
                            if (v != ORC_NONE) hist[v + ORC_DIFFS]--;
                        }
                    if (enter < w)
                        for (int yy = y0; yy <= y1; yy++) {
                            const int v = cv_value(p, img, mask, w, enter, yy);
                            if (v != ORC_NONE) hist[v + ORC_DIFFS]++;
                        }
                }
                const int u = cv_value(p, img, mask, w, x, y);
                row[x] = u == ORC_NONE ? 0.0f : cval(p, u, hist + u, lut);
            }
        }
        free(hist);
    }
}

/* A positive c-value as an exact integer in units of 2^-24.  Positive c-values are at least 0.5 (weight 1, p0 = pm
 * = 1), so their last mantissa bit is worth at least 2^-24 and the product is exact. */
unsigned long long orc_fixed24(float c) { return (unsigned long long)((double)c * 16777216.0); }

static double fx_to_double(unsigned long long v) { return (double)v * (1.0 / 16777216.0); }

static int cmp_u32_desc(const void *a, const void *b)
{
    const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? 1 : x > y ? -1 : 0;
}

/* Top-k pooling of one scale (Cambi::finish and Cambi::topk_exact, src/cambi.cpp:137-181): with no more than k
 * positive values the mean of the k largest is their sum over k; otherwise the k-th largest value is selected (its bit
 * pattern: the values are non-negative) and the mean is (sum above it + (k - number above) * it) / k.  The sums are
 * exact (units of 2^-24) and rounded to double once. */
void orc_pool(const float *c, long long n, long long k, orc_cambi_result *r, int s)
{
    unsigned long long sum = 0, count = 0;
    for (long long i = 0; i < n; i++)
        if (c[i] > 0) {
            sum += orc_fixed24(c[i]);
            count++;
        }
    r->sum_fx[s] = sum;
    r->count[s] = count;
    r->k[s] = (unsigned long long)k;
    r->thr_bits[s] = 0;
    r->above[s] = 0;
    r->above_fx[s] = 0;
    if (count <= (unsigned long long)k) {
        r->pooled[s] = fx_to_double(sum) / (double)k;
        return;
    }
    uint32_t *bits = (uint32_t *)malloc(count * sizeof(uint32_t));
    long long m = 0;
    for (long long i = 0; i < n; i++)
        if (c[i] > 0) memcpy(&bits[m++], &c[i], 4);
    qsort(bits, (size_t)count, sizeof(uint32_t), cmp_u32_desc);
    const uint32_t thr = bits[k - 1];
    unsigned long long above = 0, above_fx = 0;
    for (long long i = 0; i < m; i++)
        if (bits[i] > thr) {
            float v;
            memcpy(&v, &bits[i], 4);
            above++;
            above_fx += orc_fixed24(v);
        }
    free(bits);
    float t;
    memcpy(&t, &thr, 4);
    r->thr_bits[s] = thr;
    r->above[s] = above;
    r->above_fx[s] = above_fx;
    r->pooled[s] = (fx_to_double(above_fx) + (double)(k - (long long)above) * (double)t) / (double)k;
}

/* The score from the pooled scales (src/cambi.cpp:176-183): weights 16, 8, 4, 2, 1 over the window area, capped at
 * 1000. */
double orc_cambi_score_of(const orc_cambi_params *p, orc_cambi_result *r)
{
    static const int kScaleWeights[ORC_SCALES] = {16, 8, 4, 2, 1};
    double total = 0;
    for (int s = 0; s < ORC_SCALES; s++) total += r->pooled[s] * kScaleWeights[s];
    const int odd = 2 * (p->window >> 1) + 1;
    const double score = total / (double)(odd * odd);
    r->score = score < 1000.0 ? score : 1000.0;
    return r->score;
}

/* CAMBI of one picture (Cambi::launch + Cambi::finish, src/cambi.cpp:88-184). `cvalues` (optional) receives the
 * c-values of the scales one after another. */
int orc_cambi(const orc_cambi_params *p, const uint16_t *img, int pitch, float *cvalues, orc_cambi_result *r, int brute,
              int threads)
{
    const long long total = scale_offset(p, ORC_SCALES);
    uint16_t *images = (uint16_t *)malloc((size_t)total * 2);
    uint8_t *masks = (uint8_t *)malloc((size_t)total);
    float *cv = cvalues ? cvalues : (float *)malloc((size_t)total * 4);
    float *lut = (float *)malloc((size_t)p->lut_size * 4);
    if (!images || !masks || !cv || !lut) return -1;
    orc_recip_lut(lut, p->lut_size);
    orc_pyramid(p, img, pitch, images, masks);
    memset(r, 0, sizeof *r);
    for (int s = 0; s < ORC_SCALES; s++) {
        const long long o = scale_offset(p, s);
        orc_cvalues(p, images + o, masks + o, p->sw[s], p->sh[s], lut, cv + o, brute, threads);
        orc_pool(cv + o, p->n[s], p->k[s], r, s);
    }
    orc_cambi_score_of(p, r);
    free(images);
    free(masks);
    free(lut);
    if (!cvalues) free(cv);
    return 0;
}

/* ------------------------------------------------------------------------------------------------------------------
 * Window energies.  window_stats (src/kernels.cu:331-428) and compare.py's bands/window_energies
 * (reference/encode_score/compare.py:77-108): fine = y - box3, mid = box3 - box9 (edges repeated), gx/gy the central
 * differences of box3 (0 on the picture's border), summed per win x win window over the area whole windows cover.
 * The box weights are 1/9 and 1/81 and the luma is an integer, so every quantity is an exact integer once scaled:
 * F = 9 fine = 9y - S3, M = 81 mid = 9 S3 - S9, GX = 9 gx = S3(x+1) - S3(x-1), GY likewise, with S3 and S9 the 3x3 and
 * 9x9 sums.  The window sums are exact in int64; `sums` holds per window
 *   [F_r^2, M_r^2, F_d^2, M_d^2, F_r F_d, M_r M_d, GX_r^2, GY_r^2, GX_r GY_r, GX_d^2, GY_d^2, GX_d GY_d]
 * and `stats` the reference's 8 figures from them in FP64 (orc_window_values).  The mid planes are M (int32), written
 * on the area the windows cover.
 * ------------------------------------------------------------------------------------------------------------------ */
static void box_sums(const uint16_t *y, int pitch, int up, int w, int h, int r, int32_t *out)
{
    int32_t *rows = (int32_t *)malloc((size_t)w * h * 4);
    for (int yy = 0; yy < h; yy++)
        for (int x = 0; x < w; x++) {
            int32_t s = 0;
            for (int d = -r; d <= r; d++) {
                int xx = x + d;
                xx = xx < 0 ? 0 : xx >= w ? w - 1 : xx;
                s += (int32_t)y[(size_t)yy * pitch + xx] << up;
            }
            rows[(size_t)yy * w + x] = s;
        }
    for (int yy = 0; yy < h; yy++)
        for (int x = 0; x < w; x++) {
            int32_t s = 0;
            for (int d = -r; d <= r; d++) {
                int y2 = yy + d;
                y2 = y2 < 0 ? 0 : y2 >= h ? h - 1 : y2;
                s += rows[(size_t)y2 * w + x];
            }
            out[(size_t)yy * w + x] = s;
        }
    free(rows);
}

void orc_window_stats(const uint16_t *yr, int pitch_r, const uint16_t *yd, int pitch_d, int up, int w, int h, int win,
                      int32_t *mid_r, int32_t *mid_d, long long *sums, double *stats)
{
    const int ww = w / win, wh = h / win, W = ww * win, H = wh * win;
    int32_t *s3[2], *s9[2];
    const uint16_t *y[2] = {yr, yd};
    const int pitch[2] = {pitch_r, pitch_d};
    int32_t *mid[2] = {mid_r, mid_d};
    for (int p = 0; p < 2; p++) {
        s3[p] = (int32_t *)malloc((size_t)w * h * 4);
        s9[p] = (int32_t *)malloc((size_t)w * h * 4);
        box_sums(y[p], pitch[p], up, w, h, 1, s3[p]);
        box_sums(y[p], pitch[p], up, w, h, 4, s9[p]);
    }
    memset(sums, 0, (size_t)ww * wh * 12 * sizeof(long long));
    for (int yy = 0; yy < H; yy++)
        for (int x = 0; x < W; x++) {
            long long F[2], M[2], GX[2], GY[2];
            const size_t i = (size_t)yy * w + x;
            for (int p = 0; p < 2; p++) {
                const int32_t Y = (int32_t)y[p][(size_t)yy * pitch[p] + x] << up;
                F[p] = 9LL * Y - s3[p][i];
                M[p] = 9LL * s3[p][i] - s9[p][i];
                GX[p] = (x > 0 && x < w - 1) ? (long long)s3[p][i + 1] - s3[p][i - 1] : 0;
                GY[p] = (yy > 0 && yy < h - 1) ? (long long)s3[p][i + w] - s3[p][i - w] : 0;
                if (mid[p]) mid[p][i] = (int32_t)M[p];
            }
            long long *o = sums + ((size_t)(yy / win) * ww + x / win) * 12;
            o[0] += F[0] * F[0];
            o[1] += M[0] * M[0];
            o[2] += F[1] * F[1];
            o[3] += M[1] * M[1];
            o[4] += F[0] * F[1];
            o[5] += M[0] * M[1];
            o[6] += GX[0] * GX[0];
            o[7] += GY[0] * GY[0];
            o[8] += GX[0] * GY[0];
            o[9] += GX[1] * GX[1];
            o[10] += GY[1] * GY[1];
            o[11] += GX[1] * GY[1];
        }
    if (stats)
        for (int i = 0; i < ww * wh; i++) orc_window_values(sums + (size_t)i * 12, stats + (size_t)i * 8);
    for (int p = 0; p < 2; p++) {
        free(s3[p]);
        free(s9[p]);
    }
}

/* The 8 figures of a window (src/kernels.cu:340-341, :421-425) from its exact sums: 0 source fine, 1 source mid,
 * 2 encode fine, 3 encode mid, 4 shared fine, 5 shared mid, 6 source oriented, 7 encode oriented.  Oriented is the
 * difference of the eigenvalues of the summed structure tensor, sqrt((Jxx - Jyy)^2 + 4 Jxy^2), with J = sums / 81. */
void orc_window_values(const long long *s, double *v)
{
    v[0] = (double)s[0] / 81.0;
    v[1] = (double)s[1] / 6561.0;
    v[2] = (double)s[2] / 81.0;
    v[3] = (double)s[3] / 6561.0;
    v[4] = (double)s[4] / 81.0;
    v[5] = (double)s[5] / 6561.0;
    for (int p = 0; p < 2; p++) {
        const double d = (double)(s[6 + 3 * p] - s[7 + 3 * p]), x = (double)s[8 + 3 * p];
        v[6 + p] = sqrt(d * d + 4.0 * (x * x)) / 81.0;
    }
}

/* persist_stats (src/kernels.cu:430-451): for a centre frame, A = mean of the source mid planes at offsets +-2, 4, 6
 * and B = mean at +-1, 3, 5, T = (A + B) / 2, and per window the sums of mid_src(centre) T, mid_enc(centre) T, A B, A^2,
 * B^2.  On the scaled integers M = 81 mid: SA = 6 * 81 A and SB = 6 * 81 B are integers and 12 * 81 T = SA + SB, so
 * `sums` holds per window the exact [M_r ST, M_d ST, SA SB, SA^2, SB^2] with ST = SA + SB.  ring[j] is the source mid
 * plane of offset j - 6. */
void orc_persist(const int32_t *const *ring, const int32_t *mid_d, int w, int h, int win, long long *sums)
{
    const int ww = w / win, wh = h / win;
    memset(sums, 0, (size_t)ww * wh * 5 * sizeof(long long));
    for (int y = 0; y < wh * win; y++)
        for (int x = 0; x < ww * win; x++) {
            const size_t i = (size_t)y * w + x;
            const long long sa = (long long)ring[0][i] + ring[2][i] + ring[4][i] + ring[8][i] + ring[10][i] + ring[12][i];
            const long long sb = (long long)ring[1][i] + ring[3][i] + ring[5][i] + ring[7][i] + ring[9][i] + ring[11][i];
            const long long st = sa + sb;
            long long *o = sums + ((size_t)(y / win) * ww + x / win) * 5;
            o[0] += (long long)ring[6][i] * st;
            o[1] += (long long)mid_d[i] * st;
            o[2] += sa * sb;
            o[3] += sa * sa;
            o[4] += sb * sb;
        }
}

/* The reference's 5 lasting-texture figures of a window: sums / (81 * 972) for the two projections onto T and
 * sums / 486^2 for the products of A and B. */
void orc_persist_values(const long long *s, double *v)
{
    v[0] = (double)s[0] / 78732.0;
    v[1] = (double)s[1] / 78732.0;
    v[2] = (double)s[2] / 236196.0;
    v[3] = (double)s[3] / 236196.0;
    v[4] = (double)s[4] / 236196.0;
}

/* ------------------------------------------------------------------------------------------------------------------
 * Selects and histograms.
 * ------------------------------------------------------------------------------------------------------------------ */
static int cmp_double(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y ? 1 : 0;
}

/* numpy's linear quantile (Hyndman-Fan type 7, the definition src/main.cpp:279-289 restates): interpolation between
 * the order statistics of ranks floor(q (n - 1)) and the one after it, in FP64 throughout.  The collect uses
 * orc_quantile_f32, the reference's float evaluation of the same definition. */
double orc_quantile_linear(const double *v, long long n, double q)
{
    double *c = (double *)malloc((size_t)n * sizeof(double));
    memcpy(c, v, (size_t)n * sizeof(double));
    qsort(c, (size_t)n, sizeof(double), cmp_double);
    const double pos = q * (double)(n - 1);
    const long long lo = (long long)pos;
    const double a = c[lo];
    double out = a;
    if (lo + 1 < n) {
        const double b = c[lo + 1];
        out = a + (b - a) * (pos - (double)lo);
    }
    free(c);
    return out;
}

static int cmp_float(const void *a, const void *b)
{
    const float x = *(const float *)a, y = *(const float *)b;
    return x < y ? -1 : x > y ? 1 : 0;
}

/* quantile (src/main.cpp:279-289) as the reference evaluates it: float keys; a = the key of rank lo = floor(q (n - 1))
 * and b = the smallest key after it; b - a is a float difference, a + (b - a)(pos - lo) is evaluated in double and
 * rounded to float. */
float orc_quantile_f32(const float *v, long long n, double q)
{
    if (n <= 0) return 0.0f;
    float *c = (float *)malloc((size_t)n * sizeof(float));
    memcpy(c, v, (size_t)n * sizeof(float));
    qsort(c, (size_t)n, sizeof(float), cmp_float);
    const double pos = q * (double)(n - 1);
    const long long lo = (long long)pos;
    const float a = c[lo];
    float out = a;
    if (lo + 1 < n) {
        const float b = c[lo + 1];
        const float d = b - a;
        out = (float)((double)a + (double)d * (pos - (double)lo));
    }
    free(c);
    return out;
}

static int cmp_u64(const void *a, const void *b)
{
    const unsigned long long x = *(const unsigned long long *)a, y = *(const unsigned long long *)b;
    return x < y ? -1 : x > y ? 1 : 0;
}

/* The key of 0-based rank `rank` in ascending order (a sort). */
unsigned long long orc_select_rank_u64(const unsigned long long *keys, long long n, long long rank)
{
    unsigned long long *c = (unsigned long long *)malloc((size_t)n * 8);
    memcpy(c, keys, (size_t)n * 8);
    qsort(c, (size_t)n, 8, cmp_u64);
    const unsigned long long out = c[rank];
    free(c);
    return out;
}

/* The canonical per-frame sum of per-window doubles: blocks of 256 consecutive values (the last padded with zeros),
 * each added as a halving tree (v[i] += v[i + s] for s = 128, 64, ..., 1), then the blocks' sums in order. */
double orc_tree_sum256(const double *v, long long n)
{
    double total = 0.0;
    for (long long b = 0; b < n; b += 256) {
        double t[256];
        for (int i = 0; i < 256; i++) t[i] = b + i < n ? v[b + i] : 0.0;
        for (int s = 128; s > 0; s >>= 1)
            for (int i = 0; i < s; i++) t[i] += t[i + s];
        total += t[0];
    }
    return total;
}

/* Histogram::add (src/stats.hpp:16-20): 0.05-point bins over [-500, 500], rounded half away from zero. */
int orc_hist_bin(double pct)
{
    const double c = pct < -500.0 ? -500.0 : pct > 500.0 ? 500.0 : pct;
    return (int)lround((c - -500.0) / 0.05);
}

void orc_hist_add(orc_hist *h, double pct)
{
    h->bins[orc_hist_bin(pct)]++;
    h->n++;
}

/* Histogram::quantile (src/stats.hpp:21-30): the first bin whose running count exceeds q (n - 1). */
double orc_hist_quantile(const orc_hist *h, double q)
{
    if (!h->n) return NAN;
    const double target = q * (double)(h->n - 1);
    uint64_t seen = 0;
    for (int b = 0; b < ORC_HIST_BINS; b++) {
        seen += h->bins[b];
        if ((double)seen > target) return -500.0 + (double)b * 0.05;
    }
    return 500.0;
}

void orc_hist_merge(orc_hist *h, const orc_hist *o)
{
    for (int b = 0; b < ORC_HIST_BINS; b++) h->bins[b] += o->bins[b];
    h->n += o->n;
}

void orc_accum_clear(orc_accum *a)
{
    memset(a, 0, sizeof *a);
    a->vmaf_min = 1e9; /* src/stats.hpp:38 */
}

/* Accum::merge (src/stats.hpp:50-64), measure fields. */
void orc_accum_merge(orc_accum *a, const orc_accum *o)
{
    a->frames += o->frames;
    a->cambi += o->cambi;
    a->cambi_source += o->cambi_source;
    a->own += o->own;
    a->mine += o->mine;
    for (int i = 0; i < 3; i++) {
        a->all_r[i] += o->all_r[i];
        a->all_d[i] += o->all_d[i];
        a->st_r[i] += o->st_r[i];
        a->st_d[i] += o->st_d[i];
        orc_hist_merge(&a->amount[i], &o->amount[i]);
    }
    for (int i = 0; i < 2; i++) {
        a->sh_all_r[i] += o->sh_all_r[i];
        a->sh_all_x[i] += o->sh_all_x[i];
        a->sh_st_r[i] += o->sh_st_r[i];
        a->sh_st_x[i] += o->sh_st_x[i];
        orc_hist_merge(&a->matching[i], &o->matching[i]);
    }
    orc_hist_merge(&a->lasting, &o->lasting);
    for (int i = 0; i < 3; i++) {
        a->psnr[i] += o->psnr[i];
        a->color_r[i] += o->color_r[i];
        a->color_d[i] += o->color_d[i];
    }
    a->vmaf += o->vmaf;
    a->vmaf_min = o->vmaf_min < a->vmaf_min ? o->vmaf_min : a->vmaf_min;
}

static double clamp100(double v) { return v < 0.0 ? 0.0 : v > 100.0 ? 100.0 : v; }
static double kept(double d, double r) { return r > 0 ? 100.0 * sqrt(d / r) : NAN; }

/* Accum::score (src/stats.hpp:66-82).  Fidelity is VMAF, which this path does not compute: pass NaN to leave it out
 * of the weighted mean, as a part that could not be measured is. */
void orc_accum_score(const orc_accum *a, double vmaf_total, orc_parts *p)
{
    p->fidelity = isnan(vmaf_total) ? NAN : clamp100(vmaf_total / (double)a->frames);
    p->texture = a->lasting.n >= 100
                     ? clamp100(orc_hist_quantile(&a->lasting, 0.5)) / 2 + clamp100(orc_hist_quantile(&a->lasting, 0.1)) / 2
                     : NAN;
    p->banding = clamp100(100.0 * (1 - (a->cambi - a->cambi_source) / (double)a->frames / 5.0));
    p->grain = clamp100(100.0 - fabs(kept(a->all_d[0], a->all_r[0]) - 100.0));
    const double v[4] = {p->fidelity, p->texture, p->banding, p->grain}, wt[4] = {0.30, 0.30, 0.25, 0.15};
    double total = 0, weight = 0;
    for (int i = 0; i < 4; i++)
        if (!isnan(v[i])) {
            total += v[i] * wt[i];
            weight += wt[i];
        }
    p->score = total / weight;
}

/* The per-frame part of collect (src/main.cpp:606-629): the structure mask (oriented >= its 0.8 quantile and > 0),
 * the frame's totals and the per-window histogram entries.  The mask compares floats as the reference does: each
 * window's oriented energy rounded once to float (the reference's per-window figures are floats) against
 * orc_quantile_f32 of them.  Canonical totals: fine and mid (and their shared parts) are the exact integer sums over
 * the windows scaled once; oriented is orc_tree_sum256 of the per-window values (0 for windows outside the mask in the
 * structured totals). */
void orc_collect_frame(orc_accum *acc, const long long *sums, int windows, uint8_t *mask_out)
{
    double *st = (double *)malloc((size_t)windows * 8 * sizeof(double));
    float *orient = (float *)calloc((size_t)windows + 1, sizeof(float));
    double *ob[4];
    for (int k = 0; k < 4; k++) ob[k] = (double *)malloc((size_t)windows * sizeof(double));
    for (int i = 0; i < windows; i++) {
        orc_window_values(sums + (size_t)i * 12, st + (size_t)i * 8);
        orient[i] = (float)st[(size_t)i * 8 + 6];
    }
    const float q = orc_quantile_f32(orient, windows, 0.8);
    long long all[6] = {0}, stc[6] = {0};
    for (int i = 0; i < windows; i++) {
        const double *s = st + (size_t)i * 8;
        const long long *z = sums + (size_t)i * 12;
        const int m = orient[i] >= q && orient[i] > 0;
        mask_out[i] = (uint8_t)m;
        for (int j = 0; j < 6; j++) {
            all[j] += z[j];
            if (m) stc[j] += z[j];
        }
        ob[0][i] = s[6];
        ob[1][i] = s[7];
        ob[2][i] = m ? s[6] : 0.0;
        ob[3][i] = m ? s[7] : 0.0;
        if (m) {
            const double er[3] = {s[0], s[1], s[6]}, ed[3] = {s[2], s[3], s[7]};
            for (int key = 0; key < 3; key++) orc_hist_add(&acc->amount[key], 100.0 * sqrt(ed[key] / fmax(er[key], 1e-9)));
            for (int key = 0; key < 2; key++) orc_hist_add(&acc->matching[key], 100.0 * s[4 + key] / fmax(er[key], 1e-9));
        }
    }
    const double div[2] = {81.0, 6561.0};
    for (int key = 0; key < 2; key++) {
        acc->all_r[key] += (double)all[key] / div[key];
        acc->all_d[key] += (double)all[2 + key] / div[key];
        acc->st_r[key] += (double)stc[key] / div[key];
        acc->st_d[key] += (double)stc[2 + key] / div[key];
        acc->sh_all_r[key] += (double)all[key] / div[key];
        acc->sh_all_x[key] += (double)all[4 + key] / div[key];
        acc->sh_st_r[key] += (double)stc[key] / div[key];
        acc->sh_st_x[key] += (double)stc[4 + key] / div[key];
    }
    acc->all_r[2] += orc_tree_sum256(ob[0], windows);
    acc->all_d[2] += orc_tree_sum256(ob[1], windows);
    acc->st_r[2] += orc_tree_sum256(ob[2], windows);
    acc->st_d[2] += orc_tree_sum256(ob[3], windows);
    for (int k = 0; k < 4; k++) free(ob[k]);
    free(orient);
    free(st);
}

/* The lasting-texture part of collect (src/main.cpp:630-641): a window of the centre frame counts when it is in that
 * frame's structure mask, the two halves agree (A B / sqrt(max(A^2 B^2, 1e-9)) > 0.5) and its projection is positive;
 * own and mine take the exact sums of the counted windows, scaled once. */
void orc_collect_lasting(orc_accum *acc, const long long *psums, const uint8_t *cmask, int windows)
{
    long long own = 0, mine = 0;
    for (int i = 0; i < windows; i++) {
        double v[5];
        orc_persist_values(psums + (size_t)i * 5, v);
        const double agree = v[2] / sqrt(fmax(v[3] * v[4], 1e-9));
        if (cmask[i] && agree > 0.5 && v[0] > 0) {
            own += psums[(size_t)i * 5];
            mine += psums[(size_t)i * 5 + 1];
            orc_hist_add(&acc->lasting, 100.0 * v[1] / v[0]);
        }
    }
    acc->own += (double)own / 78732.0;
    acc->mine += (double)mine / 78732.0;
}

/* ------------------------------------------------------------------------------------------------------------------
 * Frame ingest, PSNR and colour.
 * ------------------------------------------------------------------------------------------------------------------ */

/* sample() of src/kernels.cu:289-293: the stored word shifted right by `shift`. */
static unsigned frame_sample(const void *plane, int pitch, int x, int y, const orc_frame *f)
{
    const unsigned char *row = (const unsigned char *)plane + (size_t)y * pitch;
    return (f->bytes == 1 ? row[x] : ((const uint16_t *)row)[x]) >> f->shift;
}

/* frame_planes (src/kernels.cu:297-310), pixel by pixel: luma at (x, y), and for even x and y the chroma sample at
 * ((x + x0) / 2, (y + y0) / 2), u and v interleaved in plane u for a semi-planar frame; all shifted up by pd - depth. */
void orc_frame_planes(const orc_frame *f, int x0, int y0, int pd, uint16_t *y16, uint16_t *u, uint16_t *v, int w, int h)
{
    const int up = pd - f->depth;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            y16[(size_t)y * w + x] = (uint16_t)(frame_sample(f->y, f->ypitch, x + x0, y + y0, f) << up);
            if ((x & 1) || (y & 1)) continue;
            const int cx = (x + x0) / 2, cy = (y + y0) / 2;
            const size_t ci = (size_t)(y / 2) * (w / 2) + x / 2;
            const unsigned cu = f->semi ? frame_sample(f->u, f->cpitch, 2 * cx, cy, f) : frame_sample(f->u, f->cpitch, cx, cy, f);
            const unsigned cv = f->semi ? frame_sample(f->u, f->cpitch, 2 * cx + 1, cy, f) : frame_sample(f->v, f->cpitch, cx, cy, f);
            u[ci] = (uint16_t)(cu << up);
            v[ci] = (uint16_t)(cv << up);
        }
}

/* plane_sse (src/kernels.cu:454-465): the exact sum of squared differences. */
unsigned long long orc_plane_sse(const uint16_t *a, const uint16_t *b, long long n)
{
    unsigned long long s = 0;
    for (long long i = 0; i < n; i++) {
        const long long d = (long long)a[i] - b[i];
        s += (unsigned long long)(d * d);
    }
    return s;
}

/* plane_sse's float form: the differences of the values truncated to integers. */
unsigned long long orc_plane_sse_f(const float *a, const float *b, long long n)
{
    unsigned long long s = 0;
    for (long long i = 0; i < n; i++) {
        const long long d = (long long)a[i] - (long long)b[i];
        s += (unsigned long long)(d * d);
    }
    return s;
}

/* 2^23 sqrtf(s) (IEEE single precision, round to nearest) as an exact integer: a float of at least 1 is a multiple of
 * 2^-23. */
unsigned long long orc_strength_fixed(unsigned s)
{
    if (s == 0) return 0;
    const float f = sqrtf((float)s);
    uint32_t b;
    memcpy(&b, &f, 4);
    return (unsigned long long)((b & 0x7FFFFFu) | 0x800000u) << ((b >> 23) - 127);
}

/* color_sums (src/kernels.cu:470-484): sums of u << up, v << up and of the colour strength sqrtf(du^2 + dv^2) about
 * the neutral 512.  Canonical form: the strengths, each rounded to single precision as the reference's sqrtf does,
 * are added exactly (units of 2^-23) instead of in the reference's per-thread float partial sums, whose order
 * depends on its launch shape. */
void orc_color_sums(const uint16_t *u, const uint16_t *v, long long n, int up, unsigned long long *sums)
{
    unsigned long long su = 0, sv = 0, ss = 0;
    for (long long i = 0; i < n; i++) {
        const int uu = u[i] << up, vv = v[i] << up, du = uu - 512, dv = vv - 512;
        su += (unsigned)uu;
        sv += (unsigned)vv;
        ss += orc_strength_fixed((unsigned)(du * du + dv * dv));
    }
    sums[0] = su;
    sums[1] = sv;
    sums[2] = ss;
}

/* psnr (src/main.cpp:273-277) */
double orc_psnr(unsigned long long sse, long long n, int depth)
{
    const double mse = (double)sse / (double)n, peak = (double)((1 << depth) - 1);
    const double v = 10.0 * log10(peak * peak / (mse > 1e-16 ? mse : 1e-16)), cap = 6.0 * depth + 12;
    return v < cap ? v : cap;
}

/* ------------------------------------------------------------------------------------------------------------------
 * The whole run of one pair, as src/main.cpp's queue (:516-574) and collect (:576-643) do it frame by frame: rings of
 * 13 source mid planes and 7 encode mid planes and structure masks (src/main.cpp:240, :448, :469, :478), texture
 * from frame 12 on for the centre 6 frames back, into the centre's segment.
 * ------------------------------------------------------------------------------------------------------------------ */
struct orc_run {
    orc_run_config c;
    orc_cambi_params p;
    int ww, wh, windows;
    long frames;
    int32_t *ring[13], *emid[7];
    uint8_t *mask[7];
    long long *sums, *psums;
    int nseg;
    orc_accum *seg;
};

orc_run *orc_run_create(const orc_run_config *c)
{
    orc_run *r = (orc_run *)calloc(1, sizeof *r);
    r->c = *c;
    if (orc_cambi_params_init(&r->p, c->w, c->h, c->depth, c->window_base) != 0) {
        free(r);
        return NULL;
    }
    r->ww = c->w / c->win;
    r->wh = c->h / c->win;
    r->windows = r->ww * r->wh;
    const size_t n = (size_t)c->w * c->h;
    for (int i = 0; i < 13; i++) r->ring[i] = (int32_t *)calloc(n, 4);
    for (int i = 0; i < 7; i++) {
        r->emid[i] = (int32_t *)calloc(n, 4);
        r->mask[i] = (uint8_t *)calloc((size_t)r->windows, 1);
    }
    r->sums = (long long *)malloc((size_t)r->windows * 12 * 8);
    r->psums = (long long *)malloc((size_t)r->windows * 5 * 8);
    return r;
}

void orc_run_destroy(orc_run *r)
{
    if (!r) return;
    for (int i = 0; i < 13; i++) free(r->ring[i]);
    for (int i = 0; i < 7; i++) {
        free(r->emid[i]);
        free(r->mask[i]);
    }
    free(r->sums);
    free(r->psums);
    free(r->seg);
    free(r);
}

static orc_accum *segment_of(orc_run *r, long f)
{
    const int s = (int)(f / r->c.per_segment);
    if (s >= r->nseg) {
        r->seg = (orc_accum *)realloc(r->seg, (size_t)(s + 1) * sizeof(orc_accum));
        for (int i = r->nseg; i <= s; i++) orc_accum_clear(&r->seg[i]);
        r->nseg = s + 1;
    }
    return &r->seg[s];
}

int orc_run_frame(orc_run *r, const uint16_t *src, int src_pitch, const uint16_t *enc, int enc_pitch, double *cambi,
                  double *cambi_source)
{
    const long f = r->frames;
    orc_cambi_result cr;
    orc_cambi(&r->p, src, src_pitch, NULL, &cr, 0, r->c.threads);
    const double banding_src = cr.score;
    orc_cambi(&r->p, enc, enc_pitch, NULL, &cr, 0, r->c.threads);
    const double banding = cr.score;
    orc_window_stats(src, src_pitch, enc, enc_pitch, r->c.up, r->c.w, r->c.h, r->c.win, r->ring[f % 13], r->emid[f % 7],
                     r->sums, NULL);
    orc_accum *acc = segment_of(r, f);
    acc->frames++;
    acc->cambi += banding;
    acc->cambi_source += banding_src;
    orc_collect_frame(acc, r->sums, r->windows, r->mask[f % 7]);
    if (f >= 12) {
        const long c = f - 6;
        const int32_t *ring[13];
        for (int o = -6; o <= 6; o++) ring[o + 6] = r->ring[(c + o) % 13];
        orc_persist(ring, r->emid[c % 7], r->c.w, r->c.h, r->c.win, r->psums);
        orc_collect_lasting(segment_of(r, c), r->psums, r->mask[c % 7], r->windows);
    }
    if (cambi) *cambi = banding;
    if (cambi_source) *cambi_source = banding_src;
    r->frames++;
    return 0;
}

int orc_run_frame_yuv(orc_run *r, const uint16_t *const *src, const uint16_t *const *enc, double *cambi,
                      double *cambi_source, double *psnr)
{
    const long f = r->frames;
    const long long n = (long long)r->c.w * r->c.h, nc = n / 4;
    const int pd = r->c.depth, up = 10 - pd;
    const unsigned long long sse[3] = {orc_plane_sse(src[0], enc[0], n), orc_plane_sse(src[1], enc[1], nc),
                                       orc_plane_sse(src[2], enc[2], nc)};
    unsigned long long cr[3], cd[3];
    orc_color_sums(src[1], src[2], nc, up, cr);
    orc_color_sums(enc[1], enc[2], nc, up, cd);
    const int rc = orc_run_frame(r, src[0], r->c.w, enc[0], r->c.w, cambi, cambi_source);
    if (rc) return rc;
    orc_accum *acc = segment_of(r, f);
    for (int i = 0; i < 3; i++) {
        const double p = orc_psnr(sse[i], i ? nc : n, pd);
        acc->psnr[i] += p;
        if (psnr) psnr[i] = p;
        const double sr = i < 2 ? (double)cr[i] : ldexp((double)cr[i], -23), sd = i < 2 ? (double)cd[i] : ldexp((double)cd[i], -23);
        acc->color_r[i] += sr / (double)nc;
        acc->color_d[i] += sd / (double)nc;
    }
    return 0;
}

int orc_run_segments(const orc_run *r) { return r->nseg; }
const orc_accum *orc_run_segment(const orc_run *r, int s) { return s >= 0 && s < r->nseg ? &r->seg[s] : NULL; }
