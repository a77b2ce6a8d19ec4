// Detail, structure-tensor and lasting-texture window sums for gfx950.
//
// window_stats (src/kernels.cu:331-428; reference/encode_score/compare.py:77-108): per win x win window of a (source,
// encode) pair, the energies of fine = y - box3 and mid = box3 - box9 (edges repeated), their cross products and the
// structure tensor of box3's central differences.  Luma is an integer and the box weights are 1/9 and 1/81, so the
// kernel works on exact integers F = 9 fine, M = 81 mid, GX = 9 gx, GY = 9 gy and writes exact window sums:
//   [F_r^2, M_r^2, F_d^2, M_d^2, F_r F_d, M_r M_d, GX_r^2, GY_r^2, GX_r GY_r, GX_d^2, GY_d^2, GX_d GY_d]
// Two forms: the integer strip kernel of the batched path, and an FP64 tile kernel (the plain per-frame form, exact
// too: every value is an integer below 2^53), which also writes the reference's float outputs.
//
// persist (src/kernels.cu:430-451): per window of a centre frame, from the source's mid planes of the 6 frames either
// side, SA = sum at offsets +-2, 4, 6, SB = sum at +-1, 3, 5 (6 * 81 times the reference's A and B), ST = SA + SB
// (12 * 81 T), and the sums [M_r(c) ST, M_d(c) ST, SA SB, SA^2, SB^2].
#include "device_util.hpp"
#include "energy.hpp"

namespace es {

__device__ __forceinline__ int in_sample(const PlaneIn &in, int x, int y)
{
    const size_t i = (size_t)y * in.pitch + x;
    const unsigned v = in.bytes == 1 ? (unsigned)((const ES_GLOBAL uint8_t *)in.ptr)[i]
                                     : (unsigned)((const ES_GLOBAL uint16_t *)in.ptr)[i] >> in.rshift;
    return (int)(v << in.shift);
}

template <typename A>
__device__ __forceinline__ A reduce_lanes(A v, int width)
{
    for (int o = width >> 1; o; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// ---------------------------------------------------------------------------------------------------------------
// FP64 form: the per-frame baseline the integer form is timed against, and the reference-shaped float outputs.
// One block of 256 threads per tile of 64 columns by one row of windows, taken in bands of 16 rows.  For a band, both
// pictures' samples of rows -4 .. 19 and columns -4 .. 67 around it (edges repeated) are staged in LDS; the horizontal
// 3-sums (columns -1 .. 64) and 9-sums (columns 0 .. 63) of every staged row follow, then the vertical 3-sums S3 of
// rows -1 .. 16.  The box sums are integers; every per-pixel value (F = 9 y - S3, M = 9 S3 - S9, GX and GY the central
// differences of S3) and every product and sum is then taken in double, which is exact (all integers below 2^53).
// Thread t takes the column t % 64 at rows t / 64 + 4 k of each band and keeps its 12 sums over the tile's bands; the
// tile's 64 / WIN windows then add their columns in LDS.  Outputs (any may be null): the 12 sums per window, the
// reference's 8 figures per window as floats, M as int32 or mid = M / 81 as float on the windows' area.
// ---------------------------------------------------------------------------------------------------------------
constexpr int kTileCols = 64, kBand = 16;
constexpr int kStageRows = kBand + 8, kStageCols = kTileCols + 8;  // rows -4 .. 19, columns -4 .. 67
constexpr int kH3Cols = kTileCols + 2;                              // columns -1 .. 64
constexpr int kS3Rows = kBand + 2;                                  // rows -1 .. 16

struct alignas(16) TileLds {
    int ys[2][kStageRows][kStageCols];
    int h3[2][kStageRows][kH3Cols];
    int h9[2][kStageRows][kTileCols];
    int s3[2][kS3Rows][kH3Cols];
};
static_assert(sizeof(TileLds) >= 12 * 256 * sizeof(double), "the reduction reuses the tile's LDS");

// The reference's 8 figures of a window (src/kernels.cu:340-341) from its 12 exact sums.
__device__ __forceinline__ void figures_of(const double (&z)[12], float (&f)[8])
{
    f[0] = __double2float_rn(z[0] / 81.0);
    f[1] = __double2float_rn(z[1] / 6561.0);
    f[2] = __double2float_rn(z[2] / 81.0);
    f[3] = __double2float_rn(z[3] / 6561.0);
    f[4] = __double2float_rn(z[4] / 81.0);
    f[5] = __double2float_rn(z[5] / 6561.0);
#pragma unroll
    for (int p = 0; p < 2; p++) {
        const double d = z[6 + 3 * p] - z[7 + 3 * p], x = z[8 + 3 * p];
        f[6 + p] = __double2float_rn(sqrt(__dadd_rn(__dmul_rn(d, d), __dmul_rn(4.0, __dmul_rn(x, x)))) / 81.0);
    }
}

template <int WIN>
__global__ __launch_bounds__(256) void window_stats_tile_kernel(const WsPair *pairs)
{
    __shared__ TileLds L;
    const WsPair &P = pairs[blockIdx.z];
    const int t = threadIdx.x;
    const int w = P.w, h = P.h, ww = w / WIN, W = ww * WIN;
    const int x0 = blockIdx.x * kTileCols, wy = blockIdx.y;
    if (x0 >= W || wy >= h / WIN) return;  // uniform over the block
    const PlaneIn *in[2] = {&P.r, &P.d};
    const int col = t & (kTileCols - 1), X = x0 + col;
    const bool inside = X < W;
    double acc[12];
#pragma unroll
    for (int k = 0; k < 12; k++) acc[k] = 0.0;
    for (int band = 0; band < WIN / kBand; band++) {
        const int yb = wy * WIN + band * kBand;
        __syncthreads();  // the previous band's readers are done
        for (int i = t; i < 2 * kStageRows * kStageCols; i += 256) {
            const int p = i / (kStageRows * kStageCols), r = i / kStageCols % kStageRows, c = i % kStageCols;
            const int yy = min(max(yb - 4 + r, 0), h - 1), xx = min(max(x0 - 4 + c, 0), w - 1);
            L.ys[p][r][c] = in_sample(*in[p], xx, yy) << P.up;
        }
        __syncthreads();
        for (int i = t; i < 2 * kStageRows * kH3Cols; i += 256) {
            const int p = i / (kStageRows * kH3Cols), r = i / kH3Cols % kStageRows, c = i % kH3Cols;
            const int *y = L.ys[p][r];
            L.h3[p][r][c] = y[c + 2] + y[c + 3] + y[c + 4];  // column c - 1
            if (c < kTileCols) {
                int s9 = 0;
#pragma unroll
                for (int j = 0; j < 9; j++) s9 += y[c + j];  // column c
                L.h9[p][r][c] = s9;
            }
        }
        __syncthreads();
        for (int i = t; i < 2 * kS3Rows * kH3Cols; i += 256) {
            const int p = i / (kS3Rows * kH3Cols), r = i / kH3Cols % kS3Rows, c = i % kH3Cols;
            L.s3[p][r][c] = L.h3[p][r + 2][c] + L.h3[p][r + 3][c] + L.h3[p][r + 4][c];  // row r - 1
        }
        __syncthreads();
        for (int j = t / kTileCols; j < kBand; j += 256 / kTileCols) {
            const int Y = yb + j;
            double F[2], M[2], GX[2], GY[2];
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const int s3 = L.s3[p][j + 1][col + 1];
                int s9 = 0;
#pragma unroll
                for (int k = 0; k < 9; k++) s9 += L.h9[p][j + k][col];
                F[p] = 9.0 * (double)L.ys[p][j + 4][col + 4] - (double)s3;
                M[p] = 9.0 * (double)s3 - (double)s9;
                GX[p] = (X > 0 && X < w - 1) ? (double)(L.s3[p][j + 1][col + 2] - L.s3[p][j + 1][col]) : 0.0;
                GY[p] = (Y > 0 && Y < h - 1) ? (double)(L.s3[p][j + 2][col + 1] - L.s3[p][j][col + 1]) : 0.0;
            }
            if (inside) {
                const size_t o = (size_t)Y * w + X;
                if (P.mid_r) P.mid_r[o] = (int32_t)M[0];
                if (P.mid_d) P.mid_d[o] = (int32_t)M[1];
                if (P.midf_r) P.midf_r[o] = __double2float_rn(M[0] / 81.0);
                if (P.midf_d) P.midf_d[o] = __double2float_rn(M[1] / 81.0);
            }
            const double v[12] = {F[0] * F[0], M[0] * M[0], F[1] * F[1], M[1] * M[1], F[0] * F[1], M[0] * M[1],
                                  GX[0] * GX[0], GY[0] * GY[0], GX[0] * GY[0], GX[1] * GX[1], GY[1] * GY[1], GX[1] * GY[1]};
#pragma unroll
            for (int k = 0; k < 12; k++) acc[k] += v[k];
        }
    }
    // the tile's windows: each column's 4 row groups, then each window's WIN columns (exact: integers in double)
    __syncthreads();
    double(*red)[256] = reinterpret_cast<double(*)[256]>(&L);
#pragma unroll
    for (int k = 0; k < 12; k++) red[k][t] = inside ? acc[k] : 0.0;
    __syncthreads();
    constexpr int kWins = kTileCols / WIN;
    __shared__ double win_sum[kWins][12];
    if (t < kWins * 12) {
        const int wi = t / 12, k = t % 12;
        double s = 0.0;
        for (int c = wi * WIN; c < (wi + 1) * WIN; c++) s += red[k][c] + red[k][c + 64] + red[k][c + 128] + red[k][c + 192];
        win_sum[wi][k] = s;
    }
    __syncthreads();
    if (t < kWins) {
        const int wx = x0 / WIN + t;
        if (wx < ww) {
            const size_t o = (size_t)wy * ww + wx;
            double z[12];
#pragma unroll
            for (int k = 0; k < 12; k++) z[k] = win_sum[t][k];
            if (P.sums) {
#pragma unroll
                for (int k = 0; k < 12; k++) P.sums[o * 12 + k] = (long long)z[k];
            }
            if (P.vals) {
                float f[8];
                figures_of(z, f);
#pragma unroll
                for (int k = 0; k < 8; k++) P.vals[o * 8 + k] = f[k];
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Integer form.  One wave per strip of 64 columns and kWsStrip rows, on its own (64-thread blocks, no barrier between
// waves).  Row r is staged through LDS with 4 columns of margin on either side; the next row's samples are loaded
// while the current one is processed.  Per picture a lane keeps, in registers, the horizontal 3-sums (centred on its
// column and on either neighbour), 9-sums and samples of the last rows in rings whose slots are fixed by r mod 9 (the
// row loop is unrolled 9 times, so the rings rotate by renaming).  The sums are taken as soon as their rows are in:
// F = 9 y - S3 of row r - 1, GX / GY of row r - 2 (GY needs S3 of the row below), M = 9 S3 - S9 of row r - 4.  No
// product mixes those groups, so each group keeps its own 16-row partial sums (32-bit, exact: |F|, |GX|, |GY| <=
// 9 * 1023, 16 of their products < 2^31; M's products are 64-bit) and flushes them when its row closes a block.
// ---------------------------------------------------------------------------------------------------------------
constexpr int kWsStrip = 64;


// Samples of a plane of BYTES bytes per sample (1 or 2) shifted to the pipeline depth, (word >> rshift) << lsh with
// lsh the plane's shift plus the pair's; kShift false: no shifts (16-bit words holding the samples as they are).  The
// plane is read as a buffer: the lane's column offset is fixed and the row's offset is one scalar.
template <int BYTES, bool kShift>
struct WsRead {
    __amdgpu_buffer_rsrc_t rs;
    unsigned step;  // bytes per row
    int h, rshift, lsh;
    __device__ WsRead(const PlaneIn &in, int h_, int up)
        : rs(buffer_over(in.ptr, (unsigned long long)in.pitch * h_ * BYTES)), step((unsigned)in.pitch * BYTES), h(h_),
          rshift(in.rshift), lsh(in.shift + up) {}
    __device__ unsigned col(int x) const { return (unsigned)x * BYTES; }
    __device__ unsigned row(int y) const { return (unsigned)min(max(y, 0), h - 1) * step; }
    __device__ int load(unsigned c, unsigned r) const
    {
        unsigned v = BYTES == 1 ? (unsigned)__builtin_amdgcn_raw_buffer_load_b8(rs, c, r, 0)
                                : (unsigned)__builtin_amdgcn_raw_buffer_load_b16(rs, c, r, 0);
        if (kShift) v = (v >> rshift) << lsh;
        return (int)v;
    }
};

typedef int WsRows[2][2][72 + 64];  // [buffer][picture][column - (x0 - 4)]; 72 .. 135 take unused halo stores
typedef long long WsRed[6][64];     // a group's per-lane sums on their way to the window sums

// One strip of the pair; BR and BD are the source's and the encode's bytes per sample, kShift whether either plane's
// samples are shifted.
template <int WIN, int BR, int BD, bool kShift>
__device__ __forceinline__ void ws_strip(const WsPair &P, WsRows &rowbuf, WsRed &red, unsigned bx, unsigned by)
{
    static_assert(kWsStrip % WIN == 0 && WIN % 16 == 0, "windows tile the strip in 16-row blocks");
    const int lane = threadIdx.x;
    const int w = P.w, h = P.h, ww = w / WIN, H = h / WIN * WIN, W = ww * WIN;
    const int x0 = (int)bx * 64, y0 = (int)by * kWsStrip;
    if (x0 >= W || y0 >= H) return;  // uniform over the block
    const int y1 = min(y0 + kWsStrip, H), rend = y1 + 3;  // rows r = y0 - 4 .. rend
    const int x = x0 + lane;
    const bool inside = x < W;
    // this lane's sample column and, for lanes 0 .. 7, a margin column (x0 - 4 .. x0 - 1, x0 + 64 .. x0 + 67)
    const int xa = min(x, w - 1);
    const int xh = lane < 4 ? max(x0 - 4 + lane, 0) : min(x0 + 60 + lane, w - 1);
    const int hslot = lane < 4 ? lane : lane < 8 ? lane + 64 : 72 + lane;
    const WsRead<BR, kShift> rd0(P.r, h, P.up);
    const WsRead<BD, kShift> rd1(P.d, h, P.up);
    const unsigned ca[2] = {rd0.col(xa), rd1.col(xa)}, ch[2] = {rd0.col(xh), rd1.col(xh)};
    auto sample = [&](int p, unsigned c, int r) { return p == 0 ? rd0.load(c, rd0.row(r)) : rd1.load(c, rd1.row(r)); };
    // the M planes as buffers (none: an empty one, whose stores are dropped); a lane right of the covered width stores
    // past the end
    const __amdgpu_buffer_rsrc_t mid[2] = {buffer_over(P.mid_r, P.mid_r ? (unsigned long long)w * h * 4 : 0),
                                           buffer_over(P.mid_d, P.mid_d ? (unsigned long long)w * h * 4 : 0)};
    const unsigned mo = inside ? 4u * (unsigned)x : 0x80000000u;
    // samples of rows r .. r + 2 in slots (r - y0 + 4) mod 3: a row is loaded three rows before it is used (the row
    // loop is unrolled 9 times, so the slots rotate by renaming)
    int pr[3][2], ph[3][2];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int p = 0; p < 2; p++) {
            pr[k][p] = sample(p, ca[p], y0 - 4 + k);
            ph[k][p] = sample(p, ch[p], y0 - 4 + k);
        }
    // per picture: horizontal 3-sums centred on the column (h3c) and the horizontal 9-sums (h9) of the last 9 rows,
    // the differences of the 3-sums centred right and left of it (h3d) and the samples (yv) of the last 3 rows
    int h3c[2][9], h3d[2][3], h9[2][9], yv[2][3], s9[2];
#pragma unroll
    for (int p = 0; p < 2; p++) {
#pragma unroll
        for (int k = 0; k < 9; k++) h3c[p][k] = h9[p][k] = 0;
#pragma unroll
        for (int k = 0; k < 3; k++) h3d[p][k] = yv[p][k] = 0;
        s9[p] = 0;
    }
    int accF[3] = {0, 0, 0}, accG[6] = {0, 0, 0, 0, 0, 0};
    long long accM[3] = {0, 0, 0};
    // the luma squared error of the lane's column over the strip's rows: at most 64 rows of d^2 with d <= 1023 at the
    // 10-bit scale, below 2^26; lanes right of the covered width add zero
    unsigned sse = 0;
    const unsigned keep = inside && P.sse ? ~0u : 0u;
    long long winF[3] = {0, 0, 0}, winG[6] = {0, 0, 0, 0, 0, 0}, winM[3] = {0, 0, 0};  // WIN > 16 only
    // a closed 16-row block of one group: into the lane's window sums, and when the block closes a row of windows, the
    // window sums: each lane leaves its N sums in LDS, then lane k of each window's WIN lanes adds up sum k of them
    // and writes it to slot sbase + k sstride of the window
    auto flush = [&](auto &acc, auto &win, const int sbase, const int sstride, int y) {
        constexpr int N = sizeof(acc) / sizeof(acc[0]);
#pragma unroll
        for (int k = 0; k < N; k++) {
            win[k] += (long long)acc[k];
            acc[k] = 0;
        }
        if ((y + 1) % WIN) return;  // uniform
        __syncthreads();  // one wave per block: orders the LDS traffic only
#pragma unroll
        for (int k = 0; k < N; k++) {
            red[k][lane] = win[k];
            win[k] = 0;
        }
        __syncthreads();
        const int g0 = lane & ~(WIN - 1), k = lane & (WIN - 1);
        if (k < N) {
            long long t = 0;
#pragma unroll
            for (int j = 0; j < WIN; j += 2) {
                const longlong2 v = *(const longlong2 *)&red[k][g0 + j];
                t += v.x + v.y;
            }
            if (x0 + g0 < W)
                ((ES_GLOBAL long long *)P.sums)[((size_t)(y / WIN) * ww + (x0 + g0) / WIN) * 12 + sbase + k * sstride] = t;
        }
    };
    int buf = 0;
    for (int base = y0 - 4; base <= rend; base += 9) {
#pragma unroll
        for (int i = 0; i < 9; i++) {
            // whole groups of 9 rows: rows past rend read clamped samples and produce nothing
            const int r = base + i;
#pragma unroll
            for (int p = 0; p < 2; p++) {
                rowbuf[buf][p][lane + 4] = pr[i % 3][p];
                rowbuf[buf][p][hslot] = ph[i % 3][p];
            }
            if (r >= y0 && r < y1) {  // uniform
                const unsigned d = __usad((unsigned)pr[i % 3][0], (unsigned)pr[i % 3][1], 0u) & keep;
                sse = __umul24(d, d) + sse;
            }
#pragma unroll
            for (int p = 0; p < 2; p++) {
                pr[i % 3][p] = sample(p, ca[p], r + 3);
                ph[i % 3][p] = sample(p, ch[p], r + 3);
            }
            __syncthreads();  // the row is in (one wave per block: no other wave waits here)
            int F[2], GX[2], GY[2], M[2];
#pragma unroll
            for (int p = 0; p < 2; p++) {
                int v[9];
#pragma unroll
                for (int d = 0; d < 9; d++) v[d] = rowbuf[buf][p][lane + d];  // columns x - 4 .. x + 4
                const int a = v[3] + v[4], b56 = v[5] + v[6];
                const int hc = a + v[5], hd = b56 - (v[2] + v[3]);  // (v4 + v5 + v6) - (v2 + v3 + v4)
                const int h9v = (a + v[2]) + b56 + (v[0] + v[1]) + (v[7] + v[8]);
                // slots: row r - k sits in h3c / h9 slot (i - k) mod 9 and in the 3-slot rings at (i - k) mod 3
                const int c1 = h3c[p][(i + 8) % 9], c2 = h3c[p][(i + 7) % 9], c3 = h3c[p][(i + 6) % 9];
                const int c4 = h3c[p][(i + 5) % 9], c5 = h3c[p][(i + 4) % 9];
                // row r - 1: F = 9 y - S3
                F[p] = 9 * yv[p][(i + 2) % 3] - (c2 + c1 + hc);
                // row r - 2: GX = S3(x + 1) - S3(x - 1), GY = S3(y + 1) - S3(y - 1), zero on the picture's border
                const int yg = r - 2;
                GX[p] = (x > 0 && x < w - 1) ? h3d[p][0] + h3d[p][1] + h3d[p][2] : 0;
                GY[p] = (yg > 0 && yg < h - 1) ? (hc + c1) - (c3 + c4) : 0;
                // row r - 4: M = 9 S3 - S9
                s9[p] += h9v - h9[p][i];
                M[p] = 9 * (c5 + c4 + c3) - s9[p];
                h3c[p][i] = hc;
                h9[p][i] = h9v;
                h3d[p][i % 3] = hd;
                yv[p][i % 3] = v[4];
            }
            buf ^= 1;
            const int yf = r - 1, yg = r - 2, ym = r - 4;
            if (yf >= y0 && yf < y1) {
                accF[0] = __mul24(F[0], F[0]) + accF[0];
                accF[1] = __mul24(F[1], F[1]) + accF[1];
                accF[2] = __mul24(F[0], F[1]) + accF[2];
                if (yf % 16 == 15) flush(accF, winF, 0, 2, yf);
            }
            if (yg >= y0 && yg < y1) {
                accG[0] = __mul24(GX[0], GX[0]) + accG[0];
                accG[1] = __mul24(GY[0], GY[0]) + accG[1];
                accG[2] = __mul24(GX[0], GY[0]) + accG[2];
                accG[3] = __mul24(GX[1], GX[1]) + accG[3];
                accG[4] = __mul24(GY[1], GY[1]) + accG[4];
                accG[5] = __mul24(GX[1], GY[1]) + accG[5];
                if (yg % 16 == 15) flush(accG, winG, 6, 1, yg);
            }
            if (ym >= y0 && ym < y1) {
                const unsigned so = __builtin_amdgcn_readfirstlane(4u * (unsigned)ym * (unsigned)w);
#pragma unroll
                for (int p = 0; p < 2; p++) __builtin_amdgcn_raw_buffer_store_b32(M[p], mid[p], mo, so, 0);
                accM[0] += (long long)M[0] * M[0];
                accM[1] += (long long)M[1] * M[1];
                accM[2] += (long long)M[0] * M[1];
                if (ym % 16 == 15) flush(accM, winM, 1, 2, ym);
            }
        }
    }
    if (P.sse) {  // uniform
        const long long t = wave_total64((long long)sse);
        if (lane == 0) atomicAdd(P.sse, (unsigned long long)t >> (2 * P.up));
    }
}

template <int WIN>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void window_stats_int_kernel(const WsPair *pairs)
{
    __shared__ WsRows rowbuf;
    __shared__ WsRed red;
    const uint3 b = xcd_block3();
    const WsPair P = pairs[b.z];  // a copy: the stores cannot change it, so it is read once
    // one reader per pair of plane formats, chosen once per block
    if (P.r.bytes == 1) {
        if (P.d.bytes == 1) ws_strip<WIN, 1, 1, true>(P, rowbuf, red, b.x, b.y);
        else ws_strip<WIN, 1, 2, true>(P, rowbuf, red, b.x, b.y);
    } else if (P.d.bytes == 1) {
        ws_strip<WIN, 2, 1, true>(P, rowbuf, red, b.x, b.y);
    } else if ((P.r.rshift | P.r.shift | P.d.rshift | P.d.shift | P.up) == 0) {
        ws_strip<WIN, 2, 2, false>(P, rowbuf, red, b.x, b.y);  // 10-bit words, the common case
    } else {
        ws_strip<WIN, 2, 2, true>(P, rowbuf, red, b.x, b.y);
    }
}

template <int WIN>
static void launch_ws(const WsPair *pairs, int npairs, int max_w, int max_h, bool fp64, hipStream_t st)
{
    const int wcov = max_w / WIN * WIN, wh = max_h / WIN;
    if (fp64) {
        const dim3 grid((wcov + kTileCols - 1) / kTileCols, wh, npairs);
        hipLaunchKernelGGL(window_stats_tile_kernel<WIN>, grid, dim3(256), 0, st, pairs);
    } else {
        const dim3 grid((wcov + 63) / 64, (wh * WIN + kWsStrip - 1) / kWsStrip, npairs);
        hipLaunchKernelGGL(window_stats_int_kernel<WIN>, grid, dim3(64), 0, st, pairs);
    }
}

void launch_window_stats(const WsPair *pairs, int npairs, int max_w, int max_h, int win, bool fp64, hipStream_t st)
{
    if (npairs <= 0) return;
    if (win == 16) launch_ws<16>(pairs, npairs, max_w, max_h, fp64, st);
    else if (win == 32) launch_ws<32>(pairs, npairs, max_w, max_h, fp64, st);
    else if (win == 64) launch_ws<64>(pairs, npairs, max_w, max_h, fp64, st);
    else throw Error(1, "window size must be 16, 32 or 64");
    hip_check(hipGetLastError(), "window stats");
}

// ---------------------------------------------------------------------------------------------------------------
// Lasting texture.  One block of 256 threads per window; sums are exact (int64 or double on integers).
// ---------------------------------------------------------------------------------------------------------------
template <typename A, int N>
__device__ __forceinline__ void block_sums(A (&v)[N], A (*part)[4])
{
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int k = 0; k < N; k++) {
        v[k] = reduce_lanes(v[k], 64);
        if (lane == 0) part[k][wave] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = part[k][0] + part[k][1] + part[k][2] + part[k][3];
}

// Plain form: the 13 source planes and the encode plane per (pair, centre), in double, 5 sums per window.
template <int WIN>
__global__ __launch_bounds__(256) void persist_plain_kernel(const PersistJob *jobs)
{
    __shared__ double part[5][4];
    const PersistJob &J = jobs[blockIdx.z];
    const int wx = blockIdx.x, wy = blockIdx.y, ww = J.w / WIN;
    if (wx >= ww || wy >= J.h / WIN) return;
    double v[5] = {0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < WIN * WIN; i += 256) {
        const size_t k = (size_t)(wy * WIN + i / WIN) * J.w + wx * WIN + i % WIN;
        const double a = (double)J.ring[0][k] + J.ring[2][k] + J.ring[4][k] + J.ring[8][k] + J.ring[10][k] + J.ring[12][k];
        const double b = (double)J.ring[1][k] + J.ring[3][k] + J.ring[5][k] + J.ring[7][k] + J.ring[9][k] + J.ring[11][k];
        const double t = a + b;
        v[0] += (double)J.ring[6][k] * t;
        v[1] += (double)J.mid_d[k] * t;
        v[2] += a * b;
        v[3] += a * a;
        v[4] += b * b;
    }
    block_sums(v, part);
    if (threadIdx.x == 0) {
        long long *o = J.out + ((size_t)wy * ww + wx) * 5;
        for (int k = 0; k < 5; k++) o[k] = (long long)v[k];
    }
}

// Batched form: one wave per 16 x 16 patch of the area the windows cover, 4 pixels of a row per lane; the patch sums
// of a centre go to the window directly (16 x 16 windows) or by 64-bit atomic adds into cleared outputs (larger ones).
// Per centre a lane reads the source planes c - 7, c and c + 6 and the encodes' planes c; plane c - 1 is the previous
// centre's plane c, kept in registers.
template <int WIN>
__global__ __launch_bounds__(256) void lasting_kernel(const LastingJob *jobs)
{
    const uint3 b = xcd_block3();  // neighbouring patches share cache lines
    const LastingJob &J = jobs[b.z];
    const int lane = threadIdx.x & 63, wave = wave_index();
    const int ww = J.w / WIN, wh = J.h / WIN, pw = ww * (WIN / 16), ph = wh * (WIN / 16);
    const int patch = (int)b.x * 4 + wave;
    if (J.ncentres <= 0 || patch >= pw * ph) return;  // uniform over the wave
    const int px = patch % pw, py = patch / pw;
    const int x = px * 16 + (lane & 3) * 4, y = py * 16 + (lane >> 2);
    const size_t n = (size_t)J.w * J.h, i0 = (size_t)y * J.w + x;
    const int window = (py * 16 / WIN) * ww + px * 16 / WIN;
    // 4 pixels of a plane (4-byte aligned: planes of odd sizes exist)
    auto quad = [](const int32_t *p) { return make_int4(p[0], p[1], p[2], p[3]); };
    auto plane = [&](int f) { return quad(J.ring + (size_t)(f % J.slots) * n + i0); };
    int sa[4], sb[4];
    const int c0 = J.centre0;
    if (c0 > kRadius) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            sa[k] = J.sa[i0 + k];
            sb[k] = J.sb[i0 + k];
        }
    }
    // The source planes of frames c - 7 .. c + 6 of centre c, in registers: frame f in slot (f - c0 + 7) mod 14.
    // Each centre uses frames c - 7, c - 1, c and c + 6 and then loads frame c + 7 into the slot of frame c - 7, so
    // every plane is read once per patch (the row loop over centres is unrolled 14 times, so the slots are fixed).
    int4 R[14];
#pragma unroll
    for (int q = 0; q < 14; q++) R[q] = c0 - 7 + q >= 0 ? plane(c0 - 7 + q) : make_int4(0, 0, 0, 0);
    auto get = [](const int4 &v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; };
    for (int ci0 = 0; ci0 < J.ncentres; ci0 += 14) {
#pragma unroll
        for (int u = 0; u < 14; u++) {
            const int ci = ci0 + u;
            if (ci >= J.ncentres) break;
            const int c = c0 + ci;
            const int4 &mc = R[(u + 7) % 14], &m6 = R[(u + 13) % 14], &m7 = R[u], &m1 = R[(u + 6) % 14];
            int m0[4];
            if (c == kRadius) {  // the first centre of the run: the sums in full, frames 0 .. 12 in slots 1 .. 13
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    sa[k] = get(R[1], k) + get(R[3], k) + get(R[5], k) + get(R[9], k) + get(R[11], k) + get(R[13], k);
                    sb[k] = get(R[2], k) + get(R[4], k) + get(R[6], k) + get(R[8], k) + get(R[10], k) + get(R[12], k);
                    m0[k] = get(R[7], k);
                }
            } else {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    m0[k] = get(mc, k);
                    const int a = sb[k] - get(mc, k) + get(m6, k);
                    const int b = sa[k] - get(m7, k) + get(m1, k);
                    sa[k] = a;
                    sb[k] = b;
                }
            }
            // frame c + 7, the next centre's c + 6, into the slot of frame c - 7 (past the run: a valid frame, unused)
            R[u] = plane(c + 7 < c0 + J.ncentres + kRadius ? c + 7 : c);
            long long v[4] = {0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const long long st = (long long)sa[k] + sb[k];
                v[0] += m0[k] * st;
                v[1] += (long long)sa[k] * sb[k];
                v[2] += (long long)sa[k] * sa[k];
                v[3] += (long long)sb[k] * sb[k];
            }
            const int slot = J.slot0 + ci;
            {
                const long long t = wave_total4x64(v);  // value k in lane 16 k + 15
                if ((lane & 15) == 15) {
                    long long *o = J.psrc + ((size_t)slot * ww * wh + window) * 4 + (lane >> 4);
                    if (WIN == 16) *o = t;
                    else atomicAdd((unsigned long long *)o, (unsigned long long)t);
                }
            }
            for (int e = 0; e < J.nenc; e++) {
                if (c > J.last_centre[e]) continue;
                const int4 pe = quad(J.emid[e] + (size_t)(c % J.slots) * n + i0);
                long long w = 0;
#pragma unroll
                for (int k = 0; k < 4; k++) w += (long long)get(pe, k) * ((long long)sa[k] + sb[k]);
                w = wave_total64(w);
                if (lane == 0) {
                    long long *o = J.penc[e] + (size_t)slot * ww * wh + window;
                    if (WIN == 16) *o = w;
                    else atomicAdd((unsigned long long *)o, (unsigned long long)w);
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        J.sa[i0 + k] = sa[k];
        J.sb[i0 + k] = sb[k];
    }
}

// The reference-shaped lasting texture (src/kernels.cu:435-451) on float mid planes: per 16 x 16 window of the centre
// frame, A and B the means of the source's mid planes at offsets +-2, 4, 6 and +-1, 3, 5, T = (A + B) / 2, and the
// window sums [m(c) T, mid_d T, A B, A^2, B^2], in double, written as floats.  One block of 256 threads per window.
__global__ __launch_bounds__(256) void persist_f32_kernel(const PersistF32Job *jobs)
{
    __shared__ double part[5][4];
    const PersistF32Job &J = jobs[blockIdx.z];
    const int wx = blockIdx.x, wy = blockIdx.y, ww = J.w / 16;
    if (wx >= ww || wy >= J.h / 16) return;
    const size_t k = (size_t)(wy * 16 + (threadIdx.x >> 4)) * J.w + wx * 16 + (threadIdx.x & 15);
    const double a = ((double)J.ring[0][k] + J.ring[2][k] + J.ring[4][k] + J.ring[8][k] + J.ring[10][k] + J.ring[12][k]) / 6.0;
    const double b = ((double)J.ring[1][k] + J.ring[3][k] + J.ring[5][k] + J.ring[7][k] + J.ring[9][k] + J.ring[11][k]) / 6.0;
    const double t = (a + b) / 2.0;
    double v[5] = {(double)J.ring[6][k] * t, (double)J.mid_d[k] * t, a * b, a * a, b * b};
    block_sums(v, part);
    if (threadIdx.x == 0) {
        float *o = J.out + ((size_t)wy * ww + wx) * 5;
        for (int i = 0; i < 5; i++) o[i] = __double2float_rn(v[i]);
    }
}

void launch_persist_f32(const PersistF32Job *jobs, int njobs, int max_w, int max_h, hipStream_t st)
{
    if (njobs <= 0 || max_w < 16 || max_h < 16) return;
    hipLaunchKernelGGL(persist_f32_kernel, dim3(max_w / 16, max_h / 16, njobs), dim3(256), 0, st, jobs);
    hip_check(hipGetLastError(), "lasting texture");
}

template <int WIN>
static void launch_lasting_t(const LastingJob *jobs, int njobs, int max_w, int max_h, hipStream_t st)
{
    const int patches = (max_w / WIN) * (WIN / 16) * (max_h / WIN) * (WIN / 16);
    if (patches <= 0) return;
    hipLaunchKernelGGL(lasting_kernel<WIN>, dim3((patches + 3) / 4, 1, njobs), dim3(256), 0, st, jobs);
}

void launch_lasting(const LastingJob *jobs, int njobs, int max_w, int max_h, int win, hipStream_t st)
{
    if (njobs <= 0) return;
    if (win == 16) launch_lasting_t<16>(jobs, njobs, max_w, max_h, st);
    else if (win == 32) launch_lasting_t<32>(jobs, njobs, max_w, max_h, st);
    else if (win == 64) launch_lasting_t<64>(jobs, njobs, max_w, max_h, st);
    else throw Error(1, "window size must be 16, 32 or 64");
    hip_check(hipGetLastError(), "lasting texture");
}

template <int WIN>
static void launch_persist_t(const PersistJob *jobs, int njobs, int max_w, int max_h, int kind, hipStream_t st)
{
    const dim3 grid(max_w / WIN, max_h / WIN, njobs);
    (void)kind;
    hipLaunchKernelGGL(persist_plain_kernel<WIN>, grid, dim3(256), 0, st, jobs);
}

void launch_persist(const PersistJob *jobs, int njobs, int max_w, int max_h, int win, int kind, hipStream_t st)
{
    if (njobs <= 0) return;
    if (win == 16) launch_persist_t<16>(jobs, njobs, max_w, max_h, kind, st);
    else if (win == 32) launch_persist_t<32>(jobs, njobs, max_w, max_h, kind, st);
    else if (win == 64) launch_persist_t<64>(jobs, njobs, max_w, max_h, kind, st);
    else throw Error(1, "window size must be 16, 32 or 64");
    hip_check(hipGetLastError(), "lasting texture");
}

}  // namespace es
