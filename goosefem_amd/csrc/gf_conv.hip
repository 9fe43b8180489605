// Kernels of the learned warm start for the tied PCG: a periodic 2-D encoder-decoder on the
// element grid, trained and evaluated on the GPU.  No counterpart in the reference.
//
// Layout: activations NHWC bfloat16, channel counts multiples of 16 (padded with zeros); weights
// packed as [9 taps][cout padded to 32][cin] bfloat16; master weights FP32 [9][cout][cin].
// Convolutions are 3x3, stride 1, circular padding (the cell is periodic).  One wave computes a
// 32-pixel x 32*NC-channel output tile with v_mfma_f32_32x32x16_bf16 (implicit GEMM: rows =
// pixels, k = 16 input channels of one tap, columns = output channels), FP32 accumulation.
// The data gradient of a convolution is the same kernel with tap-flipped, transposed weights.
// The weight gradient contracts over pixels: tiles of 64 pixels are transposed through LDS
// so each lane feeds 8 consecutive pixels of one channel to the MFMA; per-split partial sums are
// reduced in a fixed order, so training is deterministic run to run.
#include <algorithm>
#include <cmath>

#include "gf_internal.hpp"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

using namespace gf;

namespace {

constexpr int CF_RELU = 1;  // y = max(y, 0)
constexpr int CF_ACCUM = 2; // y = y_old + result (before ReLU / mask)

__device__ inline int wrap(int i, int n)
{
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

// ---- 3x3 circular convolution, forward and data gradient ------------------------------------------
// input channels [0, c1) come from x1 (at half resolution when up1: nearest upsampling is read in
// place), [c1, c1 + c2) from x2.  Output channels [0, ysplit) go to y (row stride ysplit),
// [ysplit, cout) to y2 (row stride cout - ysplit); y32 (optional) receives FP32 [P][cout].
// D = 2: 3x3 taps t = 3 (dy + 1) + dx + 1 on [N, H, W] (Z = 1); D = 3: 3x3x3 taps
// t = 9 (dz + 1) + 3 (dy + 1) + dx + 1 on [N, Z, H, W].
template <int NC, int D>
__global__ void __launch_bounds__(256) k_conv3x3(int64_t P, int Z, int H, int W, const __bf16* __restrict__ x1, int c1,
                                                 int up1, const __bf16* __restrict__ x2, int c2,
                                                 const __bf16* __restrict__ wpk, const float* __restrict__ bias,
                                                 int cout, int flags, __bf16* __restrict__ y, int ysplit,
                                                 __bf16* __restrict__ y2, float* __restrict__ y32,
                                                 const __bf16* __restrict__ mask)
{
    const int lane = threadIdx.x & 63;
    const int64_t p0 = (static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6)) * 32;
    if (p0 >= P) return; // uniform per wave
    const int row = lane & 31, kh = lane >> 5;
    const int cin = c1 + c2;
    constexpr int COUTP = 32 * NC;
    constexpr int NTAP = D == 3 ? 27 : 9;
    const int64_t HW = static_cast<int64_t>(H) * W, ZHW = HW * Z;
    const int64_t p = p0 + row;
    const bool pv = p < P;
    const int64_t pn = pv ? p / ZHW : 0;
    const int64_t rz = pv ? p - pn * ZHW : 0;
    const int pz = static_cast<int>(rz / HW);
    const int rem = static_cast<int>(rz - static_cast<int64_t>(pz) * HW);
    const int py = rem / W, px = rem - (rem / W) * W;
    const int Z1 = (up1 && D == 3) ? Z >> 1 : Z, H1 = up1 ? H >> 1 : H, W1 = up1 ? W >> 1 : W;
    floatx16 acc[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[j][r] = 0.0f;
    bf16x8 zero;
#pragma unroll
    for (int i = 0; i < 8; ++i) zero[i] = static_cast<__bf16>(0.0f);

    for (int t = 0; t < NTAP; ++t) {
        const int t2 = D == 3 ? t % 9 : t;
        const int zz = D == 3 ? wrap(pz + t / 9 - 1, Z) : 0;
        const int yy = wrap(py + t2 / 3 - 1, H), xx = wrap(px + t2 % 3 - 1, W);
        const int z1 = (up1 && D == 3) ? zz >> 1 : zz, y1 = up1 ? yy >> 1 : yy, xq = up1 ? xx >> 1 : xx;
        const __bf16* a1 = x1 + (((pn * Z1 + z1) * H1 + y1) * W1 + xq) * c1 + 8 * kh;
        const __bf16* a2 = x2 ? x2 + (((pn * Z + zz) * H + yy) * W + xx) * c2 + 8 * kh : nullptr;
        const __bf16* wt = wpk + (static_cast<int64_t>(t) * COUTP + row) * cin + 8 * kh;
        for (int c0 = 0; c0 < cin; c0 += 16) {
            bf16x8 a = zero;
            if (pv) a = c0 < c1 ? *reinterpret_cast<const bf16x8*>(a1 + c0)
                                : *reinterpret_cast<const bf16x8*>(a2 + (c0 - c1));
#pragma unroll
            for (int j = 0; j < NC; ++j) {
                const bf16x8 b = *reinterpret_cast<const bf16x8*>(wt + static_cast<int64_t>(32 * j) * cin + c0);
                acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[j], 0, 0, 0);
            }
        }
    }
    // accumulator r of lane (row, kh): pixel p0 + 8 (r / 4) + 4 kh + r % 4, channel 32 j + row
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        const int co = 32 * j + row;
        if (co >= cout) continue;
        const float bco = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int64_t po = p0 + 8 * (r / 4) + 4 * kh + (r % 4);
            if (po >= P) continue;
            float v = acc[j][r] + bco;
            __bf16* dst = nullptr;
            if (y) dst = co < ysplit ? y + po * ysplit + co : y2 + po * (cout - ysplit) + (co - ysplit);
            if ((flags & CF_ACCUM) && dst) v += static_cast<float>(*dst);
            if (flags & CF_RELU) v = fmaxf(v, 0.0f);
            if (mask && !(static_cast<float>(mask[po * cout + co]) > 0.0f)) v = 0.0f;
            if (y32) y32[po * cout + co] = v;
            if (dst) *dst = static_cast<__bf16>(v);
        }
    }
}

// ---- weight gradient: part[s][t][co][ci] = sum over split s's pixels of g[p][co] X[p + tap t][ci] --
// block = 4 waves for one (tap, 32-wide co tile, split); wave w owns ci tiles w, w + 4 (cin <= 256).
constexpr int WG_PIX = 64;  // pixels per LDS tile
constexpr int WG_LD = 72;   // LDS row pitch (bf16), 144 B: 16-byte aligned rows
constexpr int WG_MAXC = 256;

template <int D>
__global__ void __launch_bounds__(256) k_conv3x3_wgrad(int64_t P, int Z, int H, int W, const __bf16* __restrict__ x1, int c1,
                                                       int up1, const __bf16* __restrict__ x2, int c2,
                                                       const __bf16* __restrict__ g, int cg, int nsplit,
                                                       float* __restrict__ part, int coutp, int cinp)
{
    __shared__ __align__(16) __bf16 sG[32][WG_LD];
    __shared__ __align__(16) __bf16 sX[WG_MAXC][WG_LD];
    const int nct = coutp / 32;
    const int split = blockIdx.x % nsplit;
    const int ct = (blockIdx.x / nsplit) % nct;
    const int t = blockIdx.x / (nsplit * nct);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 31, kh = lane >> 5;
    const int cin = c1 + c2;
    const int nci = (cin + 31) / 32;
    const int64_t HW = static_cast<int64_t>(H) * W, ZHW = HW * Z;
    const int Z1 = (up1 && D == 3) ? Z >> 1 : Z, H1 = up1 ? H >> 1 : H, W1 = up1 ? W >> 1 : W;
    const int t2 = D == 3 ? t % 9 : t;
    const int dz = D == 3 ? t / 9 - 1 : 0, dy = t2 / 3 - 1, dx = t2 % 3 - 1;
    floatx16 acc[2];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[k][r] = 0.0f;

    const int64_t nchunk = (P + WG_PIX - 1) / WG_PIX;
    for (int64_t ch = split; ch < nchunk; ch += nsplit) {
        const int64_t pbase = ch * WG_PIX;
        // stage g (32 channels of this co tile) and the tap-shifted input, transposed
        for (int k = threadIdx.x; k < WG_PIX * 4; k += 256) {
            const int pp = k >> 2, c8 = (k & 3) * 8;
            const int64_t p = pbase + pp;
            const int co = ct * 32 + c8;
            bf16x8 v;
            if (p < P && co < cg) v = *reinterpret_cast<const bf16x8*>(g + p * cg + co);
            else
                for (int i = 0; i < 8; ++i) v[i] = static_cast<__bf16>(0.0f);
#pragma unroll
            for (int i = 0; i < 8; ++i) sG[c8 + i][pp] = v[i];
        }
        const int nc8 = nci * 4; // 8-channel groups, padded to the 32-wide ci tiles
        for (int k = threadIdx.x; k < WG_PIX * nc8; k += 256) {
            const int pp = k / nc8, c8 = (k - (k / nc8) * nc8) * 8;
            const int64_t p = pbase + pp;
            bf16x8 v;
            if (p < P && c8 < cin) {
                const int64_t pn = p / ZHW;
                const int64_t rz = p - pn * ZHW;
                const int pz = static_cast<int>(rz / HW);
                const int rem = static_cast<int>(rz - static_cast<int64_t>(pz) * HW);
                const int py = rem / W, px = rem - (rem / W) * W;
                const int zz = D == 3 ? wrap(pz + dz, Z) : 0;
                const int yy = wrap(py + dy, H), xx = wrap(px + dx, W);
                if (c8 < c1) {
                    const int z1 = (up1 && D == 3) ? zz >> 1 : zz, y1 = up1 ? yy >> 1 : yy, xq = up1 ? xx >> 1 : xx;
                    v = *reinterpret_cast<const bf16x8*>(x1 + (((pn * Z1 + z1) * H1 + y1) * W1 + xq) * c1 + c8);
                }
                else
                    v = *reinterpret_cast<const bf16x8*>(x2 + (((pn * Z + zz) * H + yy) * W + xx) * c2 + (c8 - c1));
            }
            else
                for (int i = 0; i < 8; ++i) v[i] = static_cast<__bf16>(0.0f);
#pragma unroll
            for (int i = 0; i < 8; ++i) sX[c8 + i][pp] = v[i];
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < WG_PIX / 16; ++ks) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sG[row][ks * 16 + 8 * kh]);
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int ci_t = wave + 4 * k;
                if (ci_t < nci) {
                    const bf16x8 b = *reinterpret_cast<const bf16x8*>(&sX[ci_t * 32 + row][ks * 16 + 8 * kh]);
                    acc[k] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[k], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // acc[k][r]: co = ct*32 + 8 (r / 4) + 4 kh + r % 4, ci = (wave + 4k)*32 + row
    constexpr int NTAP = D == 3 ? 27 : 9;
    float* dst = part + (static_cast<int64_t>(split) * NTAP + t) * coutp * cinp;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int ci_t = wave + 4 * k;
        if (ci_t >= nci) continue;
        const int ci = ci_t * 32 + row;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int co = ct * 32 + 8 * (r / 4) + 4 * kh + (r % 4);
            dst[static_cast<int64_t>(co) * cinp + ci] = acc[k][r];
        }
    }
}

// dw[t][co][ci] (master shape cout x cin) = sum_s part[s][t][co][ci], splits in ascending order
__global__ void k_wgrad_reduce(int ntap, int nsplit, int coutp, int cinp, int cout, int cin, const float* __restrict__ part,
                               float* __restrict__ dw)
{
    const int64_t n = static_cast<int64_t>(ntap) * cout * cin;
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const int t = static_cast<int>(i / (static_cast<int64_t>(cout) * cin));
    const int rem = static_cast<int>(i - static_cast<int64_t>(t) * cout * cin);
    const int co = rem / cin, ci = rem - (rem / cin) * cin;
    const int64_t stride = static_cast<int64_t>(ntap) * coutp * cinp;
    const float* s = part + (static_cast<int64_t>(t) * coutp + co) * cinp + ci;
    float acc = 0.0f;
    for (int k = 0; k < nsplit; ++k) acc += s[k * stride];
    dw[i] = acc;
}

// bias gradient partials: part[s][co] = sum over split s's pixels of g[p][co]
__global__ void __launch_bounds__(256) k_bias_grad(int64_t P, const __bf16* __restrict__ g, int cg, int cout,
                                                   int nsplit, float* __restrict__ part)
{
    __shared__ float sh[256];
    const int s = blockIdx.x;
    const int per = 256 / cg; // pixel lanes per channel (threads beyond per * cg idle)
    const int co = threadIdx.x % cg, pl = threadIdx.x / cg;
    float acc = 0.0f;
    if (pl < per)
        for (int64_t p = static_cast<int64_t>(s) * per + pl; p < P; p += static_cast<int64_t>(nsplit) * per)
            acc += static_cast<float>(g[p * cg + co]);
    sh[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < cg) {
        float tot = 0.0f;
        for (int k = 0; k < per; ++k) tot += sh[k * cg + threadIdx.x];
        if (threadIdx.x < cout) part[static_cast<int64_t>(s) * cout + threadIdx.x] = tot;
    }
}

__global__ void k_reduce_rows(int nsplit, int n, const float* __restrict__ part, float* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc = 0.0f;
    for (int k = 0; k < nsplit; ++k) acc += part[static_cast<int64_t>(k) * n + i];
    out[i] = acc;
}

// out[0] = sum of part[0..n) by one workgroup: strided per-thread sums, then an LDS tree (fixed order)
__global__ void __launch_bounds__(256) k_sum_block(int64_t n, const float* __restrict__ part, float* __restrict__ out)
{
    __shared__ double sh[256];
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) acc += part[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = static_cast<float>(sh[0]);
}

// ---- multi-scale mean-squared error -----------------------------------------------------------------
// e = pred - target on grids [B][Z][H][W][nc] (Z = 1 for D = 2).  Level s of the pyramid holds the
// mean of e over cells 2^s wide, level s + 1 the mean of 2^D level-s cells.  The gradient of
// L = sum e^2 + lam sum_s (2^D)^s sum_cells (mean_s - mu)^2 (mu: the grid's mean) is
// 2 e + 2 lam sum_s (mean_s(cell of i) - mu) at fine cell i, since the centred cell means of a level
// sum to zero; and (2^D)^s sum_cells m^2 = sum_i m(cell of i)^2, so the loss is a sum over fine cells.

// level 1 from the fine error (src_fine != 0: src holds pred, src2 target) or level s + 1 from level s
template <int D>
__global__ void __launch_bounds__(256) k_err_pool(int64_t ncoarse, int Zc, int Hc, int Wc, int nc, int fine,
                                                  const float* __restrict__ src, const float* __restrict__ src2,
                                                  float* __restrict__ dst)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= ncoarse * nc) return;
    const int c = static_cast<int>(t % nc);
    int64_t r = t / nc;
    const int x = static_cast<int>(r % Wc);
    r /= Wc;
    const int y = static_cast<int>(r % Hc);
    r /= Hc;
    const int z = static_cast<int>(r % Zc);
    const int64_t b = r / Zc;
    const int Wf = 2 * Wc, Hf = 2 * Hc, Zf = D == 3 ? 2 * Zc : 1;
    float acc = 0.0f;
#pragma unroll
    for (int dz = 0; dz < (D == 3 ? 2 : 1); ++dz)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                const int64_t i =
                    ((((b * Zf + (D == 3 ? 2 * z + dz : 0)) * Hf + 2 * y + dy) * Wf) + 2 * x + dx) * nc + c;
                acc += fine ? src[i] - src2[i] : src[i];
            }
    dst[t] = acc * (D == 3 ? 0.125f : 0.25f);
}

// per (grid, channel) mean of the coarsest level (ncell cells per grid)
__global__ void __launch_bounds__(64) k_err_mean(int64_t ngrid, int ncell, int nc, const float* __restrict__ top,
                                                 float* __restrict__ mu)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= ngrid * nc) return;
    const int64_t b = t / nc;
    const int c = static_cast<int>(t % nc);
    float acc = 0.0f;
    for (int k = 0; k < ncell; ++k) acc += top[(b * ncell + k) * nc + c];
    mu[t] = acc / static_cast<float>(ncell);
}

constexpr int MS_MAX_LEVELS = 12;
struct MsLevels {
    const float* lv[MS_MAX_LEVELS]; // level s + 1 at lv[s]
};

template <int D>
__global__ void __launch_bounds__(256) k_mse_multiscale(int64_t P, int Z, int H, int W, int nc, int S, float lam,
                                                        const float* __restrict__ pred,
                                                        const float* __restrict__ target, MsLevels L,
                                                        const float* __restrict__ mu, float scale, int cg,
                                                        __bf16* __restrict__ g, float* __restrict__ part)
{
    __shared__ float sh[256];
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    float sq = 0.0f;
    if (i < P * cg) {
        const int64_t p = i / cg;
        const int c = static_cast<int>(i - p * cg);
        float gv = 0.0f;
        if (c < nc) {
            const float d = pred[p * nc + c] - target[p * nc + c];
            int64_t r = p;
            const int x = static_cast<int>(r % W);
            r /= W;
            const int y = static_cast<int>(r % H);
            r /= H;
            const int z = static_cast<int>(r % Z);
            const int64_t b = r / Z;
            const float m0 = mu[b * nc + c];
            float ms = 0.0f, msq = 0.0f;
            for (int s = 1; s <= S; ++s) {
                const int Ws = W >> s, Hs = H >> s, Zs = D == 3 ? Z >> s : 1;
                const int64_t cell = ((b * Zs + (D == 3 ? z >> s : 0)) * Hs + (y >> s)) * Ws + (x >> s);
                const float m = L.lv[s - 1][cell * nc + c] - m0;
                ms += m;
                msq += m * m;
            }
            sq = d * d + lam * msq;
            gv = scale * (2.0f * d + 2.0f * lam * ms) / static_cast<float>(P * nc);
        }
        g[i] = static_cast<__bf16>(gv);
    }
    sh[threadIdx.x] = sq;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// ---- 2x2 (D = 2) / 2x2x2 (D = 3) pooling --------------------------------------------------------------
// grids [N, Z, H, W, C] with Z = 1 for D = 2; the coarse grid halves every pooled dimension
template <int D>
__device__ inline float block_sum(const __bf16* b, int64_t sy, int64_t sz, int C)
{
    float s = static_cast<float>(b[0]) + static_cast<float>(b[C]) + static_cast<float>(b[sy]) +
              static_cast<float>(b[sy + C]);
    if constexpr (D == 3)
        s += static_cast<float>(b[sz]) + static_cast<float>(b[sz + C]) + static_cast<float>(b[sz + sy]) +
             static_cast<float>(b[sz + sy + C]);
    return s;
}

// avg: out = mean of the 2^D block of in
template <int D>
__global__ void k_avgpool2(int64_t Pout, int Zo, int Ho, int Wo, int C, const __bf16* __restrict__ in,
                           __bf16* __restrict__ out)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= Pout * C) return;
    const int c = static_cast<int>(i % C);
    const int64_t p = i / C;
    const int64_t HWo = static_cast<int64_t>(Ho) * Wo, ZHWo = HWo * Zo;
    const int64_t n = p / ZHWo, rz = p - n * ZHWo;
    const int z = static_cast<int>(rz / HWo), rem = static_cast<int>(rz - z * HWo);
    const int y = rem / Wo, x = rem - (rem / Wo) * Wo;
    const int Zi = D == 3 ? 2 * Zo : Zo, Hi = 2 * Ho, Wi = 2 * Wo;
    const int zi = D == 3 ? 2 * z : z;
    const __bf16* b = in + (((n * Zi + zi) * Hi + 2 * y) * Wi + 2 * x) * static_cast<int64_t>(C) + c;
    const int64_t sy = static_cast<int64_t>(Wi) * C, sz = sy * Hi;
    out[i] = static_cast<__bf16>((D == 3 ? 0.125f : 0.25f) * block_sum<D>(b, sy, sz, C));
}

// fine-grid gradient from a coarse one: din[fine] (+)= w * dout[coarse]; then masked by act > 0
// (w = 1 / 2^D for average pooling)
template <int D>
__global__ void k_unpool2(int64_t Pin, int Zi, int Hi, int Wi, int C, float w, const __bf16* __restrict__ dout,
                          int accum, const __bf16* __restrict__ act, __bf16* __restrict__ din)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= Pin * C) return;
    const int c = static_cast<int>(i % C);
    const int64_t p = i / C;
    const int64_t HWi = static_cast<int64_t>(Hi) * Wi, ZHWi = HWi * Zi;
    const int64_t n = p / ZHWi, rz = p - n * ZHWi;
    const int z = static_cast<int>(rz / HWi), rem = static_cast<int>(rz - z * HWi);
    const int y = rem / Wi, x = rem - (rem / Wi) * Wi;
    const int Zo = D == 3 ? Zi / 2 : Zi, Ho = Hi / 2, Wo = Wi / 2;
    const int zo = D == 3 ? z / 2 : z;
    float v = w * static_cast<float>(dout[(((n * Zo + zo) * Ho + y / 2) * Wo + x / 2) * static_cast<int64_t>(C) + c]);
    if (accum) v += static_cast<float>(din[i]);
    if (act && !(static_cast<float>(act[i]) > 0.0f)) v = 0.0f;
    din[i] = static_cast<__bf16>(v);
}

// coarse gradient of nearest upsampling: out[coarse] = sum of the 2^D fine block; masked by act > 0
template <int D>
__global__ void k_sumpool2(int64_t Pout, int Zo, int Ho, int Wo, int C, const __bf16* __restrict__ in,
                           const __bf16* __restrict__ act, __bf16* __restrict__ out)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= Pout * C) return;
    const int c = static_cast<int>(i % C);
    const int64_t p = i / C;
    const int64_t HWo = static_cast<int64_t>(Ho) * Wo, ZHWo = HWo * Zo;
    const int64_t n = p / ZHWo, rz = p - n * ZHWo;
    const int z = static_cast<int>(rz / HWo), rem = static_cast<int>(rz - z * HWo);
    const int y = rem / Wo, x = rem - (rem / Wo) * Wo;
    const int Zi = D == 3 ? 2 * Zo : Zo, Hi = 2 * Ho, Wi = 2 * Wo;
    const int zi = D == 3 ? 2 * z : z;
    const __bf16* b = in + (((n * Zi + zi) * Hi + 2 * y) * Wi + 2 * x) * static_cast<int64_t>(C) + c;
    const int64_t sy = static_cast<int64_t>(Wi) * C, sz = sy * Hi;
    float sum = block_sum<D>(b, sy, sz, C);
    if (act && !(static_cast<float>(act[i]) > 0.0f)) sum = 0.0f;
    out[i] = static_cast<__bf16>(sum);
}

// ---- weights ------------------------------------------------------------------------------------------
// master w [9][cout][cin] FP32 -> forward pack [9][coutp][cin_s] and data-gradient pack
// [9][cinp][cout_s] (tap-flipped, transposed), zero padded
__global__ void k_pack(int ntap, int cout, int cin, const float* __restrict__ w, int coutp, int cin_s,
                       __bf16* __restrict__ wf, int cinp, int cout_s, __bf16* __restrict__ wb)
{
    const int64_t nf = static_cast<int64_t>(ntap) * coutp * cin_s, nb = static_cast<int64_t>(ntap) * cinp * cout_s;
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i < nf) {
        const int t = static_cast<int>(i / (static_cast<int64_t>(coutp) * cin_s));
        const int rem = static_cast<int>(i - static_cast<int64_t>(t) * coutp * cin_s);
        const int co = rem / cin_s, ci = rem - (rem / cin_s) * cin_s;
        const float v = (co < cout && ci < cin) ? w[(static_cast<int64_t>(t) * cout + co) * cin + ci] : 0.0f;
        wf[i] = static_cast<__bf16>(v);
    }
    if (wb && i < nb) {
        const int t = static_cast<int>(i / (static_cast<int64_t>(cinp) * cout_s));
        const int rem = static_cast<int>(i - static_cast<int64_t>(t) * cinp * cout_s);
        const int ci = rem / cout_s, co = rem - (rem / cout_s) * cout_s;
        const float v = (co < cout && ci < cin) ? w[(static_cast<int64_t>(ntap - 1 - t) * cout + co) * cin + ci] : 0.0f;
        wb[i] = static_cast<__bf16>(v);
    }
}

// Adam with bias correction (Kingma & Ba 2015), FP32 master weights
__global__ void k_adam(int64_t n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---- loss -----------------------------------------------------------------------------------------------
// pred, target [P][nc] FP32 -> g [P][cg] bf16 (channels >= nc zero) = scale * 2 (pred - target) / (P nc);
// part[block] = sum of squared differences over the block (ordered tree in LDS)
__global__ void __launch_bounds__(256) k_mse(int64_t P, int nc, const float* __restrict__ pred,
                                             const float* __restrict__ target, float scale, int cg,
                                             __bf16* __restrict__ g, float* __restrict__ part)
{
    __shared__ float sh[256];
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    float sq = 0.0f;
    if (i < P * cg) {
        const int64_t p = i / cg;
        const int c = static_cast<int>(i - p * cg);
        float d = 0.0f;
        if (c < nc) {
            d = pred[p * nc + c] - target[p * nc + c];
            sq = d * d;
        }
        g[i] = static_cast<__bf16>(c < nc ? scale * 2.0f * d / static_cast<float>(P * nc) : 0.0f);
    }
    sh[threadIdx.x] = sq;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// ---- element grid <-> tied unknowns -----------------------------------------------------------------------
// x_u[idx[k]] = gamma * (scale * (grid[k] - grid[k % nc]) + affine[k]) for idx[k] >= 0 (grid[c] of
// node 0 is the origin's prediction; the fluctuation is defined up to a translation)
__global__ void k_grid_to_dofs(int64_t n, int nc, const float* __restrict__ grid, const int32_t* __restrict__ idx,
                               const double* __restrict__ affine, double scale, double gamma, double* __restrict__ xu)
{
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= n) return;
    const int32_t d = idx[k];
    if (d < 0) return;
    const double w = static_cast<double>(grid[k]) - static_cast<double>(grid[k % nc]);
    xu[d] = gamma * (scale * w + affine[k]);
}

// grid[k] = (x_u[idx[k]] / gamma - affine[k]) / scale  (0 where idx < 0)
__global__ void k_dofs_to_grid(int64_t n, const double* __restrict__ xu, const int32_t* __restrict__ idx,
                               const double* __restrict__ affine, double scale, double gamma, float* __restrict__ grid)
{
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= n) return;
    const int32_t d = idx[k];
    const double w = d >= 0 ? xu[d] / gamma - affine[k] : 0.0;
    grid[k] = static_cast<float>(w / scale);
}

// subtract the per-channel mean of one grid [n][nc] (one block, ordered reduction)
__global__ void __launch_bounds__(256) k_center(int64_t n, int nc, float* __restrict__ grid)
{
    __shared__ double sh[256];
    for (int c = 0; c < nc; ++c) {
        double acc = 0.0;
        for (int64_t p = threadIdx.x; p < n; p += 256) acc += grid[p * nc + c];
        sh[threadIdx.x] = acc;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
            __syncthreads();
        }
        const float mean = static_cast<float>(sh[0] / static_cast<double>(n));
        __syncthreads();
        for (int64_t p = threadIdx.x; p < n; p += 256) grid[p * nc + c] -= mean;
        __syncthreads();
    }
}

// FP64 -> bf16 network input: x[p][cs] = (channel 0) scale * (log(G[p]) - shift), others 0
__global__ void k_input_logG(int64_t P, const double* __restrict__ G, double shift, double scale, int cs,
                             __bf16* __restrict__ x)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= P * cs) return;
    const int64_t p = i / cs;
    const int c = static_cast<int>(i - p * cs);
    x[i] = static_cast<__bf16>(c == 0 ? static_cast<float>(scale * (log(G[p]) - shift)) : 0.0f);
}

void check_conv_shapes(int64_t P, int Z, int H, int W, int c1, int up1, int c2, int D)
{
    GF_REQUIRE(P > 0 && Z > 0 && H > 0 && W > 0 && P % (static_cast<int64_t>(Z) * H * W) == 0,
               "P must be a multiple of Z*H*W");
    GF_REQUIRE(D == 3 || Z == 1, "2-D grids have Z = 1");
    GF_REQUIRE(!up1 || D == 2 || Z % 2 == 0, "upsampled input needs even Z");
    GF_REQUIRE(c1 > 0 && c1 % 16 == 0 && c2 >= 0 && c2 % 16 == 0, "channel counts must be multiples of 16");
    GF_REQUIRE(c1 + c2 <= WG_MAXC, "at most 256 input channels");
    GF_REQUIRE(!up1 || (H % 2 == 0 && W % 2 == 0), "upsampled input needs even H, W");
}


template <int D>
void conv_impl(const void* x1, int c1, int up1, const void* x2, int c2, int64_t P, int Z, int H, int W,
               const void* wpk, int coutp, const float* bias, int cout, int flags, void* y, int ysplit, void* y2,
               float* y32, const void* mask, void* stream)
{
    check_conv_shapes(P, Z, H, W, c1, up1, c2, D);
    GF_REQUIRE(x1 && wpk && (c2 == 0 || x2), "null input");
    GF_REQUIRE(aligned16(x1) && aligned16(x2) && aligned16(wpk), "inputs and weights must be 16-byte aligned");
    GF_REQUIRE(coutp == 32 || coutp == 64 || coutp == 128 || coutp == 256, "coutp must be 32, 64, 128 or 256");
    GF_REQUIRE(cout > 0 && cout <= coutp, "cout must be in (0, coutp]");
    GF_REQUIRE(y || y32, "no output");
    GF_REQUIRE(!y || (ysplit > 0 && ysplit <= cout && (ysplit == cout || y2)), "bad output split");
    hipStream_t s = S(stream);
    const unsigned grid = static_cast<unsigned>((P + 127) / 128);
    auto X1 = static_cast<const __bf16*>(x1);
    auto X2 = static_cast<const __bf16*>(x2);
    auto Wp = static_cast<const __bf16*>(wpk);
    auto Y = static_cast<__bf16*>(y);
    auto Y2 = static_cast<__bf16*>(y2);
    auto M = static_cast<const __bf16*>(mask);
    const int ys = y ? ysplit : cout;
    switch (coutp) {
    case 32:
        k_conv3x3<1, D><<<grid, 256, 0, s>>>(P, Z, H, W, X1, c1, up1, X2, c2, Wp, bias, cout, flags, Y, ys, Y2, y32, M);
        break;
    case 64:
        k_conv3x3<2, D><<<grid, 256, 0, s>>>(P, Z, H, W, X1, c1, up1, X2, c2, Wp, bias, cout, flags, Y, ys, Y2, y32, M);
        break;
    case 128:
        k_conv3x3<4, D><<<grid, 256, 0, s>>>(P, Z, H, W, X1, c1, up1, X2, c2, Wp, bias, cout, flags, Y, ys, Y2, y32, M);
        break;
    default:
        k_conv3x3<8, D><<<grid, 256, 0, s>>>(P, Z, H, W, X1, c1, up1, X2, c2, Wp, bias, cout, flags, Y, ys, Y2, y32, M);
    }
    check_launch();
}

void wgrad_work(int ntap, int64_t P, int cg, int cin, int* nsplit, int64_t* work_floats)
{
    GF_REQUIRE(P > 0 && cg > 0 && cin > 0 && nsplit && work_floats, "bad argument");
    const int64_t nchunk = (P + WG_PIX - 1) / WG_PIX;
    const int coutp = (cg + 31) / 32 * 32, cinp = (cin + 31) / 32 * 32;
    // about 2048 blocks over (tap, co tile, split), at least 4 chunks per split
    const int ns = static_cast<int>(
        std::min<int64_t>(std::max<int64_t>(nchunk / 4, 1), std::max(1, 2048 / (ntap * coutp / 32))));
    *nsplit = ns;
    *work_floats = static_cast<int64_t>(ns) * ntap * coutp * cinp + static_cast<int64_t>(ns) * cg;
}

template <int D>
void wgrad_impl(const void* x1, int c1, int up1, const void* x2, int c2, int64_t P, int Z, int H, int W, const void* g,
                int cg, int cout, int cin, float* work, int64_t work_floats, float* dw, float* db, void* stream)
{
    constexpr int NTAP = D == 3 ? 27 : 9;
    check_conv_shapes(P, Z, H, W, c1, up1, c2, D);
    GF_REQUIRE(x1 && g && work && dw && (c2 == 0 || x2), "null argument");
    GF_REQUIRE(aligned16(x1) && aligned16(x2) && aligned16(g), "inputs must be 16-byte aligned");
    GF_REQUIRE(cg > 0 && cg % 16 == 0 && cg <= 256, "g channels must be a multiple of 16, at most 256");
    GF_REQUIRE(cout > 0 && cout <= cg && cin > 0 && cin <= c1 + c2, "master shape exceeds the tensors");
    int ns = 0;
    int64_t need = 0;
    wgrad_work(NTAP, P, cg, c1 + c2, &ns, &need);
    GF_REQUIRE(work_floats >= need, "workspace too small (gf_conv*_wgrad_work)");
    const int coutp = (cg + 31) / 32 * 32, cinp = (c1 + c2 + 31) / 32 * 32;
    hipStream_t s = S(stream);
    const unsigned grid = static_cast<unsigned>(NTAP * (coutp / 32) * ns);
    k_conv3x3_wgrad<D><<<grid, 256, 0, s>>>(P, Z, H, W, static_cast<const __bf16*>(x1), c1, up1,
                                            static_cast<const __bf16*>(x2), c2, static_cast<const __bf16*>(g), cg, ns,
                                            work, coutp, cinp);
    const int64_t n = static_cast<int64_t>(NTAP) * cout * cin;
    k_wgrad_reduce<<<nblocks(n, 256), 256, 0, s>>>(NTAP, ns, coutp, cinp, cout, cin, work, dw);
    if (db) {
        float* bpart = work + static_cast<int64_t>(ns) * NTAP * coutp * cinp;
        k_bias_grad<<<ns, 256, 0, s>>>(P, static_cast<const __bf16*>(g), cg, cout, ns, bpart);
        k_reduce_rows<<<nblocks(cout, 256), 256, 0, s>>>(ns, cout, bpart, db);
    }
    check_launch();
}

template <int D>
void pool_impl(int kind, const void* in, int64_t N, int Z, int H, int W, int C, float w, int accum, const void* act,
               void* out, void* stream)
{
    GF_REQUIRE(in && out && N > 0 && Z > 0 && H % 2 == 0 && W % 2 == 0 && C > 0 && (D == 2 ? Z == 1 : Z % 2 == 0),
               "bad argument (even pooled dimensions)");
    hipStream_t s = S(stream);
    const int Zo = D == 3 ? Z / 2 : Z;
    const int64_t Pfine = N * Z * H * W, Pcoarse = N * Zo * (H / 2) * (W / 2);
    auto I = static_cast<const __bf16*>(in);
    auto A = static_cast<const __bf16*>(act);
    auto O = static_cast<__bf16*>(out);
    if (kind == 0)
        k_avgpool2<D><<<nblocks(Pcoarse * C, 256), 256, 0, s>>>(Pcoarse, Zo, H / 2, W / 2, C, I, O);
    else if (kind == 1)
        k_unpool2<D><<<nblocks(Pfine * C, 256), 256, 0, s>>>(Pfine, Z, H, W, C, w, I, accum, A, O);
    else
        k_sumpool2<D><<<nblocks(Pcoarse * C, 256), 256, 0, s>>>(Pcoarse, Zo, H / 2, W / 2, C, I, A, O);
    check_launch();
}

void pack_impl(int ntap, int cout, int cin, const float* w, int coutp, int cin_s, void* wf, int cinp, int cout_s,
               void* wb, void* stream)
{
    GF_REQUIRE(w && wf && cout <= coutp && cin <= cin_s, "bad argument");
    GF_REQUIRE(!wb || (cin <= cinp && cout <= cout_s), "bad transposed shape");
    const int64_t n = std::max(static_cast<int64_t>(ntap) * coutp * cin_s,
                               wb ? static_cast<int64_t>(ntap) * cinp * cout_s : int64_t(0));
    k_pack<<<nblocks(n, 256), 256, 0, S(stream)>>>(ntap, cout, cin, w, coutp, cin_s, static_cast<__bf16*>(wf), cinp,
                                                  cout_s, static_cast<__bf16*>(wb));
    check_launch();
}

} // namespace

extern "C" {

int gf_conv3x3(const void* x1, int c1, int up1, const void* x2, int c2, int64_t P, int H, int W, const void* wpk,
               int coutp, const float* bias, int cout, int flags, void* y, int ysplit, void* y2, float* y32,
               const void* mask, void* stream)
{
    return guard([&] {
        conv_impl<2>(x1, c1, up1, x2, c2, P, 1, H, W, wpk, coutp, bias, cout, flags, y, ysplit, y2, y32, mask, stream);
    });
}

int gf_conv3x3x3(const void* x1, int c1, int up1, const void* x2, int c2, int64_t P, int Z, int H, int W,
                 const void* wpk, int coutp, const float* bias, int cout, int flags, void* y, int ysplit, void* y2,
                 float* y32, const void* mask, void* stream)
{
    return guard([&] {
        conv_impl<3>(x1, c1, up1, x2, c2, P, Z, H, W, wpk, coutp, bias, cout, flags, y, ysplit, y2, y32, mask, stream);
    });
}

int gf_conv3x3_wgrad_work(int64_t P, int cg, int cin, int* nsplit, int64_t* work_floats)
{
    return guard([&] { wgrad_work(9, P, cg, cin, nsplit, work_floats); });
}

int gf_conv3x3x3_wgrad_work(int64_t P, int cg, int cin, int* nsplit, int64_t* work_floats)
{
    return guard([&] { wgrad_work(27, P, cg, cin, nsplit, work_floats); });
}

int gf_conv3x3_wgrad(const void* x1, int c1, int up1, const void* x2, int c2, int64_t P, int H, int W, const void* g,
                     int cg, int cout, int cin, float* work, int64_t work_floats, float* dw, float* db, void* stream)
{
    return guard([&] {
        wgrad_impl<2>(x1, c1, up1, x2, c2, P, 1, H, W, g, cg, cout, cin, work, work_floats, dw, db, stream);
    });
}

int gf_conv3x3x3_wgrad(const void* x1, int c1, int up1, const void* x2, int c2, int64_t P, int Z, int H, int W,
                       const void* g, int cg, int cout, int cin, float* work, int64_t work_floats, float* dw, float* db,
                       void* stream)
{
    return guard([&] {
        wgrad_impl<3>(x1, c1, up1, x2, c2, P, Z, H, W, g, cg, cout, cin, work, work_floats, dw, db, stream);
    });
}

int gf_avgpool2(const void* in, int64_t N, int H, int W, int C, void* out, void* stream)
{
    return guard([&] { pool_impl<2>(0, in, N, 1, H, W, C, 0.0f, 0, nullptr, out, stream); });
}

int gf_unpool2(const void* dout, int64_t N, int H, int W, int C, float w, int accum, const void* act, void* din,
               void* stream)
{
    return guard([&] { pool_impl<2>(1, dout, N, 1, H, W, C, w, accum, act, din, stream); });
}

int gf_sumpool2(const void* in, int64_t N, int H, int W, int C, const void* act, void* out, void* stream)
{
    return guard([&] { pool_impl<2>(2, in, N, 1, H, W, C, 0.0f, 0, act, out, stream); });
}

int gf_avgpool2x2x2(const void* in, int64_t N, int Z, int H, int W, int C, void* out, void* stream)
{
    return guard([&] { pool_impl<3>(0, in, N, Z, H, W, C, 0.0f, 0, nullptr, out, stream); });
}

int gf_unpool2x2x2(const void* dout, int64_t N, int Z, int H, int W, int C, float w, int accum, const void* act,
                   void* din, void* stream)
{
    return guard([&] { pool_impl<3>(1, dout, N, Z, H, W, C, w, accum, act, din, stream); });
}

int gf_sumpool2x2x2(const void* in, int64_t N, int Z, int H, int W, int C, const void* act, void* out, void* stream)
{
    return guard([&] { pool_impl<3>(2, in, N, Z, H, W, C, 0.0f, 0, act, out, stream); });
}

int gf_conv_pack(int cout, int cin, const float* w, int coutp, int cin_s, void* wf, int cinp, int cout_s, void* wb,
                 void* stream)
{
    return guard([&] { pack_impl(9, cout, cin, w, coutp, cin_s, wf, cinp, cout_s, wb, stream); });
}

int gf_conv_pack3(int cout, int cin, const float* w, int coutp, int cin_s, void* wf, int cinp, int cout_s, void* wb,
                  void* stream)
{
    return guard([&] { pack_impl(27, cout, cin, w, coutp, cin_s, wf, cinp, cout_s, wb, stream); });
}

int gf_adam(int64_t n, float* w, const float* g, float* m, float* v, float lr, float beta1, float beta2, float eps,
            int64_t step, void* stream)
{
    return guard([&] {
        GF_REQUIRE(n > 0 && w && g && m && v && step > 0, "bad argument");
        const float c1 = 1.0f - std::pow(beta1, static_cast<float>(step));
        const float c2 = 1.0f - std::pow(beta2, static_cast<float>(step));
        k_adam<<<nblocks(n, 256), 256, 0, S(stream)>>>(n, w, g, m, v, lr, beta1, beta2, eps, c1, c2);
        check_launch();
    });
}

int gf_mse(int64_t P, int nc, const float* pred, const float* target, float scale, int cg, void* g, float* part,
           float* loss_sum, void* stream)
{
    return guard([&] {
        GF_REQUIRE(P > 0 && nc > 0 && nc <= cg && pred && target && g && part && loss_sum, "bad argument");
        const int64_t nb = nblocks(P * cg, 256);
        hipStream_t s = S(stream);
        k_mse<<<nb, 256, 0, s>>>(P, nc, pred, target, scale, cg, static_cast<__bf16*>(g), part);
        k_sum_block<<<1, 256, 0, s>>>(nb, part, loss_sum);
        check_launch();
    });
}

namespace {
// floats of the pyramid (levels 1..S), the means and the loss partials
int64_t mse_multiscale_floats(int dim, int64_t B, int Z, int H, int W, int nc, int S, int cg)
{
    int64_t n = 0;
    for (int s = 1; s <= S; ++s) n += B * (dim == 3 ? (Z >> s) : 1) * (H >> s) * (W >> s) * nc;
    const int64_t P = B * Z * H * W;
    return n + B * nc + (P * cg + 255) / 256;
}
} // namespace

int gf_mse_multiscale_work(int dim, int64_t B, int Z, int H, int W, int nc, int S, int cg, int64_t* nfloats)
{
    return guard([&] {
        GF_REQUIRE(nfloats && (dim == 2 || dim == 3) && B > 0 && nc > 0 && cg >= nc && S >= 1 && S <= MS_MAX_LEVELS,
                   "bad argument");
        *nfloats = mse_multiscale_floats(dim, B, dim == 3 ? Z : 1, H, W, nc, S, cg);
    });
}

int gf_mse_multiscale(int dim, int64_t B, int Z, int H, int W, int nc, int levels, float lam, const float* pred,
                      const float* target, float scale, int cg, void* g, float* work, int64_t nwork, float* loss_sum,
                      void* stream)
{
    return guard([&] {
        GF_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
        if (dim == 2) Z = 1;
        GF_REQUIRE(B > 0 && Z > 0 && H > 0 && W > 0 && nc > 0 && cg >= nc && levels >= 1 && levels <= MS_MAX_LEVELS,
                   "bad argument");
        const int f = 1 << levels;
        GF_REQUIRE(H % f == 0 && W % f == 0 && (dim == 2 || Z % f == 0), "2^S must divide every grid dimension");
        GF_REQUIRE(pred && target && g && work && loss_sum, "null argument");
        GF_REQUIRE(nwork >= mse_multiscale_floats(dim, B, Z, H, W, nc, levels, cg), "workspace too small");
        hipStream_t s = S(stream);
        MsLevels L{};
        float* w = work;
        for (int k = 1; k <= levels; ++k) {
            const int Zc = dim == 3 ? Z >> k : 1, Hc = H >> k, Wc = W >> k;
            const int64_t ncoarse = B * Zc * Hc * Wc;
            const float* src = k == 1 ? pred : L.lv[k - 2];
            if (dim == 3)
                k_err_pool<3><<<nblocks(ncoarse * nc, 256), 256, 0, s>>>(ncoarse, Zc, Hc, Wc, nc, k == 1, src, target, w);
            else
                k_err_pool<2><<<nblocks(ncoarse * nc, 256), 256, 0, s>>>(ncoarse, Zc, Hc, Wc, nc, k == 1, src, target, w);
            L.lv[k - 1] = w;
            w += ncoarse * nc;
        }
        float* mu = w;
        w += B * nc;
        const int ncell = (dim == 3 ? Z >> levels : 1) * (H >> levels) * (W >> levels);
        k_err_mean<<<nblocks(B * nc, 64), 64, 0, s>>>(B, ncell, nc, L.lv[levels - 1], mu);
        const int64_t P = B * Z * H * W, nb = nblocks(P * cg, 256);
        if (dim == 3)
            k_mse_multiscale<3><<<nb, 256, 0, s>>>(P, Z, H, W, nc, levels, lam, pred, target, L, mu, scale, cg,
                                                   static_cast<__bf16*>(g), w);
        else
            k_mse_multiscale<2><<<nb, 256, 0, s>>>(P, Z, H, W, nc, levels, lam, pred, target, L, mu, scale, cg,
                                                   static_cast<__bf16*>(g), w);
        k_sum_block<<<1, 256, 0, s>>>(nb, w, loss_sum);
        check_launch();
    });
}

int gf_grid_to_dofs(int64_t n, int nc, const float* grid, const int32_t* idx, const double* affine, double scale,
                    double gamma, double* xu, void* stream)
{
    return guard([&] {
        GF_REQUIRE(n > 0 && nc > 0 && grid && idx && affine && xu, "bad argument");
        k_grid_to_dofs<<<nblocks(n, 256), 256, 0, S(stream)>>>(n, nc, grid, idx, affine, scale, gamma, xu);
        check_launch();
    });
}

int gf_dofs_to_grid(int64_t n, int nc, const double* xu, const int32_t* idx, const double* affine, double scale,
                    double gamma, int center, float* grid, void* stream)
{
    return guard([&] {
        GF_REQUIRE(n > 0 && nc > 0 && n % nc == 0 && xu && idx && affine && grid && gamma != 0.0, "bad argument");
        hipStream_t s = S(stream);
        k_dofs_to_grid<<<nblocks(n, 256), 256, 0, s>>>(n, xu, idx, affine, scale, gamma, grid);
        if (center) k_center<<<1, 256, 0, s>>>(n / nc, nc, grid);
        check_launch();
    });
}

int gf_input_logG(int64_t P, const double* G, double shift, double scale, int cs, void* x, void* stream)
{
    return guard([&] {
        GF_REQUIRE(P > 0 && G && x && cs > 0, "bad argument");
        k_input_logG<<<nblocks(P * cs, 256), 256, 0, S(stream)>>>(P, G, shift, scale, cs, static_cast<__bf16*>(x));
        check_launch();
    });
}

} // extern "C"
