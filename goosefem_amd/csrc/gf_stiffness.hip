// Hex8 element stiffness Int_gradN_dot_tensor4_dot_gradNT_dV (include/GooseFEM/Element.h:1131-1176):
//   K(m*3+j, n*3+k) = sum_q sum_{i,l} dNx(q,m,i) C_q(i,j,k,l) dNx(q,n,l) dV_q
//
// GF_BF16X3 (matrix cores).  Write kk = (q, j2, i) (72 values) and
//   T[kk][(n,k)] = dV_q sum_l C_q(i, j2, k, l) dNx(q,n,l)                      (72 x 24, VALU, FP32)
//   A[(m,j)][kk] = delta(j, j2) dNx(q,m,i)                                    (24 x 72)
// so that K = A T: one 24 x 72 by 72 x 24 product per element, evaluated on
// v_mfma_f32_32x32x16_bf16 (tile padded to 32 x 32, contraction to 80).  Both operands are split
// x = hi + lo with hi = bf16(x), lo = bf16(x - hi); the product accumulates hi*hi + hi*lo + lo*hi
// in FP32 (the dropped lo*lo term and the bf16 rounding of lo are ~2^-16 relative).
// One wavefront owns one element at a time: C_q (8 x 81), dNx and dV are staged in LDS with
// coalesced 16-byte loads (the next element's loads are issued before the current one is computed);
// each MFMA lane evaluates the T entries it feeds, the 15 MFMAs accumulate the 32 x 32 tile, and the
// 24 x 24 result is stored straight from the accumulators as contiguous FP64 rows.
//
// GF_F64: the same per-entry accumulation order as the reference loop (q, i, l for each (j, k)),
// with one wavefront per element reading C and dNx from LDS (each thread owns one (m, n) block).
#include "gf_internal.hpp"

using namespace gf;

namespace gf {
const double* quad_dNx(gf_quad q);
const double* quad_vol(gf_quad q);
int64_t quad_nelem(gf_quad q);
int quad_nip(gf_quad q);
int quad_etype(gf_quad q);
} // namespace gf

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int WAVES = 4;           // wavefronts per workgroup
constexpr int NQ = 8, NN = 8;      // quadrature points, nodes
constexpr int NC = NQ * 81;        // tensor4 doubles per element
constexpr int NG = NQ * NN * 3;    // dNx doubles per element
constexpr int KKP = 80;            // contraction length 72, padded to 5 x 16

struct alignas(16) WaveLDS {
    float C[NC];
    float g[NG];
    float vol[NQ];
};

// hi/lo split of eight values, converted in pairs (one packed conversion per pair and part)
__device__ inline void split8(const float (&x)[8], bf16x8& hi, bf16x8& lo)
{
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const floatx2 v = {x[2 * p], x[2 * p + 1]};
        const bf16x2 h = __builtin_convertvector(v, bf16x2);
        const floatx2 r = v - __builtin_convertvector(h, floatx2);
        const bf16x2 l = __builtin_convertvector(r, bf16x2);
        hi[2 * p] = h[0];
        hi[2 * p + 1] = h[1];
        lo[2 * p] = l[0];
        lo[2 * p + 1] = l[1];
    }
}

// one element's C, dNx, dV in registers (16-byte loads), so that the next element's loads are in
// flight while the current one is computed.  C is FP64 (the reference's qtensor4) or FP32.
// Loads are unconditional (lanes past the end of an array re-read its last entry, which
// store_stage ignores), so that no branch hides them from the compiler's wait counting.
constexpr int NG2 = NG / 2, RG = (NG2 + 63) / 64;

template <class TC>
struct StageC;

template <>
struct StageC<double> {
    static constexpr int N = NC / 2, R = (N + 63) / 64; // double2 loads
    double2 v[R];
    __device__ void load(const double* __restrict__ C, int64_t e, int lane)
    {
        const double2* p = reinterpret_cast<const double2*>(C + e * NC);
#pragma unroll
        for (int k = 0; k < R; ++k) v[k] = p[min(lane + 64 * k, N - 1)];
    }
    __device__ void store(float* dst, int lane) const
    {
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int t = lane + 64 * k;
            if (t < N) {
                dst[2 * t] = static_cast<float>(v[k].x);
                dst[2 * t + 1] = static_cast<float>(v[k].y);
            }
        }
    }
};

template <>
struct StageC<float> {
    static constexpr int N = NC / 4, R = (N + 63) / 64; // 16-byte loads of 4 floats
    floatx4 v[R];
    __device__ void load(const float* __restrict__ C, int64_t e, int lane)
    {
        const floatx4* p = reinterpret_cast<const floatx4*>(C + e * NC);
#pragma unroll
        for (int k = 0; k < R; ++k) v[k] = p[min(lane + 64 * k, N - 1)];
    }
    __device__ void store(float* dst, int lane) const
    {
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int t = lane + 64 * k;
            if (t < N) *reinterpret_cast<floatx4*>(dst + 4 * t) = v[k];
        }
    }
};

template <class TC>
struct Stage {
    StageC<TC> c;
    double2 g[RG];
    double v;
};

template <class TC>
__device__ inline void load_stage(Stage<TC>& st, const TC* __restrict__ C, const double* __restrict__ dNx,
                                  const double* __restrict__ vol, int64_t e, int lane)
{
    st.c.load(C, e, lane);
    const double2* ge = reinterpret_cast<const double2*>(dNx + e * NG);
#pragma unroll
    for (int k = 0; k < RG; ++k) st.g[k] = ge[min(lane + 64 * k, NG2 - 1)];
    st.v = vol[e * NQ + min(lane, NQ - 1)];
}

template <class TC>
__device__ inline void store_stage(const Stage<TC>& st, WaveLDS& L, int lane)
{
    st.c.store(L.C, lane);
#pragma unroll
    for (int k = 0; k < RG; ++k) {
        const int t = lane + 64 * k;
        if (t < NG2) {
            L.g[2 * t] = static_cast<float>(st.g[k].x);
            L.g[2 * t + 1] = static_cast<float>(st.g[k].y);
        }
    }
    if (lane < NQ) L.vol[lane] = static_cast<float>(st.v);
}

// The 32 x 32 FP32 tile whose top-left 24 x 24 block is K_e of the element staged in L.
__device__ __forceinline__ floatx16 element_tile(const WaveLDS& L, int lane)
{
    // MFMA: lane l supplies A row (l % 32) and B column (l % 32) at contraction slots
    // 8 (l / 32) + s of each 16-wide chunk.  The contraction index kk = (q, j2, i) is laid out so
    // that each half-wave h owns quadrature points 4h..4h+3: slot s of chunk ch is
    // t = 8 ch + s, q = 4h + t / 9, (j2, i) = t % 9 (t < 36; the last 4 slots are zero).  Every
    // index below is then a compile-time constant plus a per-lane base.  Row (m, j) of A and
    // column (n, k) of T share the decomposition 3 m + j = 3 n + k, so both operands use
    // dNx(q, m, :):  A = delta(j, j2) dNx(q, m, i),  T = sum_l C_q(i, j2, j, l) dV_q dNx(q, m, l).
    const int row = lane % 32, h = lane / 32;
    const int j = row % 3, m = row < 24 ? row / 3 : 7;
    const bool live = row < 24; // rows 24..31 of the 32 x 32 tile are padding
    const float* Cb = L.C + h * 4 * 81 + j * 3;
    const float* gb = L.g + (h * 4 * NN + m) * 3;
    const float* vb = L.vol + h * 4;
    floatx16 acc = {};
    float gq[3] = {0.0f, 0.0f, 0.0f}, gv[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ch = 0; ch < KKP / 16; ++ch) {
        float av[8], bv[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int t = ch * 8 + s;
            av[s] = 0.0f;
            bv[s] = 0.0f;
            if (t < 36) {
                const int ql = t / 9, r = t % 9, j2 = r / 3, i = r % 3;
                if (r == 0) {
                    const float v = vb[ql];
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        gq[c] = live ? gb[ql * 24 + c] : 0.0f;
                        gv[c] = gq[c] * v;
                    }
                }
                const float* Cq = Cb + ql * 81 + (i * 3 + j2) * 9;
                bv[s] = fmaf(Cq[0], gv[0], fmaf(Cq[1], gv[1], Cq[2] * gv[2]));
                av[s] = j == j2 ? gq[i] : 0.0f;
            }
        }
        bf16x8 ahi, alo, bhi, blo;
        split8(av, ahi, alo);
        split8(bv, bhi, blo);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bhi, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, blo, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, bhi, acc, 0, 0, 0);
    }
    return acc;
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Store K_e: the tile's 24 x 24 block goes through LDS (into the staged C, which is no longer read)
// so that each lane writes 9 doubles with full-wave contiguous stores and no lane is masked off.
// Accumulator layout: value r of lane l is K[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
template <class TK>
__device__ __forceinline__ void store_tile(WaveLDS& L, const floatx16& acc, TK* __restrict__ Ke, int lane)
{
    const int row = lane % 32, h = lane / 32;
    float* Ks = L.C;
    wave_sync();
    if (row < 24) {
#pragma unroll
        for (int r = 0; r < 12; ++r) Ks[(8 * (r / 4) + 4 * h + (r % 4)) * 24 + row] = acc[r];
    }
    wave_sync();
#pragma unroll
    for (int k = 0; k < 9; ++k) Ke[lane + 64 * k] = static_cast<TK>(Ks[lane + 64 * k]);
}

// One wavefront per element, grid-stride.  The loop is rotated (store K_e of element e, then
// wait for element e + stride's loads, which were issued before those stores) so that waiting for
// the next element's data never waits for the previous element's stores.
template <class TC, class TK>
__global__ void __launch_bounds__(64 * WAVES) k_stiffness_bf16x3(int64_t nelem, const TC* __restrict__ C,
                                                                  const double* __restrict__ dNx,
                                                                  const double* __restrict__ vol,
                                                                  TK* __restrict__ elemmat)
{
    __shared__ WaveLDS lds[WAVES];
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    WaveLDS& L = lds[w];
    const int64_t stride = static_cast<int64_t>(gridDim.x) * WAVES;
    int64_t e = static_cast<int64_t>(blockIdx.x) * WAVES + w;
    if (e >= nelem) return; // uniform per wave; the kernel has no workgroup barrier
    Stage<TC> st;
    load_stage(st, C, dNx, vol, e, lane);
    store_stage(st, L, lane);
    wave_sync();
    load_stage(st, C, dNx, vol, min(e + stride, nelem - 1), lane);
    floatx16 acc = element_tile(L, lane);
    for (;;) {
        store_tile(L, acc, elemmat + e * 576, lane);
        e += stride;
        if (e >= nelem) break;
        wave_sync();
        store_stage(st, L, lane);
        wave_sync();
        load_stage(st, C, dNx, vol, min(e + stride, nelem - 1), lane);
        acc = element_tile(L, lane);
    }
}

// FP64 in reference order, one wavefront per element, C and dNx from LDS.  STREAM: the element
// matrix goes through the wave's C area in LDS and leaves as one contiguous 16-byte stream
// (elemmat 16-byte aligned) instead of nine strided scalar stores per lane.
template <bool STREAM>
__global__ void __launch_bounds__(64 * WAVES) k_stiffness_f64(int64_t nelem, const double* __restrict__ C,
                                                               const double* __restrict__ dNx,
                                                               const double* __restrict__ vol,
                                                               double* __restrict__ elemmat)
{
    static_assert(NC >= 576, "the element matrix must fit the C area");
    __shared__ __align__(16) double sC[WAVES][NC];
    __shared__ double sg[WAVES][NG];
    __shared__ double sv[WAVES][NQ];
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int m = lane / 8, n = lane % 8;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * WAVES;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * WAVES + w; e < nelem; e += stride) {
        for (int t = lane; t < NC; t += 64) sC[w][t] = C[e * NC + t];
        for (int t = lane; t < NG; t += 64) sg[w][t] = dNx[e * NG + t];
        if (lane < NQ) sv[w][lane] = vol[e * NQ + lane];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double acc[3][3];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[j][k] = 0.0;
        for (int q = 0; q < NQ; ++q) {
            const double* g = sg[w] + q * 24;
            const double* Cq = sC[w] + q * 81;
            const double v = sv[w][q];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
#pragma unroll
                    for (int k = 0; k < 3; ++k)
#pragma unroll
                        for (int l = 0; l < 3; ++l)
                            acc[j][k] += g[m * 3 + i] * Cq[((i * 3 + j) * 3 + k) * 3 + l] * g[n * 3 + l] * v;
        }
        double* Ke = elemmat + e * 576;
        if constexpr (STREAM) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); // every lane is done reading C
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int k = 0; k < 3; ++k) sC[w][(m * 3 + j) * 24 + n * 3 + k] = acc[j][k];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            typedef double d2 __attribute__((ext_vector_type(2)));
            for (int t = lane; t < 288; t += 64)
                __builtin_nontemporal_store(reinterpret_cast<const d2*>(sC[w])[t], reinterpret_cast<d2*>(Ke) + t);
        }
        else {
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int k = 0; k < 3; ++k) Ke[(m * 3 + j) * 24 + n * 3 + k] = acc[j][k];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// out[e] = Ke[e] ue[e]  (batched element GEMV), one thread per element row; an FP32 Ke is widened
// to FP64 before each product
template <class TE>
__global__ void k_elemmat_dot(int64_t nrows, int K, const TE* __restrict__ Ke, const double* __restrict__ ue,
                              double* __restrict__ out)
{
    int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= nrows) return;
    const int64_t e = t / K;
    const TE* row = Ke + t * K;
    const double* x = ue + e * K;
    double s = 0.0;
    for (int c = 0; c < K; ++c) s += static_cast<double>(row[c]) * x[c];
    out[t] = s;
}

// the same product for K = 8 or 24: each wave streams 64 rows (64 K contiguous entries), issuing
// all of its 16-byte loads before using any, multiplies by the element vector (FP64 products),
// stages the products in LDS and sums each row in column order (the order of k_elemmat_dot, so the
// results are identical).  Ke and ue must be 16-byte aligned.
template <int K, class TE, int RW>
__global__ void __launch_bounds__(256) k_elemmat_dot_rows(int64_t nrows, const TE* __restrict__ Ke,
                                                         const double* __restrict__ ue, double* __restrict__ out)
{
    // RW rows per wave: their RW K contiguous entries are loaded by all 64 lanes, 16 bytes at a
    // time, the products go to LDS, and lane l < RW sums row l in column order
    constexpr int VW = 16 / static_cast<int>(sizeof(TE)); // entries per 16-byte load
    static_assert(K % VW == 0, "whole loads per row");
    static_assert((RW * K) % (VW * 64) == 0, "whole loads per wave");
    typedef TE vec __attribute__((ext_vector_type(VW)));
    constexpr int RP = K + 1; // LDS row pitch
    constexpr int NV = RW * K / (VW * 64);
    __shared__ double sh[4][RW * RP];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int64_t r0 = (static_cast<int64_t>(blockIdx.x) * 4 + w) * RW;
    if (r0 >= nrows) return; // uniform per wave
    const int64_t base = r0 * K, total = nrows * K;
    const vec* KV = reinterpret_cast<const vec*>(Ke + base);
    vec kv[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const int loc = VW * (k * 64 + l);
        kv[k] = base + loc < total ? KV[k * 64 + l] : vec{};
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const int loc = VW * (k * 64 + l); // a multiple of VW, so loc .. loc + VW - 1 lie in one row
        const int64_t f = base + loc;
        double p[VW];
#pragma unroll
        for (int v = 0; v < VW; ++v) p[v] = 0.0;
        if (f < total) {
            const int64_t e = f / (K * K);
            const int c = static_cast<int>(f % K);
#pragma unroll
            for (int v = 0; v < VW; ++v) p[v] = static_cast<double>(kv[k][v]) * ue[e * K + c + v];
        }
        double* d = sh[w] + (loc / K) * RP + loc % K;
#pragma unroll
        for (int v = 0; v < VW; ++v) d[v] = p[v];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int64_t t = r0 + l;
    if (l < RW && t < nrows) {
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < K; ++c) acc += sh[w][l * RP + c];
        out[t] = acc;
    }
}

unsigned grid_for(int64_t nelem)
{
    int64_t g = (nelem + WAVES - 1) / WAVES;
    return static_cast<unsigned>(std::min<int64_t>(g, 256 * 64));
}

} // namespace

namespace gf {

int stiffness_bf16x3(gf_quad q, const double* C, double* elemmat, hipStream_t s)
{
    return guard([&] {
        GF_REQUIRE(quad_etype(q) == GF_HEX8 && quad_nip(q) == NQ, "GF_BF16X3 needs Hex8 with 8 points");
        GF_REQUIRE(aligned16(C) && aligned16(elemmat),
                   "GF_BF16X3 needs 16-byte aligned qtensor4 and elemmat arrays");
        const int64_t ne = quad_nelem(q);
        if (ne == 0) return;
        k_stiffness_bf16x3<double, double><<<grid_for(ne), 64 * WAVES, 0, s>>>(ne, C, quad_dNx(q), quad_vol(q), elemmat);
        check_launch();
    });
}

int stiffness_bf16x3_f32(gf_quad q, const float* C, float* elemmat, hipStream_t s)
{
    return guard([&] {
        GF_REQUIRE(quad_etype(q) == GF_HEX8 && quad_nip(q) == NQ, "GF_BF16X3 needs Hex8 with 8 points");
        GF_REQUIRE(aligned16(C), "GF_BF16X3 needs a 16-byte aligned qtensor4 array");
        const int64_t ne = quad_nelem(q);
        if (ne == 0) return;
        k_stiffness_bf16x3<float, float><<<grid_for(ne), 64 * WAVES, 0, s>>>(ne, C, quad_dNx(q), quad_vol(q), elemmat);
        check_launch();
    });
}

int stiffness_f64_hex8(gf_quad q, const double* C, double* elemmat, hipStream_t s)
{
    return guard([&] {
        GF_REQUIRE(quad_etype(q) == GF_HEX8 && quad_nip(q) == NQ, "Hex8 with 8 points");
        const int64_t ne = quad_nelem(q);
        if (ne == 0) return;
        if (aligned16(elemmat))
            k_stiffness_f64<true><<<grid_for(ne), 64 * WAVES, 0, s>>>(ne, C, quad_dNx(q), quad_vol(q), elemmat);
        else
            k_stiffness_f64<false><<<grid_for(ne), 64 * WAVES, 0, s>>>(ne, C, quad_dNx(q), quad_vol(q), elemmat);
        check_launch();
    });
}

} // namespace gf

namespace {

template <class TE>
int elemmat_dot(int64_t nelem, int K, const TE* elemmat, const double* elemvec, double* out, void* stream)
{
    return guard([&] {
        GF_REQUIRE(elemmat && elemvec && out && K > 0, "invalid argument");
        const int64_t n = nelem * K;
        if (n == 0) return;
        constexpr int RW = 32;                                          // rows per wave
        const unsigned gw = static_cast<unsigned>((n + 4 * RW - 1) / (4 * RW)); // 4 waves per block
        const bool vec = aligned16(elemmat) && aligned16(elemvec);
        if (K == 24 && vec)
            k_elemmat_dot_rows<24, TE, RW><<<gw, 256, 0, S(stream)>>>(n, elemmat, elemvec, out);
        else if (K == 8 && vec)
            k_elemmat_dot_rows<8, TE, RW><<<gw, 256, 0, S(stream)>>>(n, elemmat, elemvec, out);
        else
            k_elemmat_dot<TE><<<nblocks(n, 256), 256, 0, S(stream)>>>(n, K, elemmat, elemvec, out);
        check_launch();
    });
}

} // namespace

extern "C" int gf_elemmat_dot(int64_t nelem, int K, const double* elemmat, const double* elemvec, double* out,
                              void* stream)
{
    return elemmat_dot(nelem, K, elemmat, elemvec, out, stream);
}

extern "C" int gf_elemmat_dot_f32(int64_t nelem, int K, const float* elemmat, const double* elemvec, double* out,
                                  void* stream)
{
    return elemmat_dot(nelem, K, elemmat, elemvec, out, stream);
}

extern "C" int gf_quad_int_gradN_dot_tensor4_dot_gradNT_dV_f32(gf_quad q, const float* qtensor4, float* elemmat,
                                                               void* stream)
{
    return guard([&] {
        GF_REQUIRE(q && qtensor4 && elemmat, "null argument");
        const int rc = gf::stiffness_bf16x3_f32(q, qtensor4, elemmat, S(stream));
        if (rc != GF_OK) throw Error(rc, gf_last_error());
    });
}
