// Wavefront and workgroup reductions for CDNA (64-lane wavefronts, GFX9 DPP).
//
// wave_sum adds the 64 lanes of a wavefront with DPP lane moves, no LDS: two quad permutations,
// the half-row and row mirrors (every lane of a 16-lane row then holds the row's sum), and the
// GFX9 row broadcasts 15 and 31, which carry the running row sums into rows 1 - 3, so that lane 63
// holds the total.  The additions run in the same order on every call, so sums are deterministic
// (run to run and across launch shapes that keep the same lane assignment).
//
// block_sum adds the per-wave totals of a workgroup of NTH threads in wave order through a small
// LDS array.  Both are the deterministic reductions of the PCG and partitioned-solver kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

namespace gf {

// DPP control codes (AMD GCN3 / CDNA ISA, DPP_CTRL field)
enum : int {
    DPP_QUAD_1032 = 0xB1,        // quad_perm [1, 0, 3, 2]
    DPP_QUAD_2301 = 0x4E,        // quad_perm [2, 3, 0, 1]
    DPP_ROW_MIRROR = 0x140,      // lane i <- lane 15 - i within a row of 16
    DPP_ROW_HALF_MIRROR = 0x141, // lane i <- lane 7 - i within a half row of 8
    DPP_ROW_BCAST15 = 0x142,     // lane 15 of each row -> every lane of the next row
    DPP_ROW_BCAST31 = 0x143,     // lane 31 -> every lane of rows 2 and 3
};

// v moved across lanes by one DPP pattern; lanes whose row is masked off receive 0.
template <int CTRL, int ROW_MASK = 0xF>
__device__ inline double dpp_move(double v)
{
    int32_t w[2];
    std::memcpy(w, &v, sizeof(v));
    w[0] = __builtin_amdgcn_update_dpp(0, w[0], CTRL, ROW_MASK, 0xF, false);
    w[1] = __builtin_amdgcn_update_dpp(0, w[1], CTRL, ROW_MASK, 0xF, false);
    double r;
    std::memcpy(&r, w, sizeof(r));
    return r;
}

// Sum over the wavefront; the result is valid in lane 63 (read it with wave_total).
__device__ inline double wave_sum_lane63(double v)
{
    v += dpp_move<DPP_QUAD_1032>(v);
    v += dpp_move<DPP_QUAD_2301>(v);
    v += dpp_move<DPP_ROW_HALF_MIRROR>(v);
    v += dpp_move<DPP_ROW_MIRROR>(v);
    v += dpp_move<DPP_ROW_BCAST15, 0xA>(v);
    v += dpp_move<DPP_ROW_BCAST31, 0xC>(v);
    return v;
}

// Sum over the wavefront, returned in every lane.
__device__ inline double wave_sum(double v)
{
    v = wave_sum_lane63(v);
    int32_t w[2];
    std::memcpy(w, &v, sizeof(v));
    w[0] = __builtin_amdgcn_readlane(w[0], 63);
    w[1] = __builtin_amdgcn_readlane(w[1], 63);
    double r;
    std::memcpy(&r, w, sizeof(r));
    return r;
}

// Sum over a workgroup of NTH threads (a multiple of 64), returned in every thread.  Every thread
// of the workgroup must call it (it synchronises).
template <int NTH>
__device__ inline double block_sum(double v)
{
    static_assert(NTH % 64 == 0 && NTH <= 1024, "NTH: a multiple of 64, at most 1024");
    constexpr int NW = NTH / 64;
    __shared__ double wsum[NW];
    v = wave_sum(v);
    if (NW == 1) return v;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wsum[wave] = v;
    __syncthreads();
    double t = wsum[0];
    for (int k = 1; k < NW; ++k) t += wsum[k];
    __syncthreads();
    return t;
}

// Three sums over a workgroup at once (one LDS pass, two barriers instead of six); each value is
// added in the same order as block_sum adds it.
template <int NTH>
__device__ inline void block_sum3(double& a, double& b, double& c)
{
    static_assert(NTH % 64 == 0 && NTH <= 1024, "NTH: a multiple of 64, at most 1024");
    constexpr int NW = NTH / 64;
    __shared__ double wsum[3][NW];
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    if (NW == 1) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        wsum[0][wave] = a;
        wsum[1][wave] = b;
        wsum[2][wave] = c;
    }
    __syncthreads();
    double ta = wsum[0][0], tb = wsum[1][0], tc = wsum[2][0];
    for (int k = 1; k < NW; ++k) {
        ta += wsum[0][k];
        tb += wsum[1][k];
        tc += wsum[2][k];
    }
    __syncthreads();
    a = ta;
    b = tb;
    c = tc;
}

} // namespace gf
