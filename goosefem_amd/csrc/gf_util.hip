// Streaming probes: the measured HBM ceilings that the bandwidth-bound kernels' roofline fractions
// are reported against next to the datasheet peak.  No counterpart in the reference (measurement
// utility).  Each workgroup streams one contiguous 16 KiB chunk per array with full-wave
// contiguous 16-byte accesses (4 per thread, 256 threads), one pass over the arrays.
//   mode 0  copy   c = a          (1 read : 1 write)
//   mode 1  copy   c = a          non-temporal loads and stores
//   mode 2  add    c = a + b      (2 reads : 1 write), non-temporal
//   mode 3  read   c[block] = sum of the block's chunk of a   (read only)
#include "gf_internal.hpp"

using namespace gf;

namespace {

typedef double dvec2 __attribute__((ext_vector_type(2)));
constexpr int TPB = 256, U = 4, CHUNK = TPB * U; // dvec2 per workgroup

template <bool NT>
__device__ inline dvec2 ld(const dvec2* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

template <bool NT>
__device__ inline void st(dvec2 v, dvec2* p)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <int MODE>
__global__ void __launch_bounds__(TPB) k_stream(int64_t n2, const dvec2* __restrict__ a, const dvec2* __restrict__ b,
                                                dvec2* __restrict__ c)
{
    constexpr bool NT = MODE != 0;
    const int64_t base = blockIdx.x * static_cast<int64_t>(CHUNK) + threadIdx.x;
    if (base + (U - 1) * TPB < n2) {
        dvec2 v[U];
#pragma unroll
        for (int k = 0; k < U; ++k) v[k] = ld<NT>(a + base + k * TPB);
        if constexpr (MODE == 2) {
            dvec2 w[U];
#pragma unroll
            for (int k = 0; k < U; ++k) w[k] = ld<NT>(b + base + k * TPB);
#pragma unroll
            for (int k = 0; k < U; ++k) v[k] += w[k];
        }
        if constexpr (MODE == 3) {
            dvec2 t = v[0] + v[1] + v[2] + v[3];
            if (t.x == 12345.678) c[blockIdx.x] = t; // keeps the loads; the probe's data never matches
        }
        else {
#pragma unroll
            for (int k = 0; k < U; ++k) st<NT>(v[k], c + base + k * TPB);
        }
        return;
    }
    for (int k = 0; k < U; ++k) {
        const int64_t i = base + k * TPB;
        if (i >= n2) break;
        dvec2 v = ld<NT>(a + i);
        if constexpr (MODE == 2) v += ld<NT>(b + i);
        if constexpr (MODE == 3) {
            if (v.x == 12345.678) c[blockIdx.x] = v;
        }
        else st<NT>(v, c + i);
    }
}

} // namespace

extern "C" {

int gf_stream_probe(int mode, const double* a, const double* b, double* c, int64_t n, void* stream)
{
    return guard([&] {
        GF_REQUIRE(mode >= 0 && mode <= 3, "mode 0 (copy), 1 (copy, non-temporal), 2 (add), 3 (read)");
        GF_REQUIRE(a && c && (mode != 2 || b) && n >= 0 && n % 2 == 0, "a, c (and b for add) non-null; n even");
        GF_REQUIRE(aligned16(a) && aligned16(c) && (mode != 2 || aligned16(b)), "16-byte aligned arrays");
        if (n == 0) return;
        const int64_t n2 = n / 2;
        const int64_t grid = (n2 + CHUNK - 1) / CHUNK;
        GF_REQUIRE(grid < (int64_t(1) << 31), "n too large");
        auto A = reinterpret_cast<const dvec2*>(a);
        auto B = reinterpret_cast<const dvec2*>(b);
        auto C = reinterpret_cast<dvec2*>(c);
        const unsigned g = static_cast<unsigned>(grid);
        switch (mode) {
            case 0: k_stream<0><<<g, TPB, 0, S(stream)>>>(n2, A, B, C); break;
            case 1: k_stream<1><<<g, TPB, 0, S(stream)>>>(n2, A, B, C); break;
            case 2: k_stream<2><<<g, TPB, 0, S(stream)>>>(n2, A, B, C); break;
            default: k_stream<3><<<g, TPB, 0, S(stream)>>>(n2, A, B, C); break;
        }
        check_launch();
    });
}

} // extern "C"
