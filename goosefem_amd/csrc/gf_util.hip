// Streaming copy: the measured HBM ceiling that the bandwidth-bound kernels' roofline fractions are
// reported against next to the datasheet peak.  No counterpart in the reference (measurement
// utility).  Each thread moves 4 x 16 bytes per grid-stride step with full-wave contiguous 16-byte
// loads and stores; mode 1 uses non-temporal (streaming) loads and stores.
#include "gf_internal.hpp"

using namespace gf;

namespace {

typedef double dvec2 __attribute__((ext_vector_type(2)));

template <bool NT>
__global__ void __launch_bounds__(256) k_copy_stream(int64_t n2, const dvec2* __restrict__ src, dvec2* __restrict__ dst)
{
    constexpr int U = 4;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * 256;
    int64_t i = blockIdx.x * static_cast<int64_t>(256) + threadIdx.x;
    for (; i + (U - 1) * stride < n2; i += U * stride) {
        dvec2 v[U];
#pragma unroll
        for (int k = 0; k < U; ++k) v[k] = NT ? __builtin_nontemporal_load(src + i + k * stride) : src[i + k * stride];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            if (NT) __builtin_nontemporal_store(v[k], dst + i + k * stride);
            else dst[i + k * stride] = v[k];
        }
    }
    for (; i < n2; i += stride) {
        const dvec2 v = NT ? __builtin_nontemporal_load(src + i) : src[i];
        if (NT) __builtin_nontemporal_store(v, dst + i);
        else dst[i] = v;
    }
}

} // namespace

extern "C" {

int gf_copy_stream(const double* src, double* dst, int64_t n, int mode, void* stream)
{
    return guard([&] {
        GF_REQUIRE(src && dst && n >= 0 && n % 2 == 0, "src, dst non-null; n >= 0 and even");
        GF_REQUIRE(aligned16(src) && aligned16(dst), "16-byte aligned buffers");
        GF_REQUIRE(mode == 0 || mode == 1, "mode 0 (plain) or 1 (non-temporal)");
        if (n == 0) return;
        const int64_t n2 = n / 2;
        // 8 workgroups per CU of 256 CUs, grid-stride over the rest
        const unsigned grid = static_cast<unsigned>(std::min<int64_t>(nblocks(n2, 256), 2048));
        if (mode) k_copy_stream<true><<<grid, 256, 0, S(stream)>>>(n2, reinterpret_cast<const dvec2*>(src),
                                                                   reinterpret_cast<dvec2*>(dst));
        else k_copy_stream<false><<<grid, 256, 0, S(stream)>>>(n2, reinterpret_cast<const dvec2*>(src),
                                                               reinterpret_cast<dvec2*>(dst));
        check_launch();
    });
}

} // extern "C"
