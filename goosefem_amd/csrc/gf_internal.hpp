// Internal helpers shared by the goosefem_amd translation units: error reporting, HIP checks,
// device buffers, launch geometry.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/goosefem_amd.h"

namespace gf {

void set_error(const std::string& msg);

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define GF_REQUIRE(cond, msg)                                                                       \
    do {                                                                                            \
        if (!(cond)) throw ::gf::Error(GF_EINVAL, std::string(__func__) + ": " + (msg));            \
    } while (0)

#define GF_HIP(call)                                                                                \
    do {                                                                                            \
        hipError_t _e = (call);                                                                     \
        if (_e != hipSuccess)                                                                       \
            throw ::gf::Error(GF_EHIP, std::string(#call) + ": " + hipGetErrorString(_e));          \
    } while (0)

// Runs `body`, converting exceptions into a status code + thread-local message.
template <class F>
inline int guard(F&& body)
{
    try {
        body();
        return GF_OK;
    }
    catch (const Error& e) {
        set_error(e.what());
        return e.code;
    }
    catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return GF_ENOMEM;
    }
    catch (const std::exception& e) {
        set_error(e.what());
        return GF_EINVAL;
    }
}

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Owning device buffer.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept
    {
        release();
        p = o.p;
        n = o.n;
        o.p = nullptr;
        o.n = 0;
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count)
    {
        release();
        n = count;
        if (count == 0) return;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            n = 0;
            throw Error(GF_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        }
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T* h, size_t count)
    {
        if (count != n) alloc(count);
        if (count) GF_HIP(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
    }
    void upload(const std::vector<T>& h) { upload(h.data(), h.size()); }
};

inline unsigned nblocks(int64_t n, int bs) { return static_cast<unsigned>((n + bs - 1) / bs); }

inline void check_launch() { GF_HIP(hipGetLastError()); }

inline bool aligned16(const void* p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; }

// Ordered gather sum: src[idx[j]] over j in [b, e), added in ascending j (the reference's serial
// scatter order) in double.  The loads of NB entries are issued together -- NB indices, clamped to
// the list so that every address is in bounds, then NB values -- and only then added, the valid ones
// in order; a loop that loads one entry per iteration waits twice per entry.  Same bits as that loop.
template <int NB, class T>
__device__ inline double ordered_gather_sum(int32_t b, int32_t e, const int32_t* __restrict__ idx,
                                            const T* __restrict__ src)
{
    double s = 0.0;
    for (int32_t j = b; j < e; j += NB) {
        int32_t k[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) k[q] = idx[min(j + q, e - 1)];
        T v[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) v[q] = src[k[q]];
#pragma unroll
        for (int q = 0; q < NB; ++q)
            if (j + q < e) s += static_cast<double>(v[q]);
    }
    return s;
}

// Host-side parallel loop over [0, n) in contiguous chunks (set-up code only).
template <class F>
inline void parallel_for(int64_t n, F&& f)
{
    int nt = static_cast<int>(std::min<int64_t>(std::max(1u, std::thread::hardware_concurrency()), 16));
    if (n < 4096 || nt == 1) {
        for (int64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    int64_t chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        int64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([b, e, &f] {
            for (int64_t i = b; i < e; ++i) f(i);
        });
    }
    for (auto& t : th) t.join();
}

// Hex8 stiffness on MFMA (BF16 hi/lo split, FP32 accumulation), gf_stiffness.hip
int stiffness_bf16x3(gf_quad q, const double* qtensor4, double* elemmat, hipStream_t s);
int stiffness_bf16x3_f32(gf_quad q, const float* qtensor4, float* elemmat, hipStream_t s);
int stiffness_f64_hex8(gf_quad q, const double* qtensor4, double* elemmat, hipStream_t s);

// Narrow an int64 host index array to int32 (device form), checking the range.
std::vector<int32_t> narrow(const int64_t* a, size_t n, const char* what);


// Pinned host copies of a solver's device convergence state, one per chunk parity, with the events
// that mark their arrival: the host waits for chunk k - 1's flag while chunk k is already queued, so
// the GPU is not left idle between chunks (a chunk launched after convergence returns at once).
struct StatePipe {
    int32_t* h = nullptr; // [2][2]: done flag, iteration count
    hipEvent_t ev[2] = {nullptr, nullptr};
    StatePipe()
    {
        GF_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
        GF_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
        GF_HIP(hipHostMalloc(reinterpret_cast<void**>(&h), 4 * sizeof(int32_t), hipHostMallocDefault));
    }
    ~StatePipe()
    {
        if (h) (void)hipHostFree(h);
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    StatePipe(const StatePipe&) = delete;
    StatePipe& operator=(const StatePipe&) = delete;
    void post(const int32_t* dstate, int k, hipStream_t s)
    {
        GF_HIP(hipMemcpyAsync(h + 2 * (k & 1), dstate, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        GF_HIP(hipEventRecord(ev[k & 1], s));
    }
    const int32_t* wait(int k)
    {
        GF_HIP(hipEventSynchronize(ev[k & 1]));
        return h + 2 * (k & 1);
    }
};

} // namespace gf

// Opaque handle definitions shared across translation units.
struct gf_vector_s {
    int64_t nelem = 0, nnode = 0, ndof = 0;
    int nne = 0, ndim = 0;
    gf::DBuf<int32_t> conn, dofs;
    // DOF-centric gathers, entries ascending = the reference's serial loop order:
    //   elemvec entries k = (e*nne + m)*ndim + i with dofs(conn(e,m), i) == d
    gf::DBuf<int32_t> dof_eptr, dof_eidx;
    //   nodevec entries j = n*ndim + i with dofs(n, i) == d
    gf::DBuf<int32_t> dof_nptr, dof_nidx;
    //   node-centric: elemvec entries (e*nne + m) with conn(e,m) == n (ascending)
    gf::DBuf<int32_t> node_eptr, node_eidx;
    // host copies (int64) for set-up code that builds derived structures
    std::vector<int64_t> conn_h, dofs_h;
};
