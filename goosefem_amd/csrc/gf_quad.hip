// Element quadrature on the device: Element::Quad4::Quadrature (ElementQuad4.h),
// Element::Quad4::QuadraturePlanar (ElementQuad4Planar.h), Element::Hex8::Quadrature
// (ElementHex8.h), with the generic loops of Element.h:892-1229.
//
// FP64 kernels evaluate every expression in the reference's order with no FMA contraction (the
// library is built with -ffp-contract=off), so they reproduce the reference's FP64 bits.  FP32
// kernels run the per-quadrature-point work in single precision on FP32 copies of the geometry.
//
// Launch shape: one thread per (element, quadrature point) or per (element, node); consecutive
// threads touch consecutive elements, so the [nelem, ...] row-major streams are read and written
// in long contiguous runs.
#include <cmath>

#include <type_traits>

#include "gf_internal.hpp"

using namespace gf;

struct gf_quad_s {
    int et = 0, nne = 0, ndim = 0, tdim = 0, nip = 0;
    int64_t nelem = 0;
    double thick = 1.0;
    std::vector<double> xi, w, N, dNxi;
    DBuf<double> x, N_d, dNxi_d, w_d, dNx, vol;
    DBuf<float> dNx32, vol32;
    bool have32 = false;
};

namespace {

constexpr int BS = 256;

inline void shape_functions(int et, int nip, const double* xi, double* N, double* dNxi)
{
    // ElementQuad4.h:287-305, ElementQuad4Planar.h:89-109, ElementHex8.h:265-304
    if (et != GF_HEX8) {
        for (int q = 0; q < nip; ++q) {
            const double x0 = xi[q * 2 + 0], x1 = xi[q * 2 + 1];
            N[q * 4 + 0] = 0.25 * (1.0 - x0) * (1.0 - x1);
            N[q * 4 + 1] = 0.25 * (1.0 + x0) * (1.0 - x1);
            N[q * 4 + 2] = 0.25 * (1.0 + x0) * (1.0 + x1);
            N[q * 4 + 3] = 0.25 * (1.0 - x0) * (1.0 + x1);
            double* d = dNxi + q * 8;
            d[0] = -0.25 * (1.0 - x1);
            d[2] = +0.25 * (1.0 - x1);
            d[4] = +0.25 * (1.0 + x1);
            d[6] = -0.25 * (1.0 + x1);
            d[1] = -0.25 * (1.0 - x0);
            d[3] = -0.25 * (1.0 + x0);
            d[5] = +0.25 * (1.0 + x0);
            d[7] = +0.25 * (1.0 - x0);
        }
        return;
    }
    static const int s[8][3] = {{-1, -1, -1}, {+1, -1, -1}, {+1, +1, -1}, {-1, +1, -1},
                                {-1, -1, +1}, {+1, -1, +1}, {+1, +1, +1}, {-1, +1, +1}};
    for (int q = 0; q < nip; ++q) {
        const double* x = xi + q * 3;
        for (int m = 0; m < 8; ++m) {
            // 0.125 * (1 +- x0) * (1 +- x1) * (1 +- x2), factors in the reference's order
            double f0 = s[m][0] < 0 ? (1.0 - x[0]) : (1.0 + x[0]);
            double f1 = s[m][1] < 0 ? (1.0 - x[1]) : (1.0 + x[1]);
            double f2 = s[m][2] < 0 ? (1.0 - x[2]) : (1.0 + x[2]);
            N[q * 8 + m] = 0.125 * f0 * f1 * f2;
            dNxi[q * 24 + m * 3 + 0] = (s[m][0] < 0 ? -0.125 : +0.125) * f1 * f2;
            dNxi[q * 24 + m * 3 + 1] = (s[m][1] < 0 ? -0.125 : +0.125) * f0 * f2;
            dNxi[q * 24 + m * 3 + 2] = (s[m][2] < 0 ? -0.125 : +0.125) * f0 * f1;
        }
    }
}

inline void default_rule(int et, int rule, std::vector<double>& xi, std::vector<double>& w)
{
    const double a = rule == GF_NODAL ? 1.0 : 1.0 / std::sqrt(3.0);
    if (et != GF_HEX8) {
        const int s[4][2] = {{-1, -1}, {+1, -1}, {+1, +1}, {-1, +1}};
        xi.resize(8);
        w.assign(4, 1.0);
        for (int q = 0; q < 4; ++q)
            for (int d = 0; d < 2; ++d) xi[q * 2 + d] = s[q][d] < 0 ? -a : +a;
    }
    else {
        const int s[8][3] = {{-1, -1, -1}, {+1, -1, -1}, {+1, +1, -1}, {-1, +1, -1},
                             {-1, -1, +1}, {+1, -1, +1}, {+1, +1, +1}, {-1, +1, +1}};
        xi.resize(24);
        w.assign(8, 1.0);
        for (int q = 0; q < 8; ++q)
            for (int d = 0; d < 3; ++d) xi[q * 3 + d] = s[q][d] < 0 ? -a : +a;
    }
}

// ---- geometry: compute_dN ----------------------------------------------------------------------
// ElementQuad4.h:493-532, ElementQuad4Planar.h:280-319 (x thick), Element.h:1178-1229 (Hex8, with
// the chain rule over j)

__global__ void k_compute_dN_quad4(int64_t n, int nip, int planar, double thick, const double* __restrict__ x,
                                   const double* __restrict__ dNxi, const double* __restrict__ w,
                                   double* __restrict__ dNx, double* __restrict__ vol)
{
    int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n) return;
    int64_t e = t / nip;
    int q = static_cast<int>(t - e * nip);
    const double* xe = x + e * 8;
    const double* g = dNxi + q * 8;
    double J00 = g[0] * xe[0] + g[2] * xe[2] + g[4] * xe[4] + g[6] * xe[6];
    double J01 = g[0] * xe[1] + g[2] * xe[3] + g[4] * xe[5] + g[6] * xe[7];
    double J10 = g[1] * xe[0] + g[3] * xe[2] + g[5] * xe[4] + g[7] * xe[6];
    double J11 = g[1] * xe[1] + g[3] * xe[3] + g[5] * xe[5] + g[7] * xe[7];
    double det = J00 * J11 - J01 * J10; // detail.h:17-37
    double I00 = J11 / det, I01 = -1.0 * J01 / det, I10 = -1.0 * J10 / det, I11 = J00 / det;
    double* o = dNx + t * 8;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        o[m * 2 + 0] = I00 * g[m * 2 + 0] + I01 * g[m * 2 + 1];
        o[m * 2 + 1] = I10 * g[m * 2 + 0] + I11 * g[m * 2 + 1];
    }
    vol[t] = planar ? w[q] * det * thick : w[q] * det;
}

// dN/dx [8][3] and det J of one Hex8 point: J = dN/dxi^T x_e, its cofactor inverse (detail.h:40-70),
// dN/dx = J^-1 dN/dxi, in the generic compute_dN order (Element.h:1178-1229, with the Q1 chain rule
// corrected).  Shared by the set-up kernel and the gradient kernel that recomputes the geometry, so
// both produce the same bits.
__device__ inline void hex8_point_geometry(const double* __restrict__ xe, const double* __restrict__ g, double* o,
                                           double& det)
{
    double A[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) A[i][j] = 0.0;
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) A[i][j] += g[m * 3 + i] * xe[m * 3 + j];
    det = (A[0][0] * A[1][1] * A[2][2] + A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[1][0] * A[2][1]) -
          (A[0][2] * A[1][1] * A[2][0] + A[0][1] * A[1][0] * A[2][2] + A[0][0] * A[1][2] * A[2][1]);
    double B[3][3];
    B[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det;
    B[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
    B[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
    B[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det;
    B[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
    B[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
    B[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det;
    B[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
    B[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < 3; ++j) s += B[i][j] * g[m * 3 + j];
            o[m * 3 + i] = s;
        }
}

__global__ void k_compute_dN_hex8(int64_t n, int nip, const double* __restrict__ x, const double* __restrict__ dNxi,
                                  const double* __restrict__ w, double* __restrict__ dNx, double* __restrict__ vol)
{
    int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n) return;
    int64_t e = t / nip;
    int q = static_cast<int>(t - e * nip);
    double det;
    hex8_point_geometry(x + e * 24, dNxi + q * 24, dNx + t * 24, det);
    vol[t] = w[q] * det;
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b)
{
    int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t < n) b[t] = static_cast<float>(a[t]);
}

// ---- LDS staging of contiguous slices ---------------------------------------------------------------
// The quadrature kernels read and write per-point / per-element records (dN/dx rows, element
// vectors, tensors) of 8 to 72 bytes.  A workgroup moves its contiguous slice of each array between
// HBM and LDS with full-wave contiguous 16-byte accesses (when the array is 16-byte aligned; element
// by element otherwise) and the threads compute on LDS, so no global access is strided per lane.
template <class T, int NT, int KB>
__device__ inline void stage_in(T* __restrict__ lds, const T* __restrict__ g, int count)
{
    // KB: 16-byte loads in flight per thread before their LDS writes (the call site's typical count)
    constexpr int VW = 16 / static_cast<int>(sizeof(T));
    typedef T V __attribute__((ext_vector_type(VW)));
    int done = 0;
    if ((reinterpret_cast<uintptr_t>(g) & 15) == 0) {
        const int nv = count / VW;
        for (int base = 0; base < nv; base += KB * NT) {
            V r[KB];
            // unconditional loads (an index past the end re-reads the last chunk), so that all KB
            // are issued before the first wait
#pragma unroll
            for (int k = 0; k < KB; ++k) {
                const int i = base + k * NT + threadIdx.x;
                r[k] = reinterpret_cast<const V*>(g)[i < nv ? i : nv - 1];
            }
#pragma unroll
            for (int k = 0; k < KB; ++k) {
                const int i = base + k * NT + threadIdx.x;
                if (i < nv) reinterpret_cast<V*>(lds)[i] = r[k];
            }
        }
        done = nv * VW;
    }
    for (int i = done + threadIdx.x; i < count; i += NT) lds[i] = g[i];
}

template <class T, int NT>
__device__ inline void stage_out(T* __restrict__ g, const T* __restrict__ lds, int count)
{
    constexpr int VW = 16 / static_cast<int>(sizeof(T));
    typedef T V __attribute__((ext_vector_type(VW)));
    int done = 0;
    if ((reinterpret_cast<uintptr_t>(g) & 15) == 0) {
        const int nv = count / VW;
        for (int i = threadIdx.x; i < nv; i += NT)
            __builtin_nontemporal_store(reinterpret_cast<const V*>(lds)[i], reinterpret_cast<V*>(g) + i);
        done = nv * VW;
    }
    for (int i = done + threadIdx.x; i < count; i += NT) g[i] = lds[i];
}

// ---- gradients: mode 0 GradN_vector, 1 GradN_vector_T, 2 SymGradN_vector -------------------------

// GradN_vector (mode 0), GradN_vector_T (1), SymGradN_vector (2): 2-D node vectors (Quad4 tdim 2 /
// planar tdim 3, ElementQuad4.h:348-434, ElementQuad4Planar.h:129-221) and Hex8 (the generic loops of
// Element.h:924-1022), in the reference's expression order.  A workgroup of QB threads owns QB
// consecutive points (e, q); LDS holds their dN/dx rows [QB][nne nd], the element vectors of the
// elements they span, and the output tensors [QB][TD TD].
template <int ND>
inline size_t grad_lds_bytes(int qb, int nip, int td, size_t sz)
{
    const int nn = ND == 2 ? 8 : 24;
    return (static_cast<size_t>(qb) * nn + static_cast<size_t>(qb) * td * td + (qb / nip + 2) * static_cast<size_t>(nn)) * sz;
}

template <class T, int TD, int mode>
__global__ void __launch_bounds__(256) k_grad_2d_lds(int64_t n, int nip, const T* __restrict__ dNx,
                                                     const T* __restrict__ elemvec, T* __restrict__ qtensor)
{
    constexpr int QB = 256;
    extern __shared__ __align__(16) unsigned char smem[];
    T* s_o = reinterpret_cast<T*>(smem); // [QB][TD*TD]
    T* s_g = s_o + QB * TD * TD;           // [QB][8]
    T* s_u = s_g + QB * 8;           // [elements][8]
    const int64_t t0 = static_cast<int64_t>(blockIdx.x) * QB;
    const int nq = static_cast<int>(min<int64_t>(QB, n - t0));
    const int64_t e0 = t0 / nip, e1 = (t0 + nq - 1) / nip;
    stage_in<T, QB, 8 / (16 / sizeof(T))>(s_g, dNx + t0 * 8, nq * 8);
    stage_in<T, QB, 1>(s_u, elemvec + e0 * 8, static_cast<int>(e1 - e0 + 1) * 8);
    __syncthreads();
    if (threadIdx.x < nq) {
        const int64_t e = (t0 + threadIdx.x) / nip;
        const T* u = s_u + (e - e0) * 8;
        const T* g = s_g + threadIdx.x * 8;
        T a00 = g[0] * u[0] + g[2] * u[2] + g[4] * u[4] + g[6] * u[6];
        T a01 = g[0] * u[1] + g[2] * u[3] + g[4] * u[5] + g[6] * u[7];
        T a10 = g[1] * u[0] + g[3] * u[2] + g[5] * u[4] + g[7] * u[6];
        T a11 = g[1] * u[1] + g[3] * u[3] + g[5] * u[5] + g[7] * u[7];
        T r[TD * TD];
#pragma unroll
        for (int k = 0; k < TD * TD; ++k) r[k] = T(0);
        if constexpr (mode == 0) {
            r[0] = a00;
            r[1] = a01;
            r[TD] = a10;
            r[TD + 1] = a11;
        }
        else if constexpr (mode == 1) {
            r[0] = a00;
            r[TD] = a01;
            r[1] = a10;
            r[TD + 1] = a11;
        }
        else {
            r[0] = a00;
            r[TD + 1] = a11;
            r[1] = T(0.5) * (g[0] * u[1] + g[2] * u[3] + g[4] * u[5] + g[6] * u[7] + g[1] * u[0] + g[3] * u[2] +
                             g[5] * u[4] + g[7] * u[6]);
            r[TD] = r[1];
        }
        T* o = s_o + threadIdx.x * TD * TD;
#pragma unroll
        for (int k = 0; k < TD * TD; ++k) o[k] = r[k];
    }
    __syncthreads();
    stage_out<T, QB>(qtensor + t0 * TD * TD, s_o, nq * TD * TD);
}

template <class T, int mode>
__global__ void __launch_bounds__(128) k_grad_3d_lds(int64_t n, int nip, const T* __restrict__ dNx,
                                                     const T* __restrict__ elemvec, T* __restrict__ qtensor)
{
    constexpr int QB = 128;
    extern __shared__ __align__(16) unsigned char smem[];
    T* s_o = reinterpret_cast<T*>(smem); // [QB][9]
    T* s_g = s_o + QB * 9;           // [QB][24]
    T* s_u = s_g + QB * 24;           // [elements][24]
    const int64_t t0 = static_cast<int64_t>(blockIdx.x) * QB;
    const int nq = static_cast<int>(min<int64_t>(QB, n - t0));
    const int64_t e0 = t0 / nip, e1 = (t0 + nq - 1) / nip;
    stage_in<T, QB, 24 / (16 / sizeof(T))>(s_g, dNx + t0 * 24, nq * 24);
    stage_in<T, QB, 1>(s_u, elemvec + e0 * 24, static_cast<int>(e1 - e0 + 1) * 24);
    __syncthreads();
    if (threadIdx.x < nq) {
        const int64_t e = (t0 + threadIdx.x) / nip;
        const T* u = s_u + (e - e0) * 24;
        const T* g = s_g + threadIdx.x * 24;
        T r[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) r[k] = T(0);
#pragma unroll
        for (int m = 0; m < 8; ++m)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    if constexpr (mode == 0) r[i * 3 + j] += g[m * 3 + i] * u[m * 3 + j];
                    else if constexpr (mode == 1) r[j * 3 + i] += g[m * 3 + i] * u[m * 3 + j];
                    else {
                        T h = T(0.5) * g[m * 3 + i] * u[m * 3 + j];
                        r[i * 3 + j] += h;
                        r[j * 3 + i] += h;
                    }
                }
        T* o = s_o + threadIdx.x * 9;
#pragma unroll
        for (int k = 0; k < 9; ++k) o[k] = r[k];
    }
    __syncthreads();
    stage_out<T, QB>(qtensor + t0 * 9, s_o, nq * 9);
}

// Hex8 gradients with the geometry recomputed: the stored dN/dx of a Hex8 quadrature is 1536 bytes
// per element against 192 bytes of element coordinates, so this kernel reads the coordinates (kept by
// the quadrature) and forms each point's dN/dx with hex8_point_geometry, the function that produced the
// stored ones: the results are bit-identical to the stored-geometry form (FP32: the FP64 dN/dx rounded
// once, as the stored FP32 copy is).  A workgroup of QB threads owns QB consecutive points; LDS holds
// the coordinates and element vectors of the elements they span, the reference-element gradients and
// the output tensors.
constexpr int QB3 = 256;

inline size_t grad3x_lds_bytes(int nip, size_t sz)
{
    const size_t ne = QB3 / nip + 2;
    return static_cast<size_t>(QB3) * 9 * sz + ne * 24 * sz + ne * 24 * 8 + static_cast<size_t>(nip) * 24 * 8;
}

template <class T, int mode>
__global__ void __launch_bounds__(QB3) k_grad_3d_x(int64_t n, int nip, const double* __restrict__ x,
                                                   const double* __restrict__ dNxi, const T* __restrict__ elemvec,
                                                   T* __restrict__ qtensor)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t t0 = static_cast<int64_t>(blockIdx.x) * QB3;
    const int nq = static_cast<int>(min<int64_t>(QB3, n - t0));
    const int64_t e0 = t0 / nip, e1 = (t0 + nq - 1) / nip;
    const int ne = static_cast<int>(e1 - e0 + 1), nemax = QB3 / nip + 2;
    T* s_o = reinterpret_cast<T*>(smem);       // [QB3][9]
    T* s_u = s_o + QB3 * 9;                      // [nemax][24]
    double* s_x = reinterpret_cast<double*>(s_u + nemax * 24); // [nemax][24]
    double* s_d = s_x + nemax * 24;              // [nip][24]
    stage_in<T, QB3, 1>(s_u, elemvec + e0 * 24, ne * 24);
    stage_in<double, QB3, 1>(s_x, x + e0 * 24, ne * 24);
    for (int i = threadIdx.x; i < nip * 24; i += QB3) s_d[i] = dNxi[i];
    __syncthreads();
    if (threadIdx.x < nq) {
        const int64_t t = t0 + threadIdx.x, e = t / nip;
        const int q = static_cast<int>(t - e * nip);
        double gd[24], det;
        hex8_point_geometry(s_x + (e - e0) * 24, s_d + q * 24, gd, det);
        T g[24];
#pragma unroll
        for (int k = 0; k < 24; ++k) g[k] = static_cast<T>(gd[k]);
        const T* u = s_u + (e - e0) * 24;
        T r[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) r[k] = T(0);
#pragma unroll
        for (int m = 0; m < 8; ++m)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    if constexpr (mode == 0) r[i * 3 + j] += g[m * 3 + i] * u[m * 3 + j];
                    else if constexpr (mode == 1) r[j * 3 + i] += g[m * 3 + i] * u[m * 3 + j];
                    else {
                        T h = T(0.5) * g[m * 3 + i] * u[m * 3 + j];
                        r[i * 3 + j] += h;
                        r[j * 3 + i] += h;
                    }
                }
        T* o = s_o + threadIdx.x * 9;
#pragma unroll
        for (int k = 0; k < 9; ++k) o[k] = r[k];
    }
    __syncthreads();
    stage_out<T, QB3>(qtensor + t0 * 9, s_o, nq * 9);
}

// ---- interpolation / integration ----------------------------------------------------------------

// InterpQuad_vector: Quad4 override for n = 2 (ElementQuad4.h:324-346) has the same order as the
// generic loop (Element.h:892-922) when starting from 0
__global__ void k_interp(int64_t n, int nip, int nne, int nv, const double* __restrict__ N,
                         const double* __restrict__ elemvec, double* __restrict__ qvector)
{
    int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n) return;
    int64_t e = t / nip;
    int q = static_cast<int>(t - e * nip);
    const double* f = elemvec + e * nne * nv;
    for (int i = 0; i < nv; ++i) {
        double s = 0.0;
        for (int m = 0; m < nne; ++m) s += N[q * nne + m] * f[m * nv + i];
        qvector[t * nv + i] = s;
    }
}

// Int_N_vector_dV (Element.h:1024-1055): thread per (e, m)
__global__ void k_int_N_vector(int64_t n, int nip, int nne, int nv, const double* __restrict__ N,
                               const double* __restrict__ vol, const double* __restrict__ qvector,
                               double* __restrict__ elemvec)
{
    int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n) return;
    int64_t e = t / nne;
    int m = static_cast<int>(t - e * nne);
    for (int i = 0; i < nv; ++i) {
        double s = 0.0;
        for (int q = 0; q < nip; ++q) s += N[q * nne + m] * qvector[(e * nip + q) * nv + i] * vol[e * nip + q];
        elemvec[t * nv + i] = s;
    }
}

// Int_N_scalar_NT_dV (ElementQuad4.h:436-464, ElementQuad4Planar.h:223-251, ElementHex8.h:323-352):
// thread per (e, m, n) writes the nd x nd block (diagonal = accumulated scalar, else 0)
__global__ void k_int_N_scalar(int64_t n, int nip, int nne, int nd, const double* __restrict__ N,
                               const double* __restrict__ vol, const double* __restrict__ qscalar,
                               double* __restrict__ elemmat)
{
    int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n) return;
    int64_t e = t / (nne * nne);
    int mn = static_cast<int>(t - e * nne * nne);
    int m = mn / nne, nn = mn - m * nne;
    double s = 0.0;
    for (int q = 0; q < nip; ++q) s += N[q * nne + m] * qscalar[e * nip + q] * N[q * nne + nn] * vol[e * nip + q];
    int K = nne * nd;
    double* M = elemmat + e * K * K;
    for (int i = 0; i < nd; ++i)
        for (int j = 0; j < nd; ++j) M[(m * nd + i) * K + nn * nd + j] = (i == j) ? s : 0.0;
}

// Int_gradN_dot_tensor2_dV: thread per (e, m)
template <class T, int ND, int TD>
__global__ void k_int_tensor2(int64_t n, int nip, int nne, const T* __restrict__ dNx, const T* __restrict__ vol,
                              const T* __restrict__ qtensor, T* __restrict__ elemvec)
{
    int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n) return;
    int64_t e = t / nne;
    int m = static_cast<int>(t - e * nne);
    T f[ND];
#pragma unroll
    for (int j = 0; j < ND; ++j) f[j] = T(0);
    for (int q = 0; q < nip; ++q) {
        const T* g = dNx + ((e * nip + q) * nne + m) * ND;
        const T* s = qtensor + (e * nip + q) * TD * TD;
        T v = vol[e * nip + q];
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            T a = g[0] * s[0 * TD + j] + g[1] * s[1 * TD + j];
            if (ND == 3) a = a + g[2] * s[2 * TD + j];
            f[j] += a * v;
        }
    }
#pragma unroll
    for (int j = 0; j < ND; ++j) elemvec[t * ND + j] = f[j];
}

// Staged form of k_int_tensor2: a workgroup of EB NNE threads owns EB consecutive elements; LDS
// holds their dN/dx [EB nip NNE ND], dV [EB nip], tensors [EB nip TD TD] and the output element
// vectors [EB NNE ND].  Thread (e, m) sums over the points in the order of k_int_tensor2.
template <class T, int NNE, int ND, int TD, int EB>
__global__ void __launch_bounds__(EB* NNE) k_int_tensor2_lds(int64_t nelem, int nip, const T* __restrict__ dNx,
                                                             const T* __restrict__ vol, const T* __restrict__ qtensor,
                                                             T* __restrict__ elemvec)
{
    constexpr int NT = EB * NNE;
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t e0 = static_cast<int64_t>(blockIdx.x) * EB;
    const int ne = static_cast<int>(min<int64_t>(EB, nelem - e0));
    T* s_g = reinterpret_cast<T*>(smem); // EB nip NNE ND
    T* s_s = s_g + EB * nip * NNE * ND;   // EB nip TD TD
    T* s_o = s_s + EB * nip * TD * TD;    // EB NNE ND
    T* s_v = s_o + EB * NNE * ND;         // EB nip
    stage_in<T, NT, NNE * ND / (16 / sizeof(T))>(s_g, dNx + e0 * nip * NNE * ND, ne * nip * NNE * ND);
    stage_in<T, NT, (TD * TD + (16 / sizeof(T)) - 1) / (16 / sizeof(T))>(s_s, qtensor + e0 * nip * TD * TD, ne * nip * TD * TD);
    stage_in<T, NT, 1>(s_v, vol + e0 * nip, ne * nip);
    __syncthreads();
    const int el = threadIdx.x / NNE, m = threadIdx.x - el * NNE;
    if (el < ne) {
        T f[ND];
#pragma unroll
        for (int j = 0; j < ND; ++j) f[j] = T(0);
        for (int q = 0; q < nip; ++q) {
            const T* g = s_g + ((el * nip + q) * NNE + m) * ND;
            const T* sq = s_s + (el * nip + q) * TD * TD;
            T v = s_v[el * nip + q];
#pragma unroll
            for (int j = 0; j < ND; ++j) {
                T a = g[0] * sq[0 * TD + j] + g[1] * sq[1 * TD + j];
                if (ND == 3) a = a + g[2] * sq[2 * TD + j];
                f[j] += a * v;
            }
        }
#pragma unroll
        for (int j = 0; j < ND; ++j) s_o[(el * NNE + m) * ND + j] = f[j];
    }
    __syncthreads();
    stage_out<T, NT>(elemvec + e0 * NNE * ND, s_o, ne * NNE * ND);
}

template <class T, int NNE, int ND, int TD, int EB>
inline size_t int2_lds_bytes(int nip)
{
    return static_cast<size_t>(EB) * (nip * NNE * ND + nip * TD * TD + NNE * ND + nip) * sizeof(T);
}

// Int_gradN_dot_tensor4_dot_gradNT_dV FP64 (Element.h:1131-1176): thread per (e, m, n) block
template <int NNE, int ND, int TD>
__global__ void k_int_tensor4_f64(int64_t n, int nip, const double* __restrict__ dNx, const double* __restrict__ vol,
                                  const double* __restrict__ C, double* __restrict__ elemmat)
{
    int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n) return;
    int64_t e = t / (NNE * NNE);
    int mn = static_cast<int>(t - e * NNE * NNE);
    int m = mn / NNE, nn = mn - m * NNE;
    double acc[ND][ND];
#pragma unroll
    for (int j = 0; j < ND; ++j)
#pragma unroll
        for (int k = 0; k < ND; ++k) acc[j][k] = 0.0;
    for (int q = 0; q < nip; ++q) {
        const double* g = dNx + (e * nip + q) * NNE * ND;
        const double* Cq = C + (e * nip + q) * TD * TD * TD * TD;
        double v = vol[e * nip + q];
#pragma unroll
        for (int i = 0; i < ND; ++i)
#pragma unroll
            for (int j = 0; j < ND; ++j)
#pragma unroll
                for (int k = 0; k < ND; ++k)
#pragma unroll
                    for (int l = 0; l < ND; ++l)
                        acc[j][k] += g[m * ND + i] * Cq[((i * TD + j) * TD + k) * TD + l] * g[nn * ND + l] * v;
    }
    constexpr int K = NNE * ND;
    double* Ke = elemmat + e * K * K;
#pragma unroll
    for (int j = 0; j < ND; ++j)
#pragma unroll
        for (int k = 0; k < ND; ++k) Ke[(m * ND + j) * K + nn * ND + k] = acc[j][k];
}

// The same integral for Quad4 with four points (planar TD = 3, plain TD = 2), FP64, as one thread per
// (element, node m): per point the products dNx(m, i) C(i, j, k, l) are formed once and used for all
// four column nodes n, in the generic loop's expression and summation order (bit-identical to it).
// A block owns 64 elements.  The whole block first copies the 16 in-plane entries of its 256 points'
// tangents into LDS (a planar point's 81-entry record is read only where it holds them: 8.7 GB of
// HBM reads at 2048^2 against 10.8 GB when each thread loads its own, and 8 % less time); the K_e
// rows then go through the same LDS (rows padded against bank conflicts) and leave as one contiguous
// 16-byte stream: a thread writing its own 128-byte rows directly costs several times the kernel's
// time.  Needs a 16-byte aligned elemmat.
constexpr int T4_EB = 64, T4_KROW = 18;

template <int TD>
__global__ void __launch_bounds__(4 * T4_EB)
    k_int_tensor4_q4_f64(int64_t nelem, const double* __restrict__ dNx, const double* __restrict__ vol,
                         const double* __restrict__ C, double* __restrict__ elemmat)
{
    typedef double dbl2 __attribute__((ext_vector_type(2)));
    __shared__ __align__(16) double sK[4 * T4_EB * T4_KROW]; // staged C [256][16], then K_e rows
    static_assert(T4_KROW >= 16, "the staged tangents must fit");
    const int64_t e0 = blockIdx.x * static_cast<int64_t>(T4_EB);
    const int ne = static_cast<int>(nelem - e0 < T4_EB ? nelem - e0 : T4_EB);
    const int t = threadIdx.x, el = t / 4, m = t % 4;
    {
        // entry r = ((i 2 + j) 2 + k) 2 + l of point p at sK[16 p + r]
        const double* Cb = C + e0 * 4 * TD * TD * TD * TD;
        for (int w = t; w < 64 * ne; w += 4 * T4_EB) {
            const int p = w / 16, r = w % 16, i = r / 8, j = (r / 4) % 2, k = (r / 2) % 2, l = r % 2;
            sK[w] = Cb[p * TD * TD * TD * TD + ((i * TD + j) * TD + k) * TD + l];
        }
    }
    __syncthreads();
    double acc[4][2][2]; // acc[n][j][k] = K(2m + j, 2n + k)
    if (t < 4 * ne) {
        const int64_t e = e0 + el;
#pragma unroll
        for (int nn = 0; nn < 4; ++nn)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int k = 0; k < 2; ++k) acc[nn][j][k] = 0.0;
#pragma unroll 1
        for (int q = 0; q < 4; ++q) {
            const double* g = dNx + (e * 4 + q) * 8;
            const double* Cq = sK + 16 * (4 * el + q);
            const double v = vol[e * 4 + q];
            double gv[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) gv[k] = g[k];
            double gc[2][2][2][2]; // dNx(m, i) * C(i, j, k, l)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int k = 0; k < 2; ++k)
#pragma unroll
                        for (int l = 0; l < 2; ++l) gc[i][j][k][l] = gv[m * 2 + i] * Cq[((i * 2 + j) * 2 + k) * 2 + l];
#pragma unroll
            for (int nn = 0; nn < 4; ++nn)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
#pragma unroll
                        for (int k = 0; k < 2; ++k)
#pragma unroll
                            for (int l = 0; l < 2; ++l) acc[nn][j][k] += gc[i][j][k][l] * gv[nn * 2 + l] * v;
        }
    }
    __syncthreads(); // the staged tangents are consumed
    if (t < 4 * ne) {
        dbl2* row = reinterpret_cast<dbl2*>(sK + T4_KROW * t); // rows 2m, 2m + 1: 16 contiguous doubles
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int nn = 0; nn < 4; ++nn) row[4 * j + nn] = dbl2{acc[nn][j][0], acc[nn][j][1]};
    }
    __syncthreads();
    // the block's K_e are 8 * (4 ne) contiguous 16-byte pieces; piece w is piece w % 8 of row w / 8
    dbl2* K = reinterpret_cast<dbl2*>(elemmat + 64 * e0);
    for (int w = t; w < 32 * ne; w += 4 * T4_EB)
        __builtin_nontemporal_store(reinterpret_cast<const dbl2*>(sK + T4_KROW * (w / 8))[w % 8], K + w);
}

// ---- material ----------------------------------------------------------------------------------

template <class T, int TD>
__global__ void k_elastic_stress(int64_t n, const T* __restrict__ eps, const T* __restrict__ K,
                                 const T* __restrict__ G, T* __restrict__ sig)
{
    int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (p >= n) return;
    const T* e = eps + p * TD * TD;
    T* s = sig + p * TD * TD;
    T tr = (TD == 3) ? e[0] + e[4] + e[8] : e[0] + e[3];
    T k = K[p], g = G[p];
#pragma unroll
    for (int i = 0; i < TD; ++i)
#pragma unroll
        for (int j = 0; j < TD; ++j) {
            if (i == j) s[i * TD + j] = k * tr + T(2) * g * (e[i * TD + j] - tr / T(3));
            else s[i * TD + j] = T(2) * g * e[i * TD + j];
        }
}

// Tangent C_ijab = K d_ij d_ab + 2 G (0.5 (d_ia d_jb + d_ib d_ja) - d_ij d_ab / 3) per point.  Every
// point has four distinct entries: K + 2G (1 - 1/3) (i = j = a = b), K + 2G (-1/3) (i = j, a = b,
// i != a), G ({a, b} = {i, j}, i != j) and 0, each formed from the same FP64 expression as the
// per-point loop of the material law (TO = float rounds it once).  A workgroup computes the four
// entries of its TP points into LDS, then writes its contiguous slice of C with full-wave
// contiguous 16-byte streaming stores: GP points (GP TD^4 entries) are NV stores, and a small LDS
// table gives, for each store of a group, the point and entry of each of its ES lanes.
constexpr int TP = 256;

// entry code of component c = ((i TD + j) TD + a) TD + b
template <int TD>
__device__ inline int tangent_code(int c)
{
    const int i = c / (TD * TD * TD), j = (c / (TD * TD)) % TD, a = (c / TD) % TD, b = c % TD;
    if (i == j && a == b) return i == a ? 0 : 1;
    if (i != j && ((i == a && j == b) || (i == b && j == a))) return 2;
    return 3;
}

template <int TD, class TO>
__global__ void __launch_bounds__(BS) k_elastic_tangent(int64_t n, const double* __restrict__ K,
                                                        const double* __restrict__ G, TO* __restrict__ C, bool vec)
{
    static_assert(BS == TP, "one point per thread in the first phase");
    constexpr int T4 = TD * TD * TD * TD;
    constexpr int ES = 16 / static_cast<int>(sizeof(TO));  // entries per 16-byte store
    constexpr int GP = T4 % ES == 0 ? 1 : ES;               // points per group (GP T4 a multiple of ES)
    constexpr int NV = GP * T4 / ES;                        // stores per group
    constexpr int NS = TP / GP * NV;                        // stores per workgroup
    typedef TO V __attribute__((ext_vector_type(ES)));
    __shared__ TO s_val[TP][4];
    __shared__ uint16_t s_tab[NV];
    const int64_t p0 = static_cast<int64_t>(blockIdx.x) * TP;
    const int np = static_cast<int>(min<int64_t>(TP, n - p0));
    if (threadIdx.x < np) {
        const double k = K[p0 + threadIdx.x], g = G[p0 + threadIdx.x];
        const double a2_0 = 0.5 * (1.0 * 1.0 + 1.0 * 1.0) - 1.0 * 1.0 / 3.0; // i = j = a = b
        const double a2_1 = 0.5 * (0.0 * 0.0 + 0.0 * 0.0) - 1.0 * 1.0 / 3.0; // i = j != a = b
        s_val[threadIdx.x][0] = static_cast<TO>(k * 1.0 * 1.0 + 2.0 * g * a2_0);
        s_val[threadIdx.x][1] = static_cast<TO>(k * 1.0 * 1.0 + 2.0 * g * a2_1);
        s_val[threadIdx.x][2] = static_cast<TO>(k * 0.0 * 0.0 + 2.0 * g * (0.5 * (1.0 * 1.0 + 0.0 * 0.0) - 0.0 * 0.0 / 3.0));
        s_val[threadIdx.x][3] = static_cast<TO>(k * 0.0 * 0.0 + 2.0 * g * (0.5 * (0.0 * 0.0 + 0.0 * 0.0) - 0.0 * 0.0 / 3.0));
    }
    for (int v = threadIdx.x; v < NV; v += BS) {
        uint16_t t = 0;
#pragma unroll
        for (int l = 0; l < ES; ++l) {
            const int f = v * ES + l, off = f / T4;
            t |= static_cast<uint16_t>((off | (tangent_code<TD>(f - off * T4) << 2)) << (4 * l));
        }
        s_tab[v] = t;
    }
    __syncthreads();
    TO* out = C + p0 * T4;
    if (vec && np == TP) {
        V* o = reinterpret_cast<V*>(out);
        for (int sidx = threadIdx.x; sidx < NS; sidx += BS) {
            const int grp = sidx / NV, v = sidx - grp * NV;
            const unsigned t = s_tab[v];
            V w;
#pragma unroll
            for (int l = 0; l < ES; ++l) {
                const unsigned e = (t >> (4 * l)) & 15u;
                w[l] = s_val[grp * GP + (e & 3u)][e >> 2];
            }
            __builtin_nontemporal_store(w, o + sidx);
        }
    }
    else {
        for (int f = threadIdx.x; f < np * T4; f += BS) {
            const int pl = f / T4;
            out[f] = s_val[pl][tangent_code<TD>(f - pl * T4)];
        }
    }
}

void compute_geometry(gf_quad_s* q, hipStream_t s)
{
    int64_t n = q->nelem * q->nip;
    if (n == 0) return;
    if (q->et == GF_HEX8)
        k_compute_dN_hex8<<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, q->x.p, q->dNxi_d.p, q->w_d.p, q->dNx.p, q->vol.p);
    else
        k_compute_dN_quad4<<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, q->et == GF_QUAD4_PLANAR, q->thick, q->x.p,
                                                       q->dNxi_d.p, q->w_d.p, q->dNx.p, q->vol.p);
    check_launch();
    q->have32 = false;
}

void ensure32(gf_quad_s* q, hipStream_t s)
{
    if (q->have32) return;
    int64_t n1 = q->nelem * q->nip * q->nne * q->ndim, n2 = q->nelem * q->nip;
    if (q->dNx32.n != static_cast<size_t>(n1)) q->dNx32.alloc(n1);
    if (q->vol32.n != static_cast<size_t>(n2)) q->vol32.alloc(n2);
    if (n1) k_to_f32<<<nblocks(n1, BS), BS, 0, s>>>(n1, q->dNx.p, q->dNx32.p);
    if (n2) k_to_f32<<<nblocks(n2, BS), BS, 0, s>>>(n2, q->vol.p, q->vol32.p);
    check_launch();
    q->have32 = true;
}

int grad_dispatch(gf_quad q, const void* elemvec, void* qtensor, int dtype, int mode, void* stream)
{
    return guard([&] {
        GF_REQUIRE(q && elemvec && qtensor, "null argument");
        GF_REQUIRE(dtype == GF_F64 || dtype == GF_F32, "dtype must be GF_F64 or GF_F32");
        int64_t n = q->nelem * q->nip;
        if (n == 0) return;
        hipStream_t s = S(stream);
        auto launch = [&](auto tag, auto mtag, const auto* u, auto* o, const auto* g) {
            using T = decltype(tag);
            constexpr int M = decltype(mtag)::value;
            const size_t sz = sizeof(T);
            // Hex8 FP64: the geometry recomputed from the coordinates (1536 -> 192 bytes of dN/dx per
            // element); FP32 reads the stored FP32 copy (768 bytes per element), cheaper than the
            // FP64 recomputation; rules with few points per element: the stored geometry
            if (q->et == GF_HEX8 && sizeof(T) == 8 && grad3x_lds_bytes(q->nip, sz) <= 65536)
                k_grad_3d_x<T, M><<<nblocks(n, QB3), QB3, grad3x_lds_bytes(q->nip, sz), s>>>(n, q->nip, q->x.p,
                                                                                             q->dNxi_d.p, u, o);
            else if (q->et == GF_HEX8)
                k_grad_3d_lds<T, M><<<nblocks(n, 128), 128, grad_lds_bytes<3>(128, q->nip, 3, sz), s>>>(n, q->nip, g,
                                                                                                        u, o);
            else if (q->et == GF_QUAD4_PLANAR)
                k_grad_2d_lds<T, 3, M><<<nblocks(n, 256), 256, grad_lds_bytes<2>(256, q->nip, 3, sz), s>>>(n, q->nip,
                                                                                                           g, u, o);
            else
                k_grad_2d_lds<T, 2, M><<<nblocks(n, 256), 256, grad_lds_bytes<2>(256, q->nip, 2, sz), s>>>(n, q->nip,
                                                                                                           g, u, o);
        };
        auto by_mode = [&](auto tag, const auto* u, auto* o, const auto* g) {
            if (mode == 0) launch(tag, std::integral_constant<int, 0>(), u, o, g);
            else if (mode == 1) launch(tag, std::integral_constant<int, 1>(), u, o, g);
            else launch(tag, std::integral_constant<int, 2>(), u, o, g);
        };
        if (dtype == GF_F64)
            by_mode(double(0), static_cast<const double*>(elemvec), static_cast<double*>(qtensor), q->dNx.p);
        else {
            ensure32(q, s);
            by_mode(float(0), static_cast<const float*>(elemvec), static_cast<float*>(qtensor), q->dNx32.p);
        }
        check_launch();
    });
}

} // namespace

// Accessors used by other translation units (stiffness, explicit).
namespace gf {
const double* quad_dNx(gf_quad q) { return q->dNx.p; }
const double* quad_vol(gf_quad q) { return q->vol.p; }
int quad_etype(gf_quad q) { return q->et; }
int64_t quad_nelem(gf_quad q) { return q->nelem; }
int quad_nip(gf_quad q) { return q->nip; }
} // namespace gf

extern "C" {

int gf_quad_create(gf_quad* out, int etype, const double* x, int64_t nelem, int rule, const double* xi_h,
                   const double* w_h, int nip, double thick, void* stream)
{
    return guard([&] {
        GF_REQUIRE(out && x, "null argument");
        GF_REQUIRE(etype == GF_QUAD4 || etype == GF_QUAD4_PLANAR || etype == GF_HEX8, "unknown element type");
        auto q = new gf_quad_s();
        try {
            q->et = etype;
            q->nne = etype == GF_HEX8 ? 8 : 4;
            q->ndim = etype == GF_HEX8 ? 3 : 2;
            q->tdim = etype == GF_QUAD4 ? 2 : 3;
            q->nelem = nelem;
            q->thick = thick;
            if (xi_h && w_h) {
                GF_REQUIRE(nip > 0, "nip > 0 with a custom rule");
                q->xi.assign(xi_h, xi_h + nip * q->ndim);
                q->w.assign(w_h, w_h + nip);
            }
            else {
                GF_REQUIRE(rule == GF_GAUSS || rule == GF_NODAL, "rule must be GF_GAUSS or GF_NODAL");
                default_rule(etype, rule, q->xi, q->w);
            }
            q->nip = static_cast<int>(q->w.size());
            q->N.resize(q->nip * q->nne);
            q->dNxi.resize(q->nip * q->nne * q->ndim);
            shape_functions(etype, q->nip, q->xi.data(), q->N.data(), q->dNxi.data());
            q->N_d.upload(q->N);
            q->dNxi_d.upload(q->dNxi);
            q->w_d.upload(q->w);
            q->x.alloc(nelem * q->nne * q->ndim);
            if (q->x.n)
                GF_HIP(hipMemcpyAsync(q->x.p, x, q->x.n * sizeof(double), hipMemcpyDeviceToDevice, S(stream)));
            q->dNx.alloc(nelem * q->nip * q->nne * q->ndim);
            q->vol.alloc(nelem * q->nip);
            compute_geometry(q, S(stream));
        }
        catch (...) {
            delete q;
            throw;
        }
        *out = q;
    });
}

int gf_quad_destroy(gf_quad q)
{
    delete q;
    return GF_OK;
}

int gf_quad_sizes(gf_quad q, int64_t* nelem, int* nne, int* ndim, int* tdim, int* nip)
{
    return guard([&] {
        GF_REQUIRE(q, "null quadrature");
        if (nelem) *nelem = q->nelem;
        if (nne) *nne = q->nne;
        if (ndim) *ndim = q->ndim;
        if (tdim) *tdim = q->tdim;
        if (nip) *nip = q->nip;
    });
}

int gf_quad_update_x(gf_quad q, const double* x, void* stream)
{
    return guard([&] {
        GF_REQUIRE(q && x, "null argument");
        if (q->x.n)
            GF_HIP(hipMemcpyAsync(q->x.p, x, q->x.n * sizeof(double), hipMemcpyDeviceToDevice, S(stream)));
        compute_geometry(q, S(stream));
    });
}

int gf_quad_gradN(gf_quad q, const double** dNx)
{
    return guard([&] {
        GF_REQUIRE(q && dNx, "null argument");
        *dNx = q->dNx.p;
    });
}

int gf_quad_dV(gf_quad q, const double** dV)
{
    return guard([&] {
        GF_REQUIRE(q && dV, "null argument");
        *dV = q->vol.p;
    });
}

int gf_quad_interp_quad_vector(gf_quad q, const double* elemvec, int nv, double* qvector, void* stream)
{
    return guard([&] {
        GF_REQUIRE(q && elemvec && qvector && nv > 0, "invalid argument");
        int64_t n = q->nelem * q->nip;
        if (n == 0) return;
        k_interp<<<nblocks(n, BS), BS, 0, S(stream)>>>(n, q->nip, q->nne, nv, q->N_d.p, elemvec, qvector);
        check_launch();
    });
}

int gf_quad_gradN_vector(gf_quad q, const void* elemvec, void* qtensor, int dtype, void* stream)
{
    return grad_dispatch(q, elemvec, qtensor, dtype, 0, stream);
}

int gf_quad_gradN_vector_T(gf_quad q, const void* elemvec, void* qtensor, int dtype, void* stream)
{
    return grad_dispatch(q, elemvec, qtensor, dtype, 1, stream);
}

int gf_quad_symgradN_vector(gf_quad q, const void* elemvec, void* qtensor, int dtype, void* stream)
{
    return grad_dispatch(q, elemvec, qtensor, dtype, 2, stream);
}

int gf_quad_int_N_vector_dV(gf_quad q, const double* qvector, int nv, double* elemvec, void* stream)
{
    return guard([&] {
        GF_REQUIRE(q && qvector && elemvec && nv > 0, "invalid argument");
        int64_t n = q->nelem * q->nne;
        if (n == 0) return;
        k_int_N_vector<<<nblocks(n, BS), BS, 0, S(stream)>>>(n, q->nip, q->nne, nv, q->N_d.p, q->vol.p, qvector,
                                                           elemvec);
        check_launch();
    });
}

int gf_quad_int_N_scalar_NT_dV(gf_quad q, const double* qscalar, double* elemmat, void* stream)
{
    return guard([&] {
        GF_REQUIRE(q && qscalar && elemmat, "null argument");
        int64_t n = q->nelem * q->nne * q->nne;
        if (n == 0) return;
        k_int_N_scalar<<<nblocks(n, BS), BS, 0, S(stream)>>>(n, q->nip, q->nne, q->ndim, q->N_d.p, q->vol.p, qscalar,
                                                           elemmat);
        check_launch();
    });
}

int gf_quad_int_gradN_dot_tensor2_dV(gf_quad q, const void* qtensor, void* elemvec, int dtype, void* stream)
{
    return guard([&] {
        GF_REQUIRE(q && qtensor && elemvec, "null argument");
        GF_REQUIRE(dtype == GF_F64 || dtype == GF_F32, "dtype must be GF_F64 or GF_F32");
        const int64_t E = q->nelem;
        if (E == 0) return;
        hipStream_t s = S(stream);
        auto launch = [&](auto tag, const auto* in, auto* o, const auto* g, const auto* v) {
            using T = decltype(tag);
            // rules with many points (LDS need above 64 KiB): the unstaged kernel, thread per (e, m)
            const size_t need = q->et == GF_HEX8 ? int2_lds_bytes<T, 8, 3, 3, 16>(q->nip)
                                                 : int2_lds_bytes<T, 4, 2, 3, 64>(q->nip);
            if (need > 65536) {
                const int64_t n = E * q->nne;
                if (q->et == GF_HEX8) k_int_tensor2<T, 3, 3><<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, q->nne, g, v, in, o);
                else if (q->et == GF_QUAD4_PLANAR)
                    k_int_tensor2<T, 2, 3><<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, q->nne, g, v, in, o);
                else k_int_tensor2<T, 2, 2><<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, q->nne, g, v, in, o);
                return;
            }
            if (q->et == GF_HEX8)
                k_int_tensor2_lds<T, 8, 3, 3, 16><<<nblocks(E, 16), 128, int2_lds_bytes<T, 8, 3, 3, 16>(q->nip), s>>>(
                    E, q->nip, g, v, in, o);
            else if (q->et == GF_QUAD4_PLANAR)
                k_int_tensor2_lds<T, 4, 2, 3, 64><<<nblocks(E, 64), 256, int2_lds_bytes<T, 4, 2, 3, 64>(q->nip), s>>>(
                    E, q->nip, g, v, in, o);
            else
                k_int_tensor2_lds<T, 4, 2, 2, 64><<<nblocks(E, 64), 256, int2_lds_bytes<T, 4, 2, 2, 64>(q->nip), s>>>(
                    E, q->nip, g, v, in, o);
        };
        if (dtype == GF_F64)
            launch(double(0), static_cast<const double*>(qtensor), static_cast<double*>(elemvec), q->dNx.p, q->vol.p);
        else {
            ensure32(q, s);
            launch(float(0), static_cast<const float*>(qtensor), static_cast<float*>(elemvec), q->dNx32.p, q->vol32.p);
        }
        check_launch();
    });
}

int gf_quad_int_gradN_dot_tensor4_dot_gradNT_dV(gf_quad q, const double* C, double* elemmat, int precision,
                                                void* stream)
{
    return guard([&] {
        GF_REQUIRE(q && C && elemmat, "null argument");
        int64_t n = q->nelem * q->nne * q->nne;
        if (n == 0) return;
        hipStream_t s = S(stream);
        if (precision == GF_BF16X3) {
            GF_REQUIRE(q->et == GF_HEX8, "GF_BF16X3 is implemented for Hex8");
            int rc = gf::stiffness_bf16x3(q, C, elemmat, s);
            if (rc != GF_OK) throw Error(rc, gf_last_error());
            return;
        }
        GF_REQUIRE(precision == GF_F64, "precision must be GF_F64 or GF_BF16X3");
        if (q->et == GF_HEX8 && q->nip == 8) {
            int rc = gf::stiffness_f64_hex8(q, C, elemmat, s);
            if (rc != GF_OK) throw Error(rc, gf_last_error());
            return;
        }
        if (q->et != GF_HEX8 && q->nip == 4 && aligned16(elemmat)) {
            if (q->et == GF_QUAD4_PLANAR)
                k_int_tensor4_q4_f64<3><<<nblocks(q->nelem, T4_EB), 4 * T4_EB, 0, s>>>(q->nelem, q->dNx.p, q->vol.p, C,
                                                                                        elemmat);
            else
                k_int_tensor4_q4_f64<2><<<nblocks(q->nelem, T4_EB), 4 * T4_EB, 0, s>>>(q->nelem, q->dNx.p, q->vol.p, C,
                                                                                        elemmat);
        }
        else if (q->et == GF_HEX8)
            k_int_tensor4_f64<8, 3, 3><<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, q->dNx.p, q->vol.p, C, elemmat);
        else if (q->et == GF_QUAD4_PLANAR)
            k_int_tensor4_f64<4, 2, 3><<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, q->dNx.p, q->vol.p, C, elemmat);
        else
            k_int_tensor4_f64<4, 2, 2><<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, q->dNx.p, q->vol.p, C, elemmat);
        check_launch();
    });
}

int gf_material_elastic_stress(int64_t nq, int tdim, const void* eps, const void* K, const void* G, void* sig,
                               int dtype, void* stream)
{
    return guard([&] {
        GF_REQUIRE(eps && K && G && sig, "null argument");
        GF_REQUIRE(tdim == 2 || tdim == 3, "tdim must be 2 or 3");
        GF_REQUIRE(dtype == GF_F64 || dtype == GF_F32, "dtype must be GF_F64 or GF_F32");
        if (nq == 0) return;
        hipStream_t s = S(stream);
        unsigned nb = nblocks(nq, BS);
        if (dtype == GF_F64) {
            auto e = static_cast<const double*>(eps);
            auto k = static_cast<const double*>(K);
            auto g = static_cast<const double*>(G);
            auto o = static_cast<double*>(sig);
            if (tdim == 3) k_elastic_stress<double, 3><<<nb, BS, 0, s>>>(nq, e, k, g, o);
            else k_elastic_stress<double, 2><<<nb, BS, 0, s>>>(nq, e, k, g, o);
        }
        else {
            auto e = static_cast<const float*>(eps);
            auto k = static_cast<const float*>(K);
            auto g = static_cast<const float*>(G);
            auto o = static_cast<float*>(sig);
            if (tdim == 3) k_elastic_stress<float, 3><<<nb, BS, 0, s>>>(nq, e, k, g, o);
            else k_elastic_stress<float, 2><<<nb, BS, 0, s>>>(nq, e, k, g, o);
        }
        check_launch();
    });
}

int gf_material_elastic_tangent(int64_t nq, int tdim, const double* K, const double* G, double* C, void* stream)
{
    return guard([&] {
        GF_REQUIRE(K && G && C, "null argument");
        GF_REQUIRE(tdim == 2 || tdim == 3, "tdim must be 2 or 3");
        if (nq == 0) return;
        if (tdim == 3) k_elastic_tangent<3, double><<<nblocks(nq, TP), BS, 0, S(stream)>>>(nq, K, G, C, aligned16(C));
        else k_elastic_tangent<2, double><<<nblocks(nq, TP), BS, 0, S(stream)>>>(nq, K, G, C, aligned16(C));
        check_launch();
    });
}

int gf_material_elastic_tangent_f32(int64_t nq, int tdim, const double* K, const double* G, float* C, void* stream)
{
    return guard([&] {
        GF_REQUIRE(K && G && C, "null argument");
        GF_REQUIRE(tdim == 2 || tdim == 3, "tdim must be 2 or 3");
        if (nq == 0) return;
        if (tdim == 3) k_elastic_tangent<3, float><<<nblocks(nq, TP), BS, 0, S(stream)>>>(nq, K, G, C, aligned16(C));
        else k_elastic_tangent<2, float><<<nblocks(nq, TP), BS, 0, S(stream)>>>(nq, K, G, C, aligned16(C));
        check_launch();
    });
}

} // extern "C"
