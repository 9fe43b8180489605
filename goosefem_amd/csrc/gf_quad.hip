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

__global__ void k_compute_dN_hex8(int64_t n, int nip, const double* __restrict__ x, const double* __restrict__ dNxi,
                                  const double* __restrict__ w, double* __restrict__ dNx, double* __restrict__ vol)
{
    int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n) return;
    int64_t e = t / nip;
    int q = static_cast<int>(t - e * nip);
    const double* xe = x + e * 24;
    const double* g = dNxi + q * 24;
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
    // detail.h:40-70
    double det = (A[0][0] * A[1][1] * A[2][2] + A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[1][0] * A[2][1]) -
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
    double* o = dNx + t * 24;
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < 3; ++j) s += B[i][j] * g[m * 3 + j];
            o[m * 3 + i] = s;
        }
    vol[t] = w[q] * det;
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b)
{
    int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t < n) b[t] = static_cast<float>(a[t]);
}

// ---- gradients: mode 0 GradN_vector, 1 GradN_vector_T, 2 SymGradN_vector -------------------------

// 2-D node vectors (Quad4 tdim 2 / planar tdim 3): ElementQuad4.h:348-434, ElementQuad4Planar.h:129-221
template <class T, int TD>
__global__ void k_grad_2d(int64_t n, int nip, int mode, const T* __restrict__ dNx, const T* __restrict__ elemvec,
                          T* __restrict__ qtensor)
{
    int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n) return;
    int64_t e = t / nip;
    const T* u = elemvec + e * 8;
    const T* g = dNx + t * 8;
    T a00 = g[0] * u[0] + g[2] * u[2] + g[4] * u[4] + g[6] * u[6];
    T a01 = g[0] * u[1] + g[2] * u[3] + g[4] * u[5] + g[6] * u[7];
    T a10 = g[1] * u[0] + g[3] * u[2] + g[5] * u[4] + g[7] * u[6];
    T a11 = g[1] * u[1] + g[3] * u[3] + g[5] * u[5] + g[7] * u[7];
    T r[TD * TD];
#pragma unroll
    for (int k = 0; k < TD * TD; ++k) r[k] = T(0);
    if (mode == 0) {
        r[0] = a00;
        r[1] = a01;
        r[TD] = a10;
        r[TD + 1] = a11;
    }
    else if (mode == 1) {
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
    T* o = qtensor + t * TD * TD;
#pragma unroll
    for (int k = 0; k < TD * TD; ++k) o[k] = r[k];
}

// Hex8: generic loops Element.h:924-1022
template <class T>
__global__ void k_grad_3d(int64_t n, int nip, int mode, const T* __restrict__ dNx, const T* __restrict__ elemvec,
                          T* __restrict__ qtensor)
{
    int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n) return;
    int64_t e = t / nip;
    const T* u = elemvec + e * 24;
    const T* g = dNx + t * 24;
    T r[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) r[k] = T(0);
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                if (mode == 0) r[i * 3 + j] += g[m * 3 + i] * u[m * 3 + j];
                else if (mode == 1) r[j * 3 + i] += g[m * 3 + i] * u[m * 3 + j];
                else {
                    T h = T(0.5) * g[m * 3 + i] * u[m * 3 + j];
                    r[i * 3 + j] += h;
                    r[j * 3 + i] += h;
                }
            }
    T* o = qtensor + t * 9;
#pragma unroll
    for (int k = 0; k < 9; ++k) o[k] = r[k];
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

// Tangent, written as contiguous output: a workgroup owns TP consecutive points (TP TD^4
// components) and its threads stride over the components, so every store instruction covers
// contiguous memory.  Component c of a point is C(i, j, a, b) with c = ((i TD + j) TD + a) TD + b,
// the same expression as the per-point loop of the material law; TO = float rounds the FP64
// value once.
constexpr int TP = 64;

template <int TD, class TO>
__global__ void __launch_bounds__(BS) k_elastic_tangent(int64_t n, const double* __restrict__ K,
                                                        const double* __restrict__ G, TO* __restrict__ C, bool pairs)
{
    // The block's slice of C is contiguous and is written with streaming stores (not re-read).
    // FP64: two entries per 16-byte store.  FP32: each thread owns one component c and walks the
    // block's points PPI at a time, so its two constant factors are computed once and the
    // PPI * T4 active threads of a pass write one contiguous run.  (Measured on Hex8 128^3: the
    // FP64 pairs beat the component mapping, whose stores are 8 bytes; for FP32 the component
    // mapping wins, the entry arithmetic being the larger cost.)
    typedef TO V2 __attribute__((ext_vector_type(2)));
    constexpr int T4 = TD * TD * TD * TD, PPI = BS / T4;
    const int64_t p0 = static_cast<int64_t>(blockIdx.x) * TP;
    const int np = static_cast<int>(min<int64_t>(TP, n - p0));
    TO* out = C + p0 * T4;
    if constexpr (sizeof(TO) == 8) {
        auto entry = [&](int t) {
            const int pl = t / T4, c = t - pl * T4;
            const int i = c / (TD * TD * TD), j = (c / (TD * TD)) % TD, a = (c / TD) % TD, b = c % TD;
            const double k = K[p0 + pl], g = G[p0 + pl];
            const double dij = i == j, dkl = a == b, dik = i == a, djl = j == b, dil = i == b, djk = j == a;
            return static_cast<TO>(k * dij * dkl + 2.0 * g * (0.5 * (dik * djl + dil * djk) - dij * dkl / 3.0));
        };
        const int nt = np * T4;
        if (pairs) {
            for (int t = threadIdx.x; t < nt / 2; t += BS)
                __builtin_nontemporal_store(V2{entry(2 * t), entry(2 * t + 1)}, reinterpret_cast<V2*>(out) + t);
            if ((nt & 1) && threadIdx.x == 0) out[nt - 1] = entry(nt - 1);
        }
        else {
            for (int t = threadIdx.x; t < nt; t += BS) out[t] = entry(t);
        }
        return;
    }
    if (threadIdx.x >= PPI * T4) return;
    const int c = threadIdx.x % T4, q = threadIdx.x / T4;
    const int i = c / (TD * TD * TD), j = (c / (TD * TD)) % TD, a = (c / TD) % TD, b = c % TD;
    const double dij = i == j, dkl = a == b, dik = i == a, djl = j == b, dil = i == b, djk = j == a;
    const double a1 = dij * dkl, a2 = 0.5 * (dik * djl + dil * djk) - dij * dkl / 3.0;
    for (int pl = q; pl < np; pl += PPI) {
        const double k = K[p0 + pl], g = G[p0 + pl];
        // k * a1 equals k * dij * dkl (a1 is 0 or 1), so the value is the same expression's
        __builtin_nontemporal_store(static_cast<TO>(k * a1 + 2.0 * g * a2), out + pl * T4 + c);
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
        if (dtype == GF_F64) {
            auto u = static_cast<const double*>(elemvec);
            auto o = static_cast<double*>(qtensor);
            if (q->et == GF_HEX8) k_grad_3d<double><<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, mode, q->dNx.p, u, o);
            else if (q->et == GF_QUAD4_PLANAR)
                k_grad_2d<double, 3><<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, mode, q->dNx.p, u, o);
            else k_grad_2d<double, 2><<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, mode, q->dNx.p, u, o);
        }
        else {
            ensure32(q, s);
            auto u = static_cast<const float*>(elemvec);
            auto o = static_cast<float*>(qtensor);
            if (q->et == GF_HEX8) k_grad_3d<float><<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, mode, q->dNx32.p, u, o);
            else if (q->et == GF_QUAD4_PLANAR)
                k_grad_2d<float, 3><<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, mode, q->dNx32.p, u, o);
            else k_grad_2d<float, 2><<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, mode, q->dNx32.p, u, o);
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
        int64_t n = q->nelem * q->nne;
        if (n == 0) return;
        hipStream_t s = S(stream);
        if (dtype == GF_F64) {
            auto in = static_cast<const double*>(qtensor);
            auto o = static_cast<double*>(elemvec);
            if (q->et == GF_HEX8)
                k_int_tensor2<double, 3, 3><<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, q->nne, q->dNx.p, q->vol.p, in, o);
            else if (q->et == GF_QUAD4_PLANAR)
                k_int_tensor2<double, 2, 3><<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, q->nne, q->dNx.p, q->vol.p, in, o);
            else
                k_int_tensor2<double, 2, 2><<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, q->nne, q->dNx.p, q->vol.p, in, o);
        }
        else {
            ensure32(q, s);
            auto in = static_cast<const float*>(qtensor);
            auto o = static_cast<float*>(elemvec);
            if (q->et == GF_HEX8)
                k_int_tensor2<float, 3, 3><<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, q->nne, q->dNx32.p, q->vol32.p, in, o);
            else if (q->et == GF_QUAD4_PLANAR)
                k_int_tensor2<float, 2, 3><<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, q->nne, q->dNx32.p, q->vol32.p, in, o);
            else
                k_int_tensor2<float, 2, 2><<<nblocks(n, BS), BS, 0, s>>>(n, q->nip, q->nne, q->dNx32.p, q->vol32.p, in, o);
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
        if (q->et == GF_HEX8)
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
