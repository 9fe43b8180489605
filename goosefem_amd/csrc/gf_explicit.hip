// Fused velocity-Verlet step for Quad4 plane-strain explicit dynamics: the caller loop of
// docs/examples/dynamics/Elastic-VelocityVerlet/main.cpp:110-157, i.e.
//   u <- (u + dt v) + c a                       (c = dt^2 / 2)
//   fint = AssembleDofs(Int_gradN_dot_tensor2_dV(sigma(SymGradN_vector(AsElement(u)))))
//   a_new = inv_mass (fext - fint);  v <- v + (dt/2)(a + a_new);  a <- a_new
// with QuadraturePlanar geometry (ElementQuad4Planar.h:280-319, Gauss 2x2, thick 1) and the linear
// elastic law sigma = K tr(eps) I + 2 G dev(eps).
//
// Structured path (conn is a Mesh::Quad4::Regular grid, dofs = Mesh::dofs): one kernel per step.
// A workgroup owns a TX x TY tile of nodes.  It stages the updated displacement and coordinates
// of the tile plus a one-node halo in LDS, computes the forces of the (TX+1) x (TY+1) elements
// touching the tile (geometry recomputed from the coordinates), keeps them in LDS, and each
// owned node sums its <= 4 element contributions in ascending element order -- the reference's
// serial assembly order -- then applies the velocity-Verlet update.  HBM traffic per step is the
// state (u, v, a read + write), inv_mass, coordinates and per-element moduli; no elemvec,
// qtensor or connectivity stream.  State is double-buffered (read old, write new).
// Tiles are ordered so each XCD sweeps a contiguous band of the mesh (its L2 sees the halo rows
// its own previous tiles loaded).
//
// General path (any Quad4 mesh / dof numbering): two kernels per step -- element kernel (gather of
// coordinates and displacement, geometry recomputed, strain + stress + force in FP64 reference
// order, FP32, or FP64 with FMA) over the elements sorted by their lowest node, then the
// DOF-centric ordered gather + update, which also forms the next step's displacement.
#include <cmath>
#include <functional>
#include <unordered_map>
#include <cstdio>
#include <cstdlib>
#include <type_traits>

#include "gf_internal.hpp"
#include "gf_wave.hpp"

using namespace gf;

struct gf_explicit_s {
    gf_vector v = nullptr;
    int64_t ndof = 0, nelem = 0;
    double dt = 0.0;
    bool structured = false;
    int64_t nx = 0, ny = 0; // structured grid size (elements)
    DBuf<double> coor;      // nodevec [nnode, 2]
    DBuf<float> Ke, Ge;     // per element
    bool k_uniform = false; // all K_e equal: pass the scalar, skip the per-element K stream
    float k_value = 0.0f;
    bool reg_x = false;     // coordinates are exactly (i sx, j sy): generated in the kernel, not read
    double sx = 0.0, sy = 0.0;
    DBuf<double> inv_mass, fext;
    // M^-1 as a palette when it holds at most 256 distinct values (a uniform density gives a handful:
    // interior, edge and corner nodes, and 0 on prescribed DOFs): one byte per DOF and a 256-entry
    // table, read instead of 8 bytes per DOF; the values, and so the results, are the same
    DBuf<uint8_t> im_code;
    DBuf<double> im_tab;
    bool has_fext = false;
    DBuf<double> u[2], vv[2], a[2];
    int cur = 0;
    DBuf<double> fint;
    bool fint_valid = false;
    // general path: elements in locality order (sorted by their lowest node), with their
    // connectivity and moduli in that order, and the per-DOF gather lists pointing at the element
    // forces of that order (each list still in the reference's (e, m, i) order); element forces of
    // the last step, [nelem, 4, 2]
    DBuf<int32_t> gconn, geidx;
    DBuf<float> gK, gG;
    DBuf<double> fe64;
    DBuf<float> fe32;
    // fused general step (when each DOF belongs to exactly one node component): node blocks of BS
    // spatially compact nodes, each with its local nodes (its own, then the halo nodes its elements
    // reach), the elements that touch its nodes (ascending locality position) with their
    // connectivity in local node numbers and their moduli, and per own node its (element slot, m)
    // contributions in the reference's (e, m) order as slot * 4 + m
    bool blocks = false;
    int64_t nblk = 0;
    int64_t lds64 = 0, lds32 = 0;                   // largest LDS need of a block, FP64 / FP32 forces
    bool dofs_ident = false;                        // dofs(n, i) = 2 n + i
    DBuf<int32_t> bkptr, blptr, bnloc, beptr, bnptr; // own nodes, local nodes (own, then halo), element slots
    DBuf<uint16_t> blconn, bnent;              // element slots' local connectivity; node contributions
    DBuf<float> bK, bG;                        // moduli per element slot
    DBuf<double> epart; // energy reduction partials [2][NEB]
};

namespace {

typedef double dvec2 __attribute__((ext_vector_type(2)));

// streaming (non-temporal) 16-byte load of a node's two DOFs (read once per step)
__device__ inline double2 nt_load2(const double* p, int64_t n)
{
    dvec2 v = __builtin_nontemporal_load(reinterpret_cast<const dvec2*>(p) + n);
    return make_double2(v.x, v.y);
}

struct ShapeTable {
    double g[4][4][2]; // dN/dxi at the 4 Gauss points [q][m][i]
    double w[4];
};

constexpr int BS = 256;

// ---- per-element force, FP64 in the reference's expression order ------------------------------
__device__ inline void element_force_f64(const ShapeTable& T, const double x[4][2], const double u[4][2],
                                         double k, double gmod, double f[4][2])
{
#pragma unroll
    for (int m = 0; m < 4; ++m) f[m][0] = f[m][1] = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double(*g)[2] = T.g[q];
        double J00 = g[0][0] * x[0][0] + g[1][0] * x[1][0] + g[2][0] * x[2][0] + g[3][0] * x[3][0];
        double J01 = g[0][0] * x[0][1] + g[1][0] * x[1][1] + g[2][0] * x[2][1] + g[3][0] * x[3][1];
        double J10 = g[0][1] * x[0][0] + g[1][1] * x[1][0] + g[2][1] * x[2][0] + g[3][1] * x[3][0];
        double J11 = g[0][1] * x[0][1] + g[1][1] * x[1][1] + g[2][1] * x[2][1] + g[3][1] * x[3][1];
        double det = J00 * J11 - J01 * J10;
        double I00 = J11 / det, I01 = -1.0 * J01 / det, I10 = -1.0 * J10 / det, I11 = J00 / det;
        double d[4][2];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            d[m][0] = I00 * g[m][0] + I01 * g[m][1];
            d[m][1] = I10 * g[m][0] + I11 * g[m][1];
        }
        double vol = T.w[q] * det * 1.0;
        // SymGradN_vector, ElementQuad4Planar.h:191-221
        double e00 = d[0][0] * u[0][0] + d[1][0] * u[1][0] + d[2][0] * u[2][0] + d[3][0] * u[3][0];
        double e11 = d[0][1] * u[0][1] + d[1][1] * u[1][1] + d[2][1] * u[2][1] + d[3][1] * u[3][1];
        double e01 = 0.5 * (d[0][0] * u[0][1] + d[1][0] * u[1][1] + d[2][0] * u[2][1] + d[3][0] * u[3][1] +
                            d[0][1] * u[0][0] + d[1][1] * u[1][0] + d[2][1] * u[2][0] + d[3][1] * u[3][0]);
        double e22 = 0.0;
        // linear elastic law (tdim 3, plane strain)
        double tr = e00 + e11 + e22;
        double s00 = k * tr + 2.0 * gmod * (e00 - tr / 3.0);
        double s11 = k * tr + 2.0 * gmod * (e11 - tr / 3.0);
        double s01 = 2.0 * gmod * e01;
        double s10 = 2.0 * gmod * e01;
        // Int_gradN_dot_tensor2_dV, ElementQuad4Planar.h:253-278
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            f[m][0] += (d[m][0] * s00 + d[m][1] * s10) * vol;
            f[m][1] += (d[m][0] * s01 + d[m][1] * s11) * vol;
        }
    }
}

// ---- element modes ---------------------------------------------------------------------------------
// With node signs (s_m, t_m) = (-1,-1), (1,-1), (1,1), (-1,1) the bilinear shape functions give
// dN_m/dxi = s_m (1 + t_m eta) / 4 and dN_m/deta = t_m (1 + s_m xi) / 4, so for a nodal field w
//   sum_m dN_m/dxi w_m = (A_w + eta H_w) / 4,   sum_m dN_m/deta w_m = (B_w + xi H_w) / 4,
// A_w = -w0 + w1 + w2 - w3, B_w = -w0 - w1 + w2 + w3, H_w = w0 - w1 + w2 - w3.  The modes (here
// A / 4, B / 4, H / 4) are formed from the FP64 nodal values in FP64, as sums of two differences
// (no cancellation of large coordinates or rigid displacements), and rounded once to the body's
// arithmetic type T.  For the FP32 bodies this single rounding replaces four (differences, then
// two FP32 sums); it is the dominant error term of the nodal forces, whose element
// contributions cancel.
template <class T>
__device__ inline void element_modes(const double w[4][2], T A[2], T B[2], T H[2])
{
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        A[i] = static_cast<T>(0.25 * ((w[1][i] - w[0][i]) + (w[2][i] - w[3][i])));
        B[i] = static_cast<T>(0.25 * ((w[2][i] - w[1][i]) + (w[3][i] - w[0][i])));
        H[i] = static_cast<T>(0.25 * ((w[0][i] - w[1][i]) + (w[2][i] - w[3][i])));
    }
}

__device__ inline float fmaT(float a, float b, float c) { return fmaf(a, b, c); }
__device__ inline double fmaT(double a, double b, double c) { return fma(a, b, c); }

// ---- per-element force, algebraic form (FP32, or FP64 with FMA) -----------------------------------
// The 2x2 Gauss quadrature of the bilinear element through the element modes: at each point J, its
// inverse (one division), the displacement gradient and the stress come from the modes, and the
// nodal force f(m,j) = sum_q dV_q sum_k dN_m/dxi_k (J^-T sigma)_kj collapses to three per-element
// sums.  Same quadrature as the reference's per-node loops (ElementQuad4Planar.h:191-278), not the
// same expression order.  The Gauss weights are 1 (w[] kept for generality).
template <class T>
__device__ inline void element_force_alg(const ShapeTable& Tb, const double x[4][2], const double u[4][2], T k,
                                         T gmod, T f[4][2])
{
    T Ax[2], Bx[2], Hx[2], Au[2], Bu[2], Hu[2];
    element_modes(x, Ax, Bx, Hx);
    element_modes(u, Au, Bu, Hu);
    const T g2 = T(2) * gmod;
    // accumulated P_q = dV_q J^-T sigma: S0 = sum P(0,.), S1 = sum P(1,.), SH = sum (eta P(0,.) + xi P(1,.))
    T S0[2] = {T(0), T(0)}, S1[2] = {T(0), T(0)}, SH[2] = {T(0), T(0)};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        // Gauss points (-a,-a), (a,-a), (a,a), (-a,a) with a = 1/sqrt(3)
        const T a = T(0.57735026918962576451);
        const T xi = (q == 1 || q == 2) ? a : -a;
        const T eta = (q >= 2) ? a : -a;
        // J(i,j) = dx_j / dxi_i
        const T J00 = fmaT(eta, Hx[0], Ax[0]), J01 = fmaT(eta, Hx[1], Ax[1]);
        const T J10 = fmaT(xi, Hx[0], Bx[0]), J11 = fmaT(xi, Hx[1], Bx[1]);
        const T det = fmaT(J00, J11, -J01 * J10);
        T rdet;
        if constexpr (sizeof(T) == 4) rdet = __builtin_amdgcn_rcpf(det);
        else rdet = T(1) / det;
        const T I00 = J11 * rdet, I01 = -J01 * rdet, I10 = -J10 * rdet, I11 = J00 * rdet;
        // du_j / dxi_k
        const T d0x = fmaT(eta, Hu[0], Au[0]), d0y = fmaT(eta, Hu[1], Au[1]);
        const T d1x = fmaT(xi, Hu[0], Bu[0]), d1y = fmaT(xi, Hu[1], Bu[1]);
        // grad(i,j) = sum_k Jinv(i,k) du_j/dxi_k
        const T gxx = fmaT(I00, d0x, I01 * d1x), gxy = fmaT(I00, d0y, I01 * d1y);
        const T gyx = fmaT(I10, d0x, I11 * d1x), gyy = fmaT(I10, d0y, I11 * d1y);
        const T tr = gxx + gyy;
        const T kt = k * tr, t3 = tr * (T(1) / T(3));
        const T vol = static_cast<T>(Tb.w[q]) * det;
        const T sxx = fmaT(g2, gxx - t3, kt) * vol;
        const T syy = fmaT(g2, gyy - t3, kt) * vol;
        const T sxy = gmod * (gxy + gyx) * vol;
        // P(k,j) = sum_i Jinv(i,k) sigma(i,j) dV
        const T P0x = fmaT(I00, sxx, I10 * sxy), P0y = fmaT(I00, sxy, I10 * syy);
        const T P1x = fmaT(I01, sxx, I11 * sxy), P1y = fmaT(I01, sxy, I11 * syy);
        S0[0] += P0x;
        S0[1] += P0y;
        S1[0] += P1x;
        S1[1] += P1y;
        SH[0] = fmaT(eta, P0x, fmaT(xi, P1x, SH[0]));
        SH[1] = fmaT(eta, P0y, fmaT(xi, P1y, SH[1]));
    }
    // f(m,j) = (s_m S0_j + t_m S1_j + s_m t_m SH_j) / 4
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const T p = T(0.25) * S0[j], r = T(0.25) * S1[j], h = T(0.25) * SH[j];
        f[0][j] = -p - r + h;
        f[1][j] = p - r - h;
        f[2][j] = p + r + h;
        f[3][j] = -p + r - h;
    }
}

// ---- per-element force, rectangular element (FP32, or FP64 with FMA) ------------------------------
// For an axis-aligned sx x sy rectangle the Jacobian is diag(sx/2, sy/2) at every Gauss point, so the
// displacement gradient is linear in (xi, eta): grad = grad_c + xi grad_xi + eta grad_eta, and so is
// the stress.  With Gauss points (+-a, +-a), sum_q xi_q = sum_q eta_q = sum_q xi_q eta_q = 0 and
// sum_q xi_q^2 = sum_q eta_q^2 = 4 a^2 = 4/3, so the sums of element_force_alg reduce to
//   S0 = 4 dV I0 sigma_c(0,.),  S1 = 4 dV I1 sigma_c(1,.),
//   SH = 4/3 dV (I0 sigma_eta(0,.) + I1 sigma_xi(1,.)),   I0 = 2/sx, I1 = 2/sy, dV = sx sy / 4,
// the same 2x2 Gauss quadrature evaluated in closed form (f as in element_force_alg).
template <class T>
__device__ inline void element_force_rect(T I0, T I1, T dv, const double u[4][2], T k, T gmod, T f[4][2])
{
    T Au[2], Bu[2], Hu[2];
    element_modes(u, Au, Bu, Hu);
    const T g2 = T(2) * gmod, kd = fmaT(g2, T(2) / T(3), k); // k + 4 G / 3
    // centre: grad(i,j) = I_i du_j/dxi_i
    const T gxx = I0 * Au[0], gxy = I0 * Au[1], gyx = I1 * Bu[0], gyy = I1 * Bu[1];
    const T tr = gxx + gyy, kt = k * tr, t3 = tr * (T(1) / T(3));
    const T sxx = fmaT(g2, gxx - t3, kt), syy = fmaT(g2, gyy - t3, kt), sxy = gmod * (gxy + gyx);
    // eta slope (only grad(0,.) depends on eta) and xi slope (only grad(1,.))
    const T sxx_e = kd * (I0 * Hu[0]), sxy_e = gmod * (I0 * Hu[1]);
    const T syy_x = kd * (I1 * Hu[1]), sxy_x = gmod * (I1 * Hu[0]);
    const T c0 = T(4) * dv * I0, c1 = T(4) * dv * I1, ch = (T(4) / T(3)) * dv;
    const T S0[2] = {c0 * sxx, c0 * sxy};
    const T S1[2] = {c1 * sxy, c1 * syy};
    const T SH[2] = {ch * fmaT(I0, sxx_e, I1 * sxy_x), ch * fmaT(I0, sxy_e, I1 * syy_x)};
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const T p = T(0.25) * S0[j], r = T(0.25) * S1[j], h = T(0.25) * SH[j];
        f[0][j] = -p - r + h;
        f[1][j] = p - r - h;
        f[2][j] = p + r + h;
        f[3][j] = -p + r - h;
    }
}

// Element bodies of the fused step: the reference-order FP64 body (bit-exact), and the algebraic /
// closed-form rectangle bodies in FP32 or in FP64 with FMA.
enum Body : int { B_REF64 = 0, B_F32 = 1, B_F32_RECT = 2, B_F64 = 3, B_F64_RECT = 4 };
constexpr bool body_rect(int b) { return b == B_F32_RECT || b == B_F64_RECT; }
constexpr bool body_f32(int b) { return b == B_F32 || b == B_F32_RECT; }

template <int BODY, class T>
__device__ inline void element_force_body(const ShapeTable& Tb, T rI0, T rI1, T rdv, const double x[4][2],
                                          const double u[4][2], float k, float gmod, T f[4][2])
{
    if constexpr (BODY == B_REF64)
        element_force_f64(Tb, x, u, static_cast<double>(k), static_cast<double>(gmod), f);
    else if constexpr (body_rect(BODY))
        element_force_rect<T>(rI0, rI1, rdv, u, static_cast<T>(k), static_cast<T>(gmod), f);
    else
        element_force_alg<T>(Tb, x, u, static_cast<T>(k), static_cast<T>(gmod), f);
}

// ---- structured fused step ----------------------------------------------------------------------
template <int TX, int TY, int BODY>
__global__ void __launch_bounds__(TX * TY) k_vv_structured(
    int nx, int ny, int ntx, int nty, int per_xcd, ShapeTable T, const double* __restrict__ coor, double sx, double sy,
    double rI0, double rI1, double rdv,
    const float* __restrict__ Ke, float Kconst, const float* __restrict__ Ge, const double* __restrict__ inv_mass,
    const uint8_t* __restrict__ im_code, const double* __restrict__ im_tab,
    const double* __restrict__ fext, double dt, double c, double hdt, const double* __restrict__ u0,
    const double* __restrict__ v0, const double* __restrict__ a0, double* __restrict__ u1, double* __restrict__ v1,
    double* __restrict__ a1, double* __restrict__ fint_out)
{
    constexpr int HX = TX + 2, HY = TY + 2;   // halo node tile
    constexpr int EX = TX + 1, EY = TY + 1;   // element tile
    constexpr int NT = TX * TY;
    constexpr int NB = (EX * EY + NT - 1) / NT; // elements per thread
    constexpr bool RECT = body_rect(BODY);
    using FT = typename std::conditional<body_f32(BODY), float, double>::type;
    __shared__ double2 s_u[HY * HX];
    __shared__ double2 s_x[RECT ? 1 : HY * HX]; // the rectangle body needs no coordinates
    __shared__ FT s_f[EY * EX][8];
    __shared__ double s_tab[256]; // the M^-1 palette (when im_code is given)

    // XCD-aware tile order: hardware places block b on XCD b % 8; give each XCD a contiguous band
    const int b = blockIdx.x;
    const int tile = (b & 7) * per_xcd + (b >> 3);
    if (tile >= ntx * nty) return; // uniform for the whole block
    const int tx = tile % ntx, ty = tile / ntx;
    const int x0 = tx * TX, y0 = ty * TY;
    const int NX = nx + 1, NY = ny + 1;

    // phase A: updated displacement and coordinates of the node tile.  Each thread loads its own
    // node (whose v, a, M^-1 and f_ext stay in registers for phase C) and threads 0..NR-1 load
    // one node each of the one-node halo ring, so no node is fetched twice by one workgroup.
    constexpr int NR = 2 * HX + 2 * TY;
    static_assert(NR <= NT, "halo ring must fit one pass of the workgroup");
    const double2 zero = make_double2(0.0, 0.0);
    const int oly = threadIdx.x / TX, olx = threadIdx.x - (threadIdx.x / TX) * TX;
    const int ogx = x0 + olx, ogy = y0 + oly;
    const bool owner = ogx < NX && ogy < NY;
    const int64_t on = static_cast<int64_t>(ogy) * NX + ogx;
    double2 o_u = zero, o_v = zero, o_a = zero, o_x = zero, o_im = zero, o_fe = zero;
    uchar2 o_code = make_uchar2(0, 0);
    if (im_code)
        for (int t = threadIdx.x; t < 256; t += NT) s_tab[t] = im_tab[t];
    if (owner) {
        o_u = reinterpret_cast<const double2*>(u0)[on];
        o_v = reinterpret_cast<const double2*>(v0)[on];
        o_a = reinterpret_cast<const double2*>(a0)[on];
        o_x = coor ? reinterpret_cast<const double2*>(coor)[on]
                   : make_double2(static_cast<double>(ogx) * sx, static_cast<double>(ogy) * sy);
        if (im_code) o_code = reinterpret_cast<const uchar2*>(im_code)[on];
        else o_im = nt_load2(inv_mass, on);
        if (fext) o_fe = nt_load2(fext, on);
    }
    int rt = -1;
    double2 h_u = zero, h_v = zero, h_a = zero, h_x = zero;
    if (threadIdx.x < NR) {
        // ring order: bottom row, top row, left column, right column
        const int r = threadIdx.x;
        int lx, ly;
        if (r < HX) {
            lx = r;
            ly = 0;
        }
        else if (r < 2 * HX) {
            lx = r - HX;
            ly = HY - 1;
        }
        else if (r < 2 * HX + TY) {
            lx = 0;
            ly = 1 + r - 2 * HX;
        }
        else {
            lx = HX - 1;
            ly = 1 + r - 2 * HX - TY;
        }
        const int gx = x0 - 1 + lx, gy = y0 - 1 + ly;
        if (gx >= 0 && gx < NX && gy >= 0 && gy < NY) {
            const int64_t n = static_cast<int64_t>(gy) * NX + gx;
            rt = ly * HX + lx;
            h_u = reinterpret_cast<const double2*>(u0)[n];
            h_v = reinterpret_cast<const double2*>(v0)[n];
            h_a = reinterpret_cast<const double2*>(a0)[n];
            h_x = coor ? reinterpret_cast<const double2*>(coor)[n]
                       : make_double2(static_cast<double>(gx) * sx, static_cast<double>(gy) * sy);
        }
    }
    double2 o_un;
    o_un.x = o_u.x + dt * o_v.x + c * o_a.x;
    o_un.y = o_u.y + dt * o_v.y + c * o_a.y;
    if (owner) {
        const int t = (oly + 1) * HX + olx + 1;
        s_u[t] = o_un;
        if constexpr (!RECT) s_x[t] = o_x; // the rectangle body needs no coordinates
    }
    if (rt >= 0) {
        double2 hn;
        hn.x = h_u.x + dt * h_v.x + c * h_a.x;
        hn.y = h_u.y + dt * h_v.y + c * h_a.y;
        s_u[rt] = hn;
        if constexpr (!RECT) s_x[rt] = h_x;
    }
    __syncthreads();

    // phase B: element forces of the element tile
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const int t = threadIdx.x + k * NT;
        if (t >= EX * EY) break;
        const int ly = t / EX, lx = t - ly * EX;
        const int ex = x0 - 1 + lx, ey = y0 - 1 + ly;
        if (ex >= 0 && ex < nx && ey >= 0 && ey < ny) {
            const int64_t e = static_cast<int64_t>(ey) * nx + ex;
            // local node (lx, ly) is global (ex, ey); conn order 0:(0,0) 1:(1,0) 2:(1,1) 3:(0,1)
            const int n0 = ly * HX + lx;
            const int nid[4] = {n0, n0 + 1, n0 + HX + 1, n0 + HX};
            double x[4][2], u[4][2];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const double2 xx = RECT ? make_double2(0.0, 0.0) : s_x[nid[m]], uu = s_u[nid[m]];
                x[m][0] = xx.x;
                x[m][1] = xx.y;
                u[m][0] = uu.x;
                u[m][1] = uu.y;
            }
            FT f[4][2];
            element_force_body<BODY, FT>(T, static_cast<FT>(rI0), static_cast<FT>(rI1), static_cast<FT>(rdv), x, u,
                                         Ke ? Ke[e] : Kconst, Ge[e], f);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                s_f[t][2 * m] = f[m][0];
                s_f[t][2 * m + 1] = f[m][1];
            }
        }
    }
    __syncthreads();

    // phase C: owned node -- ordered assembly + velocity-Verlet update
    if (owner) {
        const int lx = olx, ly = oly, gx = ogx, gy = ogy;
        // elements in ascending number: (gx-1,gy-1) slot 2, (gx,gy-1) slot 3, (gx-1,gy) slot 1,
        // (gx,gy) slot 0; local element (lx', ly') = (ex - x0 + 1, ey - y0 + 1)
        double f0 = 0.0, f1 = 0.0;
        const bool hasL = gx > 0, hasR = gx < nx, hasB = gy > 0, hasT = gy < ny;
        if (hasB && hasL) {
            const int le = ly * EX + lx;
            f0 += static_cast<double>(s_f[le][4]);
            f1 += static_cast<double>(s_f[le][5]);
        }
        if (hasB && hasR) {
            const int le = ly * EX + lx + 1;
            f0 += static_cast<double>(s_f[le][6]);
            f1 += static_cast<double>(s_f[le][7]);
        }
        if (hasT && hasL) {
            const int le = (ly + 1) * EX + lx;
            f0 += static_cast<double>(s_f[le][2]);
            f1 += static_cast<double>(s_f[le][3]);
        }
        if (hasT && hasR) {
            const int le = (ly + 1) * EX + lx + 1;
            f0 += static_cast<double>(s_f[le][0]);
            f1 += static_cast<double>(s_f[le][1]);
        }
        const double2 un = o_un;
        if (im_code) o_im = make_double2(s_tab[o_code.x], s_tab[o_code.y]);
        double2 an, vn;
        an.x = o_im.x * (o_fe.x - f0);
        an.y = o_im.y * (o_fe.y - f1);
        vn.x = o_v.x + hdt * (o_a.x + an.x);
        vn.y = o_v.y + hdt * (o_a.y + an.y);
        // plain stores: the write-back L2 merges the halves of each 128-byte line written by
        // neighbouring lanes (non-temporal stores measured ~1.5x the compulsory write bytes)
        reinterpret_cast<double2*>(u1)[on] = un;
        reinterpret_cast<double2*>(v1)[on] = vn;
        reinterpret_cast<double2*>(a1)[on] = an;
        if (fint_out) reinterpret_cast<double2*>(fint_out)[on] = make_double2(f0, f1);
    }
}

// ---- general path --------------------------------------------------------------------------------
__global__ void k_update_u(int64_t n, double dt, double c, const double* __restrict__ u0, const double* __restrict__ v0,
                           const double* __restrict__ a0, double* __restrict__ u1)
{
    int64_t d = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (d < n) u1[d] = u0[d] + dt * v0[d] + c * a0[d];
}

// One thread per element: gathers the element's coordinates (nodevec) and updated displacement
// (dofval through dofs), recomputes the geometry from the coordinates and writes the element force
// (FP64 reference order, FP32, or FP64 with FMA) as an elemvec [nelem, 4, 2].  The connectivity row of
// an element is one 16-byte load; consecutive threads read consecutive rows.
template <int BODY>
__global__ void __launch_bounds__(BS) k_element_general(int64_t nelem, ShapeTable T, const int32_t* __restrict__ conn,
                                                        const int32_t* __restrict__ dofs, const double* __restrict__ coor,
                                                        const float* __restrict__ Ke, const float* __restrict__ Ge,
                                                        const double* __restrict__ u,
                                                        typename std::conditional<body_f32(BODY), float, double>::type* __restrict__ fe)
{
    using FT = typename std::conditional<body_f32(BODY), float, double>::type;
    const int64_t e = blockIdx.x * static_cast<int64_t>(BS) + threadIdx.x;
    if (e >= nelem) return;
    const int4 cn = reinterpret_cast<const int4*>(conn)[e];
    const int nodes[4] = {cn.x, cn.y, cn.z, cn.w};
    double x[4][2], ue[4][2];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const double2 xx = reinterpret_cast<const double2*>(coor)[nodes[m]];
        const int2 dd = reinterpret_cast<const int2*>(dofs)[nodes[m]];
        x[m][0] = xx.x;
        x[m][1] = xx.y;
        ue[m][0] = u[dd.x];
        ue[m][1] = u[dd.y];
    }
    FT f[4][2];
    element_force_body<BODY, FT>(T, FT(0), FT(0), FT(0), x, ue, Ke[e], Ge[e], f);
    typedef FT v4 __attribute__((ext_vector_type(4)));
    v4* out = reinterpret_cast<v4*>(fe + e * 8);
    out[0] = v4{f[0][0], f[0][1], f[1][0], f[1][1]};
    out[1] = v4{f[2][0], f[2][1], f[3][0], f[3][1]};
}

// Ordered per-DOF gather of the element forces (the reference's serial scatter order) and the
// velocity-Verlet update of v and a; unless this is the last step of the call, also the next
// step's displacement u_next = u + dt v_new + c a_new (the displacement update of the next step,
// same expression order as k_update_u).
template <class T>
__global__ void __launch_bounds__(BS) k_node_update(int64_t ndof, const int32_t* __restrict__ ptr,
                                                    const int32_t* __restrict__ idx, const T* __restrict__ fe,
                                                    const double* __restrict__ inv_mass, const double* __restrict__ fext,
                                                    double dt, double c, double hdt, const double* __restrict__ v0,
                                                    const double* __restrict__ a0, double* __restrict__ v1,
                                                    double* __restrict__ a1, const double* __restrict__ u,
                                                    double* __restrict__ u_next, double* __restrict__ fint_out)
{
    int64_t d = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (d >= ndof) return;
    const double s = ordered_gather_sum<4>(ptr[d], ptr[d + 1], idx, fe);
    double an = inv_mass[d] * ((fext ? fext[d] : 0.0) - s);
    double vn = v0[d] + hdt * (a0[d] + an);
    v1[d] = vn;
    a1[d] = an;
    if (u_next) u_next[d] = u[d] + dt * vn + c * an;
    if (fint_out) fint_out[d] = s;
}

// The general step fused per node block.  (0) The block's local nodes -- its own BS nodes, then the
// halo nodes of its elements -- are staged in LDS: coordinates and displacements, one contiguous list
// per block.  (1) The forces of its elements (halo elements, which touch two blocks, are computed by
// each) go to LDS, computed from the staged nodes through the local connectivity.  (2) Each thread
// sums its own node's contributions from LDS in the reference's order and updates the node's two
// DOFs as k_node_update does.  No element-force array and no per-DOF gather list cross HBM.
template <int BODY, bool IDENT>
__global__ void __launch_bounds__(BS) k_explicit_blocks(
    int nblk, int per_xcd, ShapeTable T, const int32_t* __restrict__ bkptr, const int32_t* __restrict__ blptr, const int32_t* __restrict__ bnloc,
    const int32_t* __restrict__ beptr, const uint16_t* __restrict__ blconn, const float* __restrict__ bK,
    const float* __restrict__ bG, const int32_t* __restrict__ dofs, const double* __restrict__ coor,
    const int32_t* __restrict__ bnptr, const uint16_t* __restrict__ bnent, const double* __restrict__ inv_mass,
    const uint8_t* __restrict__ im_code, const double* __restrict__ im_tab,
    const double* __restrict__ fext, double dt, double c, double hdt, const double* __restrict__ v0,
    const double* __restrict__ a0, double* __restrict__ v1, double* __restrict__ a1, const double* __restrict__ u,
    double* __restrict__ u_next, double* __restrict__ fint_out)
{
    using FT = typename std::conditional<body_f32(BODY), float, double>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double s_tab[256]; // the M^-1 palette (when im_code is given)
    // XCD-aware: workgroup b runs on XCD b mod 8; each XCD takes a contiguous range of patches, which
    // are spatial neighbours (bisection order), so halo nodes hit that XCD's L2
    const int pb = static_cast<int>(blockIdx.x % 8) * per_xcd + static_cast<int>(blockIdx.x / 8);
    if (pb >= nblk) return; // uniform per workgroup
    const int32_t lb = blptr[pb], nl = blptr[pb + 1] - lb;
    const int32_t eb = beptr[pb], ne = beptr[pb + 1] - eb;
    const int32_t kb = bkptr[pb], own = bkptr[pb + 1] - kb;
    FT* s_f = reinterpret_cast<FT*>(smem);                          // [ne][4][2] element forces
    double2* s_x = reinterpret_cast<double2*>(smem + ne * 8 * sizeof(FT)); // [nl] coordinates
    double2* s_u = s_x + nl;                                        // [nl] displacements
    const int64_t k = static_cast<int64_t>(kb) + threadIdx.x;
    const bool active = static_cast<int>(threadIdx.x) < own;
    // (0) local nodes; thread t's own node is local node t
    static_assert(BS == 256, "one palette entry per thread");
    if (im_code) s_tab[threadIdx.x] = im_tab[threadIdx.x];
    int2 dme = make_int2(0, 0);
    for (int base = 0; base < nl; base += 2 * BS) {
        // two local nodes per thread and pass, their loads issued together (indices clamped)
        const int l0 = base + threadIdx.x, l1 = l0 + BS;
        const int32_t n0 = bnloc[lb + min(l0, nl - 1)], n1 = bnloc[lb + min(l1, nl - 1)];
        const double2 x0 = reinterpret_cast<const double2*>(coor)[n0], x1 = reinterpret_cast<const double2*>(coor)[n1];
        int2 d0, d1;
        double2 u0, u1;
        if constexpr (IDENT) { // dofs(n, i) = 2 n + i: a node's displacement is one 16-byte load
            d0 = make_int2(2 * n0, 2 * n0 + 1);
            d1 = make_int2(2 * n1, 2 * n1 + 1);
            u0 = reinterpret_cast<const double2*>(u)[n0];
            u1 = reinterpret_cast<const double2*>(u)[n1];
        }
        else {
            d0 = reinterpret_cast<const int2*>(dofs)[n0];
            d1 = reinterpret_cast<const int2*>(dofs)[n1];
            u0 = make_double2(u[d0.x], u[d0.y]);
            u1 = make_double2(u[d1.x], u[d1.y]);
        }
        if (l0 < nl) {
            s_x[l0] = x0;
            s_u[l0] = u0;
        }
        if (l1 < nl) {
            s_x[l1] = x1;
            s_u[l1] = u1;
        }
        if (base == 0) dme = d0;
    }
    __syncthreads();
    // the own node's operands, in flight while the elements are computed
    const int d[2] = {dme.x, dme.y};
    double im[2] = {0.0, 0.0}, vo[2] = {0.0, 0.0}, ao[2] = {0.0, 0.0}, fx[2] = {0.0, 0.0};
    int32_t j0 = 0, j1 = 0;
    if (active) {
        j0 = bnptr[k];
        j1 = bnptr[k + 1];
        if constexpr (IDENT) {
            const int n = d[0] / 2;
            double2 im2;
            if (im_code) {
                const uchar2 c2 = reinterpret_cast<const uchar2*>(im_code)[n];
                im2 = make_double2(s_tab[c2.x], s_tab[c2.y]);
            }
            else
                im2 = reinterpret_cast<const double2*>(inv_mass)[n];
            const double2 vo2 = reinterpret_cast<const double2*>(v0)[n];
            const double2 ao2 = reinterpret_cast<const double2*>(a0)[n];
            im[0] = im2.x;
            im[1] = im2.y;
            vo[0] = vo2.x;
            vo[1] = vo2.y;
            ao[0] = ao2.x;
            ao[1] = ao2.y;
            if (fext) {
                const double2 f2 = reinterpret_cast<const double2*>(fext)[n];
                fx[0] = f2.x;
                fx[1] = f2.y;
            }
        }
        else {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                im[i] = im_code ? s_tab[im_code[d[i]]] : inv_mass[d[i]];
                vo[i] = v0[d[i]];
                ao[i] = a0[d[i]];
                fx[i] = fext ? fext[d[i]] : 0.0;
            }
        }
    }
    // (1) element forces from the staged nodes
    for (int le = threadIdx.x; le < ne; le += BS) {
        const ushort4 lc = reinterpret_cast<const ushort4*>(blconn)[eb + le];
        const int ln[4] = {lc.x, lc.y, lc.z, lc.w};
        double x[4][2], ue[4][2];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const double2 xx = s_x[ln[m]], uu = s_u[ln[m]];
            x[m][0] = xx.x;
            x[m][1] = xx.y;
            ue[m][0] = uu.x;
            ue[m][1] = uu.y;
        }
        FT f[4][2];
        element_force_body<BODY, FT>(T, FT(0), FT(0), FT(0), x, ue, bK[eb + le], bG[eb + le], f);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            s_f[le * 8 + m * 2] = f[m][0];
            s_f[le * 8 + m * 2 + 1] = f[m][1];
        }
    }
    __syncthreads();
    if (!active) return;
    // (2) the own node's ordered sums and update
    double s0 = 0.0, s1 = 0.0;
    for (int32_t j = j0; j < j1; j += 4) {
        int ent[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) ent[q] = bnent[min(j + q, j1 - 1)];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (j + q < j1) {
                const FT* f = s_f + (ent[q] >> 2) * 8 + (ent[q] & 3) * 2;
                s0 += static_cast<double>(f[0]);
                s1 += static_cast<double>(f[1]);
            }
    }
    const double2 uo = s_u[threadIdx.x];
    const double sum[2] = {s0, s1}, uu[2] = {uo.x, uo.y};
    double an[2], vn[2], un[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        an[i] = im[i] * (fx[i] - sum[i]);
        vn[i] = vo[i] + hdt * (ao[i] + an[i]);
        un[i] = uu[i] + dt * vn[i] + c * an[i];
    }
    if constexpr (IDENT) {
        const int n = d[0] / 2;
        reinterpret_cast<double2*>(v1)[n] = make_double2(vn[0], vn[1]);
        reinterpret_cast<double2*>(a1)[n] = make_double2(an[0], an[1]);
        if (u_next) reinterpret_cast<double2*>(u_next)[n] = make_double2(un[0], un[1]);
        if (fint_out) reinterpret_cast<double2*>(fint_out)[n] = make_double2(sum[0], sum[1]);
    }
    else {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            v1[d[i]] = vn[i];
            a1[d[i]] = an[i];
            if (u_next) u_next[d[i]] = un[i];
            if (fint_out) fint_out[d[i]] = sum[i];
        }
    }
}

// gather list entry k = (e nne + m) nd + i of the original element e -> the same entry of e's
// position rank[e] in the locality order
__global__ void k_remap_eidx(int64_t n, const int32_t* __restrict__ idx, const int32_t* __restrict__ rank,
                             int32_t* __restrict__ out)
{
    int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (j >= n) return;
    const int32_t k = idx[j];
    out[j] = rank[k / 8] * 8 + k % 8;
}

// Energies of the current state: Ekin = 1/2 sum_d m_d v_d^2 (the example's 0.5 sqdot(mass, V),
// Elastic-VelocityVerlet/main.cpp:10-19,159) and Epot = 1/2 sum_d u_d fint_d, which for the linear
// elastic law equals sum_q W_q dV_q with W = sigma : eps / 2 on the same quadrature.  Block partials
// (DPP wavefront sums, waves in order), then one block adds them in order: deterministic.
constexpr int NEB = 512;

__global__ void __launch_bounds__(BS) k_energy_partial(int64_t n, const double* __restrict__ mass,
                                                       const double* __restrict__ v, const double* __restrict__ u,
                                                       const double* __restrict__ f, double* __restrict__ part)
{
    double ek = 0.0, ep = 0.0;
    for (int64_t d = blockIdx.x * static_cast<int64_t>(BS) + threadIdx.x; d < n; d += static_cast<int64_t>(gridDim.x) * BS) {
        ek += mass[d] * v[d] * v[d];
        ep += u[d] * f[d];
    }
    ek = block_sum<BS>(ek);
    ep = block_sum<BS>(ep);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = ek;
        part[NEB + blockIdx.x] = ep;
    }
}

__global__ void __launch_bounds__(BS) k_energy_final(const double* __restrict__ part, double* __restrict__ ekin,
                                                     double* __restrict__ epot)
{
    double ek = 0.0, ep = 0.0;
    for (int i = threadIdx.x; i < NEB; i += BS) {
        ek += part[i];
        ep += part[NEB + i];
    }
    ek = block_sum<BS>(ek);
    ep = block_sum<BS>(ep);
    if (threadIdx.x == 0) {
        *ekin = 0.5 * ek;
        *epot = 0.5 * ep;
    }
}

ShapeTable gauss_table()
{
    ShapeTable T;
    const double a = 1.0 / std::sqrt(3.0);
    const double xi[4][2] = {{-a, -a}, {+a, -a}, {+a, +a}, {-a, +a}};
    for (int q = 0; q < 4; ++q) {
        const double x0 = xi[q][0], x1 = xi[q][1];
        T.g[q][0][0] = -0.25 * (1.0 - x1);
        T.g[q][1][0] = +0.25 * (1.0 - x1);
        T.g[q][2][0] = +0.25 * (1.0 + x1);
        T.g[q][3][0] = -0.25 * (1.0 + x1);
        T.g[q][0][1] = -0.25 * (1.0 - x0);
        T.g[q][1][1] = -0.25 * (1.0 + x0);
        T.g[q][2][1] = +0.25 * (1.0 + x0);
        T.g[q][3][1] = +0.25 * (1.0 - x0);
        T.w[q] = 1.0;
    }
    return T;
}

// Is conn the Mesh::Quad4::Regular(nx, ny) connectivity and dofs = Mesh::dofs?
bool detect_structured(const gf_vector_s* v, int64_t& nx, int64_t& ny)
{
    if (v->nne != 4 || v->ndim != 2 || v->nelem == 0) return false;
    const auto& c = v->conn_h;
    nx = c[3] - 1;
    if (nx < 1 || v->nelem % nx) return false;
    ny = v->nelem / nx;
    if ((nx + 1) * (ny + 1) != v->nnode) return false;
    for (int64_t ey = 0, e = 0; ey < ny; ++ey)
        for (int64_t ex = 0; ex < nx; ++ex, ++e) {
            int64_t n0 = ey * (nx + 1) + ex;
            if (c[4 * e] != n0 || c[4 * e + 1] != n0 + 1 || c[4 * e + 2] != n0 + nx + 2 || c[4 * e + 3] != n0 + nx + 1)
                return false;
        }
    for (int64_t i = 0; i < v->nnode * 2; ++i)
        if (v->dofs_h[i] != i) return false;
    return nx < (1 << 30) && ny < (1 << 30);
}

// Tile shape (nodes per workgroup).  Defaults from a sweep on MI355X at 2048x2048 elements:
// 32x16 for FP32 element bodies (112 us/step), 32x8 for FP64 (172 us/step; the FP64 body
// needs more registers, so the smaller tile keeps more waves resident).  GF_EXPLICIT_TILE=WxH
// ("64x8", "32x8", "32x16", "64x4", "16x16") overrides both.
struct TileCfg {
    int tx, ty;
};

// Node tile per workgroup (measured on MI355X at 2048^2): the closed-form rectangle bodies 64 x 8,
// the FP32 algebraic body 32 x 16, the FP64 bodies 32 x 8.  GF_EXPLICIT_TILE=<tx>x<ty> overrides.
TileCfg tile_config(int body)
{
    static const TileCfg over = [] {
        TileCfg c{0, 0};
        if (const char* e = std::getenv("GF_EXPLICIT_TILE")) {
            int a = 0, b = 0;
            if (std::sscanf(e, "%dx%d", &a, &b) == 2) c = TileCfg{a, b};
        }
        return c;
    }();
    if (over.tx > 0) return over;
    switch (body) {
        case B_F32_RECT:
        case B_F64_RECT: return TileCfg{64, 8};
        case B_F32: return TileCfg{32, 16};
        default: return TileCfg{32, 8};
    }
}

template <int TX, int TY, int BODY>
void launch_body(unsigned grid, hipStream_t s, int nx, int ny, int ntx, int nty, int per_xcd, const ShapeTable& T,
                 gf_explicit_s* x, const double* fext, double dt, double c, double hdt, int i0, int i1, double* fo)
{
    k_vv_structured<TX, TY, BODY><<<grid, TX * TY, 0, s>>>(
        nx, ny, ntx, nty, per_xcd, T, x->reg_x ? nullptr : x->coor.p, x->sx, x->sy, 2.0 / x->sx, 2.0 / x->sy,
        x->sx * x->sy / 4.0, x->k_uniform ? nullptr : x->Ke.p, x->k_value, x->Ge.p, x->inv_mass.p,
        x->im_code.n ? x->im_code.p : nullptr, x->im_tab.p, fext, dt, c, hdt, x->u[i0].p, x->vv[i0].p, x->a[i0].p,
        x->u[i1].p, x->vv[i1].p, x->a[i1].p, fo);
}

template <int TX, int TY>
void launch_structured(int body, unsigned grid, hipStream_t s, int nx, int ny, int ntx, int nty, int per_xcd,
                       const ShapeTable& T, gf_explicit_s* x, const double* fext, double dt, double c, double hdt,
                       int i0, int i1, double* fo)
{
    switch (body) {
        case B_REF64:
            return launch_body<TX, TY, B_REF64>(grid, s, nx, ny, ntx, nty, per_xcd, T, x, fext, dt, c, hdt, i0, i1, fo);
        case B_F32:
            return launch_body<TX, TY, B_F32>(grid, s, nx, ny, ntx, nty, per_xcd, T, x, fext, dt, c, hdt, i0, i1, fo);
        case B_F32_RECT:
            return launch_body<TX, TY, B_F32_RECT>(grid, s, nx, ny, ntx, nty, per_xcd, T, x, fext, dt, c, hdt, i0, i1,
                                                   fo);
        case B_F64:
            return launch_body<TX, TY, B_F64>(grid, s, nx, ny, ntx, nty, per_xcd, T, x, fext, dt, c, hdt, i0, i1, fo);
        default:
            return launch_body<TX, TY, B_F64_RECT>(grid, s, nx, ny, ntx, nty, per_xcd, T, x, fext, dt, c, hdt, i0, i1,
                                                   fo);
    }
}

// Node blocks of the fused general step (k_explicit_blocks).  Used when every DOF belongs to
// exactly one (node, component) -- a bijection onto [0, ndof) -- so that a DOF's contributions are
// those of one node; tied meshes keep the element kernel + per-DOF gather.  The blocks come from a
// recursive coordinate bisection of the nodes: a set is split across its longer extent with a
// multiple of BS nodes on the lower side, until it holds at most BS, so every block is a compact
// patch of BS nodes (one block may hold fewer).  Each block's nodes are then put in ascending
// number so that neighbouring threads touch neighbouring DOFs.  rank[e]: position of original
// element e in the locality order; cp, kp, gp: connectivity [4] and moduli in that order.  Blocks
// whose local nodes and element forces would not fit in 64 KiB of LDS keep the two-kernel path.
void build_node_blocks(gf_explicit_s* x, gf_vector v, const double* coor_d, const std::vector<int32_t>& rank,
                       const std::vector<int32_t>& cp, const std::vector<float>& kp, const std::vector<float>& gp)
{
    const int64_t nn = v->nnode, E = v->nelem, nd = v->ndof;
    if (nd != 2 * nn || nn == 0 || E == 0) return;
    {
        std::vector<char> seen(nd, 0);
        for (int64_t k = 0; k < 2 * nn; ++k) {
            const int64_t d = v->dofs_h[k];
            if (d < 0 || d >= nd || seen[d]) return;
            seen[d] = 1;
        }
    }
    std::vector<double> xy(2 * nn);
    GF_HIP(hipMemcpy(xy.data(), coor_d, xy.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<int32_t> node(nn);
    for (int64_t n = 0; n < nn; ++n) node[n] = static_cast<int32_t>(n);
    // recursive coordinate bisection, iteratively over a stack of node ranges
    std::vector<std::pair<int64_t, int64_t>> todo{{0, nn}};
    while (!todo.empty()) {
        const auto [b, e] = todo.back();
        todo.pop_back();
        const int64_t n = e - b;
        if (n <= BS) continue;
        double l[2] = {xy[2 * node[b]], xy[2 * node[b] + 1]}, h[2] = {l[0], l[1]};
        for (int64_t k = b; k < e; ++k)
            for (int i = 0; i < 2; ++i) {
                l[i] = std::min(l[i], xy[2 * node[k] + i]);
                h[i] = std::max(h[i], xy[2 * node[k] + i]);
            }
        const int dim = (h[0] - l[0] >= h[1] - l[1]) ? 0 : 1;
        const int64_t lower = std::max<int64_t>(BS, ((n + BS) / (2 * BS)) * BS); // a multiple of BS, < n
        std::nth_element(node.begin() + b, node.begin() + b + lower, node.begin() + e, [&](int32_t p, int32_t q) {
            const double xp = xy[2 * p + dim], xq = xy[2 * q + dim];
            return xp < xq || (xp == xq && p < q);
        });
        todo.push_back({b, b + lower});
        todo.push_back({b + lower, e});
    }
    // (e, m) of every node, e ascending then m: the reference's scatter order for each component
    std::vector<int64_t> nptr(nn + 1, 0);
    for (int64_t k = 0; k < 4 * E; ++k) ++nptr[v->conn_h[k] + 1];
    for (int64_t n = 0; n < nn; ++n) nptr[n + 1] += nptr[n];
    std::vector<int32_t> nel(4 * E);
    {
        std::vector<int64_t> fill(nptr.begin(), nptr.end() - 1);
        for (int64_t k = 0; k < 4 * E; ++k) nel[fill[v->conn_h[k]]++] = static_cast<int32_t>(k); // k = e * 4 + m
    }
    std::vector<int32_t> bkptr{0}, beptr{0}, blptr{0}, bnloc, bnptr(nn + 1, 0);
    std::vector<uint16_t> bnent, blconn;
    std::vector<float> bK, bG;
    bnent.reserve(4 * E);
    blconn.reserve(4 * (E + E / 4));
    bnloc.reserve(nn + nn / 3);
    int64_t lds64 = 0, lds32 = 0;
    std::vector<int32_t> el, halo, lidx(nn, -1);
    // node[k0, k1) as one block, or bisected again while its local nodes and element forces need more
    // than LDS_CAP bytes (FP64); blocks are emitted in node-array order
    constexpr int64_t LDS_CAP = 40960;
    std::function<bool(int64_t, int64_t)> emit = [&](int64_t k0, int64_t k1) -> bool {
        el.clear();
        for (int64_t k = k0; k < k1; ++k)
            for (int64_t j = nptr[node[k]]; j < nptr[node[k] + 1]; ++j) el.push_back(rank[nel[j] / 4]);
        std::sort(el.begin(), el.end());
        el.erase(std::unique(el.begin(), el.end()), el.end());
        for (int64_t k = k0; k < k1; ++k) lidx[node[k]] = 0;
        halo.clear();
        for (int32_t r : el)
            for (int m = 0; m < 4; ++m)
                if (lidx[cp[4 * r + m]] < 0) halo.push_back(cp[4 * r + m]);
        for (int64_t k = k0; k < k1; ++k) lidx[node[k]] = -1;
        std::sort(halo.begin(), halo.end());
        halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
        const int64_t own = k1 - k0, nl = own + static_cast<int64_t>(halo.size()), ne = static_cast<int64_t>(el.size());
        if (nl * 32 + ne * 64 > LDS_CAP && own > 1) {
            double l[2] = {xy[2 * node[k0]], xy[2 * node[k0] + 1]}, h[2] = {l[0], l[1]};
            for (int64_t k = k0; k < k1; ++k)
                for (int i = 0; i < 2; ++i) {
                    l[i] = std::min(l[i], xy[2 * node[k] + i]);
                    h[i] = std::max(h[i], xy[2 * node[k] + i]);
                }
            const int dim = (h[0] - l[0] >= h[1] - l[1]) ? 0 : 1;
            const int64_t mid = k0 + own / 2;
            std::nth_element(node.begin() + k0, node.begin() + mid, node.begin() + k1, [&](int32_t p, int32_t q) {
                const double xp = xy[2 * p + dim], xq = xy[2 * q + dim];
                return xp < xq || (xp == xq && p < q);
            });
            return emit(k0, mid) && emit(mid, k1);
        }
        if (nl > 65535 || nl * 32 + ne * 64 > 65536 - 2048) return false; // beyond LDS (with the 2 KiB palette)
        // own nodes in ascending number (local number k - k0), then the halo nodes, ascending
        std::sort(node.begin() + k0, node.begin() + k1);
        for (int64_t k = k0; k < k1; ++k) lidx[node[k]] = static_cast<int32_t>(k - k0);
        for (size_t h = 0; h < halo.size(); ++h) lidx[halo[h]] = static_cast<int32_t>(own + h);
        lds64 = std::max(lds64, nl * 32 + ne * 64);
        lds32 = std::max(lds32, nl * 32 + ne * 32);
        bnloc.insert(bnloc.end(), node.begin() + k0, node.begin() + k1);
        bnloc.insert(bnloc.end(), halo.begin(), halo.end());
        blptr.push_back(static_cast<int32_t>(bnloc.size()));
        for (int32_t r : el) {
            for (int m = 0; m < 4; ++m) blconn.push_back(static_cast<uint16_t>(lidx[cp[4 * r + m]]));
            bK.push_back(kp[r]);
            bG.push_back(gp[r]);
        }
        beptr.push_back(beptr.back() + static_cast<int32_t>(ne));
        bkptr.push_back(static_cast<int32_t>(k1));
        for (int64_t k = k0; k < k1; ++k) {
            for (int64_t j = nptr[node[k]]; j < nptr[node[k] + 1]; ++j) {
                const int32_t e = nel[j] / 4, m = nel[j] % 4;
                const int32_t slot =
                    static_cast<int32_t>(std::lower_bound(el.begin(), el.end(), rank[e]) - el.begin());
                bnent.push_back(static_cast<uint16_t>(slot * 4 + m));
            }
            bnptr[k + 1] = static_cast<int32_t>(bnent.size());
        }
        for (int64_t k = k0; k < k1; ++k) lidx[node[k]] = -1;
        for (int32_t n : halo) lidx[n] = -1;
        return true;
    };
    for (int64_t k0 = 0; k0 < nn; k0 += BS)
        if (!emit(k0, std::min(nn, k0 + BS))) return;
    x->bkptr.upload(bkptr);
    x->blptr.upload(blptr);
    x->bnloc.upload(bnloc);
    x->beptr.upload(beptr);
    x->blconn.upload(blconn);
    x->bK.upload(bK);
    x->bG.upload(bG);
    x->bnptr.upload(bnptr);
    x->bnent.upload(bnent);
    x->nblk = static_cast<int64_t>(bkptr.size()) - 1;
    x->dofs_ident = true;
    for (int64_t k = 0; k < 2 * nn && x->dofs_ident; ++k) x->dofs_ident = v->dofs_h[k] == k;
    x->lds64 = lds64;
    x->lds32 = lds32;
    x->blocks = true;
}

} // namespace

extern "C" {

int gf_explicit_create(gf_explicit* out, gf_vector v, const double* coor, const float* K_e, const float* G_e,
                       const double* inv_mass, const double* fext, double dt)
{
    return guard([&] {
        GF_REQUIRE(out && v && coor && K_e && G_e && inv_mass, "null argument");
        GF_REQUIRE(v->nne == 4 && v->ndim == 2, "explicit integrator needs a Quad4 mesh (nne 4, ndim 2)");
        auto x = new gf_explicit_s();
        try {
            x->v = v;
            x->ndof = v->ndof;
            x->nelem = v->nelem;
            x->dt = dt;
            x->coor.alloc(v->nnode * 2);
            GF_HIP(hipMemcpy(x->coor.p, coor, x->coor.n * sizeof(double), hipMemcpyDeviceToDevice));
            x->Ke.alloc(v->nelem);
            x->Ge.alloc(v->nelem);
            GF_HIP(hipMemcpy(x->Ke.p, K_e, v->nelem * sizeof(float), hipMemcpyDeviceToDevice));
            {
                std::vector<float> kh(v->nelem);
                if (v->nelem) GF_HIP(hipMemcpy(kh.data(), K_e, v->nelem * sizeof(float), hipMemcpyDeviceToHost));
                x->k_uniform = v->nelem > 0 && std::all_of(kh.begin(), kh.end(), [&](float z) { return z == kh[0]; });
                x->k_value = v->nelem ? kh[0] : 0.0f;
            }
            GF_HIP(hipMemcpy(x->Ge.p, G_e, v->nelem * sizeof(float), hipMemcpyDeviceToDevice));
            x->inv_mass.alloc(v->ndof);
            GF_HIP(hipMemcpy(x->inv_mass.p, inv_mass, v->ndof * sizeof(double), hipMemcpyDeviceToDevice));
            {
                // palette of M^-1 (bit patterns, so the table returns exactly the given values)
                std::vector<double> mh(v->ndof);
                if (v->ndof) GF_HIP(hipMemcpy(mh.data(), inv_mass, v->ndof * sizeof(double), hipMemcpyDeviceToHost));
                std::vector<uint64_t> keys;
                std::unordered_map<uint64_t, uint8_t> index;
                std::vector<uint8_t> code(v->ndof);
                bool ok = v->ndof > 0;
                for (int64_t d = 0; ok && d < v->ndof; ++d) {
                    uint64_t bits;
                    std::memcpy(&bits, &mh[d], sizeof(bits));
                    auto it = index.find(bits);
                    if (it == index.end()) {
                        if (keys.size() == 256) {
                            ok = false;
                            break;
                        }
                        it = index.emplace(bits, static_cast<uint8_t>(keys.size())).first;
                        keys.push_back(bits);
                    }
                    code[d] = it->second;
                }
                if (ok) {
                    std::vector<double> tab(256, 0.0);
                    for (size_t j = 0; j < keys.size(); ++j) std::memcpy(&tab[j], &keys[j], sizeof(double));
                    x->im_code.upload(code);
                    x->im_tab.upload(tab);
                }
            }
            if (fext) {
                x->fext.alloc(v->ndof);
                GF_HIP(hipMemcpy(x->fext.p, fext, v->ndof * sizeof(double), hipMemcpyDeviceToDevice));
                x->has_fext = true;
            }
            for (int k = 0; k < 2; ++k) {
                x->u[k].alloc(v->ndof);
                x->vv[k].alloc(v->ndof);
                x->a[k].alloc(v->ndof);
            }
            for (auto* b : {&x->u[0], &x->vv[0], &x->a[0]})
                GF_HIP(hipMemset(b->p, 0, v->ndof * sizeof(double)));
            x->fint.alloc(v->ndof);
            x->structured = detect_structured(v, x->nx, x->ny);
            if (x->structured) {
                // Regular coordinates (i sx, j sy) bit for bit, as the Regular mesh builds them:
                // the kernel then generates them instead of streaming 16 bytes per node
                const int64_t N1 = x->nx + 1, nn = N1 * (x->ny + 1);
                std::vector<double> ch(static_cast<size_t>(nn) * 2);
                GF_HIP(hipMemcpy(ch.data(), coor, ch.size() * sizeof(double), hipMemcpyDeviceToHost));
                const double sx = nn > 1 ? ch[2] : 0.0, sy = x->ny > 0 ? ch[2 * N1 + 1] : 0.0;
                bool reg = sx > 0.0 && sy > 0.0;
                for (int64_t n = 0; reg && n < nn; ++n) {
                    const int64_t i = n % N1, j = n / N1;
                    reg = ch[2 * n] == static_cast<double>(i) * sx && ch[2 * n + 1] == static_cast<double>(j) * sy;
                }
                x->reg_x = reg;
                x->sx = sx;
                x->sy = sy;
            }
            if (!x->structured) {
                GF_REQUIRE(v->nelem < (int64_t(1) << 28), "general path: nelem < 2^28");
                const int64_t E = v->nelem;
                // locality order: elements sorted by their lowest node (stable)
                std::vector<int32_t> order(E), rank(E), cp(E * 4);
                std::vector<int64_t> key(E);
                for (int64_t e = 0; e < E; ++e) {
                    order[e] = static_cast<int32_t>(e);
                    key[e] = *std::min_element(v->conn_h.begin() + 4 * e, v->conn_h.begin() + 4 * e + 4);
                }
                std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
                for (int64_t j = 0; j < E; ++j) {
                    rank[order[j]] = static_cast<int32_t>(j);
                    for (int m = 0; m < 4; ++m) cp[4 * j + m] = static_cast<int32_t>(v->conn_h[4 * order[j] + m]);
                }
                x->gconn.upload(cp);
                std::vector<float> kh(E), gh(E), kp(E), gp(E);
                GF_HIP(hipMemcpy(kh.data(), K_e, E * sizeof(float), hipMemcpyDeviceToHost));
                GF_HIP(hipMemcpy(gh.data(), G_e, E * sizeof(float), hipMemcpyDeviceToHost));
                for (int64_t j = 0; j < E; ++j) {
                    kp[j] = kh[order[j]];
                    gp[j] = gh[order[j]];
                }
                x->gK.upload(kp);
                x->gG.upload(gp);
                build_node_blocks(x, v, coor, rank, cp, kp, gp);
                if (!x->blocks) {
                    DBuf<int32_t> drank;
                    drank.upload(rank);
                    const int64_t nidx = E * 8;
                    x->geidx.alloc(nidx);
                    k_remap_eidx<<<nblocks(nidx, BS), BS>>>(nidx, v->dof_eidx.p, drank.p, x->geidx.p);
                    check_launch();
                    GF_HIP(hipDeviceSynchronize());
                    x->fe64.alloc(E * 8);
                    x->fe32.alloc(E * 8);
                }
            }
            GF_HIP(hipDeviceSynchronize());
        }
        catch (...) {
            delete x;
            throw;
        }
        *out = x;
    });
}

int gf_explicit_destroy(gf_explicit x)
{
    delete x;
    return GF_OK;
}

int gf_explicit_state(gf_explicit x, double** u, double** v, double** a)
{
    return guard([&] {
        GF_REQUIRE(x, "null integrator");
        if (u) *u = x->u[x->cur].p;
        if (v) *v = x->vv[x->cur].p;
        if (a) *a = x->a[x->cur].p;
    });
}

int gf_explicit_path(gf_explicit x, int* path)
{
    return guard([&] {
        GF_REQUIRE(x && path, "null argument");
        *path = x->structured ? 0 : (x->blocks ? 1 : 2);
    });
}

int gf_explicit_is_structured(gf_explicit x, int* structured)
{
    return guard([&] {
        GF_REQUIRE(x && structured, "null argument");
        *structured = x->structured ? 1 : 0;
    });
}

int gf_explicit_fint(gf_explicit x, const double** fint)
{
    return guard([&] {
        GF_REQUIRE(x && fint, "null argument");
        if (!x->fint_valid) throw Error(GF_ESTATE, "no step run with store_fint");
        *fint = x->fint.p;
    });
}

int gf_explicit_energy(gf_explicit x, const double* mass, double* ekin, double* epot, void* stream)
{
    return guard([&] {
        GF_REQUIRE(x && mass && ekin && epot, "null argument");
        if (!x->fint_valid) throw Error(GF_ESTATE, "no step run with store_fint: f_int of the current state unknown");
        if (x->epart.n != static_cast<size_t>(2 * NEB)) x->epart.alloc(2 * NEB);
        hipStream_t s = S(stream);
        k_energy_partial<<<NEB, BS, 0, s>>>(x->ndof, mass, x->vv[x->cur].p, x->u[x->cur].p, x->fint.p, x->epart.p);
        k_energy_final<<<1, BS, 0, s>>>(x->epart.p, ekin, epot);
        check_launch();
    });
}

int gf_explicit_step(gf_explicit x, int64_t nsteps, int precision, int store_fint, void* stream)
{
    return guard([&] {
        GF_REQUIRE(x, "null integrator");
        GF_REQUIRE(precision == GF_F64 || precision == GF_F32 || precision == GF_F64_FAST,
                   "precision must be GF_F64, GF_F32 or GF_F64_FAST");
        GF_REQUIRE(nsteps >= 0, "nsteps >= 0");
        hipStream_t s = S(stream);
        const double dt = x->dt, c = 0.5 * dt * dt, hdt = 0.5 * dt;
        const double* fext = x->has_fext ? x->fext.p : nullptr;
        static const ShapeTable T = gauss_table();
        if (x->structured) {
            const int body = precision == GF_F64 ? B_REF64
                             : precision == GF_F32 ? (x->reg_x ? B_F32_RECT : B_F32)
                                                   : (x->reg_x ? B_F64_RECT : B_F64);
            const TileCfg tc = tile_config(body);
            const int TXr = tc.tx, TYr = tc.ty;
            const int nx = static_cast<int>(x->nx), ny = static_cast<int>(x->ny);
            const int ntx = (nx + 1 + TXr - 1) / TXr, nty = (ny + 1 + TYr - 1) / TYr;
            const int ntiles = ntx * nty, per_xcd = (ntiles + 7) / 8;
            const unsigned grid = static_cast<unsigned>(per_xcd * 8);
            for (int64_t k = 0; k < nsteps; ++k) {
                const int i0 = x->cur, i1 = 1 - x->cur;
                double* fo = (store_fint && k == nsteps - 1) ? x->fint.p : nullptr;
                if (TXr == 32 && TYr == 8)
                    launch_structured<32, 8>(body, grid, s, nx, ny, ntx, nty, per_xcd, T, x, fext, dt, c, hdt, i0, i1, fo);
                else if (TXr == 64 && TYr == 8)
                    launch_structured<64, 8>(body, grid, s, nx, ny, ntx, nty, per_xcd, T, x, fext, dt, c, hdt, i0, i1, fo);
                else if (TXr == 32 && TYr == 16)
                    launch_structured<32, 16>(body, grid, s, nx, ny, ntx, nty, per_xcd, T, x, fext, dt, c, hdt, i0, i1, fo);
                else if (TXr == 64 && TYr == 4)
                    launch_structured<64, 4>(body, grid, s, nx, ny, ntx, nty, per_xcd, T, x, fext, dt, c, hdt, i0, i1, fo);
                else if (TXr == 16 && TYr == 16)
                    launch_structured<16, 16>(body, grid, s, nx, ny, ntx, nty, per_xcd, T, x, fext, dt, c, hdt, i0, i1, fo);
                else
                    throw Error(GF_EINVAL, "GF_EXPLICIT_TILE must be one of 32x8, 64x8, 32x16, 64x4, 16x16");
                check_launch();
                x->cur = i1;
            }
        }
        else {
            auto v = x->v;
            const unsigned ge = nblocks(x->nelem, BS), gd = nblocks(x->ndof, BS);
            // the displacement of the first step; each node update then writes the next one
            if (nsteps > 0)
                k_update_u<<<gd, BS, 0, s>>>(x->ndof, dt, c, x->u[x->cur].p, x->vv[x->cur].p, x->a[x->cur].p,
                                            x->u[1 - x->cur].p);
            for (int64_t k = 0; k < nsteps && x->blocks; ++k) {
                const int i0 = x->cur, i1 = 1 - x->cur;
                const bool last = k == nsteps - 1;
                double* fo = (store_fint && last) ? x->fint.p : nullptr;
                const double* un = x->u[i1].p; // this step's displacement
                double* unext = last ? nullptr : x->u[i0].p;
                const int per_xcd = static_cast<int>((x->nblk + 7) / 8);
                auto run = [&](auto body) {
                    constexpr int B = decltype(body)::value;
                    const size_t lds = static_cast<size_t>(body_f32(B) ? x->lds32 : x->lds64);
                    auto kern = x->dofs_ident ? k_explicit_blocks<B, true> : k_explicit_blocks<B, false>;
                    kern<<<8 * per_xcd, BS, lds, s>>>(static_cast<int>(x->nblk), per_xcd, T, x->bkptr.p, x->blptr.p, x->bnloc.p, x->beptr.p,
                                                            x->blconn.p, x->bK.p, x->bG.p, v->dofs.p, x->coor.p,
                                                            x->bnptr.p, x->bnent.p, x->inv_mass.p,
                                                            x->im_code.n ? x->im_code.p : nullptr, x->im_tab.p, fext, dt, c, hdt,
                                                            x->vv[i0].p, x->a[i0].p, x->vv[i1].p, x->a[i1].p, un, unext,
                                                            fo);
                };
                if (precision == GF_F32) run(std::integral_constant<int, B_F32>());
                else if (precision == GF_F64) run(std::integral_constant<int, B_REF64>());
                else run(std::integral_constant<int, B_F64>());
                check_launch();
                x->cur = i1;
            }
            for (int64_t k = 0; k < nsteps && !x->blocks; ++k) {
                const int i0 = x->cur, i1 = 1 - x->cur;
                const bool last = k == nsteps - 1;
                double* fo = (store_fint && last) ? x->fint.p : nullptr;
                const double* un = x->u[i1].p; // this step's displacement
                double* unext = last ? nullptr : x->u[i0].p;
                if (precision == GF_F32) {
                    k_element_general<B_F32><<<ge, BS, 0, s>>>(x->nelem, T, x->gconn.p, v->dofs.p, x->coor.p, x->gK.p,
                                                                x->gG.p, un, x->fe32.p);
                    k_node_update<float><<<gd, BS, 0, s>>>(x->ndof, v->dof_eptr.p, x->geidx.p, x->fe32.p,
                                                           x->inv_mass.p, fext, dt, c, hdt, x->vv[i0].p, x->a[i0].p,
                                                           x->vv[i1].p, x->a[i1].p, un, unext, fo);
                }
                else {
                    if (precision == GF_F64)
                        k_element_general<B_REF64><<<ge, BS, 0, s>>>(x->nelem, T, x->gconn.p, v->dofs.p, x->coor.p,
                                                                      x->gK.p, x->gG.p, un, x->fe64.p);
                    else
                        k_element_general<B_F64><<<ge, BS, 0, s>>>(x->nelem, T, x->gconn.p, v->dofs.p, x->coor.p,
                                                                    x->gK.p, x->gG.p, un, x->fe64.p);
                    k_node_update<double><<<gd, BS, 0, s>>>(x->ndof, v->dof_eptr.p, x->geidx.p, x->fe64.p,
                                                            x->inv_mass.p, fext, dt, c, hdt, x->vv[i0].p, x->a[i0].p,
                                                            x->vv[i1].p, x->a[i1].p, un, unext, fo);
                }
                check_launch();
                x->cur = i1;
            }
        }
        if (store_fint && nsteps > 0) x->fint_valid = true;
    });
}

} // extern "C"
