// FP64 Jacobi-preconditioned conjugate gradients on the tied operator of
// MatrixPartitionedTyings (include/GooseFEM/MatrixPartitionedTyings.h:655-674, Q9):
//   K' = A_uu + A_ud C_du + C_du^T A_du + C_du^T A_dd C_du        (ACuu)
//   K'_up x_p = (A_up + A_ud C_dp + C_du^T A_dp + C_du^T A_dd C_dp) x_p   (ACup)
// The reference factorises ACuu with Eigen's SimplicialLDLT (MatrixPartitionedTyings.h:527-675);
// here the operator is applied matrix-free, for a linear elastic material:
//   expand  x = [x_u, x_p, C_du x_u + C_dp x_p]           ([u, p, d] numbering of Tyings::Periodic)
//   element f_e = Int_gradN_dot_tensor2_dV(sigma(SymGradN_vector(AsElement(x))))   (FP64, reference order)
//   gather  y = AssembleDofs(f)                          (ordered DOF-centric gather, Vector.h:723-741)
//   fold    out_u = y_u + C_du^T y_d
// Reductions are two-stage and deterministic: every block sums its grid-stride range in a fixed
// order, its wavefronts with DPP lane moves and its waves in order; consumers reduce the block
// partials the same way.  The iteration
// runs with no host synchronisation: scalars live on the device and a device flag stops the
// iteration (kernels launched after convergence return immediately).  On the structured Quad4 cell
// the iteration is the single-reduction (Chronopoulos-Gear) form in one launch; elsewhere the
// standard two-reduction form.
#include <algorithm>
#include <cmath>
#include <tuple>
#include <utility>

#include "gf_internal.hpp"
#include "gf_wave.hpp"

using namespace gf;

namespace gf {
const double* quad_dNx(gf_quad q);
const double* quad_vol(gf_quad q);
int64_t quad_nelem(gf_quad q);
int quad_nip(gf_quad q);
int quad_etype(gf_quad q);
} // namespace gf

namespace {

constexpr int BS = 256;
constexpr int NBR = 512; // blocks of the reduction kernels (fixed: partial count)

struct Csr {
    DBuf<int32_t> ptr, idx;
    DBuf<double> val;
};

} // namespace

struct gf_pcg_s {
    gf_vector v = nullptr;
    gf_quad q = nullptr;
    int et = 0, nne = 0, nd = 0, K = 0;
    int64_t nelem = 0, nnu = 0, nnp = 0, nnd = 0, nni = 0, ndof = 0;
    DBuf<double> Ke, Ge;
    Csr cdu, cdp, cduT;
    DBuf<double> xfull, fe, invdiag, x, r, z, qv, tmp;
    DBuf<double> partial;  // [4][NBR]
    DBuf<double> scal;     // [8]: rz (2 slots), bb, rr, tol2, ...
    DBuf<int32_t> state;   // [0] done, [1] iterations, [2] maxiter
    // structured periodic Quad4 operator (Regular n x n cell, periodic tyings, uniform geometry)
    bool structured = false;
    int sn = 0, ntiles = 0, ntx = 0, per_xcd = 0;
    double sh_g[32] = {}, sh_v[4] = {};
    DBuf<int32_t> sidx; // [n^d * nd] u-DOF of grid node component c, or -1
    int nty = 0;
    DBuf<double> ke8;   // Hex8: [2][24][24] element matrices per unit K and per unit G
    DBuf<double> sr[2], sw[2], ss[2]; // Quad4 single-reduction iteration: r, w, s by iteration parity
    DBuf<double> gp[2];               // general operator: the search direction by iteration parity
    DBuf<double> spart;               // its block partials [2][3][ntiles]
    DBuf<double> xprev; // the last converged solution (the previous-step start of the fallback)
    bool has_prev = false;
    int last_start = 0; // start of the last solve: GF_START_GIVEN / _PREVIOUS / _ZERO
};

namespace {

// block reduction of one value per thread (block of NTH threads) into partial[blockIdx]: DPP
// wavefront sums, then the waves in order (deterministic)
template <int NTH>
__device__ inline void block_reduce_store_n(double v, double* partial)
{
    const double t = block_sum<NTH>(v);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// sum of `count` partials in fixed order (every block computes the same value)
__device__ inline double reduce_partials(const double* partial, int count = NBR)
{
    double v = 0.0;
    for (int i = threadIdx.x; i < count; i += BS) v += partial[i];
    return block_sum<BS>(v);
}

// ---- one Gauss point of the plane-strain Quad4 element force, FP64, reference expression order ----
// ElementQuad4Planar.h:191-221 (strain), the linear elastic stress, :253-278 (force); g = dN/dx of
// the point [4][2], v = its dV; f accumulates
__device__ inline void quad4_planar_qp_force(const double* g, double v, const double u[4][2], double k, double gm,
                                             double f[4][2])
{
    double s[2][2];
    double e00 = g[0] * u[0][0] + g[2] * u[1][0] + g[4] * u[2][0] + g[6] * u[3][0];
    double e11 = g[1] * u[0][1] + g[3] * u[1][1] + g[5] * u[2][1] + g[7] * u[3][1];
    double e01 = 0.5 * (g[0] * u[0][1] + g[2] * u[1][1] + g[4] * u[2][1] + g[6] * u[3][1] +
                        g[1] * u[0][0] + g[3] * u[1][0] + g[5] * u[2][0] + g[7] * u[3][0]);
    double e22 = 0.0;
    double tr = e00 + e11 + e22;
    s[0][0] = k * tr + 2.0 * gm * (e00 - tr / 3.0);
    s[1][1] = k * tr + 2.0 * gm * (e11 - tr / 3.0);
    s[0][1] = 2.0 * gm * e01;
    s[1][0] = 2.0 * gm * e01;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        f[m][0] += (g[2 * m] * s[0][0] + g[2 * m + 1] * s[1][0]) * v;
        f[m][1] += (g[2 * m] * s[0][1] + g[2 * m + 1] * s[1][1]) * v;
    }
}

// ---- structured periodic operator: Quad4 Regular n x n cell with periodic tyings ---------------------
// Every element has the same dN/dx and dV (Regular mesh), so they are kernel arguments.  The tied
// operator K' x_u on a periodic field is the element-force assembly on the torus: node (i, j)
// collects the elements (i-1, j-1), (i, j-1), (i-1, j), (i, j) taken modulo n (MatrixPartitionedTyings
// condensation with C_du = periodic copy, MatrixPartitionedTyings.h:655-674).  sidx maps grid node
// (i, j), component c to its u-DOF, or -1 for a prescribed DOF (the origin), whose value is 0.
struct ShapeQ4 {
    double g[4][8];
    double v[4];
};

template <int TX, int TY>
__global__ void __launch_bounds__(TX * TY) k_apply_periodic_q4(int n, int ntx, int per_xcd, ShapeQ4 sh,
                                                               const int32_t* __restrict__ sidx,
                                                               const double* __restrict__ Ke,
                                                               const double* __restrict__ Ge,
                                                               const double* __restrict__ x, double* __restrict__ y,
                                                               double* __restrict__ partial,
                                                               const int32_t* __restrict__ state_c)
{
    if (state_c && state_c[0]) return; // uniform for the grid
    constexpr int HX = TX + 2, HY = TY + 2, EX = TX + 1, EY = TY + 1, NT = TX * TY;
    __shared__ double2 s_x[HY * HX];
    __shared__ double s_f[EY * EX][8];
    const int b = blockIdx.x;
    const int tile = per_xcd > 0 ? (b & 7) * per_xcd + (b >> 3) : b;
    const int tx = tile % ntx, ty = tile / ntx;
    const int x0 = tx * TX, y0 = ty * TY;
    // halo tile of the operand (periodic wrap, prescribed DOFs read as 0)
    for (int t = threadIdx.x; t < HX * HY; t += NT) {
        const int ly = t / HX, lx = t - (t / HX) * HX;
        int gi = x0 - 1 + lx, gj = y0 - 1 + ly;
        gi += gi < 0 ? n : 0;
        gi -= gi >= n ? n : 0;
        gj += gj < 0 ? n : 0;
        gj -= gj >= n ? n : 0;
        const int64_t k = (static_cast<int64_t>(gj) * n + gi) * 2;
        const int32_t d0 = sidx[k], d1 = sidx[k + 1];
        s_x[t] = make_double2(d0 >= 0 ? x[d0] : 0.0, d1 >= 0 ? x[d1] : 0.0);
    }
    __syncthreads();
    // element forces of the (TX+1) x (TY+1) elements touching the tile
    for (int t = threadIdx.x; t < EX * EY; t += NT) {
        const int ly = t / EX, lx = t - (t / EX) * EX;
        int ei = x0 - 1 + lx, ej = y0 - 1 + ly;
        ei += ei < 0 ? n : 0;
        ej += ej < 0 ? n : 0;
        const int64_t e = static_cast<int64_t>(ej) * n + ei;
        const int n0 = ly * HX + lx;
        const int nid[4] = {n0, n0 + 1, n0 + HX + 1, n0 + HX};
        double u[4][2], f[4][2];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const double2 xx = s_x[nid[m]];
            u[m][0] = xx.x;
            u[m][1] = xx.y;
            f[m][0] = f[m][1] = 0.0;
        }
        const double k = Ke[e], gm = Ge[e];
#pragma unroll
        for (int q = 0; q < 4; ++q) quad4_planar_qp_force(sh.g[q], sh.v[q], u, k, gm, f);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            s_f[t][2 * m] = f[m][0];
            s_f[t][2 * m + 1] = f[m][1];
        }
    }
    __syncthreads();
    // owned node: elements (i-1, j-1) as its node 2, (i, j-1) node 3, (i-1, j) node 1, (i, j) node 0
    const int oly = threadIdx.x / TX, olx = threadIdx.x - (threadIdx.x / TX) * TX;
    const int l0 = oly * EX + olx;
    double f0 = s_f[l0][4], f1 = s_f[l0][5];
    f0 += s_f[l0 + 1][6];
    f1 += s_f[l0 + 1][7];
    f0 += s_f[l0 + EX][2];
    f1 += s_f[l0 + EX][3];
    f0 += s_f[l0 + EX + 1][0];
    f1 += s_f[l0 + EX + 1][1];
    const int64_t k = (static_cast<int64_t>(y0 + oly) * n + x0 + olx) * 2;
    const int32_t d0 = sidx[k], d1 = sidx[k + 1];
    const double2 po = s_x[(oly + 1) * HX + olx + 1]; // the operand at the owned node
    double acc = 0.0;
    if (d0 >= 0) {
        y[d0] = f0;
        acc += po.x * f0;
    }
    if (d1 >= 0) {
        y[d1] = f1;
        acc += po.y * f1;
    }
    if (partial) block_reduce_store_n<NT>(acc, partial);
}

// Single-reduction (Chronopoulos-Gear) Jacobi-PCG, iteration `it` in one launch, on the structured
// periodic Quad4 operator.  With u = r ./ diag and s = K' p carried by recurrence:
//   beta = gam / gam_prev,  alpha = gam / (del - beta gam / alpha_prev)   (it = 0: beta = 0, alpha = gam / del)
//   p = u + beta p;  s = w + beta s;  x += alpha p;  r -= alpha s;  u = r ./ diag;  w = K' u
//   gam = (r, u),  del = (w, u),  rr = (r, r)   (block partials; the next launch reduces them)
// The same iterates as the two-reduction form in exact arithmetic, with one global reduction per
// iteration instead of two, so one launch: the tile forms the new u on its halo ring from the
// neighbours' r, w, s (double-buffered by iteration parity, since the launch writes the tile's own)
// and the owned nodes' p and x are read and written by their tile only.  u is never stored.
// Every block takes the same decisions from the same ordered partials; block 0 records gam and
// alpha (slots by parity), the iteration count and the done flag.
template <int TX, int TY>
__global__ void __launch_bounds__(TX * TY) k_cgsr_periodic_q4(
    int n, int ntx, int ntiles, int per_xcd, ShapeQ4 sh, const int32_t* __restrict__ sidx,
    const double* __restrict__ Ke, const double* __restrict__ Ge, const double* __restrict__ invd,
    const double* __restrict__ rc, const double* __restrict__ wc, const double* __restrict__ sc,
    double* __restrict__ rn, double* __restrict__ wn, double* __restrict__ sn_, double* __restrict__ p,
    double* __restrict__ x, const double* __restrict__ init_partial, const double* __restrict__ pin,
    double* __restrict__ pout, double* __restrict__ scal, int32_t* __restrict__ state, int64_t it)
{
    if (state[0]) return; // uniform for the grid
    constexpr int HX = TX + 2, HY = TY + 2, EX = TX + 1, EY = TY + 1, NT = TX * TY;
    constexpr int NH = (HX * HY + NT - 1) / NT, NE = (EX * EY + NT - 1) / NT; // halo nodes, elements per thread
    static_assert(NT == BS, "the partial sums stride by BS");
    __shared__ double2 s_x[HY * HX];                              // the new u on the halo tile
    __shared__ double s_f[EY * EX][8];                            // element forces
    __shared__ double2 s_u0[NT], s_sv[NT], s_rv[NT];              // owned nodes: old u, new s, new r
    const int b = blockIdx.x;
    const int tile = per_xcd > 0 ? (b & 7) * per_xcd + (b >> 3) : b;
    const int tx = tile % ntx, ty = tile / ntx;
    const int x0 = tx * TX, y0 = ty * TY;
    const int oly = threadIdx.x / TX, olx = threadIdx.x - (threadIdx.x / TX) * TX;

    // every load that does not depend on this iteration's scalars is issued first, so that its
    // latency overlaps the reduction of the previous launch's partials
    int32_t hd[NH][2];
    double hr[NH][2], hw[NH][2], hs[NH][2], hi[NH][2];
#pragma unroll
    for (int j = 0; j < NH; ++j) {
        const int t = threadIdx.x + j * NT;
        hd[j][0] = hd[j][1] = -1;
        if (t < HX * HY) {
            const int ly = t / HX, lx = t - (t / HX) * HX;
            int gi = x0 - 1 + lx, gj = y0 - 1 + ly;
            gi += gi < 0 ? n : 0;
            gi -= gi >= n ? n : 0;
            gj += gj < 0 ? n : 0;
            gj -= gj >= n ? n : 0;
            const int64_t k = (static_cast<int64_t>(gj) * n + gi) * 2;
            hd[j][0] = sidx[k];
            hd[j][1] = sidx[k + 1];
        }
    }
#pragma unroll
    for (int j = 0; j < NH; ++j)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int32_t d = hd[j][c] >= 0 ? hd[j][c] : 0; // clamped: the values of d < 0 are unused
            hr[j][c] = rc[d];
            hw[j][c] = wc[d];
            hs[j][c] = it == 0 ? 0.0 : sc[d];
            hi[j][c] = invd[d];
        }
    const int64_t ko = (static_cast<int64_t>(y0 + oly) * n + x0 + olx) * 2;
    const int32_t od[2] = {sidx[ko], sidx[ko + 1]};
    double op[2], ox[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int32_t d = od[c] >= 0 ? od[c] : 0;
        op[c] = it == 0 ? 0.0 : p[d];
        ox[c] = x[d];
    }
    double ek[NE], eg[NE];
#pragma unroll
    for (int j = 0; j < NE; ++j) {
        const int t = threadIdx.x + j * NT;
        ek[j] = eg[j] = 0.0;
        if (t < EX * EY) {
            const int ly = t / EX, lx = t - (t / EX) * EX;
            int ei = x0 - 1 + lx, ej = y0 - 1 + ly;
            ei += ei < 0 ? n : 0;
            ej += ej < 0 ? n : 0;
            const int64_t e = static_cast<int64_t>(ej) * n + ei;
            ek[j] = Ke[e];
            eg[j] = Ge[e];
        }
    }

    // this iteration's scalars from the ordered partials (identical in every block)
    double gam = 0.0, del = 0.0, rr = 0.0;
    if (it == 0) { // gam, rr from the initial residual pass, del from the first operator pass
        for (int i = threadIdx.x; i < NBR; i += BS) {
            gam += init_partial[i];
            rr += init_partial[NBR + i];
        }
        for (int i = threadIdx.x; i < ntiles; i += BS) del += pin[ntiles + i];
    }
    else {
        for (int i = threadIdx.x; i < ntiles; i += BS) {
            gam += pin[i];
            del += pin[ntiles + i];
            rr += pin[2 * ntiles + i];
        }
    }
    block_sum3<BS>(gam, del, rr);
    if (rr <= scal[4]) { // converged after `it` iterations
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            scal[3] = rr;
            state[0] = 1;
        }
        return;
    }
    double beta = 0.0, alpha = 0.0;
    if (it == 0) alpha = gam / del;
    else {
        beta = gam / scal[5 + ((it + 1) & 1)];
        alpha = gam / (del - beta * gam / scal[7 + ((it + 1) & 1)]);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[5 + (it & 1)] = gam;
        scal[7 + (it & 1)] = alpha;
        scal[3] = rr;
        state[1] = static_cast<int32_t>(it + 1);
    }

    // s = w + beta s, r = r - alpha s, u = r ./ diag on the halo tile (periodic wrap, prescribed
    // DOFs read as 0); the owned nodes also keep the old u, the new s and the new r in LDS
#pragma unroll
    for (int j = 0; j < NH; ++j) {
        const int t = threadIdx.x + j * NT;
        if (t >= HX * HY) continue;
        const int ly = t / HX, lx = t - (t / HX) * HX;
        double uv[2], u0[2], sv[2], rv[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            uv[c] = u0[c] = sv[c] = rv[c] = 0.0;
            if (hd[j][c] >= 0) {
                sv[c] = it == 0 ? hw[j][c] : hw[j][c] + beta * hs[j][c];
                rv[c] = hr[j][c] - alpha * sv[c];
                uv[c] = rv[c] * hi[j][c];
                u0[c] = hr[j][c] * hi[j][c];
            }
        }
        s_x[t] = make_double2(uv[0], uv[1]);
        if (lx >= 1 && lx <= TX && ly >= 1 && ly <= TY) {
            const int o = (ly - 1) * TX + lx - 1;
            s_u0[o] = make_double2(u0[0], u0[1]);
            s_sv[o] = make_double2(sv[0], sv[1]);
            s_rv[o] = make_double2(rv[0], rv[1]);
        }
    }
    __syncthreads();
    // w = K' u: element forces of the (TX+1) x (TY+1) elements touching the tile
#pragma unroll
    for (int j = 0; j < NE; ++j) {
        const int t = threadIdx.x + j * NT;
        if (t >= EX * EY) continue;
        const int ly = t / EX, lx = t - (t / EX) * EX;
        const int n0 = ly * HX + lx;
        const int nid[4] = {n0, n0 + 1, n0 + HX + 1, n0 + HX};
        double u[4][2], f[4][2];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const double2 xx = s_x[nid[m]];
            u[m][0] = xx.x;
            u[m][1] = xx.y;
            f[m][0] = f[m][1] = 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) quad4_planar_qp_force(sh.g[q], sh.v[q], u, ek[j], eg[j], f);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            s_f[t][2 * m] = f[m][0];
            s_f[t][2 * m + 1] = f[m][1];
        }
    }
    __syncthreads();
    // owned node: elements (i-1, j-1) as its node 2, (i, j-1) node 3, (i-1, j) node 1, (i, j) node 0
    const int l0 = oly * EX + olx;
    double f[2];
    f[0] = s_f[l0][4];
    f[1] = s_f[l0][5];
    f[0] += s_f[l0 + 1][6];
    f[1] += s_f[l0 + 1][7];
    f[0] += s_f[l0 + EX][2];
    f[1] += s_f[l0 + EX][3];
    f[0] += s_f[l0 + EX + 1][0];
    f[1] += s_f[l0 + EX + 1][1];
    const double2 uo = s_x[(oly + 1) * HX + olx + 1], u0o = s_u0[threadIdx.x], svo = s_sv[threadIdx.x],
                  rvo = s_rv[threadIdx.x];
    const double un[2] = {uo.x, uo.y}, u0[2] = {u0o.x, u0o.y}, sv[2] = {svo.x, svo.y}, rv[2] = {rvo.x, rvo.y};
    double g = 0.0, dl = 0.0, r2 = 0.0;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int32_t d = od[c];
        if (d < 0) continue;
        const double pv = it == 0 ? u0[c] : u0[c] + beta * op[c];
        x[d] = ox[c] + alpha * pv;
        p[d] = pv;
        sn_[d] = sv[c];
        rn[d] = rv[c];
        wn[d] = f[c];
        g += rv[c] * un[c];
        dl += f[c] * un[c];
        r2 += rv[c] * rv[c];
    }
    block_sum3<NT>(g, dl, r2);
    if (threadIdx.x == 0) {
        pout[b] = g;
        pout[ntiles + b] = dl;
        pout[2 * ntiles + b] = r2;
    }
}

// Hex8 Regular n^3 cell with periodic tyings: identical element geometry makes the element
// stiffness K_e = K_e,bulk KK + G_e KG with two fixed 24 x 24 matrices (isotropic linear elasticity,
// C = K I(x)I + 2G (I_sym - I(x)I/3)).  Each thread owns one node and sums, over its 8 elements,
// the node's 3 rows of K_e u_e; the x halo of the 8 x 8 x 4 node tile sits in LDS (per component).
template <int TX, int TY, int TZ>
__global__ void __launch_bounds__(TX * TY * TZ) k_apply_periodic_h8(int n, int ntx, int nty, const double* __restrict__ kkg,
                                                                    const int32_t* __restrict__ sidx,
                                                                    const double* __restrict__ Ke,
                                                                    const double* __restrict__ Ge,
                                                                    const double* __restrict__ x, double* __restrict__ y,
                                                                    double* __restrict__ partial,
                                                                    const int32_t* __restrict__ state)
{
    if (state && state[0]) return;
    constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2, NH = HX * HY * HZ, NT = TX * TY * TZ;
    constexpr int NJ = (NH + NT - 1) / NT; // halo nodes per thread
    __shared__ double s_x[3][NH];
    const int tile = blockIdx.x;
    const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
    const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
    // the halo tile: every DOF index of the thread's NJ nodes is loaded first (node indices clamped,
    // so all addresses are in bounds), then every value (prescribed DOFs read as 0), so the two
    // dependent loads wait once per thread rather than once per node
    int32_t dd[NJ][3];
#pragma unroll
    for (int q = 0; q < NJ; ++q) {
        const int t = min(static_cast<int>(threadIdx.x) + q * NT, NH - 1);
        const int lz = t / (HX * HY), r = t - lz * HX * HY, ly = r / HX, lx = r - (r / HX) * HX;
        int gi = x0 - 1 + lx, gj = y0 - 1 + ly, gk = z0 - 1 + lz;
        gi += gi < 0 ? n : 0;
        gi -= gi >= n ? n : 0;
        gj += gj < 0 ? n : 0;
        gj -= gj >= n ? n : 0;
        gk += gk < 0 ? n : 0;
        gk -= gk >= n ? n : 0;
        const int64_t k = ((static_cast<int64_t>(gk) * n + gj) * n + gi) * 3;
#pragma unroll
        for (int c = 0; c < 3; ++c) dd[q][c] = sidx[k + c];
    }
    double xv[NJ][3];
#pragma unroll
    for (int q = 0; q < NJ; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) xv[q][c] = x[dd[q][c] >= 0 ? dd[q][c] : 0];
#pragma unroll
    for (int q = 0; q < NJ; ++q) {
        const int t = static_cast<int>(threadIdx.x) + q * NT;
        if (t < NH)
#pragma unroll
            for (int c = 0; c < 3; ++c) s_x[c][t] = dd[q][c] >= 0 ? xv[q][c] : 0.0;
    }
    __syncthreads();
    const int olz = threadIdx.x / (TX * TY), orr = threadIdx.x - olz * TX * TY, oly = orr / TX,
              olx = orr - (orr / TX) * TX;
    // local node offsets of the Hex8 connectivity: 0 (0,0,0) 1 (1,0,0) 2 (1,1,0) 3 (0,1,0) 4..7 at z + 1
    // (OX, OY, OZ)[m] = (((m + 1) >> 1) & 1, (m >> 1) & 1, m >> 2)
    constexpr int OX[8] = {0, 1, 1, 0, 0, 1, 1, 0}, OY[8] = {0, 0, 1, 1, 0, 0, 1, 1}, OZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
    for (int el = 0; el < 8; ++el) {
        // element whose lower corner is the owned node minus (1 - e): the owned node is its local node la
        const int ex = ((el + 1) >> 1) & 1, ey = (el >> 1) & 1, ez = el >> 2;
        const int ax = 1 - ex, ay = 1 - ey, az = 1 - ez;
        const int la = az * 4 + (ay ? (ax ? 2 : 3) : (ax ? 1 : 0));
        int ei = x0 + olx - 1 + ex, ej = y0 + oly - 1 + ey, ek = z0 + olz - 1 + ez;
        ei += ei < 0 ? n : 0;
        ej += ej < 0 ? n : 0;
        ek += ek < 0 ? n : 0;
        const int64_t e = (static_cast<int64_t>(ek) * n + ej) * n + ei;
        // element nodes in the halo tile: lower corner (olx + ex, oly + ey, olz + ez)
        const int c0 = ((olz + ez) * HY + (oly + ey)) * HX + (olx + ex);
        double ue[24];
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int h = c0 + (OZ[m] * HY + OY[m]) * HX + OX[m];
#pragma unroll
            for (int c = 0; c < 3; ++c) ue[3 * m + c] = s_x[c][h];
        }
        const double k = Ke[e], gm = Ge[e];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double* rk = kkg + (la * 3 + j) * 24;
            const double* rg = kkg + 576 + (la * 3 + j) * 24;
            // fused multiply-adds (this operator is not tied to the reference's expression order)
            double sk = 0.0, sg = 0.0;
#pragma unroll
            for (int c = 0; c < 24; ++c) {
                sk = fma(rk[c], ue[c], sk);
                sg = fma(rg[c], ue[c], sg);
            }
            acc[j] = fma(k, sk, fma(gm, sg, acc[j]));
        }
    }
    const int64_t kn = ((static_cast<int64_t>(z0 + olz) * n + y0 + oly) * n + x0 + olx) * 3;
    const int ho = ((olz + 1) * HY + oly + 1) * HX + olx + 1; // the owned node in the halo tile
    double dot = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int32_t d = sidx[kn + c];
        if (d >= 0) {
            y[d] = acc[c];
            dot += s_x[c][ho] * acc[c];
        }
    }
    if (partial) block_reduce_store_n<NT>(dot, partial);
}

// ---- element force from a full dofval, FP64, reference expression order -------------------------
template <int ET>
__global__ void __launch_bounds__(128) k_elem_force(int64_t nelem, const int32_t* __restrict__ conn, const int32_t* __restrict__ dofs,
                             const double* __restrict__ dNx, const double* __restrict__ vol,
                             const double* __restrict__ Ke, const double* __restrict__ Ge,
                             const double* __restrict__ x, double* __restrict__ fe, const int32_t* __restrict__ state)
{
    if (state && state[0]) return;
    constexpr int NNE = ET == GF_HEX8 ? 8 : 4, ND = ET == GF_HEX8 ? 3 : 2;
    int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (e >= nelem) return;
    double u[NNE][ND];
#pragma unroll
    for (int m = 0; m < NNE; ++m)
#pragma unroll
        for (int i = 0; i < ND; ++i) u[m][i] = x[dofs[static_cast<int64_t>(conn[e * NNE + m]) * ND + i]];
    const double k = Ke[e], gm = Ge[e];
    double f[NNE][ND];
#pragma unroll
    for (int m = 0; m < NNE; ++m)
#pragma unroll
        for (int i = 0; i < ND; ++i) f[m][i] = 0.0;
#pragma unroll 1
    for (int q = 0; q < NNE; ++q) {
        const double* g = dNx + (e * NNE + q) * NNE * ND;
        const double v = vol[e * NNE + q];
        double s[3][3];
        if constexpr (ET == GF_HEX8) {
            // SymGradN_vector, generic loop Element.h:990-1022
            double t[3][3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) t[i][j] = 0.0;
#pragma unroll
            for (int m = 0; m < 8; ++m)
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        double h = 0.5 * g[m * 3 + i] * u[m][j];
                        t[i][j] += h;
                        t[j][i] += h;
                    }
            double tr = t[0][0] + t[1][1] + t[2][2];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    s[i][j] = (i == j) ? k * tr + 2.0 * gm * (t[i][j] - tr / 3.0) : 2.0 * gm * t[i][j];
            // ElementHex8.h:354-383
#pragma unroll
            for (int m = 0; m < 8; ++m)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    f[m][j] += (g[m * 3 + 0] * s[0][j] + g[m * 3 + 1] * s[1][j] + g[m * 3 + 2] * s[2][j]) * v;
        }
        else {
            (void)s;
            quad4_planar_qp_force(g, v, u, k, gm, f);
        }
    }
#pragma unroll
    for (int m = 0; m < NNE; ++m)
#pragma unroll
        for (int i = 0; i < ND; ++i) fe[(e * NNE + m) * ND + i] = f[m][i];
}

// element stiffness diagonal K_e(mi, mi) = sum_q dV [(K + G/3) g_i^2 + G |g|^2] (isotropic law)
template <int ET>
__global__ void k_elem_diag(int64_t nelem, const double* __restrict__ dNx, const double* __restrict__ vol,
                            const double* __restrict__ Ke, const double* __restrict__ Ge, double* __restrict__ de)
{
    constexpr int NNE = ET == GF_HEX8 ? 8 : 4, ND = ET == GF_HEX8 ? 3 : 2;
    int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= nelem * NNE) return;
    const int64_t e = t / NNE;
    const int m = static_cast<int>(t - e * NNE);
    const double a = Ke[e] + Ge[e] / 3.0, b = Ge[e];
    double d[ND];
#pragma unroll
    for (int i = 0; i < ND; ++i) d[i] = 0.0;
    for (int q = 0; q < NNE; ++q) {
        const double* g = dNx + ((e * NNE + q) * NNE + m) * ND;
        double n2 = 0.0;
#pragma unroll
        for (int i = 0; i < ND; ++i) n2 += g[i] * g[i];
#pragma unroll
        for (int i = 0; i < ND; ++i) d[i] += vol[e * NNE + q] * (a * g[i] * g[i] + b * n2);
    }
#pragma unroll
    for (int i = 0; i < ND; ++i) de[t * ND + i] = d[i];
}

__device__ inline double row_sum(const int32_t* ptr, const int32_t* idx, const double* fe, int64_t r)
{
    return ordered_gather_sum<4>(ptr[r], ptr[r + 1], idx, fe);
}

// block reduction of one value per thread into partial[blockIdx]
__device__ inline void block_reduce_store(double v, double* partial) { block_reduce_store_n<BS>(v, partial); }


// out_u = y_u + C_du^T y_d with y = ordered gather of fe; optional partial of dot(out, w)
__global__ void __launch_bounds__(BS) k_gather_fold(int64_t nnu, int64_t nni, const int32_t* __restrict__ eptr,
                                                    const int32_t* __restrict__ eidx, const double* __restrict__ fe,
                                                    const int32_t* __restrict__ tptr, const int32_t* __restrict__ tidx,
                                                    const double* __restrict__ tval, double* __restrict__ out,
                                                    const double* __restrict__ w, double* __restrict__ partial,
                                                    const int32_t* __restrict__ state)
{
    if (state && state[0]) return;
    double acc = 0.0;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(BS) + threadIdx.x; i < nnu;
         i += static_cast<int64_t>(gridDim.x) * BS) {
        double s = row_sum(eptr, eidx, fe, i);
        for (int32_t k = tptr[i]; k < tptr[i + 1]; ++k) s += tval[k] * row_sum(eptr, eidx, fe, nni + tidx[k]);
        out[i] = s;
        if (w) acc += w[i] * s;
    }
    if (partial) block_reduce_store(acc, partial);
}

// x_full = [x_u, x_p, C_du x_u + C_dp x_p] (x_u / x_p may be NULL = 0); when x_u aliases the
// u-part of x_full it is not copied
__global__ void k_expand(int64_t nnu, int64_t nnp, int64_t nnd, const double* __restrict__ xu,
                         const double* __restrict__ xp, const int32_t* __restrict__ uptr,
                         const int32_t* __restrict__ uidx, const double* __restrict__ uval,
                         const int32_t* __restrict__ pptr, const int32_t* __restrict__ pidx,
                         const double* __restrict__ pval, double* __restrict__ xfull, int copy_u,
                         const int32_t* __restrict__ state)
{
    if (state && state[0]) return;
    const int64_t n = (copy_u ? nnu : 0) + nnp + nnd;
    for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n;
         t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int64_t i = copy_u ? t : t + nnu;
        if (i < nnu) xfull[i] = xu ? xu[i] : 0.0;
        else if (i < nnu + nnp) xfull[i] = xp ? xp[i - nnu] : 0.0;
        else {
            const int64_t d = i - nnu - nnp;
            double s = 0.0;
            if (xu)
                for (int32_t k = uptr[d]; k < uptr[d + 1]; ++k) s += uval[k] * xu[uidx[k]];
            if (xp)
                for (int32_t k = pptr[d]; k < pptr[d + 1]; ++k) s += pval[k] * xp[pidx[k]];
            xfull[i] = s;
        }
    }
}

// General operator, CG iteration it >= 1: the direction update fused into the expansion.  beta =
// rz_new / rz_old from the previous update's partials (identical in every block); the u part of
// x_full is the new direction p = z + beta p_old (also written to pnew), the p part 0, and each
// dependent entry sum_k C_du(d, k) p(k) with every p(k) formed the same way (p_old and pnew are
// distinct buffers, so no block reads an entry another block is replacing).  Block 0 records rz_new
// and r.r and raises done at convergence (r.r <= tol^2 b.b).
__global__ void k_expand_dir(int64_t nnu, int64_t nnp, int64_t nnd, const double* __restrict__ z,
                             const double* __restrict__ pold, double* __restrict__ pnew,
                             const int32_t* __restrict__ uptr, const int32_t* __restrict__ uidx,
                             const double* __restrict__ uval, double* __restrict__ xfull,
                             const double* __restrict__ rz_rr, double* __restrict__ scal, int32_t* __restrict__ state,
                             int64_t it)
{
    if (state[0]) return;
    double rz_new = 0.0, rr = 0.0, unused = 0.0;
    for (int i = threadIdx.x; i < NBR; i += BS) {
        rz_new += rz_rr[i];
        rr += rz_rr[NBR + i];
    }
    block_sum3<BS>(rz_new, rr, unused);
    if (rr <= scal[4]) { // converged after `it` iterations
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            scal[3] = rr;
            state[0] = 1;
        }
        return;
    }
    const double beta = rz_new / scal[(it - 1) & 1];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[it & 1] = rz_new;
        scal[3] = rr;
    }
    const int64_t n = nnu + nnp + nnd;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(BS) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * BS) {
        if (i < nnu) {
            const double pv = z[i] + beta * pold[i];
            pnew[i] = pv;
            xfull[i] = pv;
        }
        else if (i < nnu + nnp)
            xfull[i] = 0.0;
        else {
            const int64_t d = i - nnu - nnp;
            double sd = 0.0;
            for (int32_t k = uptr[d]; k < uptr[d + 1]; ++k) {
                const int32_t c = uidx[k];
                sd += uval[k] * (z[c] + beta * pold[c]);
            }
            xfull[i] = sd;
        }
    }
}

__global__ void k_diag_fold(int64_t nnu, int64_t nni, const int32_t* __restrict__ eptr, const int32_t* __restrict__ eidx,
                            const double* __restrict__ de, const int32_t* __restrict__ tptr,
                            const int32_t* __restrict__ tidx, const double* __restrict__ tval,
                            double* __restrict__ invdiag)
{
    int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= nnu) return;
    double s = row_sum(eptr, eidx, de, i);
    for (int32_t k = tptr[i]; k < tptr[i + 1]; ++k) s += tval[k] * tval[k] * row_sum(eptr, eidx, de, nni + tidx[k]);
    invdiag[i] = 1.0 / s;
}

// r = b - q, z = r ./ d, x = x0 (taken from p), p = z; partials rz, rr, bb
__global__ void __launch_bounds__(BS) k_pcg_init(int64_t n, const double* __restrict__ b, const double* __restrict__ q,
                                                 const double* __restrict__ invd, double* __restrict__ r,
                                                 double* __restrict__ z, double* __restrict__ x, double* __restrict__ p,
                                                 double* __restrict__ partial)
{
    double rz = 0.0, rr = 0.0, bb = 0.0;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(BS) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * BS) {
        double ri = b[i] - q[i];
        double zi = ri * invd[i];
        r[i] = ri;
        z[i] = zi;
        x[i] = p[i];
        p[i] = zi;
        rz += ri * zi;
        rr += ri * ri;
        bb += b[i] * b[i];
    }
    block_reduce_store(rz, partial);
    block_reduce_store(rr, partial + NBR);
    block_reduce_store(bb, partial + 2 * NBR);
}

// finalize the initial scalars: rz -> scal[0], bb -> scal[2]; done if rr <= tol^2 bb
__global__ void k_pcg_init_scalars(const double* __restrict__ partial, double* __restrict__ scal, int32_t* __restrict__ state,
                                   double tol)
{
    double rz = reduce_partials(partial), rr = reduce_partials(partial + NBR), bb = reduce_partials(partial + 2 * NBR);
    if (threadIdx.x == 0) {
        scal[0] = rz;
        scal[2] = bb;
        scal[3] = rr;
        scal[4] = tol * tol * (bb > 0.0 ? bb : 1.0);
        state[0] = (rr <= scal[4]) ? 1 : 0;
        state[1] = 0;
    }
}

// alpha = rz / (p.q); x += alpha p; r -= alpha q; z = r ./ d; partials of rz_new, rr
__global__ void __launch_bounds__(BS) k_pcg_update(int64_t n, const double* __restrict__ pq_partial, int npq,
                                                   const double* __restrict__ scal, const double* __restrict__ p,
                                                   const double* __restrict__ q, const double* __restrict__ invd,
                                                   double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
                                                   double* __restrict__ partial, const int32_t* __restrict__ state,
                                                   int64_t it)
{
    if (state[0]) return;
    const double pq = reduce_partials(pq_partial, npq);
    const int slot = static_cast<int>(it & 1);
    const double alpha = scal[slot] / pq;
    double rz = 0.0, rr = 0.0;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(BS) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * BS) {
        x[i] += alpha * p[i];
        double ri = r[i] - alpha * q[i];
        double zi = ri * invd[i];
        r[i] = ri;
        z[i] = zi;
        rz += ri * zi;
        rr += ri * ri;
    }
    block_reduce_store(rz, partial);
    block_reduce_store(rr, partial + NBR);
    if (blockIdx.x == 0 && threadIdx.x == 0) const_cast<int32_t*>(state)[1] = static_cast<int32_t>(it + 1);
}

// beta = rz_new / rz_old; p = z + beta p.  Block 0 records rz_new in the other slot, stores the
// iteration count and raises the done flag (state[0]) at convergence or at the iteration cap.  The
// iteration number comes from the host (it equals the device count until the flag is raised, after
// which every kernel returns at once), so no block reads a value that block 0 writes in this launch
// except the flag itself, and a block that sees the flag raised skips a p update nobody uses.
__global__ void __launch_bounds__(BS) k_pcg_direction(int64_t n, const double* __restrict__ partial,
                                                      double* __restrict__ scal, const double* __restrict__ z,
                                                      double* __restrict__ p, int32_t* __restrict__ state, int64_t it)
{
    if (state[0]) return;
    const double rz_new = reduce_partials(partial), rr = reduce_partials(partial + NBR);
    const double beta = rz_new / scal[it & 1];
    for (int64_t i = blockIdx.x * static_cast<int64_t>(BS) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * BS)
        p[i] = z[i] + beta * p[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[(it + 1) & 1] = rz_new;
        scal[3] = rr;
        state[1] = static_cast<int32_t>(it + 1);
        if (rr <= scal[4] || it + 1 >= state[2]) state[0] = 1;
    }
}

__global__ void __launch_bounds__(BS) k_fill_zero(int64_t n, double* __restrict__ a)
{
    for (int64_t i = blockIdx.x * static_cast<int64_t>(BS) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * BS)
        a[i] = 0.0;
}

__global__ void __launch_bounds__(BS) k_residual(int64_t n, const double* __restrict__ b, const double* __restrict__ q,
                                                 double* __restrict__ partial)
{
    double rr = 0.0, bb = 0.0;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(BS) + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * BS) {
        double d = b[i] - q[i];
        rr += d * d;
        bb += b[i] * b[i];
    }
    block_reduce_store(rr, partial);
    block_reduce_store(bb, partial + NBR);
}

__global__ void k_residual_final(const double* __restrict__ partial, double* __restrict__ out)
{
    double rr = reduce_partials(partial), bb = reduce_partials(partial + NBR);
    if (threadIdx.x == 0) out[0] = std::sqrt(rr) / (bb > 0.0 ? std::sqrt(bb) : 1.0);
}

// the vectors and block partials of the single-reduction iteration (structured operators)
void alloc_single_reduction(gf_pcg_s* P)
{
    for (int k = 0; k < 2; ++k) {
        P->sr[k].alloc(P->nnu);
        P->sw[k].alloc(P->nnu);
        P->ss[k].alloc(P->nnu);
    }
    P->spart.alloc(6 * static_cast<size_t>(P->ntiles));
}

void upload_csr(Csr& c, const int64_t* ptr, const int64_t* idx, const double* val, int64_t nrows)
{
    int64_t nnz = ptr[nrows];
    c.ptr.upload(narrow(ptr, static_cast<size_t>(nrows + 1), "csr ptr"));
    c.idx.upload(narrow(idx, static_cast<size_t>(nnz), "csr idx"));
    c.val.upload(val, static_cast<size_t>(nnz));
    if (nnz == 0) {
        // keep valid (non-null) pointers for empty matrices
        c.idx.alloc(1);
        c.val.alloc(1);
    }
}

// y_u = K' x on the general operator from an already expanded x_full (element forces, ordered
// gather + C_du^T fold, with the partials of (x_u, y_u)); the second half of apply
void operator_from_full(gf_pcg_s* P, double* yu, double* partial, const int32_t* state, hipStream_t s)
{
    auto v = P->v;
    const unsigned ge = nblocks(P->nelem, 128);
    if (P->et == GF_HEX8)
        k_elem_force<GF_HEX8><<<ge, 128, 0, s>>>(P->nelem, v->conn.p, v->dofs.p, quad_dNx(P->q), quad_vol(P->q),
                                                 P->Ke.p, P->Ge.p, P->xfull.p, P->fe.p, state);
    else
        k_elem_force<GF_QUAD4_PLANAR><<<ge, 128, 0, s>>>(P->nelem, v->conn.p, v->dofs.p, quad_dNx(P->q),
                                                         quad_vol(P->q), P->Ke.p, P->Ge.p, P->xfull.p, P->fe.p, state);
    k_gather_fold<<<NBR, BS, 0, s>>>(P->nnu, P->nni, v->dof_eptr.p, v->dof_eidx.p, P->fe.p, P->cduT.ptr.p,
                                     P->cduT.idx.p, P->cduT.val.p, yu, P->xfull.p, partial, state);
    check_launch();
}

// y_u = K' x_u + K'_up x_p (x_u or x_p may be NULL); state == NULL: always runs
void apply(gf_pcg_s* P, const double* xu, const double* xp, double* yu, bool xu_in_full, hipStream_t s,
           const double* w, double* partial, const int32_t* state)
{
    auto v = P->v;
    if (P->structured && !xp) {
        const double* xs = xu_in_full ? P->xfull.p : xu;
        if ((!w || w == xs) && P->et == GF_HEX8) {
            k_apply_periodic_h8<8, 8, 4><<<P->ntiles, 256, 0, s>>>(P->sn, P->ntx, P->nty, P->ke8.p, P->sidx.p, P->Ke.p,
                                                                   P->Ge.p, xs, yu, w ? partial : nullptr, state);
            check_launch();
            return;
        }
        if (!w || w == xs) {
            ShapeQ4 sh;
            for (int q = 0; q < 4; ++q) {
                for (int k = 0; k < 8; ++k) sh.g[q][k] = P->sh_g[q * 8 + k];
                sh.v[q] = P->sh_v[q];
            }
            k_apply_periodic_q4<32, 8><<<P->ntiles, 256, 0, s>>>(P->sn, P->ntx, P->per_xcd, sh, P->sidx.p, P->Ke.p,
                                                                 P->Ge.p, xs, yu, w ? partial : nullptr, state);
            check_launch();
            return;
        }
    }
    const int64_t nexp = (xu_in_full ? 0 : P->nnu) + P->nnp + P->nnd;
    if (nexp > 0)
        k_expand<<<std::min<unsigned>(nblocks(nexp, BS), 4096), BS, 0, s>>>(
            P->nnu, P->nnp, P->nnd, xu_in_full ? P->xfull.p : xu, xp, P->cdu.ptr.p, P->cdu.idx.p, P->cdu.val.p,
            P->cdp.ptr.p, P->cdp.idx.p, P->cdp.val.p, P->xfull.p, xu_in_full ? 0 : 1, state);
    const unsigned ge = nblocks(P->nelem, 128);
    if (P->et == GF_HEX8)
        k_elem_force<GF_HEX8><<<ge, 128, 0, s>>>(P->nelem, v->conn.p, v->dofs.p, quad_dNx(P->q), quad_vol(P->q),
                                                 P->Ke.p, P->Ge.p, P->xfull.p, P->fe.p, state);
    else
        k_elem_force<GF_QUAD4_PLANAR><<<ge, 128, 0, s>>>(P->nelem, v->conn.p, v->dofs.p, quad_dNx(P->q),
                                                         quad_vol(P->q), P->Ke.p, P->Ge.p, P->xfull.p, P->fe.p, state);
    k_gather_fold<<<NBR, BS, 0, s>>>(P->nnu, P->nni, v->dof_eptr.p, v->dof_eidx.p, P->fe.p, P->cduT.ptr.p,
                                     P->cduT.idx.p, P->cduT.val.p, yu, w, partial, state);
    check_launch();
}

} // namespace

extern "C" {

namespace {

// Jacobi diagonal of K' (element diagonals, ordered gather, folded through C_du), on stream s
void compute_invdiag(gf_pcg P, hipStream_t s)
{
    const int64_t n = P->nelem * P->nne;
    if (P->et == GF_HEX8)
        k_elem_diag<GF_HEX8><<<nblocks(n, BS), BS, 0, s>>>(P->nelem, quad_dNx(P->q), quad_vol(P->q), P->Ke.p,
                                                         P->Ge.p, P->fe.p);
    else
        k_elem_diag<GF_QUAD4_PLANAR><<<nblocks(n, BS), BS, 0, s>>>(P->nelem, quad_dNx(P->q), quad_vol(P->q),
                                                                 P->Ke.p, P->Ge.p, P->fe.p);
    k_diag_fold<<<nblocks(P->nnu, BS), BS, 0, s>>>(P->nnu, P->nni, P->v->dof_eptr.p, P->v->dof_eidx.p, P->fe.p,
                                                  P->cduT.ptr.p, P->cduT.idx.p, P->cduT.val.p, P->invdiag.p);
    check_launch();
}

// flags[block] = 1 if any element's n values differ from element 0's by more than tol (no atomics:
// an ordered LDS OR per block, one store per block)
__global__ void __launch_bounds__(BS) k_uniform_check(int64_t nelem, int n, const double* __restrict__ data, double tol,
                                                      int32_t* __restrict__ flags)
{
    __shared__ int32_t sh[BS];
    int32_t bad = 0;
    const int64_t total = nelem * n;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(BS) + threadIdx.x; i < total;
         i += static_cast<int64_t>(gridDim.x) * BS)
        bad |= fabs(data[i] - data[i % n]) > tol ? 1 : 0;
    sh[threadIdx.x] = bad;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] |= sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) flags[blockIdx.x] = sh[0];
}

// true if every element's n values of `data` equal element 0's to 1e-12 relative; values of element 0
// are returned in first[0..n)
bool uniform_per_element(const double* data, int64_t nelem, int n, double* first)
{
    GF_HIP(hipMemcpy(first, data, n * sizeof(double), hipMemcpyDeviceToHost));
    double vmax = 0.0;
    for (int k = 0; k < n; ++k) vmax = std::max(vmax, std::abs(first[k]));
    if (!(vmax > 0.0)) return false;
    constexpr int NBLK = 1024;
    DBuf<int32_t> flags;
    flags.alloc(NBLK);
    k_uniform_check<<<NBLK, BS>>>(nelem, n, data, 1e-12 * vmax, flags.p);
    check_launch();
    std::vector<int32_t> h(NBLK);
    GF_HIP(hipMemcpy(h.data(), flags.p, NBLK * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int32_t f : h)
        if (f) return false;
    return true;
}

// Regular n x n Quad4 cell whose dependent DOFs are periodic copies (C_du rows = one unit entry of
// the partner's u-DOF, or empty when the partner is prescribed) and whose elements share one
// geometry: then K' x_u is applied by k_apply_periodic_q4.
void detect_structured(gf_pcg P, const int64_t* cdu_ptr_h, const int64_t* cdu_col_h, const double* cdu_val_h)
{
    constexpr int TX = 32, TY = 8;
    auto v = P->v;
    if (P->et != GF_QUAD4_PLANAR) return;
    const int64_t n = static_cast<int64_t>(std::llround(std::sqrt(static_cast<double>(P->nelem))));
    if (n * n != P->nelem || n % TX != 0 || n % TY != 0 || n >= (1 << 15)) return;
    const int64_t N1 = n + 1;
    if (v->nnode < N1 * N1 || v->ndim != 2) return;
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) {
            const int64_t e = j * n + i, a = j * N1 + i;
            const int64_t want[4] = {a, a + 1, a + N1 + 1, a + N1};
            for (int m = 0; m < 4; ++m)
                if (v->conn_h[e * 4 + m] != want[m]) return;
        }
    auto dof = [&](int64_t node, int c) { return v->dofs_h[node * 2 + c]; };
    // independent grid nodes: u or p; dependent edge nodes: periodic copies of their partner
    for (int64_t j = 0; j <= n; ++j)
        for (int64_t i = 0; i <= n; ++i)
            for (int c = 0; c < 2; ++c) {
                const int64_t d = dof(j * N1 + i, c);
                if (i < n && j < n) {
                    if (d < 0 || d >= P->nni) return;
                    continue;
                }
                if (d < P->nni || d >= P->ndof) return;
                const int64_t pd = dof((j % n) * N1 + (i % n), c), r = d - P->nni;
                const int64_t cnt = cdu_ptr_h[r + 1] - cdu_ptr_h[r];
                if (pd < P->nnu) {
                    if (cnt != 1 || cdu_col_h[cdu_ptr_h[r]] != pd || cdu_val_h[cdu_ptr_h[r]] != 1.0) return;
                }
                else if (cnt != 0)
                    return;
            }
    // uniform geometry: every element's dN/dx and dV equal element 0's to 1e-12 relative (a Regular
    // mesh gives the same Jacobian everywhere up to the rounding of the coordinate sums, ~1e-16)
    std::vector<double> g(32), vol(4);
    if (!uniform_per_element(quad_dNx(P->q), P->nelem, 32, g.data())) return;
    if (!uniform_per_element(quad_vol(P->q), P->nelem, 4, vol.data())) return;
    std::vector<int32_t> sidx(static_cast<size_t>(n * n * 2));
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i)
            for (int c = 0; c < 2; ++c) {
                const int64_t d = dof(j * N1 + i, c);
                sidx[(j * n + i) * 2 + c] = d < P->nnu ? static_cast<int32_t>(d) : -1;
            }
    P->sidx.alloc(sidx.size());
    GF_HIP(hipMemcpy(P->sidx.p, sidx.data(), sidx.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    for (int k = 0; k < 32; ++k) P->sh_g[k] = g[k];
    for (int k = 0; k < 4; ++k) P->sh_v[k] = vol[k];
    P->sn = static_cast<int>(n);
    P->ntx = static_cast<int>(n / TX);
    P->ntiles = static_cast<int>((n / TX) * (n / TY));
    P->per_xcd = P->ntiles % 8 == 0 ? P->ntiles / 8 : 0;
    alloc_single_reduction(P);
    P->structured = true;
}

// Hex8 Regular n x n x n cell with periodic copies as dependent DOFs and identical element
// geometry (to 1e-12): K' x_u is applied by k_apply_periodic_h8 with K_e = K KK + G KG.
void detect_structured_hex8(gf_pcg P, const int64_t* cdu_ptr_h, const int64_t* cdu_col_h, const double* cdu_val_h)
{
    constexpr int TX = 8, TY = 8, TZ = 4;
    auto v = P->v;
    if (P->et != GF_HEX8 || v->ndim != 3) return;
    const int64_t n = static_cast<int64_t>(std::llround(std::cbrt(static_cast<double>(P->nelem))));
    if (n * n * n != P->nelem || n % TX != 0 || n % TY != 0 || n % TZ != 0 || n >= 1024) return;
    const int64_t N1 = n + 1;
    if (v->nnode < N1 * N1 * N1) return;
    auto node = [&](int64_t i, int64_t j, int64_t k) { return (k * N1 + j) * N1 + i; };
    static const int OX[8] = {0, 1, 1, 0, 0, 1, 1, 0}, OY[8] = {0, 0, 1, 1, 0, 0, 1, 1}, OZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    for (int64_t k = 0; k < n; ++k)
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < n; ++i) {
                const int64_t e = (k * n + j) * n + i;
                for (int m = 0; m < 8; ++m)
                    if (v->conn_h[e * 8 + m] != node(i + OX[m], j + OY[m], k + OZ[m])) return;
            }
    auto dof = [&](int64_t nd, int c) { return v->dofs_h[nd * 3 + c]; };
    for (int64_t k = 0; k <= n; ++k)
        for (int64_t j = 0; j <= n; ++j)
            for (int64_t i = 0; i <= n; ++i)
                for (int c = 0; c < 3; ++c) {
                    const int64_t d = dof(node(i, j, k), c);
                    if (i < n && j < n && k < n) {
                        if (d < 0 || d >= P->nni) return;
                        continue;
                    }
                    if (d < P->nni || d >= P->ndof) return;
                    const int64_t pd = dof(node(i % n, j % n, k % n), c), r = d - P->nni;
                    const int64_t cnt = cdu_ptr_h[r + 1] - cdu_ptr_h[r];
                    if (pd < P->nnu) {
                        if (cnt != 1 || cdu_col_h[cdu_ptr_h[r]] != pd || cdu_val_h[cdu_ptr_h[r]] != 1.0) return;
                    }
                    else if (cnt != 0)
                        return;
                }
    constexpr int NGE = 8 * 8 * 3;
    std::vector<double> g(NGE), vol(8);
    if (!uniform_per_element(quad_dNx(P->q), P->nelem, NGE, g.data())) return;
    if (!uniform_per_element(quad_vol(P->q), P->nelem, 8, vol.data())) return;
    // KK(m*3+j, n*3+k) = sum_q dV dNx(m,j) dNx(n,k)
    // KG(m*3+j, n*3+k) = sum_q dV [dNx(m,k) dNx(n,j) + delta_jk dNx(m).dNx(n) - 2/3 dNx(m,j) dNx(n,k)]
    std::vector<double> kkg(2 * 576, 0.0);
    for (int q = 0; q < 8; ++q) {
        const double* gq = g.data() + q * 24;
        const double dv = vol[q];
        for (int m = 0; m < 8; ++m)
            for (int nn = 0; nn < 8; ++nn) {
                const double dot = gq[m * 3] * gq[nn * 3] + gq[m * 3 + 1] * gq[nn * 3 + 1] + gq[m * 3 + 2] * gq[nn * 3 + 2];
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) {
                        const int r = m * 3 + j, c = nn * 3 + k;
                        kkg[r * 24 + c] += dv * gq[m * 3 + j] * gq[nn * 3 + k];
                        kkg[576 + r * 24 + c] += dv * (gq[m * 3 + k] * gq[nn * 3 + j] + (j == k ? dot : 0.0) -
                                                       2.0 / 3.0 * gq[m * 3 + j] * gq[nn * 3 + k]);
                    }
            }
    }
    std::vector<int32_t> sidx(static_cast<size_t>(n * n * n * 3));
    for (int64_t k = 0; k < n; ++k)
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < n; ++i)
                for (int c = 0; c < 3; ++c) {
                    const int64_t d = dof(node(i, j, k), c);
                    sidx[((k * n + j) * n + i) * 3 + c] = d < P->nnu ? static_cast<int32_t>(d) : -1;
                }
    P->sidx.alloc(sidx.size());
    GF_HIP(hipMemcpy(P->sidx.p, sidx.data(), sidx.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    P->ke8.alloc(kkg.size());
    GF_HIP(hipMemcpy(P->ke8.p, kkg.data(), kkg.size() * sizeof(double), hipMemcpyHostToDevice));
    P->sn = static_cast<int>(n);
    P->ntx = static_cast<int>(n / TX);
    P->nty = static_cast<int>(n / TY);
    P->ntiles = static_cast<int>((n / TX) * (n / TY) * (n / TZ));
    P->per_xcd = 0;
    P->structured = true;
}

} // namespace

int gf_pcg_create(gf_pcg* out, gf_vector v, gf_quad q, const double* K_e, const double* G_e, int64_t nnu,
                  int64_t nnp, const int64_t* cdu_ptr_h, const int64_t* cdu_col_h, const double* cdu_val_h,
                  const int64_t* cdp_ptr_h, const int64_t* cdp_col_h, const double* cdp_val_h)
{
    return guard([&] {
        GF_REQUIRE(out && v && q && K_e && G_e && cdu_ptr_h && cdp_ptr_h, "null argument");
        const int et = quad_etype(q);
        GF_REQUIRE(et == GF_QUAD4_PLANAR || et == GF_HEX8, "PCG operator needs QuadraturePlanar or Hex8");
        GF_REQUIRE(quad_nelem(q) == v->nelem && quad_nip(q) == v->nne, "quadrature / vector mismatch (Gauss rule)");
        auto P = new gf_pcg_s();
        try {
            P->v = v;
            P->q = q;
            P->et = et;
            P->nne = v->nne;
            P->nd = v->ndim;
            P->K = v->nne * v->ndim;
            P->nelem = v->nelem;
            P->nnu = nnu;
            P->nnp = nnp;
            P->ndof = v->ndof;
            P->nni = nnu + nnp;
            P->nnd = P->ndof - P->nni;
            GF_REQUIRE(P->nnd >= 0 && nnu > 0, "sizes: ndof = nnu + nnp + nnd");
            P->Ke.alloc(P->nelem);
            P->Ge.alloc(P->nelem);
            GF_HIP(hipMemcpy(P->Ke.p, K_e, P->nelem * sizeof(double), hipMemcpyDeviceToDevice));
            GF_HIP(hipMemcpy(P->Ge.p, G_e, P->nelem * sizeof(double), hipMemcpyDeviceToDevice));
            upload_csr(P->cdu, cdu_ptr_h, cdu_col_h, cdu_val_h, P->nnd);
            upload_csr(P->cdp, cdp_ptr_h, cdp_col_h, cdp_val_h, P->nnd);
            for (int64_t k = 0; k < cdu_ptr_h[P->nnd]; ++k)
                GF_REQUIRE(cdu_col_h[k] >= 0 && cdu_col_h[k] < nnu, "Cdu column out of range");
            for (int64_t k = 0; k < cdp_ptr_h[P->nnd]; ++k)
                GF_REQUIRE(cdp_col_h[k] >= 0 && cdp_col_h[k] < nnp, "Cdp column out of range");
            // transpose of C_du: for each u row the (d, coefficient) pairs
            std::vector<int64_t> tptr(nnu + 1, 0);
            for (int64_t k = 0; k < cdu_ptr_h[P->nnd]; ++k) tptr[cdu_col_h[k] + 1]++;
            for (int64_t i = 0; i < nnu; ++i) tptr[i + 1] += tptr[i];
            std::vector<int64_t> tidx(tptr[nnu]), fill(tptr.begin(), tptr.end() - 1);
            std::vector<double> tval(tptr[nnu]);
            for (int64_t d = 0; d < P->nnd; ++d)
                for (int64_t k = cdu_ptr_h[d]; k < cdu_ptr_h[d + 1]; ++k) {
                    int64_t j = fill[cdu_col_h[k]]++;
                    tidx[j] = d;
                    tval[j] = cdu_val_h[k];
                }
            upload_csr(P->cduT, tptr.data(), tidx.data(), tval.data(), nnu);
            P->xfull.alloc(P->ndof);
            GF_HIP(hipMemset(P->xfull.p, 0, P->ndof * sizeof(double)));
            P->fe.alloc(P->nelem * P->K);
            for (auto* b : {&P->invdiag, &P->x, &P->r, &P->z, &P->qv, &P->tmp}) b->alloc(nnu);
            P->scal.alloc(12);
            P->state.alloc(4);
            detect_structured(P, cdu_ptr_h, cdu_col_h, cdu_val_h);
            detect_structured_hex8(P, cdu_ptr_h, cdu_col_h, cdu_val_h);
            P->partial.alloc(3 * NBR + std::max(NBR, P->ntiles));
            compute_invdiag(P, nullptr);
            GF_HIP(hipDeviceSynchronize());
        }
        catch (...) {
            delete P;
            throw;
        }
        *out = P;
    });
}

int gf_pcg_destroy(gf_pcg P)
{
    delete P;
    return GF_OK;
}

int gf_pcg_set_moduli(gf_pcg P, const double* K_e, const double* G_e, void* stream)
{
    return guard([&] {
        GF_REQUIRE(P && K_e && G_e, "null argument");
        hipStream_t s = S(stream);
        GF_HIP(hipMemcpyAsync(P->Ke.p, K_e, P->nelem * sizeof(double), hipMemcpyDeviceToDevice, s));
        GF_HIP(hipMemcpyAsync(P->Ge.p, G_e, P->nelem * sizeof(double), hipMemcpyDeviceToDevice, s));
        compute_invdiag(P, s);
    });
}

int gf_pcg_is_structured(gf_pcg P, int* structured)
{
    return guard([&] {
        GF_REQUIRE(P && structured, "null argument");
        *structured = P->structured ? 1 : 0;
    });
}

int gf_pcg_sizes(gf_pcg P, int64_t* nnu, int64_t* nnp, int64_t* nnd)
{
    return guard([&] {
        GF_REQUIRE(P, "null solver");
        if (nnu) *nnu = P->nnu;
        if (nnp) *nnp = P->nnp;
        if (nnd) *nnd = P->nnd;
    });
}

int gf_pcg_apply(gf_pcg P, const double* x_u, const double* x_p, double* y_u, void* stream)
{
    return guard([&] {
        GF_REQUIRE(P && y_u, "null argument");
        apply(P, x_u, x_p, y_u, false, S(stream), nullptr, nullptr, nullptr);
    });
}

int gf_pcg_expand(gf_pcg P, const double* x_u, const double* x_p, double* dofval, void* stream)
{
    return guard([&] {
        GF_REQUIRE(P && dofval, "null argument");
        const int64_t n = P->ndof;
        k_expand<<<std::min<unsigned>(nblocks(n, BS), 4096), BS, 0, S(stream)>>>(
            P->nnu, P->nnp, P->nnd, x_u, x_p, P->cdu.ptr.p, P->cdu.idx.p, P->cdu.val.p, P->cdp.ptr.p, P->cdp.idx.p,
            P->cdp.val.p, dofval, 1, nullptr);
        check_launch();
    });
}

int gf_pcg_inverse_diagonal(gf_pcg P, const double** invdiag)
{
    return guard([&] {
        GF_REQUIRE(P && invdiag, "null argument");
        *invdiag = P->invdiag.p;
    });
}

int gf_pcg_solve(gf_pcg P, const double* b_u, double* x_u, double tol, int64_t maxiter, int64_t check_every,
                 int64_t* iterations, double* relres, void* stream)
{
    return guard([&] {
        GF_REQUIRE(P && b_u && x_u, "null argument");
        GF_REQUIRE(tol > 0.0 && maxiter > 0 && maxiter < INT32_MAX, "tol > 0, 0 < maxiter < 2^31");
        if (check_every <= 0) check_every = 32;
        hipStream_t s = S(stream);
        const int64_t n = P->nnu;
        // Initial residual of a start (nullptr: zero) -> host (||r0||^2, ||b||^2).  The start goes to
        // the u-part of x_full, where the search direction lives during the iteration.
        auto init = [&](const double* x0) {
            if (x0) GF_HIP(hipMemcpyAsync(P->xfull.p, x0, n * sizeof(double), hipMemcpyDeviceToDevice, s));
            else k_fill_zero<<<std::min<unsigned>(nblocks(n, BS), 4096), BS, 0, s>>>(n, P->xfull.p);
            apply(P, nullptr, nullptr, P->qv.p, true, s, nullptr, nullptr, nullptr);
            k_pcg_init<<<NBR, BS, 0, s>>>(n, b_u, P->qv.p, P->invdiag.p, P->r.p, P->z.p, P->x.p, P->xfull.p,
                                          P->partial.p);
            k_pcg_init_scalars<<<1, BS, 0, s>>>(P->partial.p, P->scal.p, P->state.p, tol);
            check_launch();
            double sc[2];
            GF_HIP(hipMemcpyAsync(sc, P->scal.p + 2, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
            GF_HIP(hipStreamSynchronize(s));
            return std::make_pair(sc[1], sc[0]); // (rr, bb)
        };
        // Fallback of a bad start (a learned prediction gone wrong: NaN, huge, garbage): a start whose
        // residual is not finite or larger than ||b|| (worse than a zero guess) is replaced by the
        // previous solve's solution, when that one does better than zero, else by zero.
        auto [rr, bb] = init(x_u);
        if (!std::isfinite(bb))
            throw Error(GF_EINVAL, "gf_pcg_solve: the right-hand side is not finite (or its squared norm overflows)");
        P->last_start = GF_START_GIVEN;
        if (!(std::isfinite(rr) && rr <= bb)) {
            bool ok = false;
            if (P->has_prev) {
                std::tie(rr, bb) = init(P->xprev.p);
                ok = std::isfinite(rr) && rr <= bb;
                P->last_start = GF_START_PREVIOUS;
            }
            if (!ok) {
                std::tie(rr, bb) = init(nullptr);
                P->last_start = GF_START_ZERO;
            }
        }
        int32_t cap = static_cast<int32_t>(maxiter);
        GF_HIP(hipMemcpyAsync(P->state.p + 2, &cap, sizeof(int32_t), hipMemcpyHostToDevice, s));
        int32_t st[2] = {0, 0};
        ShapeQ4 sh;
        for (int q = 0; q < 4; ++q) {
            for (int k = 0; k < 8; ++k) sh.g[q][k] = P->sh_g[q * 8 + k];
            sh.v[q] = P->sh_v[q];
        }
        const bool done0 = rr <= tol * tol * (bb > 0.0 ? bb : 1.0);
        // Quad4: one launch per iteration (single-reduction CG).  Hex8 keeps three launches
        // (operator, update, direction): on its 3-D halo of 2.34 nodes per owned node, forming the
        // direction in the operator (z and p_prev read) measured 1.25x slower per solve, and the
        // single-reduction form (r, w, s and the diagonal read) 2.7x slower.
        const bool single = P->structured && P->et == GF_QUAD4_PLANAR;
        const bool hexs = P->structured && P->et == GF_HEX8, general = !P->structured;
        if (general && !done0) {
            // the first direction p0 = z0 (in x_full's u part) is the p_old of iteration 1
            if (P->gp[0].n != static_cast<size_t>(n)) {
                P->gp[0].alloc(n);
                P->gp[1].alloc(n);
            }
            GF_HIP(hipMemcpyAsync(P->gp[0].p, P->xfull.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        }
        if (single && !done0) {
            // r0 is in r, u0 = r0 ./ diag in the u-part of xfull: w0 = K' u0 with the partials of
            // (u0, w0) in the 'del' slot of the first launch's input
            GF_HIP(hipMemcpyAsync(P->sr[0].p, P->r.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
            apply(P, nullptr, nullptr, P->sw[0].p, true, s, P->xfull.p, P->spart.p + P->ntiles, nullptr);
        }
        StatePipe pipe;
        int chunk = 0;
        for (int64_t it = 0; it < maxiter && !done0;) {
            for (int64_t k = 0; single && k < check_every && it < maxiter; ++k, ++it) {
                const int i0 = static_cast<int>(it & 1), i1 = 1 - i0;
                k_cgsr_periodic_q4<32, 8><<<P->ntiles, 256, 0, s>>>(
                    P->sn, P->ntx, P->ntiles, P->per_xcd, sh, P->sidx.p, P->Ke.p, P->Ge.p, P->invdiag.p, P->sr[i0].p,
                    P->sw[i0].p, P->ss[i0].p, P->sr[i1].p, P->sw[i1].p, P->ss[i1].p, P->z.p, P->x.p, P->partial.p,
                    P->spart.p + 3 * i0 * P->ntiles, P->spart.p + 3 * i1 * P->ntiles, P->scal.p, P->state.p, it);
            }
            for (int64_t k = 0; hexs && k < check_every && it < maxiter; ++k, ++it) {
                apply(P, nullptr, nullptr, P->qv.p, true, s, P->xfull.p, P->partial.p + 3 * NBR, P->state.p);
                k_pcg_update<<<NBR, BS, 0, s>>>(n, P->partial.p + 3 * NBR, P->ntiles, P->scal.p, P->xfull.p, P->qv.p,
                                                P->invdiag.p, P->x.p, P->r.p, P->z.p, P->partial.p, P->state.p, it);
                k_pcg_direction<<<NBR, BS, 0, s>>>(n, P->partial.p, P->scal.p, P->z.p, P->xfull.p, P->state.p, it);
            }
            for (int64_t k = 0; general && k < check_every && it < maxiter; ++k, ++it) {
                // four launches: the expansion forming the search direction (after the first
                // iteration), element forces, ordered gather + fold with the p.q partials, update
                if (it == 0)
                    apply(P, nullptr, nullptr, P->qv.p, true, s, P->xfull.p, P->partial.p + 3 * NBR, P->state.p);
                else {
                    k_expand_dir<<<std::min<unsigned>(nblocks(P->ndof, BS), 4096), BS, 0, s>>>(
                        P->nnu, P->nnp, P->nnd, P->z.p, P->gp[(it + 1) & 1].p, P->gp[it & 1].p, P->cdu.ptr.p,
                        P->cdu.idx.p, P->cdu.val.p, P->xfull.p, P->partial.p, P->scal.p, P->state.p, it);
                    operator_from_full(P, P->qv.p, P->partial.p + 3 * NBR, P->state.p, s);
                }
                k_pcg_update<<<NBR, BS, 0, s>>>(n, P->partial.p + 3 * NBR, NBR, P->scal.p, P->xfull.p, P->qv.p,
                                                P->invdiag.p, P->x.p, P->r.p, P->z.p, P->partial.p, P->state.p, it);
            }
            check_launch();
            // the host looks at the previous chunk's flag while this chunk is queued
            pipe.post(P->state.p, chunk, s);
            if (chunk > 0 && pipe.wait(chunk - 1)[0]) break;
            ++chunk;
        }
        // true residual ||b - K' x|| / ||b||
        GF_HIP(hipMemcpyAsync(x_u, P->x.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        GF_HIP(hipMemcpyAsync(P->xfull.p, P->x.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        apply(P, nullptr, nullptr, P->qv.p, true, s, nullptr, nullptr, nullptr);
        k_residual<<<NBR, BS, 0, s>>>(n, b_u, P->qv.p, P->partial.p);
        k_residual_final<<<1, BS, 0, s>>>(P->partial.p, P->tmp.p);
        double res = 0.0;
        GF_HIP(hipMemcpyAsync(&res, P->tmp.p, sizeof(double), hipMemcpyDeviceToHost, s));
        GF_HIP(hipMemcpyAsync(st, P->state.p, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        GF_HIP(hipStreamSynchronize(s));
        if (iterations) *iterations = st[1];
        if (relres) *relres = res;
        if (!(res <= tol * 10.0)) throw Error(GF_ENOTCONV, "PCG did not reach the tolerance (true residual " +
                                                               std::to_string(res) + ")");
        // keep the solution as the previous-step start of a later fallback
        if (P->xprev.n != static_cast<size_t>(n)) P->xprev.alloc(n);
        GF_HIP(hipMemcpyAsync(P->xprev.p, P->x.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        P->has_prev = true;
    });
}

int gf_pcg_start_used(gf_pcg P, int* start)
{
    return guard([&] {
        GF_REQUIRE(P && start, "null argument");
        *start = P->last_start;
    });
}

} // extern "C"
