// Element::Quad4::QuadratureAxisymmetric (include/GooseFEM/ElementQuad4Axisymmetric.h:33-467) on the
// device.  Nodal coordinates are (z, r); tensors are in (r, theta, z) order.  Of the reference's
// B-matrix [nelem, nip, nne, 3, 3, 3] only five components per node are non-zero (:401-438):
//   B(m,r,r,r) = B(m,r,z,z) = dN/dr,  B(m,t,t,r) = N / r_q,  B(m,z,r,r) = B(m,z,z,z) = dN/dz,
// so the geometry is kept as G[e, q, m] = (dN/dr, N / r_q, dN/dz) (12 doubles per point instead of
// 108) and dV = w J 2 pi r_q.  Every kernel keeps the reference's FP64 expression and summation
// order (no contraction), so results equal the reference's bit for bit.  The stiffness runs the
// 25 products of each node pair over the five non-zero slots a (outer) and b (inner), slot =
// (i, j, value, nodal component): (r,r: dN/dr, u_r) (t,t: N/r, u_r) (z,z: dN/dz, u_z)
// (r,z: dN/dr, u_z) (z,r: dN/dz, u_r), and K(m, c_a; n, c_b) += B_a(m) C(i_a, j_a, j_b, i_b) B_b(n) dV
// (:298-398, in the listed order).
#include "gf_internal.hpp"

using namespace gf;

namespace {

constexpr int MAXQ = 16;

struct AxiRule {
    int nip;
    double N[MAXQ][4], dNxi[MAXQ][4][2], w[MAXQ];
};

} // namespace

struct gf_axi_s {
    int64_t nelem = 0;
    AxiRule rule;
    DBuf<double> G, vol;
};

namespace {

constexpr int BS = 256;

__global__ void k_axi_geometry(int64_t nelem, AxiRule R, const double* __restrict__ x, double* __restrict__ G,
                               double* __restrict__ vol)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= nelem * R.nip) return;
    const int64_t e = t / R.nip;
    const int q = static_cast<int>(t - e * R.nip);
    const double* xe = x + 8 * e;
    double xv[8];
    for (int k = 0; k < 8; ++k) xv[k] = xe[k];
    const double(*g)[2] = R.dNxi[q];
    const double* n = R.N[q];
    const double J00 = g[0][0] * xv[0] + g[1][0] * xv[2] + g[2][0] * xv[4] + g[3][0] * xv[6];
    const double J01 = g[0][0] * xv[1] + g[1][0] * xv[3] + g[2][0] * xv[5] + g[3][0] * xv[7];
    const double J10 = g[0][1] * xv[0] + g[1][1] * xv[2] + g[2][1] * xv[4] + g[3][1] * xv[6];
    const double J11 = g[0][1] * xv[1] + g[1][1] * xv[3] + g[2][1] * xv[5] + g[3][1] * xv[7];
    // cofactor inverse (include/GooseFEM/detail.h:17-70)
    const double det = J00 * J11 - J01 * J10;
    const double I00 = J11 / det, I01 = -1.0 * J01 / det, I10 = -1.0 * J10 / det, I11 = J00 / det;
    const double rq = n[0] * xv[1] + n[1] * xv[3] + n[2] * xv[5] + n[3] * xv[7];
    double* o = G + 12 * t;
    for (int m = 0; m < 4; ++m) {
        o[3 * m + 0] = I10 * g[m][0] + I11 * g[m][1];
        o[3 * m + 1] = 1.0 / rq * n[m];
        o[3 * m + 2] = I00 * g[m][0] + I01 * g[m][1];
    }
    vol[t] = R.w[q] * det * 2.0 * M_PI * rq;
}

// gradN_vector (mode 0, :140-169), gradN_vector_T (1, :171-200), symGradN_vector (2, :202-233)
__global__ void k_axi_grad(int64_t n, int nip, int mode, const double* __restrict__ G, const double* __restrict__ u_,
                           double* __restrict__ out)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n) return;
    const int64_t e = t / nip;
    const double* b = G + 12 * t;
    const double* u = u_ + 8 * e; // (u_z, u_r) per node
    const double rr = b[0] * u[1] + b[3] * u[3] + b[6] * u[5] + b[9] * u[7];
    const double tt = b[1] * u[1] + b[4] * u[3] + b[7] * u[5] + b[10] * u[7];
    const double zz = b[2] * u[0] + b[5] * u[2] + b[8] * u[4] + b[11] * u[6];
    double r[9] = {rr, 0.0, 0.0, 0.0, tt, 0.0, 0.0, 0.0, zz};
    if (mode == 2) {
        const double s = 0.5 * (b[0] * u[0] + b[3] * u[2] + b[6] * u[4] + b[9] * u[6] + b[2] * u[1] + b[5] * u[3] +
                                b[8] * u[5] + b[11] * u[7]);
        r[2] = s;
        r[6] = s;
    }
    else {
        const double dzr = b[0] * u[0] + b[3] * u[2] + b[6] * u[4] + b[9] * u[6];
        const double drz = b[2] * u[1] + b[5] * u[3] + b[8] * u[5] + b[11] * u[7];
        r[mode == 0 ? 2 : 6] = dzr;
        r[mode == 0 ? 6 : 2] = drz;
    }
    double* o = out + 9 * t;
    for (int k = 0; k < 9; ++k) o[k] = r[k];
}

// one thread per (element, node m, node n): the 2 x 2 block K(m *, n *) of the element matrix

// int_N_scalar_NT_dV (:235-266)
__global__ void k_axi_int_N(int64_t nelem, AxiRule R, const double* __restrict__ vol, const double* __restrict__ qs,
                            double* __restrict__ M)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= nelem * 16) return;
    const int64_t e = t / 16;
    const int m = static_cast<int>(t % 16) / 4, k = static_cast<int>(t % 4);
    double a = 0.0, c = 0.0;
    for (int q = 0; q < R.nip; ++q) {
        const double v = vol[e * R.nip + q], rho = qs[e * R.nip + q];
        a += R.N[q][m] * rho * R.N[q][k] * v;
        c += R.N[q][m] * rho * R.N[q][k] * v;
    }
    double* Me = M + 64 * e;
    Me[(2 * m) * 8 + 2 * k] = a;
    Me[(2 * m) * 8 + 2 * k + 1] = 0.0;
    Me[(2 * m + 1) * 8 + 2 * k] = 0.0;
    Me[(2 * m + 1) * 8 + 2 * k + 1] = c;
}

// int_gradN_dot_tensor2_dV (:268-296): one thread per (element, node)
__global__ void k_axi_int_t2(int64_t nelem, int nip, const double* __restrict__ G, const double* __restrict__ vol,
                             const double* __restrict__ sig, double* __restrict__ f)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= nelem * 4) return;
    const int64_t e = t / 4;
    const int m = static_cast<int>(t % 4);
    double f0 = 0.0, f1 = 0.0;
    for (int q = 0; q < nip; ++q) {
        const double* b = G + 12 * (e * nip + q) + 3 * m;
        const double* s = sig + 9 * (e * nip + q);
        const double v = vol[e * nip + q];
        f0 += v * (b[2] * s[8] + b[0] * s[2]);
        f1 += v * (b[0] * s[0] + b[1] * s[4] + b[2] * s[6]);
    }
    f[8 * e + 2 * m] = f0;
    f[8 * e + 2 * m + 1] = f1;
}

// int_gradN_dot_tensor4_dot_gradNT_dV (:298-398)
__global__ void k_axi_int_t4(int64_t nelem, int nip, const double* __restrict__ G, const double* __restrict__ vol,
                             const double* __restrict__ C, double* __restrict__ K)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= nelem * 16) return;
    const int64_t e = t / 16;
    const int m = static_cast<int>(t % 16) / 4, n = static_cast<int>(t % 4);
    constexpr int si[5] = {0, 1, 2, 0, 2}, sj[5] = {0, 1, 2, 2, 0}, sv[5] = {0, 1, 2, 0, 2}, sc[5] = {1, 1, 0, 0, 1};
    double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int q = 0; q < nip; ++q) {
        const double* bm = G + 12 * (e * nip + q) + 3 * m;
        const double* bn = G + 12 * (e * nip + q) + 3 * n;
        const double* c = C + 81 * (e * nip + q);
        const double v = vol[e * nip + q];
#pragma unroll
        for (int a = 0; a < 5; ++a)
#pragma unroll
            for (int b = 0; b < 5; ++b)
                acc[sc[a]][sc[b]] += bm[sv[a]] * c[((si[a] * 3 + sj[a]) * 3 + sj[b]) * 3 + si[b]] * bn[sv[b]] * v;
    }
    double* Ke = K + 64 * e;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) Ke[(2 * m + i) * 8 + 2 * n + j] = acc[i][j];
}

void compute(gf_axi m, const double* x, hipStream_t s)
{
    const int64_t n = m->nelem * m->rule.nip;
    if (n) k_axi_geometry<<<nblocks(n, BS), BS, 0, s>>>(m->nelem, m->rule, x, m->G.p, m->vol.p);
    check_launch();
}

} // namespace

extern "C" {

int gf_axi_create(gf_axi* out, int64_t nelem, int nip, const double* x, const double* xi_h, const double* w_h,
                  void* stream)
{
    return guard([&] {
        GF_REQUIRE(out && (x || nelem == 0) && xi_h && w_h, "null argument");
        GF_REQUIRE(nelem >= 0 && nip >= 1 && nip <= MAXQ, "0 <= nelem, 1 <= nip <= 16");
        auto m = new gf_axi_s();
        try {
            m->nelem = nelem;
            AxiRule& R = m->rule;
            R.nip = nip;
            for (int q = 0; q < nip; ++q) {
                const double a = xi_h[2 * q], b = xi_h[2 * q + 1];
                // ElementQuad4Axisymmetric.h:74-96 (the Quad4 shape functions)
                R.N[q][0] = 0.25 * (1.0 - a) * (1.0 - b);
                R.N[q][1] = 0.25 * (1.0 + a) * (1.0 - b);
                R.N[q][2] = 0.25 * (1.0 + a) * (1.0 + b);
                R.N[q][3] = 0.25 * (1.0 - a) * (1.0 + b);
                R.dNxi[q][0][0] = -0.25 * (1.0 - b);
                R.dNxi[q][1][0] = +0.25 * (1.0 - b);
                R.dNxi[q][2][0] = +0.25 * (1.0 + b);
                R.dNxi[q][3][0] = -0.25 * (1.0 + b);
                R.dNxi[q][0][1] = -0.25 * (1.0 - a);
                R.dNxi[q][1][1] = -0.25 * (1.0 + a);
                R.dNxi[q][2][1] = +0.25 * (1.0 + a);
                R.dNxi[q][3][1] = +0.25 * (1.0 - a);
                R.w[q] = w_h[q];
            }
            m->G.alloc(12 * nelem * nip);
            m->vol.alloc(nelem * nip);
            compute(m, x, S(stream));
        }
        catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

int gf_axi_destroy(gf_axi m)
{
    delete m;
    return GF_OK;
}

int gf_axi_update_x(gf_axi m, const double* x, void* stream)
{
    return guard([&] {
        GF_REQUIRE(m && (x || m->nelem == 0), "null argument");
        compute(m, x, S(stream));
    });
}

int gf_axi_geometry(gf_axi m, const double** G, const double** dV)
{
    return guard([&] {
        GF_REQUIRE(m, "null argument");
        if (G) *G = m->G.p;
        if (dV) *dV = m->vol.p;
    });
}

int gf_axi_grad(gf_axi m, int mode, const double* elemvec, double* qtensor, void* stream)
{
    return guard([&] {
        GF_REQUIRE(m && elemvec && qtensor, "null argument");
        GF_REQUIRE(mode >= 0 && mode <= 2, "mode: 0 gradN_vector, 1 gradN_vector_T, 2 symGradN_vector");
        const int64_t n = m->nelem * m->rule.nip;
        if (n) k_axi_grad<<<nblocks(n, BS), BS, 0, S(stream)>>>(n, m->rule.nip, mode, m->G.p, elemvec, qtensor);
        check_launch();
    });
}

int gf_axi_int_N_scalar_NT_dV(gf_axi m, const double* qscalar, double* elemmat, void* stream)
{
    return guard([&] {
        GF_REQUIRE(m && qscalar && elemmat, "null argument");
        if (m->nelem)
            k_axi_int_N<<<nblocks(16 * m->nelem, BS), BS, 0, S(stream)>>>(m->nelem, m->rule, m->vol.p, qscalar, elemmat);
        check_launch();
    });
}

int gf_axi_int_gradN_dot_tensor2_dV(gf_axi m, const double* qtensor, double* elemvec, void* stream)
{
    return guard([&] {
        GF_REQUIRE(m && qtensor && elemvec, "null argument");
        if (m->nelem)
            k_axi_int_t2<<<nblocks(4 * m->nelem, BS), BS, 0, S(stream)>>>(m->nelem, m->rule.nip, m->G.p, m->vol.p,
                                                                         qtensor, elemvec);
        check_launch();
    });
}

int gf_axi_int_gradN_dot_tensor4_dot_gradNT_dV(gf_axi m, const double* qtensor, double* elemmat, void* stream)
{
    return guard([&] {
        GF_REQUIRE(m && qtensor && elemmat, "null argument");
        if (m->nelem)
            k_axi_int_t4<<<nblocks(16 * m->nelem, BS), BS, 0, S(stream)>>>(m->nelem, m->rule.nip, m->G.p, m->vol.p,
                                                                          qtensor, elemmat);
        check_launch();
    });
}

} // extern "C"
