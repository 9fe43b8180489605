// J2 elasto-plasticity with linear isotropic hardening (small strain, 3-d tensors) per quadrature
// point: the material law the statics examples call through
// GMatElastoPlastic.Cartesian3d.LinearHardening2d (docs/examples/statics/periodic_elastoplastic.py:
// 71-104; python-gmatelastoplastic >= 0.3.1, environment.yaml:16).  GMat is not part of the
// reference; the law is the published radial return with its consistent tangent (de Souza Neto,
// Peric, Owen 2008, Box 7.3 and eq. 7.120):
//   Ee = Eps - Epsp_t;  sigm = K tr(Ee);  s = 2G dev(Ee);  q = sqrt(3/2 s:s)
//   phi = q - (sigy0 + H epsp_t);  if phi > 0:  dg = phi / (3G + H)
//   Sig = sigm I + f s,  f = 1 - 3G dg / q;   Epsp = Epsp_t + (3/2 dg / q) s;   epsp = epsp_t + dg
//   C = K I(x)I + 2G f I_d + c n(x)n,  c = 6G^2 (dg / q - 1 / (3G + H)),  n = s / |s|
// (f = 1, c = 0, dg = 0 when phi <= 0).  The committed state (Epsp_t, epsp_t) changes only in
// increment(), as in GMat, so Newton iterations within an increment re-evaluate from it.
//
// Layout: Eps, Sig, Epsp [nq, 3, 3] and C [nq, 3, 3, 3, 3] row-major FP64 (the quadrature
// kernels' qtensor layout); K, G, sigy0, H, epsp [nq].  One block handles TP points: the tensors
// are staged through LDS so every global access is a contiguous, coalesced stream, and the 81
// tangent entries per point are written by the whole block from per-point scalars kept in LDS,
// two entries per 16-byte streaming store (the outputs are not re-read by this kernel).
// The arithmetic keeps the expression order of the oracle's restatement (FP64, no contraction).
#include "gf_internal.hpp"

using namespace gf;

struct gf_plastic_s {
    int64_t nq = 0;
    DBuf<double> K, G, sy0, H;
    DBuf<double> Epsp, Epsp_t, epsp, epsp_t;
};

namespace {

constexpr int TP = 128, BS = 256; // points and threads per block (the copy loops use every thread)
typedef double dbl2 __attribute__((ext_vector_type(2)));

__global__ void __launch_bounds__(BS)
    k_plastic_refresh(int64_t nq, const double* __restrict__ Eps, const double* __restrict__ Kp,
                      const double* __restrict__ Gp, const double* __restrict__ Sy, const double* __restrict__ Hp,
                      const double* __restrict__ Epsp_t, const double* __restrict__ epsp_t,
                      double* __restrict__ Epsp, double* __restrict__ epsp, double* __restrict__ Sig,
                      double* __restrict__ C, bool pairs)
{
    __shared__ double sE[TP * 9], sP[TP * 9], sN[TP * 9], sK[TP], sF[TP], sC[TP];
    const int64_t p0 = blockIdx.x * static_cast<int64_t>(TP);
    const int np = static_cast<int>(nq - p0 < TP ? nq - p0 : TP);
    const int t = threadIdx.x;
    const double* gE = Eps + 9 * p0;
    const double* gP = Epsp_t + 9 * p0;
    for (int i = t; i < np * 9; i += BS) {
        sE[i] = gE[i];
        sP[i] = gP[i];
    }
    __syncthreads();
    if (t < np) {
        const int64_t p = p0 + t;
        double* e = sE + 9 * t;
        double* ep = sP + 9 * t;
        double* n = sN + 9 * t;
        double ee[9], s[9];
        for (int k = 0; k < 9; ++k) ee[k] = e[k] - ep[k];
        const double tr = ee[0] + ee[4] + ee[8];
        const double epsm = tr / 3.0;
        const double k_ = Kp[p], g = Gp[p], h = Hp[p], ept = epsp_t[p];
        double ss = 0.0;
        for (int k = 0; k < 9; ++k) {
            s[k] = 2.0 * g * (k % 4 == 0 ? ee[k] - epsm : ee[k]);
            ss += s[k] * s[k];
        }
        const double q = sqrt(1.5 * ss);
        const double phi = q - (Sy[p] + h * ept);
        double f = 1.0, c = 0.0, dg = 0.0;
        if (phi > 0.0) {
            dg = phi / (3.0 * g + h);
            f = 1.0 - 3.0 * g * dg / q;
            c = 6.0 * g * g * (dg / q - 1.0 / (3.0 * g + h));
            const double nrm = sqrt(ss);
            for (int k = 0; k < 9; ++k) n[k] = s[k] / nrm;
        }
        else {
            for (int k = 0; k < 9; ++k) n[k] = 0.0;
        }
        const double sigm = k_ * tr;
        const double a = 1.5 * dg / q;
        for (int k = 0; k < 9; ++k) {
            e[k] = f * s[k] + (k % 4 == 0 ? sigm : 0.0);
            if (phi > 0.0) ep[k] = ep[k] + a * s[k];
        }
        epsp[p] = ept + dg;
        sK[t] = k_;
        sF[t] = 2.0 * g * f;
        sC[t] = c;
    }
    __syncthreads();
    // outputs: pairs of doubles with streaming (non-temporal) stores where the bases allow it; the
    // block's slices of Sig, Epsp and C are contiguous
    double* oS = Sig + 9 * p0;
    double* oP = Epsp + 9 * p0;
    const int n9 = np * 9;
    if (pairs) {
        for (int i = t; i < n9 / 2; i += BS) {
            __builtin_nontemporal_store(dbl2{sE[2 * i], sE[2 * i + 1]}, reinterpret_cast<dbl2*>(oS) + i);
            __builtin_nontemporal_store(dbl2{sP[2 * i], sP[2 * i + 1]}, reinterpret_cast<dbl2*>(oP) + i);
        }
        if ((n9 & 1) && t == 0) {
            oS[n9 - 1] = sE[n9 - 1];
            oP[n9 - 1] = sP[n9 - 1];
        }
    }
    else {
        for (int i = t; i < n9; i += BS) {
            oS[i] = sE[i];
            oP[i] = sP[i];
        }
    }
    if (!C) return;
    // tangent: two entries per 16-byte streaming store over the block's contiguous slice (faster
    // than one 8-byte store per entry, even with a component per thread)
    double* oC = C + 81 * p0;
    auto entry = [&](int i) {
        const int pl = i / 81, r = i - 81 * pl, ij = r / 9, kl = r - 9 * ij;
        const int ii = ij / 3, jj = ij - 3 * ii, aa = kl / 3, bb = kl - 3 * aa;
        const double dij = ii == jj, dkl = aa == bb, dik = ii == aa, djl = jj == bb, dil = ii == bb, djk = jj == aa;
        return sK[pl] * dij * dkl + sF[pl] * (0.5 * (dik * djl + dil * djk) - dij * dkl / 3.0) +
               sC[pl] * (sN[9 * pl + ij] * sN[9 * pl + kl]);
    };
    const int n81 = np * 81;
    if (pairs) {
        for (int i = t; i < n81 / 2; i += BS)
            __builtin_nontemporal_store(dbl2{entry(2 * i), entry(2 * i + 1)}, reinterpret_cast<dbl2*>(oC) + i);
        if ((n81 & 1) && t == 0) oC[n81 - 1] = entry(n81 - 1);
    }
    else {
        for (int i = t; i < n81; i += BS) oC[i] = entry(i);
    }
}

// ---- the law fused into the Quad4 plane-strain element kernels -----------------------------------
// One pass of a Newton iteration of the statics examples (periodic_elastoplastic.py:87-104):
//   Eps = symGradN_vector(ue); refresh(); fe = int_gradN_dot_tensor2_dV(Sig);
//   Ke = int_gradN_dot_tensor4_dot_gradNT_dV(C)
// without writing Eps or C to memory.  A block owns EB elements: first one thread per quadrature
// point forms the strain, runs the return map (updating Epsp, epsp) and leaves in LDS the point's
// geometry, the in-plane stress and the four scalars that define its tangent; then one thread per
// (element, node m) forms the force fe(m, *) and the rows K(2m .. 2m + 1, *): per quadrature point it
// forms the 16 in-plane tangent entries C(i, j, k, l) and the products dNx(m, i) C(i, j, k, l) once
// and uses them for all four column nodes n.  The per-point tensors written (Epsp, and Sig when
// asked for) and the element matrices go out through LDS, so the block writes each as one contiguous,
// coalesced stream; a thread writing its own 72- or 128-byte record directly costs up to 2.5x the
// pass time.  Every value follows the expression and summation order of the unfused calls
// (Element.h:1102-1176 loops for the integrals; contraction is off), so the results are theirs bit
// for bit.
constexpr int EB = 64, NB = 256; // elements and threads per block (EB * 4 quadrature points = NB)
constexpr int KROW = 18;         // doubles per staged (element, m) row: 16 entries + 2 of bank padding
// LDS carve-up (doubles): sG[NB][8], sV[NB], sS[NB][4], sT[NB][3], sNn[NB][4]
constexpr int L_G = 0, L_V = L_G + 8 * NB, L_S = L_V + NB, L_T = L_S + 4 * NB, L_N = L_T + 3 * NB,
              L_END = L_N + 4 * NB;
static_assert(KROW * NB <= L_END, "the staged K rows must fit the point arrays");
static_assert(9 * NB <= L_END - L_S, "a staged point tensor must fit the sS, sT, sNn arrays");

// the block's slice of a per-point [., 3, 3] tensor, staged in R = smem + L_S, as one stream
__device__ inline void stream_point_tensor(const double* R, double* dst, int ne, int t)
{
    dbl2* d = reinterpret_cast<dbl2*>(dst);
    for (int w = t; w < 18 * ne; w += NB) __builtin_nontemporal_store(reinterpret_cast<const dbl2*>(R)[w], d + w);
}

__global__ void __launch_bounds__(NB)
    k_plastic_newton_q4(int64_t nelem, const double* __restrict__ dNx, const double* __restrict__ vol,
                        const double* __restrict__ ue, const double* __restrict__ Kp, const double* __restrict__ Gp,
                        const double* __restrict__ Sy, const double* __restrict__ Hp,
                        const double* __restrict__ Epsp_t, const double* __restrict__ epsp_t,
                        double* __restrict__ Epsp, double* __restrict__ epsp, double* __restrict__ fe,
                        double* __restrict__ Ke, double* __restrict__ Sig)
{
    __shared__ __align__(16) double smem[L_END];
    double(*sG)[8] = reinterpret_cast<double(*)[8]>(smem + L_G);
    double* sV = smem + L_V;
    double(*sS)[4] = reinterpret_cast<double(*)[4]>(smem + L_S);
    double(*sT)[3] = reinterpret_cast<double(*)[3]>(smem + L_T);
    double(*sNn)[4] = reinterpret_cast<double(*)[4]>(smem + L_N);
    double* R = smem + L_S; // staging for the point tensors, before sS, sT, sNn are written
    const int64_t e0 = blockIdx.x * static_cast<int64_t>(EB);
    const int ne = static_cast<int>(nelem - e0 < EB ? nelem - e0 : EB);
    const int t = threadIdx.x;
    double sg[9], kT[3], kN[4];
    if (t < 4 * ne) {
        const int64_t p = 4 * e0 + t; // quadrature point (element-major)
        const double* g = dNx + 8 * p;
        const double* u = ue + 8 * (e0 + t / 4);
        double gv[8], uv[8];
        for (int k = 0; k < 8; ++k) {
            gv[k] = g[k];
            uv[k] = u[k];
            sG[t][k] = gv[k];
        }
        const double v = vol[p];
        sV[t] = v;
        // plane-strain symmetric gradient (the planar symGradN_vector)
        double ee[9];
        for (int k = 0; k < 9; ++k) ee[k] = 0.0;
        ee[0] = gv[0] * uv[0] + gv[2] * uv[2] + gv[4] * uv[4] + gv[6] * uv[6];
        ee[4] = gv[1] * uv[1] + gv[3] * uv[3] + gv[5] * uv[5] + gv[7] * uv[7];
        ee[1] = 0.5 * (gv[0] * uv[1] + gv[2] * uv[3] + gv[4] * uv[5] + gv[6] * uv[7] + gv[1] * uv[0] + gv[3] * uv[2] +
                       gv[5] * uv[4] + gv[7] * uv[6]);
        ee[3] = ee[1];
        // return map, as in k_plastic_refresh
        const double* ept_ = Epsp_t + 9 * p;
        double ep[9];
        for (int k = 0; k < 9; ++k) {
            ep[k] = ept_[k];
            ee[k] = ee[k] - ep[k];
        }
        const double tr = ee[0] + ee[4] + ee[8];
        const double epsm = tr / 3.0;
        const double k_ = Kp[p], gm = Gp[p], h = Hp[p], ept = epsp_t[p];
        double s[9], ss = 0.0;
        for (int k = 0; k < 9; ++k) {
            s[k] = 2.0 * gm * (k % 4 == 0 ? ee[k] - epsm : ee[k]);
            ss += s[k] * s[k];
        }
        const double q = sqrt(1.5 * ss);
        const double phi = q - (Sy[p] + h * ept);
        double f = 1.0, c = 0.0, dg = 0.0, n[4] = {0.0, 0.0, 0.0, 0.0};
        if (phi > 0.0) {
            dg = phi / (3.0 * gm + h);
            f = 1.0 - 3.0 * gm * dg / q;
            c = 6.0 * gm * gm * (dg / q - 1.0 / (3.0 * gm + h));
            const double nrm = sqrt(ss);
            n[0] = s[0] / nrm;
            n[1] = s[1] / nrm;
            n[2] = s[3] / nrm;
            n[3] = s[4] / nrm;
        }
        const double sigm = k_ * tr;
        const double a = 1.5 * dg / q;
        for (int k = 0; k < 9; ++k) R[9 * t + k] = phi > 0.0 ? ep[k] + a * s[k] : ep[k];
        epsp[p] = ept + dg;
        for (int k = 0; k < 9; ++k) sg[k] = f * s[k] + (k % 4 == 0 ? sigm : 0.0);
        kT[0] = k_;
        kT[1] = 2.0 * gm * f;
        kT[2] = c;
        for (int k = 0; k < 4; ++k) kN[k] = n[k];
    }
    __syncthreads();
    stream_point_tensor(R, Epsp + 36 * e0, ne, t);
    if (Sig) { // uniform over the block
        __syncthreads();
        if (t < 4 * ne)
            for (int k = 0; k < 9; ++k) R[9 * t + k] = sg[k];
        __syncthreads();
        stream_point_tensor(R, Sig + 36 * e0, ne, t);
    }
    __syncthreads(); // R is free: the in-plane stress and the tangent scalars take its place
    if (t < 4 * ne) {
        sS[t][0] = sg[0];
        sS[t][1] = sg[1];
        sS[t][2] = sg[3];
        sS[t][3] = sg[4];
        for (int k = 0; k < 3; ++k) sT[t][k] = kT[k];
        for (int k = 0; k < 4; ++k) sNn[t][k] = kN[k];
    }
    __syncthreads();
    const int el = t / 4, m = t % 4;
    double acc[4][2][2]; // acc[n][j][k] = K(2m + j, 2n + k)
    if (t < 4 * ne) {
        for (int nn = 0; nn < 4; ++nn)
            for (int j = 0; j < 2; ++j)
                for (int k = 0; k < 2; ++k) acc[nn][j][k] = 0.0;
        double f0 = 0.0, f1 = 0.0;
        for (int qp = 0; qp < 4; ++qp) {
            const int r = 4 * el + qp;
            const double* g = sG[r];
            const double v = sV[r], K0 = sT[r][0], F0 = sT[r][1], C0 = sT[r][2];
            f0 += (g[m * 2 + 0] * sS[r][0] + g[m * 2 + 1] * sS[r][2]) * v;
            f1 += (g[m * 2 + 0] * sS[r][1] + g[m * 2 + 1] * sS[r][3]) * v;
            double gc[2][2][2][2]; // dNx(m, i) * C(i, j, k, l)
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j)
                    for (int k = 0; k < 2; ++k)
                        for (int l = 0; l < 2; ++l) {
                            const double dij = i == j, dkl = k == l, dik = i == k, djl = j == l, dil = i == l, djk = j == k;
                            const double C = K0 * dij * dkl + F0 * (0.5 * (dik * djl + dil * djk) - dij * dkl / 3.0) +
                                             C0 * (sNn[r][2 * i + j] * sNn[r][2 * k + l]);
                            gc[i][j][k][l] = g[m * 2 + i] * C;
                        }
            for (int nn = 0; nn < 4; ++nn)
                for (int i = 0; i < 2; ++i)
                    for (int j = 0; j < 2; ++j)
                        for (int k = 0; k < 2; ++k)
                            for (int l = 0; l < 2; ++l) acc[nn][j][k] += gc[i][j][k][l] * g[nn * 2 + l] * v;
        }
        __builtin_nontemporal_store(dbl2{f0, f1}, reinterpret_cast<dbl2*>(fe + 8 * (e0 + el)) + m);
    }
    __syncthreads(); // every thread is done with the point arrays
    if (t < 4 * ne) { // rows 2m and 2m + 1 of the element matrix: 16 contiguous doubles
        dbl2* row = reinterpret_cast<dbl2*>(smem + KROW * t);
        for (int j = 0; j < 2; ++j)
            for (int nn = 0; nn < 4; ++nn) row[4 * j + nn] = dbl2{acc[nn][j][0], acc[nn][j][1]};
    }
    __syncthreads();
    // the block's K is 8 * (4 ne) contiguous 16-byte pieces; piece w is piece w % 8 of row w / 8
    dbl2* K = reinterpret_cast<dbl2*>(Ke + 64 * e0);
    for (int w = t; w < 32 * ne; w += NB)
        __builtin_nontemporal_store(reinterpret_cast<const dbl2*>(smem + KROW * (w / 8))[w % 8], K + w);
}

} // namespace

extern "C" {

int gf_plastic_create(gf_plastic* out, int64_t nq, const double* K, const double* G, const double* sigy0,
                      const double* H, void* stream)
{
    return guard([&] {
        GF_REQUIRE(out && K && G && sigy0 && H, "null argument");
        GF_REQUIRE(nq > 0, "nq must be positive");
        auto M = new gf_plastic_s();
        try {
            hipStream_t s = S(stream);
            M->nq = nq;
            const size_t b1 = nq * sizeof(double), b9 = 9 * b1;
            DBuf<double>* par[4] = {&M->K, &M->G, &M->sy0, &M->H};
            const double* src[4] = {K, G, sigy0, H};
            for (int k = 0; k < 4; ++k) {
                par[k]->alloc(nq);
                GF_HIP(hipMemcpyAsync(par[k]->p, src[k], b1, hipMemcpyDeviceToDevice, s));
            }
            M->Epsp.alloc(9 * nq);
            M->Epsp_t.alloc(9 * nq);
            M->epsp.alloc(nq);
            M->epsp_t.alloc(nq);
            GF_HIP(hipMemsetAsync(M->Epsp.p, 0, b9, s));
            GF_HIP(hipMemsetAsync(M->Epsp_t.p, 0, b9, s));
            GF_HIP(hipMemsetAsync(M->epsp.p, 0, b1, s));
            GF_HIP(hipMemsetAsync(M->epsp_t.p, 0, b1, s));
            GF_HIP(hipStreamSynchronize(s));
        }
        catch (...) {
            delete M;
            throw;
        }
        *out = M;
    });
}

int gf_plastic_destroy(gf_plastic m)
{
    delete m;
    return GF_OK;
}

int gf_plastic_refresh(gf_plastic m, const double* Eps, double* Sig, double* C, void* stream)
{
    return guard([&] {
        GF_REQUIRE(m && Eps && Sig, "null argument");
        hipStream_t s = S(stream);
        // 16-byte stores need 16-byte aligned bases (each block's offset is a multiple of 16 bytes)
        const bool pairs = aligned16(Sig) && aligned16(m->Epsp.p) && (!C || aligned16(C));
        k_plastic_refresh<<<nblocks(m->nq, TP), BS, 0, s>>>(m->nq, Eps, m->K.p, m->G.p, m->sy0.p, m->H.p, m->Epsp_t.p,
                                                            m->epsp_t.p, m->Epsp.p, m->epsp.p, Sig, C, pairs);
        check_launch();
    });
}

int gf_plastic_increment(gf_plastic m, void* stream)
{
    return guard([&] {
        GF_REQUIRE(m, "null argument");
        hipStream_t s = S(stream);
        GF_HIP(hipMemcpyAsync(m->Epsp_t.p, m->Epsp.p, 9 * m->nq * sizeof(double), hipMemcpyDeviceToDevice, s));
        GF_HIP(hipMemcpyAsync(m->epsp_t.p, m->epsp.p, m->nq * sizeof(double), hipMemcpyDeviceToDevice, s));
    });
}

int gf_plastic_state(gf_plastic m, const double** Epsp, const double** epsp, const double** Epsp_t,
                     const double** epsp_t)
{
    return guard([&] {
        GF_REQUIRE(m, "null argument");
        if (Epsp) *Epsp = m->Epsp.p;
        if (epsp) *epsp = m->epsp.p;
        if (Epsp_t) *Epsp_t = m->Epsp_t.p;
        if (epsp_t) *epsp_t = m->epsp_t.p;
    });
}

int gf_plastic_newton_q4(gf_plastic m, gf_quad q, const double* elemvec, double* fe, double* Ke, double* Sig,
                         void* stream)
{
    return guard([&] {
        GF_REQUIRE(m && q && elemvec && fe && Ke, "null argument");
        int64_t nelem = 0;
        int nne = 0, ndim = 0, tdim = 0, nip = 0;
        int rc = gf_quad_sizes(q, &nelem, &nne, &ndim, &tdim, &nip);
        if (rc != GF_OK) throw Error(rc, gf_last_error());
        GF_REQUIRE(nne == 4 && ndim == 2 && tdim == 3 && nip == 4,
                   "needs a Quad4 QuadraturePlanar with four quadrature points");
        GF_REQUIRE(m->nq == 4 * nelem, "the material must hold one point per quadrature point");
        GF_REQUIRE(aligned16(fe) && aligned16(Ke) && (!Sig || aligned16(Sig)), "fe, Ke and Sig must be 16-byte aligned");
        const double *dNx = nullptr, *dV = nullptr;
        gf_quad_gradN(q, &dNx);
        gf_quad_dV(q, &dV);
        if (nelem)
            k_plastic_newton_q4<<<nblocks(nelem, EB), NB, 0, S(stream)>>>(nelem, dNx, dV, elemvec, m->K.p, m->G.p, m->sy0.p,
                                                                          m->H.p, m->Epsp_t.p, m->epsp_t.p, m->Epsp.p,
                                                                          m->epsp.p, fe, Ke, Sig);
        check_launch();
    });
}

} // extern "C"
