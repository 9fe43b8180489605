// VectorPartitioned (include/GooseFEM/VectorPartitioned.h) and MatrixDiagonalPartitioned
// (include/GooseFEM/MatrixDiagonalPartitioned.h) on the device: the DOFs of a Vector split into
// unknown (iiu = setdiff1d(dofs, iip), ascending) and prescribed (iip, in the caller's order), with
// part = Mesh::Reorder({iiu, iip}).apply(dofs) (Mesh.h:2384-2454) numbering the unknowns first.
//
// Every operation is a gather or scatter through iiu / iip / part, or a per-DOF product, so the
// FP64 results are the reference's bit for bit.  The "overwrite entries that occur more than once"
// conversions (AsDofs_u / AsDofs_p from nodevec or elemvec) take, for each DOF, the last source in
// the reference's serial loop order: the last entry of the Vector's ascending DOF-to-source lists.
// The diagonal matrix keeps the unpartitioned diagonal (assembled by gf_diag in the reference's
// (e, m, i) order, which is also the order of MatrixDiagonalPartitioned::assemble per DOF).
#include <algorithm>

#include "gf_internal.hpp"
#include "gf_wave.hpp"

using namespace gf;

struct gf_vecpart_s {
    gf_vector v = nullptr;
    int64_t nnu = 0, nnp = 0;
    std::vector<int64_t> iiu_h, iip_h, part_h;
    DBuf<int32_t> iiu, iip, part; // part [nnode * ndim]
    DBuf<int32_t> holes;          // DOF numbers below ndof that occur in neither iiu nor iip
};

struct gf_diagpart_s {
    gf_vecpart p = nullptr;
    gf_diag d = nullptr;
};

// MatrixPartitioned: the four blocks uu, up, pu, pp as CSR in the partitioned numbering (rows and
// columns ascending), each entry a copy of the unpartitioned CSR entry it comes from (src).
struct gf_matpart_s {
    gf_vecpart p = nullptr;
    gf_matrix full = nullptr;
    int64_t nrows[4] = {0, 0, 0, 0}, nnz[4] = {0, 0, 0, 0};
    DBuf<int32_t> rowptr[4], col[4], src[4];
    DBuf<double> val[4];
    DBuf<double> work; // x_u, x_p, b_u, b_p, then the PCG vectors
    DBuf<double> red;  // reduction partials and scalars
    DBuf<int32_t> state;
};

// MatrixPartitionedTyings: the unpartitioned CSR K over the tied DOF numbering [u | p | d] and the
// tying matrices C_du, C_dp (rows d) with C_ud = C_du^T (rows u), all CSR on the device.
struct gf_mattyings_s {
    gf_vector v = nullptr;
    gf_matrix K = nullptr;
    int64_t nnu = 0, nnp = 0, nnd = 0, ndof = 0;
    DBuf<int32_t> du_rp, du_cl, dp_rp, dp_cl, ud_rp, ud_cl;
    DBuf<double> du_v, dp_v, ud_v;
    DBuf<double> work; // x, r, z, p, q, invd, rhs [nnu]; xfull, yfull, kdiag [ndof]; p2 [nnu]
    DBuf<double> red;
    DBuf<int32_t> state;
};

namespace {

constexpr int BS = 256;

// dst[k] = src[idx[k]]
__global__ void k_gather(int64_t n, const int32_t* __restrict__ idx, const double* __restrict__ src,
                         double* __restrict__ dst)
{
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k < n) dst[k] = src[idx[k]];
}

__global__ void k_zero(int64_t n, double* __restrict__ a)
{
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k < n) a[k] = 0.0;
}

// a[idx[k]] = 0
__global__ void k_zero_idx(int64_t n, const int32_t* __restrict__ idx, double* __restrict__ a)
{
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k < n) a[idx[k]] = 0.0;
}

// dst[idx[k]] = src[k]
__global__ void k_scatter(int64_t n, const int32_t* __restrict__ idx, const double* __restrict__ src,
                          double* __restrict__ dst)
{
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k < n) dst[idx[k]] = src[k];
}

// out[k] = last source of DOF idx[k] (0 if the DOF has none): the overwrite-in-loop-order result
__global__ void k_gather_last(int64_t n, const int32_t* __restrict__ idx, const int32_t* __restrict__ ptr,
                              const int32_t* __restrict__ lst, const double* __restrict__ src,
                              double* __restrict__ out)
{
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= n) return;
    const int32_t d = idx[k], b = ptr[d], e = ptr[d + 1];
    out[k] = e > b ? src[lst[e - 1]] : 0.0;
}

// nodevec[j] = part[j] < nnu ? u[part[j]] : p[part[j] - nnu]   (j = m * ndim + i)
__global__ void k_node_from_part(int64_t n, int64_t nnu, const int32_t* __restrict__ part,
                                 const double* __restrict__ u, const double* __restrict__ p,
                                 double* __restrict__ nodevec)
{
    const int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (j >= n) return;
    const int64_t d = part[j];
    nodevec[j] = d < nnu ? u[d] : p[d - nnu];
}

// elemvec[(e nne + m) ndim + i] from the node (conn(e, m)) entry of the partitioned values
__global__ void k_elem_from_part(int64_t n, int nne, int nd, int64_t nnu, const int32_t* __restrict__ conn,
                                 const int32_t* __restrict__ part, const double* __restrict__ u,
                                 const double* __restrict__ p, double* __restrict__ elemvec)
{
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= n) return;
    const int64_t em = k / nd;
    const int i = static_cast<int>(k - em * nd);
    const int64_t d = part[static_cast<int64_t>(conn[em]) * nd + i];
    elemvec[k] = d < nnu ? u[d] : p[d - nnu];
}

// dest[j] = src[j] where part[j] is unknown (which 0) or prescribed (which 1)
__global__ void k_copy_part(int64_t n, int64_t nnu, int which, const int32_t* __restrict__ part,
                            const double* __restrict__ src, double* __restrict__ dest)
{
    const int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (j >= n) return;
    const bool is_u = part[j] < nnu;
    if (is_u == (which == 0)) dest[j] = src[j];
}

// b[k] = A[idx[k]] * x[k] on the partitioned values (Dot_u, Dot_p, Reaction_p), or
// x[k] = inv[idx[k]] * b[k] (Solve_u)
__global__ void k_mul_part(int64_t n, const int32_t* __restrict__ idx, const double* __restrict__ A,
                           const double* __restrict__ x, double* __restrict__ b)
{
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k < n) b[k] = A[idx[k]] * x[k];
}

// dofval: b[r] = A[r] x[r] for r = idx[k] (only the selected DOFs are written)
__global__ void k_mul_sel_dofval(int64_t n, const int32_t* __restrict__ idx, const double* __restrict__ A,
                                 const double* __restrict__ x, double* __restrict__ b)
{
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= n) return;
    const int32_t r = idx[k];
    b[r] = A[r] * x[r];
}

// nodevec: b[j] = A[dofs[j]] x[j] where part[j] is unknown (which 0) or prescribed (which 1)
__global__ void k_mul_sel_nodevec(int64_t n, int64_t nnu, int which, const int32_t* __restrict__ part,
                                  const int32_t* __restrict__ dofs, const double* __restrict__ A,
                                  const double* __restrict__ x, double* __restrict__ b)
{
    const int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (j >= n) return;
    const bool is_u = part[j] < nnu;
    if (is_u == (which == 0)) b[j] = A[dofs[j]] * x[j];
}

// block CSR values from the unpartitioned ones
__global__ void k_take(int64_t n, const int32_t* __restrict__ src, const double* __restrict__ full,
                       double* __restrict__ out)
{
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k < n) out[k] = full[src[k]];
}

// out[r] = init[r] + sign * (sum over A's row r, then B's row r, of value * x[column]), one running
// sum over ascending columns of A then of B: the order of Eigen's "dst = A xa; dst += B xb" (and
// "dst = init; dst -= B xb") for a column-major sparse matrix times a vector.  A, B or init may be
// absent (null).
__global__ void k_spmv2(int64_t nrows, const double* __restrict__ init, double sign, const int32_t* __restrict__ ra,
                        const int32_t* __restrict__ ca, const double* __restrict__ va, const double* __restrict__ xa,
                        const int32_t* __restrict__ rb, const int32_t* __restrict__ cb, const double* __restrict__ vb,
                        const double* __restrict__ xb, double* __restrict__ out)
{
    const int64_t r = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (r >= nrows) return;
    double s = init ? init[r] : 0.0;
    if (ra)
        for (int32_t k = ra[r]; k < ra[r + 1]; ++k) s += va[k] * (sign * xa[ca[k]]);
    if (rb)
        for (int32_t k = rb[r]; k < rb[r + 1]; ++k) s += vb[k] * (sign * xb[cb[k]]);
    out[r] = s;
}

// nodevec b(m, i) from the partitioned b_u / b_p: unknown entries (which 0), prescribed (1) or all (2)
__global__ void k_node_from_part_sel(int64_t n, int64_t nnu, int which, const int32_t* __restrict__ part,
                                     const double* __restrict__ u, const double* __restrict__ p,
                                     double* __restrict__ nodevec)
{
    const int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (j >= n) return;
    const int64_t d = part[j];
    if (d < nnu) {
        if (which != 1) nodevec[j] = u[d];
    }
    else if (which != 0)
        nodevec[j] = p[d - nnu];
}

// ---- Jacobi-PCG on a CSR block ----------------------------------------------------------------
// Two launches per iteration: the product (which also forms the search direction from the
// previous iteration's z and p) and the x / r / z update.
// Deterministic two-stage reductions: NPB blocks each store one partial sum, every consumer sums
// the NPB partials in index order.  scal: [0], [1] rz by iteration parity, [2] rr, [3] bb,
// [4] tol^2 bb.  state: [0] done, [1] iterations, [2] iteration cap.  Every kernel returns at once
// when done is set, so launches queued past convergence leave the result alone.
constexpr int PB = 256, NPB = 256;

// the workgroup's sum (DPP wavefront sums, waves added in order) stored by thread 0
__device__ inline void block_sum_store(double v, double* out)
{
    const double t = block_sum<PB>(v);
    if (threadIdx.x == 0) *out = t;
}

__device__ inline double sum_partials(const double* partial)
{
    double s = 0.0;
    for (int b = 0; b < NPB; ++b) s += partial[b];
    return s;
}

// The sum of count partials for a whole block: thread t adds partials t, t + PB, ... in turn, then
// the fixed-order workgroup sum; every thread of the block must call it.  Deterministic, with
// count / PB loads per thread instead of count.
__device__ inline double block_sum_partials(const double* partial, int count)
{
    double v = 0.0;
    for (int b = threadIdx.x; b < count; b += PB) v += partial[b];
    return block_sum<PB>(v);
}

// start (x given): r = b - A x, z = r / diag, p = z; partials of r.z, r.r, b.b
__global__ void __launch_bounds__(PB) k_cg_init(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ cl,
                                                const double* __restrict__ vl, const double* __restrict__ b,
                                                const double* __restrict__ x, const double* __restrict__ invd,
                                                double* __restrict__ r, double* __restrict__ z, double* __restrict__ p,
                                                double* __restrict__ partial)
{
    double arz = 0.0, arr = 0.0, abb = 0.0;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(PB) + threadIdx.x; i < n; i += static_cast<int64_t>(NPB) * PB) {
        double s = 0.0;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s += vl[k] * x[cl[k]];
        const double ri = b[i] - s;
        const double zi = invd[i] * ri;
        r[i] = ri;
        z[i] = zi;
        p[i] = zi;
        arz += ri * zi;
        arr += ri * ri;
        abb += b[i] * b[i];
    }
    block_sum_store(arz, partial + blockIdx.x);
    block_sum_store(arr, partial + NPB + blockIdx.x);
    block_sum_store(abb, partial + 2 * NPB + blockIdx.x);
}

__global__ void k_cg_init_scal(const double* __restrict__ partial, double* __restrict__ scal, int32_t* __restrict__ state,
                               double tol, int32_t maxiter)
{
    if (threadIdx.x != 0) return;
    const double rz = sum_partials(partial), rr = sum_partials(partial + NPB), bb = sum_partials(partial + 2 * NPB);
    scal[0] = rz;
    scal[2] = rr;
    scal[3] = bb;
    scal[4] = tol * tol * (bb > 0.0 ? bb : 1.0);
    // 2: non-finite right-hand side, 3: non-finite residual of the start (both stop the kernels)
    state[0] = !isfinite(bb) ? 2 : (!isfinite(rr) ? 3 : (rr <= scal[4] ? 1 : 0));
    state[1] = 0;
    state[2] = maxiter;
}

// q = A p and the per-block partials of p.q.  LPR lanes share a row (entries j, j + LPR, ...),
// combined in a fixed order; the grid covers every row, so the gathers of p have enough waves in
// flight to hide their latency.
constexpr int LPR = 4;

__global__ void __launch_bounds__(PB) k_cg_apply(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ cl,
                                                 const double* __restrict__ vl, const double* __restrict__ p,
                                                 double* __restrict__ q, double* __restrict__ partial,
                                                 const int32_t* __restrict__ state)
{
    if (state[0]) return;
    const int64_t g = blockIdx.x * static_cast<int64_t>(PB) + threadIdx.x;
    const int64_t r = g / LPR;
    const int j = static_cast<int>(g % LPR);
    double s = 0.0;
    if (r < n)
        for (int32_t k = rp[r] + j; k < rp[r + 1]; k += LPR) s += vl[k] * p[cl[k]];
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    double c = 0.0;
    if (r < n && j == 0) {
        q[r] = s;
        c = p[r] * s;
    }
    block_sum_store(c, partial + blockIdx.x);
}

// Iterations after the first: the direction update fused into the product.  beta = rz_new / rz_old
// from the previous update's partials; every gathered p(c) is formed as z(c) + beta p_prev(c) (the
// value the separate direction pass would have stored), the block's own rows are written to p_new,
// and q = A p_new with the partials of p_new.q.  Block 0 also stores rz_new, r.r, the iteration
// count, and raises done at convergence (r.r <= tol^2 b.b) or at the iteration cap.
__global__ void __launch_bounds__(PB)
    k_cg_apply_dir(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ cl,
                   const double* __restrict__ vl, const double* __restrict__ z, const double* __restrict__ pprev,
                   double* __restrict__ pnew, double* __restrict__ q, const double* __restrict__ rz_rr,
                   double* __restrict__ scal, double* __restrict__ partial, int32_t* __restrict__ state, int64_t it)
{
    if (state[0]) return;
    const double rz_old = scal[(it - 1) & 1];
    const double rz_new = block_sum_partials(rz_rr, NPB);
    const double beta = rz_old != 0.0 ? rz_new / rz_old : 0.0;
    if (blockIdx.x == 0) {
        const double rr = block_sum_partials(rz_rr + NPB, NPB);
        if (threadIdx.x == 0) {
            scal[it & 1] = rz_new;
            scal[2] = rr;
            state[1] = static_cast<int32_t>(it);
            if (rr <= scal[4] || it >= state[2]) state[0] = 1;
        }
    }
    const int64_t g = blockIdx.x * static_cast<int64_t>(PB) + threadIdx.x;
    const int64_t r = g / LPR;
    const int j = static_cast<int>(g % LPR);
    // the lane's entries k = rp[r] + j, + LPR, ... in batches of NB: NB column indices and values
    // (positions clamped to the row, so every address is in bounds), then the NB gathered operands,
    // then the products added in entry order; a loop that loads one entry per pass waits twice per
    // entry.  Same bits as that loop.
    constexpr int NB = 5;
    double sum = 0.0;
    if (r < n) {
        const int32_t e = rp[r + 1];
        for (int32_t k = rp[r] + j; k < e; k += NB * LPR) {
            int32_t c[NB];
            double v[NB], zz[NB], pp[NB];
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                const int32_t kk = min(k + q * LPR, e - 1);
                c[q] = cl[kk];
                v[q] = vl[kk];
            }
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                zz[q] = z[c[q]];
                pp[q] = pprev[c[q]];
            }
#pragma unroll
            for (int q = 0; q < NB; ++q)
                if (k + q * LPR < e) sum += v[q] * (zz[q] + beta * pp[q]);
        }
    }
    sum += __shfl_xor(sum, 1);
    sum += __shfl_xor(sum, 2);
    double contrib = 0.0;
    if (r < n && j == 0) {
        const double pr = z[r] + beta * pprev[r];
        pnew[r] = pr;
        q[r] = sum;
        contrib = pr * sum;
    }
    block_sum_store(contrib, partial + blockIdx.x);
}

// alpha = rz / p.q; x += alpha p; r -= alpha q; z = r / diag; partials of r.z and r.r
__global__ void __launch_bounds__(PB) k_cg_update(int64_t n, const double* __restrict__ pq_partial, int npq,
                                                  const double* __restrict__ scal, const double* __restrict__ invd,
                                                  const double* __restrict__ p, const double* __restrict__ q,
                                                  double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
                                                  double* __restrict__ partial, const int32_t* __restrict__ state,
                                                  int64_t it)
{
    if (state[0]) return;
    const double pq = block_sum_partials(pq_partial, npq);
    const double alpha = pq != 0.0 ? scal[it & 1] / pq : 0.0;
    double arz = 0.0, arr = 0.0;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(PB) + threadIdx.x; i < n; i += static_cast<int64_t>(NPB) * PB) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        const double zi = invd[i] * ri;
        r[i] = ri;
        z[i] = zi;
        arz += ri * zi;
        arr += ri * ri;
    }
    block_sum_store(arz, partial + blockIdx.x);
    block_sum_store(arr, partial + NPB + blockIdx.x);
}

// |b - A x|^2 partials and b.b partials
__global__ void __launch_bounds__(PB) k_cg_residual(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ cl,
                                                    const double* __restrict__ vl, const double* __restrict__ b,
                                                    const double* __restrict__ x, double* __restrict__ partial)
{
    double arr = 0.0, abb = 0.0;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(PB) + threadIdx.x; i < n; i += static_cast<int64_t>(NPB) * PB) {
        double s = 0.0;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s += vl[k] * x[cl[k]];
        const double ri = b[i] - s;
        arr += ri * ri;
        abb += b[i] * b[i];
    }
    block_sum_store(arr, partial + blockIdx.x);
    block_sum_store(abb, partial + NPB + blockIdx.x);
}

__global__ void k_cg_relres(const double* __restrict__ partial, double* __restrict__ out)
{
    if (threadIdx.x != 0) return;
    const double rr = sum_partials(partial), bb = sum_partials(partial + NPB);
    out[0] = sqrt(rr / (bb > 0.0 ? bb : 1.0));
}

// inverse diagonal of a CSR block
__global__ void k_csr_invdiag(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ cl,
                              const double* __restrict__ vl, double* __restrict__ invd)
{
    const int64_t r = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (r >= n) return;
    double d = 0.0;
    for (int32_t k = rp[r]; k < rp[r + 1]; ++k)
        if (cl[k] == r) d = vl[k];
    invd[r] = 1.0 / d;
}

// ---- tied systems: expand, fold and a CG on the condensed operator --------------------------------

// xfull = (x_u, x_p, C_du x_u + C_dp x_p); x_u or x_p may be absent (zero)
__global__ void k_ty_expand(int64_t nnu, int64_t nnp, int64_t nnd, const double* __restrict__ xu,
                            const double* __restrict__ xp, const int32_t* __restrict__ du_rp,
                            const int32_t* __restrict__ du_cl, const double* __restrict__ du_v,
                            const int32_t* __restrict__ dp_rp, const int32_t* __restrict__ dp_cl,
                            const double* __restrict__ dp_v, double* __restrict__ xfull)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    const int64_t ni = nnu + nnp;
    if (t < nnu) xfull[t] = xu ? xu[t] : 0.0;
    else if (t < ni) xfull[t] = xp ? xp[t - nnu] : 0.0;
    else if (t < ni + nnd) {
        const int64_t d = t - ni;
        double s = 0.0;
        if (xu)
            for (int32_t k = du_rp[d]; k < du_rp[d + 1]; ++k) s += du_v[k] * xu[du_cl[k]];
        if (xp)
            for (int32_t k = dp_rp[d]; k < dp_rp[d + 1]; ++k) s += dp_v[k] * xp[dp_cl[k]];
        xfull[t] = s;
    }
}

// q_u = y_u + C_ud y_d (the condensed product) and the per-block partials of p.q (p may be null)
__global__ void __launch_bounds__(PB) k_ty_fold(int64_t nnu, int64_t nni, const double* __restrict__ y,
                                                const int32_t* __restrict__ ud_rp, const int32_t* __restrict__ ud_cl,
                                                const double* __restrict__ ud_v, const double* __restrict__ p,
                                                double* __restrict__ q, double* __restrict__ partial)
{
    double acc = 0.0;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(PB) + threadIdx.x; i < nnu; i += static_cast<int64_t>(NPB) * PB) {
        double s = y[i];
        for (int32_t k = ud_rp[i]; k < ud_rp[i + 1]; ++k) s += ud_v[k] * y[nni + ud_cl[k]];
        q[i] = s;
        if (p) acc += p[i] * s;
    }
    if (partial) block_sum_store(acc, partial + blockIdx.x);
}

// The condensed product and its p.q partials in one launch: q_u = (K x)_u + C_ud (K x)_d, where
// the product rows of the dependent DOFs tied to u are formed where they are folded (a periodic
// cell ties a handful of them to each boundary DOF and none to the interior).  LPR lanes per row
// (lane-strided products, then shuffle sums in a fixed order); the partials are per workgroup.
__global__ void __launch_bounds__(PB) k_ty_spmv_fold(int64_t nnu, int64_t nni, const int32_t* __restrict__ krp,
                                                     const int32_t* __restrict__ kcl, const double* __restrict__ kv,
                                                     const double* __restrict__ x, const int32_t* __restrict__ ud_rp,
                                                     const int32_t* __restrict__ ud_cl, const double* __restrict__ ud_v,
                                                     const double* __restrict__ p, double* __restrict__ q,
                                                     double* __restrict__ partial, const int32_t* __restrict__ state)
{
    if (state[0]) return;
    const int64_t g = blockIdx.x * static_cast<int64_t>(PB) + threadIdx.x;
    const int64_t i = g / LPR;
    const int j = static_cast<int>(g % LPR);
    auto rowdot = [&](int64_t r) { // the same value in every lane of the row's group
        double s = 0.0;
        for (int32_t k = krp[r] + j; k < krp[r + 1]; k += LPR) s += kv[k] * x[kcl[k]];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        return s;
    };
    double c = 0.0;
    if (i < nnu) { // uniform over the row's lanes
        double s = rowdot(i);
        for (int32_t k = ud_rp[i]; k < ud_rp[i + 1]; ++k) s += ud_v[k] * rowdot(nni + ud_cl[k]);
        if (j == 0) {
            q[i] = s;
            c = p[i] * s;
        }
    }
    block_sum_store(c, partial + blockIdx.x);
}

// the diagonal of each CSR row (0 where the row has none)
__global__ void k_csr_diag(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ cl,
                           const double* __restrict__ vl, double* __restrict__ dg)
{
    const int64_t r = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (r >= n) return;
    double d = 0.0;
    for (int32_t k = rp[r]; k < rp[r + 1]; ++k)
        if (cl[k] == r) d = vl[k];
    dg[r] = d;
}

// Jacobi preconditioner of the condensed operator: 1 / (K_ii + sum_d C_ud(i, d)^2 K_dd)
__global__ void k_ty_invdiag(int64_t nnu, int64_t nni, const double* __restrict__ kd, const int32_t* __restrict__ ud_rp,
                             const int32_t* __restrict__ ud_cl, const double* __restrict__ ud_v,
                             double* __restrict__ invd)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= nnu) return;
    double d = kd[i];
    for (int32_t k = ud_rp[i]; k < ud_rp[i + 1]; ++k) d += ud_v[k] * ud_v[k] * kd[nni + ud_cl[k]];
    invd[i] = d != 0.0 ? 1.0 / d : 1.0;
}

// rhs = b_u + C_ud b_d - t  (t: the condensed product of the prescribed part)
__global__ void k_ty_rhs(int64_t nnu, int64_t nni, const double* __restrict__ bfull, const int32_t* __restrict__ ud_rp,
                         const int32_t* __restrict__ ud_cl, const double* __restrict__ ud_v, const double* __restrict__ t,
                         double* __restrict__ rhs)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= nnu) return;
    double s = bfull[i];
    for (int32_t k = ud_rp[i]; k < ud_rp[i + 1]; ++k) s += ud_v[k] * bfull[nni + ud_cl[k]];
    rhs[i] = s - t[i];
}

// y += A x for a CSR A (rows in order, entries in order)
__global__ void k_csr_addmv(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ cl,
                            const double* __restrict__ vl, const double* __restrict__ x, double* __restrict__ y)
{
    const int64_t r = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (r >= n) return;
    double s = y[r];
    for (int32_t k = rp[r]; k < rp[r + 1]; ++k) s += vl[k] * x[cl[k]];
    y[r] = s;
}

// |b - q|^2 and |b|^2 partials (q = A x formed by the caller)
__global__ void __launch_bounds__(PB) k_cg_residual_vec(int64_t n, const double* __restrict__ b,
                                                        const double* __restrict__ q, double* __restrict__ partial)
{
    double arr = 0.0, abb = 0.0;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(PB) + threadIdx.x; i < n; i += static_cast<int64_t>(NPB) * PB) {
        const double ri = b[i] - q[i];
        arr += ri * ri;
        abb += b[i] * b[i];
    }
    block_sum_store(arr, partial + blockIdx.x);
    block_sum_store(abb, partial + NPB + blockIdx.x);
}

// Iterations after the first: the direction update fused into the expansion of the operand.  beta =
// rz_new / rz_old from the previous update's partials (identical in every block); the u part of the
// full vector is the new direction p = z + beta p_prev (also written to pnew), the prescribed part 0,
// and each dependent entry sum_k C_du(d, k) p(k) with every p(k) formed the same way, so the values
// are the ones a separate direction pass would have stored.  Block 0 stores rz_new, r.r and the
// iteration count, and raises done at convergence (r.r <= tol^2 b.b).
__global__ void __launch_bounds__(PB) k_ty_expand_dir(int64_t nnu, int64_t nnp, int64_t nnd,
                                                      const double* __restrict__ z, const double* __restrict__ pprev,
                                                      double* __restrict__ pnew, const int32_t* __restrict__ du_rp,
                                                      const int32_t* __restrict__ du_cl, const double* __restrict__ du_v,
                                                      double* __restrict__ xfull, const double* __restrict__ partial,
                                                      double* __restrict__ scal, int32_t* __restrict__ state, int64_t it)
{
    if (state[0]) return;
    const double rz_new = block_sum_partials(partial, NPB);
    const double rr = block_sum_partials(partial + NPB, NPB);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[2] = rr;
        state[1] = static_cast<int32_t>(it); // the updates done
    }
    if (rr <= scal[4]) { // converged after `it` iterations
        if (blockIdx.x == 0 && threadIdx.x == 0) state[0] = 1;
        return;
    }
    const double rz_old = scal[(it - 1) & 1];
    const double beta = rz_old != 0.0 ? rz_new / rz_old : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) scal[it & 1] = rz_new;
    const int64_t t = blockIdx.x * static_cast<int64_t>(PB) + threadIdx.x;
    const int64_t ni = nnu + nnp;
    if (t < nnu) {
        const double pv = z[t] + beta * pprev[t];
        pnew[t] = pv;
        xfull[t] = pv;
    }
    else if (t < ni)
        xfull[t] = 0.0;
    else if (t < ni + nnd) {
        const int64_t d = t - ni;
        double sd = 0.0;
        for (int32_t k = du_rp[d]; k < du_rp[d + 1]; ++k) {
            const int32_t c = du_cl[k];
            sd += du_v[k] * (z[c] + beta * pprev[c]);
        }
        xfull[t] = sd;
    }
}

// start from x = 0: r = z... (r = rhs, z = r / diag, p = z); partials of r.z, r.r, b.b
__global__ void __launch_bounds__(PB) k_ty_init(int64_t n, const double* __restrict__ b, const double* __restrict__ invd,
                                                double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
                                                double* __restrict__ p, double* __restrict__ partial)
{
    double arz = 0.0, arr = 0.0;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(PB) + threadIdx.x; i < n; i += static_cast<int64_t>(NPB) * PB) {
        const double ri = b[i], zi = invd[i] * ri;
        x[i] = 0.0;
        r[i] = ri;
        z[i] = zi;
        p[i] = zi;
        arz += ri * zi;
        arr += ri * ri;
    }
    block_sum_store(arz, partial + blockIdx.x);
    block_sum_store(arr, partial + NPB + blockIdx.x);
    block_sum_store(arr, partial + 2 * NPB + blockIdx.x);
}

} // namespace

extern "C" {

int gf_vecpart_create(gf_vecpart* out, gf_vector v, const int64_t* iip_h, int64_t nnp)
{
    return guard([&] {
        GF_REQUIRE(out && v && (iip_h || nnp == 0) && nnp >= 0, "invalid argument");
        auto P = new gf_vecpart_s();
        try {
            P->v = v;
            const int64_t ndof = v->ndof;
            std::vector<char> is_p(ndof, 0);
            for (int64_t k = 0; k < nnp; ++k) {
                GF_REQUIRE(iip_h[k] >= 0 && iip_h[k] < ndof, "iip entry outside [0, ndof)");
                GF_REQUIRE(!is_p[iip_h[k]], "iip entries must be unique");
                is_p[iip_h[k]] = 1;
            }
            P->iip_h.assign(iip_h, iip_h + nnp);
            // iiu = setdiff1d(dofs, iip): the DOFs that occur in dofs and not in iip, ascending
            std::vector<char> seen(ndof, 0);
            for (int64_t d : v->dofs_h) seen[d] = 1;
            for (int64_t d = 0; d < ndof; ++d)
                if (seen[d] && !is_p[d]) P->iiu_h.push_back(d);
            P->nnu = static_cast<int64_t>(P->iiu_h.size());
            P->nnp = nnp;
            // Reorder({iiu, iip}): renum[iiu[k]] = k, renum[iip[k]] = nnu + k
            std::vector<int64_t> renum(ndof, -1);
            for (int64_t k = 0; k < P->nnu; ++k) renum[P->iiu_h[k]] = k;
            for (int64_t k = 0; k < nnp; ++k) renum[iip_h[k]] = P->nnu + k;
            P->part_h.resize(v->dofs_h.size());
            for (size_t j = 0; j < v->dofs_h.size(); ++j) P->part_h[j] = renum[v->dofs_h[j]];
            std::vector<int64_t> holes;
            for (int64_t d = 0; d < ndof; ++d)
                if (renum[d] < 0) holes.push_back(d);
            P->holes.upload(narrow(holes.data(), holes.size(), "holes"));
            P->iiu.upload(narrow(P->iiu_h.data(), P->iiu_h.size(), "iiu"));
            P->iip.upload(narrow(P->iip_h.data(), P->iip_h.size(), "iip"));
            P->part.upload(narrow(P->part_h.data(), P->part_h.size(), "part"));
        }
        catch (...) {
            delete P;
            throw;
        }
        *out = P;
    });
}

int gf_vecpart_destroy(gf_vecpart p)
{
    delete p;
    return GF_OK;
}

int gf_vecpart_sizes(gf_vecpart p, int64_t* nnu, int64_t* nnp)
{
    return guard([&] {
        GF_REQUIRE(p, "null argument");
        if (nnu) *nnu = p->nnu;
        if (nnp) *nnp = p->nnp;
    });
}

int gf_vecpart_index(gf_vecpart p, int64_t* iiu_h, int64_t* iip_h, int64_t* part_h)
{
    return guard([&] {
        GF_REQUIRE(p, "null argument");
        if (iiu_h) std::copy(p->iiu_h.begin(), p->iiu_h.end(), iiu_h);
        if (iip_h) std::copy(p->iip_h.begin(), p->iip_h.end(), iip_h);
        if (part_h) std::copy(p->part_h.begin(), p->part_h.end(), part_h);
    });
}

int gf_vecpart_as_dofs(gf_vecpart p, int kind, const double* in, double* dofval_u, double* dofval_p, void* stream)
{
    return guard([&] {
        GF_REQUIRE(p && in, "null argument");
        GF_REQUIRE(kind >= 0 && kind <= 2, "kind must be 0 (dofval), 1 (nodevec) or 2 (elemvec)");
        auto v = p->v;
        hipStream_t s = S(stream);
        const int32_t* ptr = kind == 1 ? v->dof_nptr.p : v->dof_eptr.p;
        const int32_t* lst = kind == 1 ? v->dof_nidx.p : v->dof_eidx.p;
        const struct {
            double* out;
            int64_t n;
            const int32_t* idx;
        } parts[2] = {{dofval_u, p->nnu, p->iiu.p}, {dofval_p, p->nnp, p->iip.p}};
        for (const auto& q : parts) {
            if (!q.out || q.n == 0) continue;
            if (kind == 0) k_gather<<<nblocks(q.n, BS), BS, 0, s>>>(q.n, q.idx, in, q.out);
            else k_gather_last<<<nblocks(q.n, BS), BS, 0, s>>>(q.n, q.idx, ptr, lst, in, q.out);
            check_launch();
        }
    });
}

int gf_vecpart_from_partitioned(gf_vecpart p, int kind, const double* dofval_u, const double* dofval_p, double* out,
                                void* stream)
{
    return guard([&] {
        GF_REQUIRE(p && out && (dofval_u || p->nnu == 0) && (dofval_p || p->nnp == 0), "null argument");
        GF_REQUIRE(kind >= 0 && kind <= 2, "kind must be 0 (dofval), 1 (nodevec) or 2 (elemvec)");
        auto v = p->v;
        hipStream_t s = S(stream);
        if (kind == 0) {
            // dofsFromParitioned zero-fills first (VectorPartitioned.h:303): DOF numbers that occur in
            // neither iiu nor iip (absent from dofs) stay 0
            if (v->ndof) k_zero<<<nblocks(v->ndof, BS), BS, 0, s>>>(v->ndof, out);
            if (p->nnu) k_scatter<<<nblocks(p->nnu, BS), BS, 0, s>>>(p->nnu, p->iiu.p, dofval_u, out);
            if (p->nnp) k_scatter<<<nblocks(p->nnp, BS), BS, 0, s>>>(p->nnp, p->iip.p, dofval_p, out);
        }
        else if (kind == 1) {
            const int64_t n = v->nnode * v->ndim;
            if (n) k_node_from_part<<<nblocks(n, BS), BS, 0, s>>>(n, p->nnu, p->part.p, dofval_u, dofval_p, out);
        }
        else {
            const int64_t n = v->nelem * v->nne * v->ndim;
            if (n)
                k_elem_from_part<<<nblocks(n, BS), BS, 0, s>>>(n, v->nne, v->ndim, p->nnu, v->conn.p, p->part.p,
                                                               dofval_u, dofval_p, out);
        }
        check_launch();
    });
}

int gf_vecpart_copy(gf_vecpart p, int which, const double* nodevec_src, double* nodevec_dest, void* stream)
{
    return guard([&] {
        GF_REQUIRE(p && nodevec_src && nodevec_dest, "null argument");
        GF_REQUIRE(which == 0 || which == 1, "which must be 0 (unknown) or 1 (prescribed)");
        const int64_t n = p->v->nnode * p->v->ndim;
        if (n) k_copy_part<<<nblocks(n, BS), BS, 0, S(stream)>>>(n, p->nnu, which, p->part.p, nodevec_src, nodevec_dest);
        check_launch();
    });
}

int gf_diagpart_create(gf_diagpart* out, gf_vecpart p)
{
    return guard([&] {
        GF_REQUIRE(out && p, "null argument");
        auto D = new gf_diagpart_s();
        D->p = p;
        const int rc = gf_diag_create(&D->d, p->v);
        if (rc != GF_OK) {
            delete D;
            throw Error(rc, gf_last_error());
        }
        *out = D;
    });
}

int gf_diagpart_destroy(gf_diagpart d)
{
    if (d) gf_diag_destroy(d->d);
    delete d;
    return GF_OK;
}

int gf_diagpart_diag(gf_diagpart d, gf_diag* out)
{
    return guard([&] {
        GF_REQUIRE(d && out, "null argument");
        *out = d->d;
    });
}

int gf_diagpart_set(gf_diagpart d, const double* A, void* stream)
{
    return guard([&] {
        GF_REQUIRE(d && A, "null argument");
        int rc = gf_diag_set(d->d, A, stream);
        if (rc != GF_OK) throw Error(rc, gf_last_error());
        // set() copies only the unknown and prescribed entries (MatrixDiagonalPartitioned.h:121-135)
        const double* Ad = nullptr;
        rc = gf_diag_data(d->d, &Ad);
        if (rc != GF_OK) throw Error(rc, gf_last_error());
        const int64_t n = static_cast<int64_t>(d->p->holes.n);
        if (n) k_zero_idx<<<nblocks(n, BS), BS, 0, S(stream)>>>(n, d->p->holes.p, const_cast<double*>(Ad));
        check_launch();
    });
}

int gf_diagpart_data_part(gf_diagpart d, double* A_uu, double* A_pp, void* stream)
{
    return guard([&] {
        GF_REQUIRE(d, "null argument");
        const double* A = nullptr;
        int rc = gf_diag_data(d->d, &A);
        if (rc != GF_OK) throw Error(rc, gf_last_error());
        hipStream_t s = S(stream);
        auto p = d->p;
        if (A_uu && p->nnu) k_gather<<<nblocks(p->nnu, BS), BS, 0, s>>>(p->nnu, p->iiu.p, A, A_uu);
        if (A_pp && p->nnp) k_gather<<<nblocks(p->nnp, BS), BS, 0, s>>>(p->nnp, p->iip.p, A, A_pp);
        check_launch();
    });
}

int gf_diagpart_dot_part(gf_diagpart d, int which, const double* x, double* b, void* stream)
{
    return guard([&] {
        GF_REQUIRE(d && x && b, "null argument");
        GF_REQUIRE(which == 0 || which == 1, "which must be 0 (unknown) or 1 (prescribed)");
        const double* A = nullptr;
        int rc = gf_diag_data(d->d, &A);
        if (rc != GF_OK) throw Error(rc, gf_last_error());
        auto p = d->p;
        const int64_t n = which == 0 ? p->nnu : p->nnp;
        if (n) k_mul_part<<<nblocks(n, BS), BS, 0, S(stream)>>>(n, which == 0 ? p->iiu.p : p->iip.p, A, x, b);
        check_launch();
    });
}

int gf_diagpart_solve_u(gf_diagpart d, const double* b_u, double* x_u, void* stream)
{
    return guard([&] {
        GF_REQUIRE(d && b_u && x_u, "null argument");
        const double* inv = nullptr;
        int rc = gf_diag_inverse(d->d, &inv, stream);
        if (rc != GF_OK) throw Error(rc, gf_last_error());
        auto p = d->p;
        if (p->nnu) k_mul_part<<<nblocks(p->nnu, BS), BS, 0, S(stream)>>>(p->nnu, p->iiu.p, inv, b_u, x_u);
        check_launch();
    });
}

int gf_diagpart_apply_sel(gf_diagpart d, int op, int kind, const double* in, double* out, void* stream)
{
    return guard([&] {
        GF_REQUIRE(d && in && out, "null argument");
        GF_REQUIRE(op == 0 || op == 1, "op must be 0 (Solve) or 1 (Reaction)");
        GF_REQUIRE(kind == 0 || kind == 1, "kind must be 0 (dofval) or 1 (nodevec)");
        const double* M = nullptr;
        int rc = op == 0 ? gf_diag_inverse(d->d, &M, stream) : gf_diag_data(d->d, &M);
        if (rc != GF_OK) throw Error(rc, gf_last_error());
        auto p = d->p;
        auto v = p->v;
        hipStream_t s = S(stream);
        if (kind == 0) {
            const int64_t n = op == 0 ? p->nnu : p->nnp;
            if (n) k_mul_sel_dofval<<<nblocks(n, BS), BS, 0, s>>>(n, op == 0 ? p->iiu.p : p->iip.p, M, in, out);
        }
        else {
            const int64_t n = v->nnode * v->ndim;
            if (n) k_mul_sel_nodevec<<<nblocks(n, BS), BS, 0, s>>>(n, p->nnu, op, p->part.p, v->dofs.p, M, in, out);
        }
        check_launch();
    });
}

} // extern "C"

// ---- MatrixPartitioned ----------------------------------------------------------------------------

namespace {

// work layout (doubles): x_u [nnu], x_p [nnp], b_u [nnu], b_p [nnp], then the PCG vectors
// x, r, z, p, q, invdiag, rhs, second p [nnu each]
struct MatPartWork {
    double *xu, *xp, *bu, *bp, *cx, *cr, *cz, *cp, *cq, *invd, *rhs, *cp2;
};

MatPartWork work_of(gf_matpart A)
{
    const int64_t nu = A->p->nnu, np = A->p->nnp;
    double* w = A->work.p;
    MatPartWork W;
    W.xu = w;
    W.xp = W.xu + nu;
    W.bu = W.xp + np;
    W.bp = W.bu + nu;
    W.cx = W.bp + np;
    W.cr = W.cx + nu;
    W.cz = W.cr + nu;
    W.cp = W.cz + nu;
    W.cq = W.cp + nu;
    W.invd = W.cq + nu;
    W.rhs = W.invd + nu;
    W.cp2 = W.rhs + nu;
    return W;
}

// B_u (which 0) = A_uu x_u + A_up x_p, or B_p (which 1) = A_pu x_u + A_pp x_p
void matpart_product(gf_matpart A, int which, const double* xu, const double* xp, double* out, hipStream_t s)
{
    const int b0 = which == 0 ? 0 : 2, b1 = b0 + 1;
    const int64_t n = A->nrows[b0];
    if (!n) return;
    k_spmv2<<<nblocks(n, BS), BS, 0, s>>>(n, nullptr, 1.0, A->rowptr[b0].p, A->col[b0].p, A->val[b0].p, xu,
                                          A->rowptr[b1].p, A->col[b1].p, A->val[b1].p, xp, out);
    check_launch();
}

void vecpart_as_dofs(gf_vecpart p, int kind, const double* in, double* u, double* pp, void* stream)
{
    const int rc = gf_vecpart_as_dofs(p, kind, in, u, pp, stream);
    if (rc != GF_OK) throw Error(rc, gf_last_error());
}

// x_u = A_uu^-1 rhs by Jacobi-PCG from the x_u given (warm) or from zero
void matpart_cg(gf_matpart A, const double* rhs, double* xu, bool warm, double tol, int64_t maxiter, int64_t* iterations,
                double* relres, hipStream_t s)
{
    GF_REQUIRE(tol > 0.0 && maxiter > 0 && maxiter < INT32_MAX, "tol > 0, 0 < maxiter < 2^31");
    const int64_t n = A->nrows[0];
    if (n == 0) {
        if (iterations) *iterations = 0;
        if (relres) *relres = 0.0;
        return;
    }
    auto W = work_of(A);
    double* partial = A->red.p;
    double* scal = partial + 3 * NPB;
    int32_t* state = A->state.p;
    const int32_t *rp = A->rowptr[0].p, *cl = A->col[0].p;
    const double* vl = A->val[0].p;
    int32_t st[2] = {0, 0};
    auto init = [&](bool from_x) {
        if (from_x)
            GF_HIP(hipMemcpyAsync(W.cx, xu, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        else
            k_zero<<<nblocks(n, BS), BS, 0, s>>>(n, W.cx);
        k_cg_init<<<NPB, PB, 0, s>>>(n, rp, cl, vl, rhs, W.cx, W.invd, W.cr, W.cz, W.cp, partial);
        k_cg_init_scal<<<1, 64, 0, s>>>(partial, scal, state, tol, static_cast<int32_t>(maxiter));
        check_launch();
        GF_HIP(hipMemcpyAsync(st, state, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        GF_HIP(hipStreamSynchronize(s));
    };
    init(warm);
    if (st[0] == 2) throw Error(GF_EINVAL, "non-finite right-hand side");
    if (st[0] == 3) init(false); // a start with non-finite entries: solve from zero instead
    constexpr int64_t check_every = 16;
    const unsigned napply = nblocks(LPR * n, PB);
    double* pq_part = partial + 3 * NPB + 8; // one p.q partial per apply block
    double* pb[2] = {W.cp, W.cp2};
    int64_t it = 0;
    StatePipe pipe;
    int chunk = 0;
    while (!st[0] && it < maxiter) {
        for (int64_t k = 0; k < check_every && it < maxiter; ++k, ++it) {
            if (it == 0)
                k_cg_apply<<<napply, PB, 0, s>>>(n, rp, cl, vl, pb[0], W.cq, pq_part, state);
            else
                k_cg_apply_dir<<<napply, PB, 0, s>>>(n, rp, cl, vl, W.cz, pb[(it - 1) & 1], pb[it & 1], W.cq, partial,
                                                     scal, pq_part, state, it);
            k_cg_update<<<NPB, PB, 0, s>>>(n, pq_part, static_cast<int>(napply), scal, W.invd, pb[it & 1], W.cq, W.cx,
                                           W.cr, W.cz, partial, state, it);
        }
        if (it == maxiter) // the convergence test and iteration count of the last update
            k_cg_apply_dir<<<napply, PB, 0, s>>>(n, rp, cl, vl, W.cz, pb[(it - 1) & 1], pb[it & 1], W.cq, partial, scal,
                                                 pq_part, state, it);
        check_launch();
        // the host looks at the previous chunk's flag while this chunk is queued
        pipe.post(state, chunk, s);
        if (chunk > 0 && pipe.wait(chunk - 1)[0]) st[0] = 1;
        ++chunk;
    }
    // the true residual ||rhs - A_uu x_u|| / ||rhs||
    k_cg_residual<<<NPB, PB, 0, s>>>(n, rp, cl, vl, rhs, W.cx, partial);
    k_cg_relres<<<1, 64, 0, s>>>(partial, scal + 5);
    check_launch();
    GF_HIP(hipMemcpyAsync(xu, W.cx, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    double res = 0.0;
    GF_HIP(hipMemcpyAsync(&res, scal + 5, sizeof(double), hipMemcpyDeviceToHost, s));
    GF_HIP(hipMemcpyAsync(st, state, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GF_HIP(hipStreamSynchronize(s));
    if (iterations) *iterations = st[1];
    if (relres) *relres = res;
    if (!(res <= 10.0 * tol))
        throw Error(GF_ENOTCONV, "PCG on A_uu did not reach the tolerance (true residual " + std::to_string(res) + ")");
}

// block values from the unpartitioned CSR, and the Jacobi preconditioner of A_uu
void matpart_split(gf_matpart A, hipStream_t s)
{
    const int32_t *rp = nullptr, *cl = nullptr;
    const double* val = nullptr;
    gf_matrix_csr(A->full, &rp, &cl, &val);
    for (int b = 0; b < 4; ++b) {
        const int64_t n = A->nnz[b];
        if (n) k_take<<<nblocks(n, BS), BS, 0, s>>>(n, A->src[b].p, val, A->val[b].p);
    }
    const int64_t nu = A->nrows[0];
    if (nu) k_csr_invdiag<<<nblocks(nu, BS), BS, 0, s>>>(nu, A->rowptr[0].p, A->col[0].p, A->val[0].p, work_of(A).invd);
    check_launch();
}

} // namespace

extern "C" {

int gf_matpart_create(gf_matpart* out, gf_vecpart p)
{
    return guard([&] {
        GF_REQUIRE(out && p, "null argument");
        auto A = new gf_matpart_s();
        try {
            A->p = p;
            int rc = gf_matrix_create(&A->full, p->v);
            if (rc != GF_OK) throw Error(rc, gf_last_error());
            int64_t nrows = 0, nnz = 0;
            gf_matrix_sizes(A->full, &nrows, &nnz);
            const int32_t *drp = nullptr, *dcl = nullptr;
            const double* dval = nullptr;
            gf_matrix_csr(A->full, &drp, &dcl, &dval);
            std::vector<int32_t> rp(nrows + 1), cl(nnz);
            GF_HIP(hipMemcpy(rp.data(), drp, (nrows + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
            if (nnz) GF_HIP(hipMemcpy(cl.data(), dcl, nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
            // renum: DOF -> partitioned number (unknowns first), -1 for holes
            const int64_t nu = p->nnu, np = p->nnp;
            std::vector<int64_t> renum(nrows, -1);
            for (int64_t k = 0; k < nu; ++k) renum[p->iiu_h[k]] = k;
            for (int64_t k = 0; k < np; ++k) renum[p->iip_h[k]] = nu + k;
            // one row of the unpartitioned CSR per partitioned row
            std::vector<int64_t> row_of(nu + np, -1);
            for (int64_t d = 0; d < nrows; ++d) {
                if (renum[d] >= 0)
                    row_of[renum[d]] = d;
                else
                    GF_REQUIRE(rp[d + 1] == rp[d], "a DOF number outside iiu and iip has matrix entries");
            }
            A->nrows[0] = A->nrows[1] = nu;
            A->nrows[2] = A->nrows[3] = np;
            std::vector<int32_t> brp[4], bcl[4], bsrc[4];
            for (int b = 0; b < 4; ++b) brp[b].assign(A->nrows[b] + 1, 0);
            std::vector<std::pair<int64_t, int32_t>> ent; // (partitioned column, unpartitioned position)
            for (int64_t pr = 0; pr < nu + np; ++pr) {
                const int64_t d = row_of[pr];
                const int rb = pr < nu ? 0 : 2;
                const int64_t lr = pr < nu ? pr : pr - nu;
                ent.clear();
                if (d >= 0)
                    for (int32_t k = rp[d]; k < rp[d + 1]; ++k) ent.emplace_back(renum[cl[k]], k);
                std::sort(ent.begin(), ent.end());
                for (auto& e : ent) {
                    const int b = rb + (e.first < nu ? 0 : 1);
                    bcl[b].push_back(static_cast<int32_t>(e.first < nu ? e.first : e.first - nu));
                    bsrc[b].push_back(e.second);
                }
                brp[rb][lr + 1] = static_cast<int32_t>(bcl[rb].size());
                brp[rb + 1][lr + 1] = static_cast<int32_t>(bcl[rb + 1].size());
            }
            for (int b = 0; b < 4; ++b) {
                A->nnz[b] = static_cast<int64_t>(bcl[b].size());
                A->rowptr[b].upload(brp[b]);
                A->col[b].upload(bcl[b]);
                A->src[b].upload(bsrc[b]);
                A->val[b].alloc(A->nnz[b]);
            }
            A->work.alloc(2 * (nu + np) + 8 * nu);
            A->red.alloc(3 * NPB + 8 + nblocks(LPR * nu, PB));
            A->state.alloc(4);
        }
        catch (...) {
            if (A->full) gf_matrix_destroy(A->full);
            delete A;
            throw;
        }
        *out = A;
    });
}

int gf_matpart_destroy(gf_matpart A)
{
    if (A && A->full) gf_matrix_destroy(A->full);
    delete A;
    return GF_OK;
}

int gf_matpart_sizes(gf_matpart A, int64_t* nnz)
{
    return guard([&] {
        GF_REQUIRE(A && nnz, "null argument");
        for (int b = 0; b < 4; ++b) nnz[b] = A->nnz[b];
    });
}

int gf_matpart_matrix(gf_matpart A, gf_matrix* full)
{
    return guard([&] {
        GF_REQUIRE(A && full, "null argument");
        *full = A->full;
    });
}

int gf_matpart_csr(gf_matpart A, int block, const int32_t** rowptr, const int32_t** col, const double** val)
{
    return guard([&] {
        GF_REQUIRE(A, "null matrix");
        GF_REQUIRE(block >= 0 && block < 4, "block must be 0 (uu), 1 (up), 2 (pu) or 3 (pp)");
        if (rowptr) *rowptr = A->rowptr[block].p;
        if (col) *col = A->col[block].p;
        if (val) *val = A->val[block].p;
    });
}

int gf_matpart_assemble(gf_matpart A, const double* elemmat, void* stream)
{
    return guard([&] {
        GF_REQUIRE(A && elemmat, "null argument");
        const int rc = gf_matrix_assemble(A->full, elemmat, stream);
        if (rc != GF_OK) throw Error(rc, gf_last_error());
        matpart_split(A, S(stream));
    });
}

int gf_matpart_assemble_f32(gf_matpart A, const float* elemmat, void* stream)
{
    return guard([&] {
        GF_REQUIRE(A && elemmat, "null argument");
        const int rc = gf_matrix_assemble_f32(A->full, elemmat, stream);
        if (rc != GF_OK) throw Error(rc, gf_last_error());
        matpart_split(A, S(stream));
    });
}

int gf_matpart_dot_part(gf_matpart A, int which, const double* x_u, const double* x_p, double* b, void* stream)
{
    return guard([&] {
        GF_REQUIRE(A && (x_u || A->p->nnu == 0) && (x_p || A->p->nnp == 0) && (b || A->nrows[2 * which] == 0),
                   "null argument");
        GF_REQUIRE(which == 0 || which == 1, "which must be 0 (B_u) or 1 (B_p)");
        matpart_product(A, which, x_u, x_p, b, S(stream));
    });
}

int gf_matpart_apply(gf_matpart A, int op, int kind, const double* x, double* b, void* stream)
{
    return guard([&] {
        GF_REQUIRE(A && x && b, "null argument");
        GF_REQUIRE(op == 0 || op == 1, "op must be 0 (Dot) or 1 (Reaction)");
        GF_REQUIRE(kind == 0 || kind == 1, "kind must be 0 (dofval) or 1 (nodevec)");
        auto p = A->p;
        auto W = work_of(A);
        hipStream_t s = S(stream);
        vecpart_as_dofs(p, kind, x, W.xu, W.xp, stream);
        if (op == 0) matpart_product(A, 0, W.xu, W.xp, W.bu, s);
        matpart_product(A, 1, W.xu, W.xp, W.bp, s);
        if (kind == 0) {
            if (op == 0) {
                const int rc = gf_vecpart_from_partitioned(p, 0, W.bu, W.bp, b, stream);
                if (rc != GF_OK) throw Error(rc, gf_last_error());
            }
            else if (p->nnp)
                k_scatter<<<nblocks(p->nnp, BS), BS, 0, s>>>(p->nnp, p->iip.p, W.bp, b);
        }
        else {
            const int64_t n = p->v->nnode * p->v->ndim;
            if (n) k_node_from_part_sel<<<nblocks(n, BS), BS, 0, s>>>(n, p->nnu, op == 0 ? 2 : 1, p->part.p, W.bu, W.bp, b);
        }
        check_launch();
    });
}

int gf_matpart_solve_u(gf_matpart A, const double* b_u, const double* x_p, double* x_u, int warm, double tol,
                       int64_t maxiter, int64_t* iterations, double* relres, void* stream)
{
    return guard([&] {
        GF_REQUIRE(A && ((b_u && x_u) || A->p->nnu == 0) && (x_p || A->p->nnp == 0), "null argument");
        auto W = work_of(A);
        hipStream_t s = S(stream);
        const int64_t nu = A->nrows[0];
        // rhs = b_u - A_up x_p
        if (nu)
            k_spmv2<<<nblocks(nu, BS), BS, 0, s>>>(nu, b_u, -1.0, nullptr, nullptr, nullptr, nullptr, A->rowptr[1].p,
                                                   A->col[1].p, A->val[1].p, x_p, W.rhs);
        check_launch();
        matpart_cg(A, W.rhs, x_u, warm != 0, tol, maxiter, iterations, relres, s);
    });
}

int gf_matpart_solve(gf_matpart A, int kind, const double* b, double* x, int warm, double tol, int64_t maxiter,
                     int64_t* iterations, double* relres, void* stream)
{
    return guard([&] {
        GF_REQUIRE(A && b && x, "null argument");
        GF_REQUIRE(kind == 0 || kind == 1, "kind must be 0 (dofval) or 1 (nodevec)");
        auto p = A->p;
        auto W = work_of(A);
        hipStream_t s = S(stream);
        const int64_t nu = A->nrows[0];
        // B_u from b, X_u (the warm start) and X_p from x; rhs = B_u - A_up X_p
        vecpart_as_dofs(p, kind, b, W.bu, nullptr, stream);
        vecpart_as_dofs(p, kind, x, W.xu, W.xp, stream);
        if (nu)
            k_spmv2<<<nblocks(nu, BS), BS, 0, s>>>(nu, W.bu, -1.0, nullptr, nullptr, nullptr, nullptr, A->rowptr[1].p,
                                                   A->col[1].p, A->val[1].p, W.xp, W.rhs);
        check_launch();
        matpart_cg(A, W.rhs, W.xu, warm != 0, tol, maxiter, iterations, relres, s);
        if (kind == 0) {
            if (nu) k_scatter<<<nblocks(nu, BS), BS, 0, s>>>(nu, p->iiu.p, W.xu, x);
        }
        else {
            const int64_t n = p->v->nnode * p->v->ndim;
            if (n) k_node_from_part_sel<<<nblocks(n, BS), BS, 0, s>>>(n, p->nnu, 0, p->part.p, W.xu, W.xp, x);
        }
        check_launch();
    });
}

// ---- MatrixPartitionedTyings -------------------------------------------------------------------

static void to_i32(const int64_t* a, size_t n, std::vector<int32_t>& out, const char* what)
{
    out.resize(n);
    for (size_t i = 0; i < n; ++i) {
        GF_REQUIRE(a[i] >= 0 && a[i] < INT32_MAX, std::string(what) + " entry outside the int32 range");
        out[i] = static_cast<int32_t>(a[i]);
    }
}

int gf_mattyings_create(gf_mattyings* out, gf_vector v, int64_t nnu, int64_t nnp, int64_t nnd, const int64_t* du_rp,
                        const int64_t* du_cl, const double* du_v, const int64_t* dp_rp, const int64_t* dp_cl,
                        const double* dp_v)
{
    return guard([&] {
        GF_REQUIRE(out && v && du_rp && dp_rp, "null argument");
        GF_REQUIRE(nnu >= 0 && nnp >= 0 && nnd >= 0 && nnu + nnp + nnd == v->ndof,
                   "nnu + nnp + nnd must equal the number of DOFs (numbering [u | p | d])");
        auto A = new gf_mattyings_s();
        try {
            A->v = v;
            A->nnu = nnu;
            A->nnp = nnp;
            A->nnd = nnd;
            A->ndof = v->ndof;
            int rc = gf_matrix_create(&A->K, v);
            if (rc != GF_OK) throw Error(rc, gf_last_error());
            std::vector<int32_t> rp, cl;
            // C_du, C_dp as given (rows d, columns ascending within a row as produced by the tyings)
            to_i32(du_rp, nnd + 1, rp, "C_du rowptr");
            A->du_rp.upload(rp);
            const int64_t nz_du = du_rp[nnd];
            to_i32(du_cl, nz_du, cl, "C_du col");
            for (int32_t c : cl) GF_REQUIRE(c < nnu, "C_du column outside [0, nnu)");
            A->du_cl.upload(cl);
            A->du_v.upload(du_v, nz_du);
            to_i32(dp_rp, nnd + 1, rp, "C_dp rowptr");
            A->dp_rp.upload(rp);
            const int64_t nz_dp = dp_rp[nnd];
            to_i32(dp_cl, nz_dp, cl, "C_dp col");
            for (int32_t c : cl) GF_REQUIRE(c < nnp, "C_dp column outside [0, nnp)");
            A->dp_cl.upload(cl);
            A->dp_v.upload(dp_v, nz_dp);
            // C_ud = C_du^T: rows u, entries in ascending d
            std::vector<int32_t> trp(nnu + 1, 0), tcl(nz_du);
            std::vector<double> tv(nz_du);
            for (int64_t k = 0; k < nz_du; ++k) ++trp[du_cl[k] + 1];
            for (int64_t i = 0; i < nnu; ++i) trp[i + 1] += trp[i];
            std::vector<int32_t> fill(trp.begin(), trp.end() - 1);
            for (int64_t d = 0; d < nnd; ++d)
                for (int64_t k = du_rp[d]; k < du_rp[d + 1]; ++k) {
                    const int32_t at = fill[du_cl[k]]++;
                    tcl[at] = static_cast<int32_t>(d);
                    tv[at] = du_v[k];
                }
            A->ud_rp.upload(trp);
            A->ud_cl.upload(tcl);
            A->ud_v.upload(tv);
            A->work.alloc(8 * nnu + 3 * A->ndof);
            A->red.alloc(3 * NPB + 8 + nblocks(LPR * nnu, PB)); // + the fused product's p.q partials
            A->state.alloc(4);
        }
        catch (...) {
            if (A->K) gf_matrix_destroy(A->K);
            delete A;
            throw;
        }
        *out = A;
    });
}

int gf_mattyings_destroy(gf_mattyings A)
{
    if (A && A->K) gf_matrix_destroy(A->K);
    delete A;
    return GF_OK;
}

int gf_mattyings_matrix(gf_mattyings A, gf_matrix* K)
{
    return guard([&] {
        GF_REQUIRE(A && K, "null argument");
        *K = A->K;
    });
}

int gf_mattyings_assemble(gf_mattyings A, const double* elemmat, void* stream)
{
    return guard([&] {
        GF_REQUIRE(A && elemmat, "null argument");
        const int rc = gf_matrix_assemble(A->K, elemmat, stream);
        if (rc != GF_OK) throw Error(rc, gf_last_error());
        const int32_t *rp = nullptr, *cl = nullptr;
        const double* val = nullptr;
        gf_matrix_csr(A->K, &rp, &cl, &val);
        hipStream_t s = S(stream);
        double* kd = A->work.p + 7 * A->nnu + 2 * A->ndof;
        double* invd = A->work.p + 5 * A->nnu;
        k_csr_diag<<<nblocks(A->ndof, BS), BS, 0, s>>>(A->ndof, rp, cl, val, kd);
        if (A->nnu)
            k_ty_invdiag<<<nblocks(A->nnu, BS), BS, 0, s>>>(A->nnu, A->nnu + A->nnp, kd, A->ud_rp.p, A->ud_cl.p, A->ud_v.p,
                                                             invd);
        check_launch();
    });
}

// the condensed product of a full-vector expansion: q_u = [K (x_u, x_p, C_du x_u + C_dp x_p)]_u +
// C_ud [K (...)]_d; p (for the partials) may be null
static void ty_product(gf_mattyings A, const double* xu, const double* xp, double* q, const double* p, double* partial,
                       hipStream_t s)
{
    double* xfull = A->work.p + 7 * A->nnu;
    double* yfull = xfull + A->ndof;
    k_ty_expand<<<nblocks(A->ndof, BS), BS, 0, s>>>(A->nnu, A->nnp, A->nnd, xu, xp, A->du_rp.p, A->du_cl.p, A->du_v.p,
                                                     A->dp_rp.p, A->dp_cl.p, A->dp_v.p, xfull);
    check_launch();
    const int rc = gf_matrix_dot_dofval(A->K, xfull, yfull, s);
    if (rc != GF_OK) throw Error(rc, gf_last_error());
    k_ty_fold<<<NPB, PB, 0, s>>>(A->nnu, A->nnu + A->nnp, yfull, A->ud_rp.p, A->ud_cl.p, A->ud_v.p, p, q, partial);
    check_launch();
}

int gf_mattyings_solve(gf_mattyings A, const double* b, double* x, double tol, int64_t maxiter, int64_t* iterations,
                       double* relres, void* stream)
{
    return guard([&] {
        GF_REQUIRE(A && b && x, "null argument");
        GF_REQUIRE(tol > 0.0 && maxiter > 0 && maxiter < INT32_MAX, "tol > 0, 0 < maxiter < 2^31");
        hipStream_t s = S(stream);
        const int64_t nu = A->nnu, ni = A->nnu + A->nnp, nd = A->ndof;
        double* w = A->work.p;
        double *cx = w, *cr = w + nu, *cz = w + 2 * nu, *cp = w + 3 * nu, *cq = w + 4 * nu, *invd = w + 5 * nu,
               *rhs = w + 6 * nu, *xfull = w + 7 * nu;
        double* partial = A->red.p;
        double* scal = partial + 3 * NPB;
        int32_t* state = A->state.p;
        // b, x as dofval (nodevec -> dofval, overwrite in loop order); X_p is read from xin
        DBuf<double> bfull(nd), xin(nd);
        int rc = gf_vector_as_dofs_nodevec(A->v, b, bfull.p, stream);
        if (rc != GF_OK) throw Error(rc, gf_last_error());
        rc = gf_vector_as_dofs_nodevec(A->v, x, xin.p, stream);
        if (rc != GF_OK) throw Error(rc, gf_last_error());
        const double* Xp = xin.p + nu;
        // rhs = B_u + C_ud B_d - ACup X_p
        ty_product(A, nullptr, Xp, cq, nullptr, nullptr, s);
        if (nu) k_ty_rhs<<<nblocks(nu, BS), BS, 0, s>>>(nu, ni, bfull.p, A->ud_rp.p, A->ud_cl.p, A->ud_v.p, cq, rhs);
        check_launch();
        int32_t st[2] = {0, 0};
        if (nu) {
            k_ty_init<<<NPB, PB, 0, s>>>(nu, rhs, invd, cx, cr, cz, cp, partial);
            k_cg_init_scal<<<1, 64, 0, s>>>(partial, scal, state, tol, static_cast<int32_t>(maxiter));
            check_launch();
            GF_HIP(hipMemcpyAsync(st, state, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
            GF_HIP(hipStreamSynchronize(s));
            if (st[0] == 2) throw Error(GF_EINVAL, "non-finite right-hand side");
            constexpr int64_t check_every = 16;
            // three launches per iteration: the operand expansion forming the search direction (by
            // iteration parity in p, p2), the condensed product with the p.q partials, the update
            double* pb[2] = {cp, w + 7 * nu + 3 * nd};
            const unsigned gexp = nblocks(nd, PB), gprod = nblocks(LPR * nu, PB);
            double* pq_part = partial + 3 * NPB + 8;
            const int32_t *krp = nullptr, *kcl = nullptr;
            const double* kv = nullptr;
            rc = gf_matrix_csr(A->K, &krp, &kcl, &kv);
            if (rc != GF_OK) throw Error(rc, gf_last_error());
            auto expand_dir = [&](int64_t it) {
                k_ty_expand_dir<<<gexp, PB, 0, s>>>(nu, A->nnp, A->nnd, cz, pb[(it + 1) & 1], pb[it & 1], A->du_rp.p,
                                                     A->du_cl.p, A->du_v.p, xfull, partial, scal, state, it);
            };
            StatePipe pipe;
            int chunk = 0;
            for (int64_t it = 0; !st[0] && it < maxiter;) {
                for (int64_t k = 0; k < check_every && it < maxiter; ++k, ++it) {
                    if (it == 0)
                        k_ty_expand<<<nblocks(nd, BS), BS, 0, s>>>(nu, A->nnp, A->nnd, cp, nullptr, A->du_rp.p,
                                                                    A->du_cl.p, A->du_v.p, A->dp_rp.p, A->dp_cl.p,
                                                                    A->dp_v.p, xfull);
                    else
                        expand_dir(it);
                    k_ty_spmv_fold<<<gprod, PB, 0, s>>>(nu, nu + A->nnp, krp, kcl, kv, xfull, A->ud_rp.p, A->ud_cl.p,
                                                        A->ud_v.p, pb[it & 1], cq, pq_part, state);
                    k_cg_update<<<NPB, PB, 0, s>>>(nu, pq_part, static_cast<int>(gprod), scal, invd, pb[it & 1], cq, cx,
                                                   cr, cz, partial, state, it);
                }
                if (it == maxiter) expand_dir(it); // the convergence test and count of the last update
                check_launch();
                // the host looks at the previous chunk's flag while this chunk is queued
                pipe.post(state, chunk, s);
                if (chunk > 0 && pipe.wait(chunk - 1)[0]) st[0] = 1;
                ++chunk;
            }
            // true residual of the condensed system
            ty_product(A, cx, nullptr, cq, nullptr, nullptr, s);
            k_cg_residual_vec<<<NPB, PB, 0, s>>>(nu, rhs, cq, partial);
            k_cg_relres<<<1, 64, 0, s>>>(partial, scal + 5);
            check_launch();
        }
        // x = AsNode(X_u, X_p, C_du X_u + C_dp X_p): u and d entries solved, p entries as given
        k_ty_expand<<<nblocks(nd, BS), BS, 0, s>>>(nu, A->nnp, A->nnd, cx, Xp, A->du_rp.p, A->du_cl.p, A->du_v.p,
                                                    A->dp_rp.p, A->dp_cl.p, A->dp_v.p, xfull);
        check_launch();
        rc = gf_vector_as_node_dofval(A->v, xfull, x, stream);
        if (rc != GF_OK) throw Error(rc, gf_last_error());
        double res = 0.0;
        if (nu) GF_HIP(hipMemcpyAsync(&res, scal + 5, sizeof(double), hipMemcpyDeviceToHost, s));
        GF_HIP(hipMemcpyAsync(st, state, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        GF_HIP(hipStreamSynchronize(s));
        if (iterations) *iterations = nu ? st[1] : 0;
        if (relres) *relres = res;
        if (!(res <= 10.0 * tol))
            throw Error(GF_ENOTCONV, "PCG on the condensed system did not reach the tolerance (true residual " +
                                         std::to_string(res) + ")");
    });
}

int gf_csr_addmv(int64_t nrows, const int32_t* rowptr, const int32_t* col, const double* val, const double* x, double* y,
                 void* stream)
{
    return guard([&] {
        GF_REQUIRE(rowptr && (x || nrows == 0) && (y || nrows == 0), "null argument");
        if (nrows) k_csr_addmv<<<nblocks(nrows, BS), BS, 0, S(stream)>>>(nrows, rowptr, col, val, x, y);
        check_launch();
    });
}

} // extern "C"
