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
