// GooseFEM::Vector on the device (include/GooseFEM/Vector.h).
//
// The reference scatter-adds serially in (e, m, i) loop order (Vector.h:723-741) and overwrites
// with "last write wins" (Vector.h:590-609, :634-653).  Here every conversion is a gather: each
// output entry owns a CSR row of its sources, sorted in that same loop order, so
//   - sums start at 0.0 and add the sources in ascending order: FP64 bit-identical to the serial
//     reference loop, no atomics, deterministic;
//   - overwrites take the last source of the row.
// Rows are built once on the host at construction (counting sort, O(nelem * nne * ndim)).
#include "gf_internal.hpp"

using namespace gf;

namespace {

constexpr int BS = 256;

// CSR of "value -> ascending list of flat source indices" by counting sort.
void build_csr(const std::vector<int64_t>& key, int64_t nkeys, std::vector<int32_t>& ptr,
               std::vector<int32_t>& idx)
{
    ptr.assign(static_cast<size_t>(nkeys + 1), 0);
    for (int64_t k : key) ptr[k + 1]++;
    for (int64_t i = 0; i < nkeys; ++i) ptr[i + 1] += ptr[i];
    idx.resize(key.size());
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    for (size_t j = 0; j < key.size(); ++j) idx[fill[key[j]]++] = static_cast<int32_t>(j);
}

template <int ND>
__global__ void k_as_element_nodevec(int64_t n, int nne, const int32_t* __restrict__ conn,
                                     const double* __restrict__ nodevec, double* __restrict__ elemvec)
{
    int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= n) return;
    int64_t em = k / ND;
    int i = static_cast<int>(k - em * ND);
    elemvec[k] = nodevec[static_cast<int64_t>(conn[em]) * ND + i];
}

__global__ void k_as_element_nodevec_g(int64_t n, int nd, const int32_t* __restrict__ conn,
                                       const double* __restrict__ nodevec, double* __restrict__ elemvec)
{
    int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= n) return;
    int64_t em = k / nd;
    int i = static_cast<int>(k - em * nd);
    elemvec[k] = nodevec[static_cast<int64_t>(conn[em]) * nd + i];
}

__global__ void k_as_element_dofval(int64_t n, int nd, const int32_t* __restrict__ conn,
                                    const int32_t* __restrict__ dofs, const double* __restrict__ dofval,
                                    double* __restrict__ elemvec)
{
    int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= n) return;
    int64_t em = k / nd;
    int i = static_cast<int>(k - em * nd);
    elemvec[k] = dofval[dofs[static_cast<int64_t>(conn[em]) * nd + i]];
}

// last source of each row (0 for an empty row)
__global__ void k_row_last(int64_t nrows, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                           const double* __restrict__ src, double* __restrict__ dst)
{
    int64_t r = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (r >= nrows) return;
    int32_t b = ptr[r], e = ptr[r + 1];
    dst[r] = (e > b) ? src[idx[e - 1]] : 0.0;
}

// ordered row sum
template <class T>
__global__ void k_row_sum(int64_t nrows, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                          const T* __restrict__ src, double* __restrict__ dst)
{
    int64_t r = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (r >= nrows) return;
    dst[r] = ordered_gather_sum<4>(ptr[r], ptr[r + 1], idx, src);
}

// AsNode(elemvec): per node component, the last (e, m) of the node's row
__global__ void k_as_node_elemvec(int64_t nnode, int nd, const int32_t* __restrict__ ptr,
                                  const int32_t* __restrict__ idx, const double* __restrict__ elemvec,
                                  double* __restrict__ nodevec)
{
    int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= nnode * nd) return;
    int64_t n = k / nd;
    int i = static_cast<int>(k - n * nd);
    int32_t b = ptr[n], e = ptr[n + 1];
    nodevec[k] = (e > b) ? elemvec[static_cast<int64_t>(idx[e - 1]) * nd + i] : 0.0;
}

__global__ void k_as_node_dofval(int64_t n, const int32_t* __restrict__ dofs, const double* __restrict__ dofval,
                                 double* __restrict__ nodevec)
{
    int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= n) return;
    nodevec[k] = dofval[dofs[k]];
}

// AssembleNode(elemvec) fused: each node component sums its DOF's row (tied nodes receive the
// DOF sum, Vector.h:746-757)
template <class T>
__global__ void k_assemble_node(int64_t n, const int32_t* __restrict__ dofs, const int32_t* __restrict__ ptr,
                                const int32_t* __restrict__ idx, const T* __restrict__ elemvec,
                                double* __restrict__ nodevec)
{
    int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= n) return;
    int32_t d = dofs[k];
    nodevec[k] = ordered_gather_sum<4>(ptr[d], ptr[d + 1], idx, elemvec);
}

} // namespace

extern "C" {

int gf_vector_create(gf_vector* out, const int64_t* conn_h, int64_t nelem, int nne, const int64_t* dofs_h,
                     int64_t nnode, int ndim)
{
    return guard([&] {
        GF_REQUIRE(out && conn_h && dofs_h, "null argument");
        GF_REQUIRE(nelem >= 0 && nne > 0 && nnode > 0 && ndim > 0, "sizes");
        GF_REQUIRE(nelem * nne * ndim < INT32_MAX, "elemvec too large for int32 gathers");
        auto v = new gf_vector_s();
        try {
            v->nelem = nelem;
            v->nne = nne;
            v->nnode = nnode;
            v->ndim = ndim;
            v->conn_h.assign(conn_h, conn_h + nelem * nne);
            v->dofs_h.assign(dofs_h, dofs_h + nnode * ndim);
            int64_t mx = -1, mn = INT64_MAX;
            for (auto d : v->dofs_h) {
                mx = std::max(mx, d);
                mn = std::min(mn, d);
            }
            GF_REQUIRE(mn >= 0, "negative dof");
            v->ndof = mx + 1;
            GF_REQUIRE(v->ndof <= nnode * ndim, "ndof <= nnode * ndim (Vector.h:46)");
            for (auto n : v->conn_h) GF_REQUIRE(n >= 0 && n < nnode, "conn out of range (Vector.h:45)");
            v->conn.upload(narrow(conn_h, v->conn_h.size(), "conn"));
            v->dofs.upload(narrow(dofs_h, v->dofs_h.size(), "dofs"));
            std::vector<int32_t> ptr, idx;
            // DOF <- elemvec entries
            std::vector<int64_t> key(static_cast<size_t>(nelem * nne * ndim));
            for (int64_t em = 0; em < nelem * nne; ++em)
                for (int i = 0; i < ndim; ++i) key[em * ndim + i] = v->dofs_h[v->conn_h[em] * ndim + i];
            build_csr(key, v->ndof, ptr, idx);
            v->dof_eptr.upload(ptr);
            v->dof_eidx.upload(idx);
            // DOF <- nodevec entries
            build_csr(v->dofs_h, v->ndof, ptr, idx);
            v->dof_nptr.upload(ptr);
            v->dof_nidx.upload(idx);
            // node <- (e, m)
            build_csr(v->conn_h, nnode, ptr, idx);
            v->node_eptr.upload(ptr);
            v->node_eidx.upload(idx);
        }
        catch (...) {
            delete v;
            throw;
        }
        *out = v;
    });
}

int gf_vector_destroy(gf_vector v)
{
    delete v;
    return GF_OK;
}

int gf_vector_sizes(gf_vector v, int64_t* nelem, int* nne, int64_t* nnode, int* ndim, int64_t* ndof)
{
    return guard([&] {
        GF_REQUIRE(v, "null vector");
        if (nelem) *nelem = v->nelem;
        if (nne) *nne = v->nne;
        if (nnode) *nnode = v->nnode;
        if (ndim) *ndim = v->ndim;
        if (ndof) *ndof = v->ndof;
    });
}

int gf_vector_device_indices(gf_vector v, const int32_t** conn, const int32_t** dofs)
{
    return guard([&] {
        GF_REQUIRE(v, "null vector");
        if (conn) *conn = v->conn.p;
        if (dofs) *dofs = v->dofs.p;
    });
}

int gf_vector_as_element_nodevec(gf_vector v, const double* nodevec, double* elemvec, void* stream)
{
    return guard([&] {
        GF_REQUIRE(v && nodevec && elemvec, "null argument");
        int64_t n = v->nelem * v->nne * v->ndim;
        if (n == 0) return;
        if (v->ndim == 2)
            k_as_element_nodevec<2><<<nblocks(n, BS), BS, 0, S(stream)>>>(n, v->nne, v->conn.p, nodevec, elemvec);
        else if (v->ndim == 3)
            k_as_element_nodevec<3><<<nblocks(n, BS), BS, 0, S(stream)>>>(n, v->nne, v->conn.p, nodevec, elemvec);
        else
            k_as_element_nodevec_g<<<nblocks(n, BS), BS, 0, S(stream)>>>(n, v->ndim, v->conn.p, nodevec, elemvec);
        check_launch();
    });
}

int gf_vector_as_element_dofval(gf_vector v, const double* dofval, double* elemvec, void* stream)
{
    return guard([&] {
        GF_REQUIRE(v && dofval && elemvec, "null argument");
        int64_t n = v->nelem * v->nne * v->ndim;
        if (n == 0) return;
        k_as_element_dofval<<<nblocks(n, BS), BS, 0, S(stream)>>>(n, v->ndim, v->conn.p, v->dofs.p, dofval, elemvec);
        check_launch();
    });
}

int gf_vector_as_dofs_nodevec(gf_vector v, const double* nodevec, double* dofval, void* stream)
{
    return guard([&] {
        GF_REQUIRE(v && nodevec && dofval, "null argument");
        k_row_last<<<nblocks(v->ndof, BS), BS, 0, S(stream)>>>(v->ndof, v->dof_nptr.p, v->dof_nidx.p, nodevec, dofval);
        check_launch();
    });
}

int gf_vector_as_dofs_elemvec(gf_vector v, const double* elemvec, double* dofval, void* stream)
{
    return guard([&] {
        GF_REQUIRE(v && elemvec && dofval, "null argument");
        k_row_last<<<nblocks(v->ndof, BS), BS, 0, S(stream)>>>(v->ndof, v->dof_eptr.p, v->dof_eidx.p, elemvec, dofval);
        check_launch();
    });
}

int gf_vector_as_node_dofval(gf_vector v, const double* dofval, double* nodevec, void* stream)
{
    return guard([&] {
        GF_REQUIRE(v && dofval && nodevec, "null argument");
        int64_t n = v->nnode * v->ndim;
        k_as_node_dofval<<<nblocks(n, BS), BS, 0, S(stream)>>>(n, v->dofs.p, dofval, nodevec);
        check_launch();
    });
}

int gf_vector_as_node_elemvec(gf_vector v, const double* elemvec, double* nodevec, void* stream)
{
    return guard([&] {
        GF_REQUIRE(v && elemvec && nodevec, "null argument");
        int64_t n = v->nnode * v->ndim;
        k_as_node_elemvec<<<nblocks(n, BS), BS, 0, S(stream)>>>(v->nnode, v->ndim, v->node_eptr.p, v->node_eidx.p,
                                                              elemvec, nodevec);
        check_launch();
    });
}

int gf_vector_assemble_dofs_nodevec(gf_vector v, const double* nodevec, double* dofval, void* stream)
{
    return guard([&] {
        GF_REQUIRE(v && nodevec && dofval, "null argument");
        k_row_sum<double><<<nblocks(v->ndof, BS), BS, 0, S(stream)>>>(v->ndof, v->dof_nptr.p, v->dof_nidx.p,
                                                                    nodevec, dofval);
        check_launch();
    });
}

int gf_vector_assemble_dofs_elemvec(gf_vector v, const void* elemvec, int dtype, double* dofval, void* stream)
{
    return guard([&] {
        GF_REQUIRE(v && elemvec && dofval, "null argument");
        GF_REQUIRE(dtype == GF_F64 || dtype == GF_F32, "dtype must be GF_F64 or GF_F32");
        if (dtype == GF_F64)
            k_row_sum<double><<<nblocks(v->ndof, BS), BS, 0, S(stream)>>>(
                v->ndof, v->dof_eptr.p, v->dof_eidx.p, static_cast<const double*>(elemvec), dofval);
        else
            k_row_sum<float><<<nblocks(v->ndof, BS), BS, 0, S(stream)>>>(
                v->ndof, v->dof_eptr.p, v->dof_eidx.p, static_cast<const float*>(elemvec), dofval);
        check_launch();
    });
}

int gf_vector_assemble_node_elemvec(gf_vector v, const void* elemvec, int dtype, double* nodevec, void* stream)
{
    return guard([&] {
        GF_REQUIRE(v && elemvec && nodevec, "null argument");
        GF_REQUIRE(dtype == GF_F64 || dtype == GF_F32, "dtype must be GF_F64 or GF_F32");
        int64_t n = v->nnode * v->ndim;
        if (dtype == GF_F64)
            k_assemble_node<double><<<nblocks(n, BS), BS, 0, S(stream)>>>(
                n, v->dofs.p, v->dof_eptr.p, v->dof_eidx.p, static_cast<const double*>(elemvec), nodevec);
        else
            k_assemble_node<float><<<nblocks(n, BS), BS, 0, S(stream)>>>(
                n, v->dofs.p, v->dof_eptr.p, v->dof_eidx.p, static_cast<const float*>(elemvec), nodevec);
        check_launch();
    });
}

} // extern "C"
